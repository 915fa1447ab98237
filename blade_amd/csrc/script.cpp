// script.cpp - the reference's input-script language over the blade_* API
// (blade_interpretter, io/io.h:69-70).
//
// The language follows doc/MAIN.txt and the reference's interpreter
// (io/io.cxx:288-312, io/control.cxx, io/variables.cxx, io/calculate.cxx,
// system/selections.cxx) and the command sets of the classes it drives
// (system/parameters.cxx:103-130, system/structure.cxx, system/
// coordinates.cxx:70-330, msld/msld.cu:150-470, run/run.cu:220-420): a line
// is a class name and its keywords; text after '!' is a comment; {name} is
// replaced by a variable's value, innermost braces first; variables
// calculate and the if / elseif / while conditions take prefix
// expressions; stream runs another file in place.  Each command ends in
// the same blade_* setters a library caller uses.  An error is recorded on
// the system (blade_system_error) and ends the script.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/blade.h"
#include "setup.h"

namespace blade {
void api_report_error(System *s, const std::string &m);
const Structure *api_structure(System *s);
bool api_has_parameters(System *s);
bool api_has_coordinates(System *s);
bool api_has_msld(System *s);
int api_reset_device(System *s);
void *&api_script_state(System *s, void (*release)(void *));
}  // namespace blade

namespace {

using blade::AtomRec;

const double kPicosecond = 20.454828284385908;    // AKMA time units per ps
const double kAtmosphere = 0.000014584319218586;  // kcal/mol/A^3 per atm
const double kBoltz = 0.001987191;                // kcal/mol/K

struct ScriptError {
  std::string what;
};

[[noreturn]] void fail(const std::string &m) { throw ScriptError{m}; }

// whitespace tokens up to a '!' comment
std::vector<std::string> split(const std::string &line) {
  std::vector<std::string> out;
  std::istringstream is(line.substr(0, line.find('!')));
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

struct Tokens {
  std::vector<std::string> t;
  size_t k = 0;
  bool more() const { return k < t.size(); }
  std::string peek() const { return more() ? t[k] : std::string(); }
  std::string next() { return more() ? t[k++] : std::string(); }
  std::string need(const char *what) {
    if (!more()) fail(std::string("missing ") + what);
    return t[k++];
  }
  double real(const char *what) {
    const std::string s = need(what);
    char *e = nullptr;
    const double v = std::strtod(s.c_str(), &e);
    if (!e || *e || e == s.c_str()) fail(std::string("expected a number for ") + what + ", got " + s);
    return v;
  }
  long integer(const char *what) {
    const std::string s = need(what);
    char *e = nullptr;
    const long v = std::strtol(s.c_str(), &e, 10);
    if (!e || *e || e == s.c_str()) fail(std::string("expected an integer for ") + what + ", got " + s);
    return v;
  }
  bool boolean(const char *what) {
    const std::string s = need(what);
    if (s == "on" || s == "true" || s == "yes" || s == "1" || s == "T") return true;
    if (s == "off" || s == "false" || s == "no" || s == "0" || s == "F") return false;
    fail(std::string("expected on/off for ") + what + ", got " + s);
  }
  // the rest of the line, space-joined
  std::string rest() {
    std::string r;
    while (more()) r += (r.empty() ? "" : " ") + next();
    return r;
  }
};

// prefix expressions: operators before their operands (io/calculate.cxx)
double evaluate(Tokens &t) {
  const std::string op = t.need("an expression");
  auto two = [&](auto f) {
    const double a = evaluate(t);
    const double b = evaluate(t);
    return f(a, b);
  };
  if (op == "+") return two([](double a, double b) { return a + b; });
  if (op == "-") return two([](double a, double b) { return a - b; });
  if (op == "*") return two([](double a, double b) { return a * b; });
  if (op == "/") return two([](double a, double b) { return a / b; });
  if (op == "==") return two([](double a, double b) { return (double)(a == b); });
  if (op == ">=") return two([](double a, double b) { return (double)(a >= b); });
  if (op == "<=") return two([](double a, double b) { return (double)(a <= b); });
  if (op == ">") return two([](double a, double b) { return (double)(a > b); });
  if (op == "<") return two([](double a, double b) { return (double)(a < b); });
  if (op == "&&") return two([](double a, double b) { return (double)(a != 0 && b != 0); });
  if (op == "||") return two([](double a, double b) { return (double)(a != 0 || b != 0); });
  if (op == "not") return (double)(evaluate(t) == 0);
  if (op == "sqrt") return std::sqrt(evaluate(t));
  if (op == "exp") return std::exp(evaluate(t));
  if (op == "log") return std::log(evaluate(t));
  if (op == "sin") return std::sin(evaluate(t));
  if (op == "cos") return std::cos(evaluate(t));
  if (op == "tan") return std::tan(evaluate(t));
  if (op == "floor") return std::floor(evaluate(t));
  if (op == "erf") return std::erf(evaluate(t));
  if (op == "pi") return M_PI;
  if (op == "eq") {
    const std::string a = t.need("a string"), b = t.need("a string");
    return (double)(a == b);
  }
  if (op == "-f") {
    std::ifstream f(t.need("a file name"));
    return (double)(bool)f;
  }
  char *e = nullptr;
  const double v = std::strtod(op.c_str(), &e);
  if (!e || *e || e == op.c_str()) fail("unrecognized token " + op + " in an expression");
  return v;
}

// run variables with the reference's defaults (run/run.cu:25-75)
struct RunVars {
  long nsteps = 5000;
  double dt_ps = 0.001, T = 300, gamma_ps = 1.0;
  std::string fnmXTC = "default.xtc", fnmLMD = "default.lmd", fnmNRG = "default.nrg", fnmCPI, fnmCPO = "default.cpt";
  int freqXTC = 1000, freqLMD = 10, freqNRG = 10;
  bool hrLMD = true;
  double invBetaEwald = 3.2, rCut = 10, rSwitch = 8.5, gridSpace = 1.0;
  int vdw = BM_VFSWITCH, elec = BM_PME, grid[3] = {-1, -1, -1}, order = 6;
  double shakeTol = 2e-7;
  int freqNPT = 50;
  double volumeFluctuation = 100, pressure_atm = 1;
  // the parameters a device system is built with
  std::vector<double> physics() const {
    return {dt_ps, T, gamma_ps, invBetaEwald, rCut, rSwitch, gridSpace, (double)vdw, (double)elec,
            (double)grid[0], (double)grid[1], (double)grid[2], (double)order, shakeTol, (double)freqNPT,
            volumeFluctuation, pressure_atm};
  }
};

// msld flags the script sets one by one (msld/msld.cu:25-90 defaults)
struct MsldFlags {
  double gamma_ps = 1.0, fnex = 5.5, kRestraint = 59.2;
  bool softCore = false, softCore14 = false, fix = false;
  int ewaldType = 2;
  double softBondRadius = 1.0, softBondExponent = 2, softNotBondExponent = 1;
  double kCR1 = 0, kCR2 = 0, kCR3 = 0, q0CR3 = 0, wCR3 = 0;
  bool piecewise = false;
  double wellWidth = 1.0, wellK = 100.0, edgeK = 0, edgeN = 2, edgeAlpha = 10, edgePhi = 0;
  int collSites = 0, indeSites = 0;
};

struct Frame {
  std::string type;   // "if", "elseif", "else" or "while"
  size_t line;        // the line of the while (re-evaluated at endwhile)
};

struct File {
  std::vector<std::string> lines;
  size_t pc = 0;
  std::vector<Frame> stack;
};

// what a system's scripts share from one blade_interpretter call to the
// next (the reference keeps it in its System: variables, selections, the
// Run and Msld objects)
struct ScriptState {
  std::map<std::string, std::string> vars;
  std::map<std::string, std::vector<char>> sel;
  size_t selLimit = 100;
  RunVars run;
  MsldFlags ms;
  std::vector<double> builtWith;   // run physics of the current device system
  bool outputsOpen = false;
  bool cpiRead = false;
  long velDraws = 0;
};

class Interp {
 public:
  Interp(System *s, ScriptState &st)
      : s_(s), vars_(st.vars), sel_(st.sel), selLimit_(st.selLimit), run_(st.run), ms_(st.ms),
        builtWith_(st.builtWith), outputsOpen_(st.outputsOpen), cpiRead_(st.cpiRead), velDraws_(st.velDraws) {}
  void run_file(const std::string &path);

 private:
  System *s_;
  std::map<std::string, std::string> &vars_;
  std::map<std::string, std::vector<char>> &sel_;
  size_t &selLimit_;
  RunVars &run_;
  MsldFlags &ms_;
  std::vector<double> &builtWith_;
  bool &outputsOpen_;
  bool &cpiRead_;
  long &velDraws_;
  std::vector<File> files_;

  std::string substitute(const std::string &line) const;
  void exec(const std::string &line);
  void skip_block(bool to_else);
  void cmd_variables(Tokens &t);
  void cmd_parameters(Tokens &t);
  void cmd_structure(Tokens &t);
  void cmd_selection(Tokens &t);
  void cmd_coordinates(Tokens &t);
  void cmd_msld(Tokens &t);
  void cmd_run(Tokens &t);
  std::vector<char> select(Tokens &t);
  const std::vector<char> &named(const std::string &name) const;
  void push_msld_flags();
  void prepare_run();
  void log(const std::string &m) const { blade_log(m.c_str()); }
  int natoms() const {
    const blade::Structure *st = blade::api_structure(s_);
    return st ? (int)st->atoms.size() : 0;
  }
  void check_api(const char *what) const {
    const char *e = blade_system_error(s_);
    if (e && *e) fail(std::string(what) + ": " + e);
  }
};

std::string Interp::substitute(const std::string &line) const {
  std::string l = line;
  for (;;) {
    const size_t bang = l.find('!');
    const size_t close = l.find('}');
    const size_t open = close == std::string::npos ? std::string::npos : l.rfind('{', close);
    const bool hasOpen = l.substr(0, bang).find('{') != std::string::npos;
    if (close == std::string::npos || (bang != std::string::npos && close > bang)) {
      if (hasOpen) fail("mismatched curly braces in \"" + line + "\"");
      return l;
    }
    if (open == std::string::npos) fail("mismatched curly braces in \"" + line + "\"");
    const std::string key = l.substr(open + 1, close - open - 1);
    auto it = vars_.find(key);
    if (it == vars_.end()) fail("unrecognized variable {" + key + "} in \"" + line + "\"");
    l = l.substr(0, open) + it->second + l.substr(close + 1);
  }
}

void Interp::run_file(const std::string &path) {
  std::ifstream in(path);
  if (!in) fail("cannot open script " + path);
  File f;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    f.lines.push_back(line);
  }
  if (files_.size() > 64) fail("stream nesting deeper than 64 files");
  files_.push_back(std::move(f));
  while (files_.back().pc < files_.back().lines.size()) {
    File &cur = files_.back();
    const std::string raw = cur.lines[cur.pc++];
    exec(substitute(raw));
  }
  if (!files_.back().stack.empty()) fail("end of " + path + " inside an if or while block");
  files_.pop_back();
}

// from the line after an if/elseif/while whose condition is false: to the
// next elseif/else/endif of the same depth (to_else, left unread) or past
// the matching endwhile
void Interp::skip_block(bool to_else) {
  File &f = files_.back();
  int depth = 0;
  while (f.pc < f.lines.size()) {
    const std::vector<std::string> t = split(f.lines[f.pc]);
    const std::string w = t.empty() ? "" : t[0];
    if (depth == 0) {
      if (to_else && (w == "elseif" || w == "else" || w == "endif")) return;
      if (!to_else && w == "endwhile") {
        f.pc++;
        return;
      }
    }
    if (w == "if" || w == "while") depth++;
    if (w == "endif" || w == "endwhile") depth--;
    f.pc++;
  }
  fail(to_else ? "file ended inside an if block" : "file ended inside a while block");
}

void Interp::exec(const std::string &line) {
  Tokens t{split(line)};
  if (!t.more()) return;
  const std::string c = t.next();
  File &f = files_.back();
  if (c == "variables") return cmd_variables(t);
  if (c == "if") {
    f.stack.push_back({"if", f.pc - 1});
    bool cond = evaluate(t) != 0;
    while (!cond) {
      skip_block(true);
      Tokens u{split(substitute(f.lines[f.pc++]))};
      const std::string w = u.next();
      if (w == "elseif") cond = evaluate(u) != 0;
      else if (w == "else") cond = true;
      else {   // endif
        f.stack.pop_back();
        return;
      }
    }
    return;
  }
  if (c == "elseif" || c == "else") {
    // a branch already ran: go past the endif
    if (f.stack.empty() || f.stack.back().type != "if") fail("unexpected " + c);
    for (;;) {
      skip_block(true);
      const std::string w = split(f.lines[f.pc++])[0];
      if (w == "endif") break;
    }
    f.stack.pop_back();
    return;
  }
  if (c == "endif") {
    if (f.stack.empty() || f.stack.back().type != "if") fail("unexpected endif");
    f.stack.pop_back();
    return;
  }
  if (c == "while") {
    if (evaluate(t) != 0) f.stack.push_back({"while", f.pc - 1});
    else skip_block(false);
    return;
  }
  if (c == "endwhile") {
    if (f.stack.empty() || f.stack.back().type != "while") fail("unexpected endwhile");
    f.pc = f.stack.back().line;   // the while line again
    f.stack.pop_back();
    return;
  }
  if (c == "stream") return run_file(t.need("a file name"));
  if (c == "verbose") {
    blade_set_verbose(s_, (int)t.integer("verbose level"));
    return;
  }
  if (c == "parameters") return cmd_parameters(t);
  if (c == "structure") return cmd_structure(t);
  if (c == "selection") return cmd_selection(t);
  if (c == "coordinates") return cmd_coordinates(t);
  if (c == "msld") return cmd_msld(t);
  if (c == "run") return cmd_run(t);
  if (c == "help") {
    log("help> classes: variables if elseif else endif while endwhile stream verbose parameters structure "
        "selection coordinates msld run\n");
    return;
  }
  if (c == "arrest") return;   // a pause for attaching a debugger: nothing to wait for here
  fail("unrecognized command " + c);
}

void Interp::cmd_variables(Tokens &t) {
  const std::string w = t.need("a variables directive");
  if (w == "set" || w == "setupper" || w == "setlower") {
    const std::string key = t.need("a variable name");
    std::string v = t.rest();
    for (auto &ch : v) {
      if (w == "setupper") ch = (char)std::toupper((unsigned char)ch);
      if (w == "setlower") ch = (char)std::tolower((unsigned char)ch);
    }
    vars_[key] = v;
  } else if (w == "calculate") {
    const std::string key = t.need("a variable name"), type = t.need("int or real");
    const double v = evaluate(t);
    char buf[64];
    if (type == "int") snprintf(buf, sizeof buf, "%d", (int)std::floor(v + 0.5));
    else if (type == "real") snprintf(buf, sizeof buf, "%g", v);
    else fail("variables calculate takes int or real, not " + type);
    vars_[key] = buf;
  } else if (w == "reset") {
    vars_.clear();
  } else if (w == "print") {
    for (auto &kv : vars_) log("variables print> " + kv.first + " = " + kv.second + "\n");
  } else {
    fail("unrecognized token after variables: " + w);
  }
}

void Interp::cmd_parameters(Tokens &t) {
  const std::string w = t.need("a parameters directive");
  if (w == "reset") {
    blade_init_parameters(s_);
  } else if (w == "file") {
    if (!blade::api_has_parameters(s_)) blade_init_parameters(s_);
    const std::string path = t.need("a parameter file");
    if (blade_read_parameter_file(s_, path.c_str())) check_api("parameters file");
  } else if (w == "print") {
    log("parameters print> read\n");
  } else {
    fail("unrecognized token after parameters: " + w);
  }
}

void Interp::cmd_structure(Tokens &t) {
  const std::string w = t.need("a structure directive");
  if (w == "reset") {
    blade_init_structure(s_);
  } else if (w == "file") {
    std::string path = t.need("a PSF file");
    if (path == "psf") path = t.need("a PSF file");
    if (blade_read_psf(s_, path.c_str())) check_api("structure file");
    sel_.clear();
  } else if (w == "shake") {
    const std::string m = t.need("hbond or none");
    if (m == "hbond") blade_add_shake(s_, 1);
    else if (m == "none") blade_add_shake(s_, 0);
    else fail("structure shake takes hbond or none, not " + m);
  } else if (w == "print") {
    log("structure print> " + std::to_string(natoms()) + " atoms\n");
  } else {
    fail("structure " + w + " is not supported by this interpreter");
  }
}

const std::vector<char> &Interp::named(const std::string &name) const {
  auto it = sel_.find(name);
  if (it == sel_.end()) fail("selection " + name + " not found");
  return it->second;
}

// selection expressions, operators first (system/selections.cxx:100-240)
std::vector<char> Interp::select(Tokens &t) {
  static const std::set<std::string> known = {"not", "and", "or", "none", "all", "hydrogen", "segid",
                                              "resid", "resids", "residrange", "resname", "atomname",
                                              "atomnames", "atom", "atomsearch", "selection", ""};
  const blade::Structure *st = blade::api_structure(s_);
  if (!st) fail("selections need a structure");
  const std::vector<AtomRec> &A = st->atoms;
  const size_t N = A.size();
  std::vector<char> s(N, 0);
  const std::string w = t.need("a selection token");
  auto each = [&](auto pred) {
    for (size_t i = 0; i < N; i++) s[i] = pred(A[i]) ? 1 : 0;
  };
  // a '%' in a pattern matches any one character; "*" matches anything
  auto match = [](const std::string &a, const std::string &p) {
    if (p == "*") return true;
    if (a.size() != p.size() || p.empty()) return false;
    for (size_t k = 0; k < p.size(); k++)
      if (p[k] != '%' && p[k] != a[k]) return false;
    return true;
  };
  if (w == "not") {
    s = select(t);
    for (auto &v : s) v = !v;
  } else if (w == "and" || w == "or") {
    s = select(t);
    const std::vector<char> b = select(t);
    for (size_t i = 0; i < N; i++) s[i] = w == "and" ? (s[i] && b[i]) : (s[i] || b[i]);
  } else if (w == "none") {
  } else if (w == "all") {
    each([](const AtomRec &) { return true; });
  } else if (w == "hydrogen") {
    each([](const AtomRec &a) { return !a.atomTypeName.empty() && a.atomTypeName[0] == 'H'; });
  } else if (w == "segid") {
    const std::string v = t.need("a segid");
    each([&](const AtomRec &a) { return a.segName == v; });
  } else if (w == "resid") {
    const std::string v = t.need("a resid");
    each([&](const AtomRec &a) { return a.resIdx == v; });
  } else if (w == "resids" || w == "atomnames") {
    std::set<std::string> vs;
    while (!known.count(t.peek())) vs.insert(t.next());
    if (w == "resids") each([&](const AtomRec &a) { return vs.count(a.resIdx) > 0; });
    else each([&](const AtomRec &a) { return vs.count(a.atomName) > 0; });
  } else if (w == "residrange") {
    const long a0 = t.integer("a first resid"), a1 = t.integer("a last resid");
    each([&](const AtomRec &a) {
      const long r = std::strtol(a.resIdx.c_str(), nullptr, 10);
      return r >= a0 && r <= a1;
    });
  } else if (w == "resname") {
    const std::string v = t.need("a resname");
    each([&](const AtomRec &a) { return a.resName == v; });
  } else if (w == "atomname") {
    const std::string v = t.need("an atom name");
    each([&](const AtomRec &a) { return a.atomName == v; });
  } else if (w == "atom" || w == "atomsearch") {
    const std::string sg = t.need("a segid"), rs = t.need("a resid"), an = t.need("an atom name");
    if (w == "atom")
      each([&](const AtomRec &a) { return a.segName == sg && a.resIdx == rs && a.atomName == an; });
    else
      each([&](const AtomRec &a) { return match(a.segName, sg) && match(a.resIdx, rs) && match(a.atomName, an); });
  } else if (w == "selection") {
    s = named(t.need("a selection name"));
  } else {
    fail("unrecognized selection token " + w);
  }
  return s;
}

void Interp::cmd_selection(Tokens &t) {
  const std::string w = t.need("a selection directive");
  if (w == "define") {
    const std::string name = t.need("a selection name");
    if (!sel_.count(name) && sel_.size() >= selLimit_)
      fail("selection limit reached; raise it with selection limit");
    sel_[name] = select(t);
  } else if (w == "delete") {
    const std::string name = t.need("a selection name");
    if (!sel_.erase(name)) fail("selection " + name + " not found");
  } else if (w == "count") {
    const std::vector<char> &s = named(t.need("a selection name"));
    long n = 0;
    for (char v : s) n += v != 0;
    vars_[t.need("a variable name")] = std::to_string(n);
  } else if (w == "limit") {
    selLimit_ = (size_t)t.integer("a selection limit");
  } else if (w == "reset") {
    sel_.clear();
  } else if (w == "print") {
    for (auto &kv : sel_) {
      long n = 0;
      for (char v : kv.second) n += v != 0;
      log("selection print> " + kv.first + " contains " + std::to_string(n) + " atoms\n");
    }
  } else {
    fail("unrecognized token after selection: " + w);
  }
}

void Interp::cmd_coordinates(Tokens &t) {
  const std::string w = t.need("a coordinates directive");
  if (w == "reset") {
    blade_init_coordinates(s_, natoms());
  } else if (w == "file") {
    const std::string type = t.need("crd or pdb"), path = t.need("a coordinate file");
    int rc;
    if (type == "crd") rc = blade_read_crd(s_, path.c_str());
    else if (type == "pdb") rc = blade_read_pdb(s_, path.c_str());
    else fail("coordinates file takes crd or pdb, not " + type);
    if (rc) check_api("coordinates file");
    if (blade_get_device_system(s_)) {
      // a running system continues from the coordinates just read
      if (blade::api_reset_device(s_)) check_api("coordinates file");
      builtWith_.clear();
    }
  } else if (w == "box") {
    static const char *names[] = {"cubi", "tetr", "orth", "mono", "tric", "hexa", "rhom", "octa", "rhdo"};
    const std::string nm = t.need("a box type");
    int id = -1;
    for (int k = 0; k < 9; k++)
      if (nm == names[k]) id = k;
    if (id < 0) fail("unrecognized box name " + nm);
    double v[6];
    for (double &x : v) x = t.real("a box length or angle");
    if (!blade::api_has_coordinates(s_)) blade_init_coordinates(s_, natoms());
    blade_add_coordinates_box(s_, id, v[0], v[1], v[2], v[3], v[4], v[5]);
  } else if (w == "velocity") {
    // Maxwell-Boltzmann velocities at T (system/coordinates.cxx:308-320)
    // from Philox normals keyed by the atom and axis
    const double T = t.real("a temperature");
    const blade::Structure *st = blade::api_structure(s_);
    if (!st) fail("coordinates velocity needs a structure");
    if (!blade::api_has_coordinates(s_)) blade_init_coordinates(s_, natoms());
    const uint64_t seed = 0x5645ull ^ (uint64_t)(++velDraws_) * 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < st->atoms.size(); i++) {
      double v[3];
      for (int k = 0; k < 3; k++) {
        const double u1 = bm_uniform(seed, (long)i, 2u * (unsigned)k + 1u);
        const double u2 = bm_uniform(seed, (long)i, 2u * (unsigned)k + 2u);
        const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
        const double m = st->atoms[i].mass;
        v[k] = m > 0 ? std::sqrt(kBoltz * T / m) * g : 0.0;
      }
      blade_add_coordinates_velocity(s_, (int)i + 1, v[0], v[1], v[2]);
    }
  } else if (w == "print") {
    log("coordinates print> " + std::to_string(natoms()) + " atoms\n");
  } else {
    fail("unrecognized token after coordinates: " + w);
  }
}

void Interp::push_msld_flags() {
  blade_add_msld_flags(s_, ms_.gamma_ps / kPicosecond, ms_.fnex, ms_.softCore, ms_.softCore14, ms_.ewaldType,
                       ms_.kRestraint, ms_.softBondRadius, ms_.softBondExponent, ms_.softNotBondExponent, ms_.fix);
  blade_add_msld_charges(s_, ms_.kCR1, ms_.kCR2, ms_.kCR3, ms_.q0CR3, ms_.wCR3);
  blade_set_msld_piecewise_constraint(s_, ms_.piecewise, ms_.wellWidth, ms_.wellK);
  blade_set_msld_thetaedgebias(s_, ms_.edgeK, ms_.edgeN, ms_.edgeAlpha, ms_.edgePhi);
}

void Interp::cmd_msld(Tokens &t) {
  const std::string w = t.need("an msld directive");
  if (w == "nblocks") {
    blade_init_msld(s_, (int)t.integer("a block count"));
    ms_ = MsldFlags();
    push_msld_flags();
    return;
  }
  if (w == "reset") {
    blade_init_msld(s_, 0);
    ms_ = MsldFlags();
    return;
  }
  if (!blade::api_has_msld(s_)) fail("msld nblocks comes before other msld commands");
  if (w == "call") {
    const long b = t.integer("a block");
    const std::vector<char> &s = named(t.need("a selection name"));
    for (size_t i = 0; i < s.size(); i++)
      if (s[i]) blade_add_msld_atomassignment(s_, (int)i + 1, (int)b + 1);
  } else if (w == "initialize") {
    const long b = t.integer("a block"), site = t.integer("a site");
    const double th = t.real("theta"), tv = t.real("theta velocity"), tm = t.real("theta mass");
    const double fb = t.real("fixed bias"), q = t.real("block charge");
    blade_add_msld_initialconditions(s_, (int)b + 1, (int)site + 1, th, tv, tm, fb, q);
  } else if (w == "gamma") {
    ms_.gamma_ps = t.real("gamma");
  } else if (w == "fnex") {
    ms_.fnex = t.real("fnex");
  } else if (w == "bias") {
    const long i = t.integer("a block"), j = t.integer("a block"), type = t.integer("a bias type");
    const double l0 = t.real("l0"), k = t.real("k");
    const long n = t.integer("n");
    blade_add_msld_bias(s_, (int)i + 1, (int)j + 1, (int)type, l0, k, (int)n);
  } else if (w == "thetabias") {
    const std::string kind = t.need("collective or independent"), what = t.need("nsites, set or auto");
    const bool coll = kind == "collective";
    if (!coll && kind != "independent") fail("thetabias takes collective or independent, not " + kind);
    if (what == "nsites") {
      (coll ? ms_.collSites : ms_.indeSites) = (int)t.integer("a site count") + 1;
      return;
    }
    const int count = coll ? ms_.collSites : ms_.indeSites;
    if (!count) fail("thetabias " + kind + " nsites comes before set or auto");
    long i0 = t.integer("a site");
    long i1 = i0 + 1;
    if (i0 == 0) {
      i0 = 1;
      i1 = count;
    }
    if (coll) {
      if (what != "set") fail("thetabias collective takes nsites or set");
      const double k = t.real("k"), n = t.real("N");
      for (long i = i0; i < i1; i++) blade_add_msld_thetacollbias(s_, count, (int)i + 1, k, n);
    } else {
      double k = 0;
      if (what == "set") {
        k = t.real("k");
      } else if (what == "auto") {
        // the bias that flattens a site of Ns substituents at T
        // (msld/msld.cu:303-318, a fixed-point iteration)
        const double ns = t.real("Ns"), T = t.real("T");
        k = 1;
        for (int it = 0; it < 50; it++) {
          k = 0.5 * std::log(0.25 * k * ns * ns * M_PI / 2);
          if (!(k > 0)) k = 0;
        }
        k *= kBoltz * T;
      } else {
        fail("thetabias independent takes nsites, set or auto");
      }
      for (long i = i0; i < i1; i++) blade_add_msld_thetaindebias(s_, count, (int)i + 1, k);
    }
    return;
  } else if (w == "removescaling") {
    int sc[6] = {1, 1, 1, 1, 1, 1};
    static const char *names[] = {"bond", "urey", "angle", "dihe", "impr", "cmap"};
    std::string nm;
    while (!(nm = t.next()).empty()) {
      int id = -1;
      for (int k = 0; k < 6; k++)
        if (nm == names[k]) id = k;
      if (id < 0) fail("removescaling takes bond, urey, angle, dihe, impr or cmap, not " + nm);
      sc[id] = 0;
    }
    blade_add_msld_termscaling(s_, sc[0], sc[1], sc[2], sc[3], sc[4], sc[5]);
    return;
  } else if (w == "softcore") {
    ms_.softCore = t.boolean("softcore");
  } else if (w == "softcore14") {
    ms_.softCore14 = t.boolean("softcore14");
  } else if (w == "ewaldtype") {
    ms_.ewaldType = (int)t.integer("an Ewald type");
    if (ms_.ewaldType < 1 || ms_.ewaldType > 3) fail("msld ewaldtype must be 1, 2 or 3");
  } else if (w == "parameter") {
    const std::string p = t.need("a parameter name");
    const double v = t.real("a parameter value");
    if (p == "krestraint") ms_.kRestraint = v;
    else if (p == "kchargerestraint1") ms_.kCR1 = v;
    else if (p == "kchargerestraint2") ms_.kCR2 = v;
    else if (p == "kchargerestraint3") ms_.kCR3 = v;
    else if (p == "q0chargerestraint3") ms_.q0CR3 = v;
    else if (p == "wchargerestraint3") ms_.wCR3 = v;
    else if (p == "softbondradius") ms_.softBondRadius = v;
    else if (p == "softbondexponent") ms_.softBondExponent = v;
    else if (p == "softnotbondexponent") ms_.softNotBondExponent = v;
    else if (p == "restscaling") fail("msld parameter restscaling (REST scaling) is not supported");
    else fail("unrecognized msld parameter " + p);
  } else if (w == "restrain" || w == "atomrestraint") {
    const std::string name = t.need("a selection name");
    if (name == "reset") fail("msld restrain reset is not supported by this interpreter");
    const std::vector<char> &s = named(name);
    bool any = false;
    for (size_t i = 0; i < s.size(); i++) {
      if (!s[i]) continue;
      if (!any) blade_add_msld_atomrestraint(s_);
      any = true;
      blade_add_msld_atomrestraint_element(s_, (int)i + 1);
    }
    if (!any && w == "atomrestraint") blade_add_msld_atomrestraint(s_);
  } else if (w == "softbond") {
    const std::vector<char> &s = named(t.need("a selection name"));
    std::vector<int> at;
    for (size_t i = 0; i < s.size(); i++)
      if (s[i]) at.push_back((int)i + 1);
    if (at.size() != 2) fail("a soft bond selection holds 2 atoms, not " + std::to_string(at.size()));
    blade_add_msld_softbond(s_, at[0], at[1]);
  } else if (w == "fix") {
    ms_.fix = t.boolean("fix");
  } else if (w == "piecewise") {
    ms_.piecewise = t.boolean("piecewise");
  } else if (w == "well_width") {
    ms_.wellWidth = t.real("well width");
  } else if (w == "well_k") {
    ms_.wellK = t.real("well k");
  } else if (w == "N_target") {
    ms_.edgeN = t.real("N target");
  } else if (w == "imp_m") {
    ms_.edgeK = t.real("edge bias k");
  } else if (w == "imp_alpha") {
    ms_.edgeAlpha = t.real("edge bias alpha");
  } else if (w == "imp_xi") {
    ms_.edgePhi = t.real("edge bias xi");
  } else if (w == "rest") {
    fail("msld rest (REST scaling of dihedrals) is not supported");
  } else if (w == "print") {
    for (auto &kv : sel_) log("msld print> selection " + kv.first + "\n");
    return;
  } else {
    fail("unrecognized token after msld: " + w);
  }
  push_msld_flags();
}

// the run parameters into the system; a device system built with other
// physics parameters is rebuilt from its current state, as the reference
// builds its state again at every dynamics call (run/run.cu:572-622)
void Interp::prepare_run() {
  if (!blade::api_has_msld(s_)) fail("msld nblocks comes before run (use msld nblocks 0 for plain dynamics)");
  const std::vector<double> phys = run_.physics();
  if (blade_get_device_system(s_) && phys != builtWith_) {
    if (blade::api_reset_device(s_)) check_api("run");
  }
  blade_init_run(s_);
  blade_add_run_flags(s_, run_.gamma_ps / kPicosecond, 1.0 / run_.invBetaEwald, run_.rCut, run_.rSwitch, run_.vdw,
                      run_.elec, run_.gridSpace, run_.grid[0], run_.grid[1], run_.grid[2], run_.order,
                      run_.shakeTol);
  blade_add_run_dynopts(s_, 0, 0, (int)run_.nsteps, run_.dt_ps * kPicosecond, run_.T, run_.freqNPT,
                        run_.volumeFluctuation, run_.pressure_atm * kAtmosphere);
  if (!blade_get_device_system(s_) && !run_.fnmCPI.empty() && !cpiRead_) {
    if (blade_read_checkpoint(s_, run_.fnmCPI.c_str())) check_api("fnmcpi");
    cpiRead_ = true;
  }
  builtWith_ = phys;
}

void Interp::cmd_run(Tokens &t) {
  const std::string w = t.need("a run directive");
  if (w == "setvariable") {
    const std::string v = t.need("a run variable");
    if (v == "dt") run_.dt_ps = t.real("dt");
    else if (v == "nsteps") run_.nsteps = t.integer("nsteps");
    else if (v == "fnmxtc") run_.fnmXTC = t.need("a file name"), outputsOpen_ = false;
    else if (v == "fnmlmd") run_.fnmLMD = t.need("a file name"), outputsOpen_ = false;
    else if (v == "fnmnrg") run_.fnmNRG = t.need("a file name"), outputsOpen_ = false;
    else if (v == "fnmcpi") run_.fnmCPI = t.need("a file name");
    else if (v == "fnmcpo") run_.fnmCPO = t.need("a file name");
    else if (v == "freqxtc") run_.freqXTC = (int)t.integer("freqxtc"), outputsOpen_ = false;
    else if (v == "freqlmd") run_.freqLMD = (int)t.integer("freqlmd"), outputsOpen_ = false;
    else if (v == "freqnrg") run_.freqNRG = (int)t.integer("freqnrg"), outputsOpen_ = false;
    else if (v == "hrlmd") run_.hrLMD = t.boolean("hrlmd"), outputsOpen_ = false;
    else if (v == "prettyxtc") (void)t.boolean("prettyxtc");   // frames keep the wrapped positions
    else if (v == "T") run_.T = t.real("T");
    else if (v == "gamma") run_.gamma_ps = t.real("gamma");
    else if (v == "invbetaewald") run_.invBetaEwald = t.real("invbetaewald");
    else if (v == "rcut") run_.rCut = t.real("rcut");
    else if (v == "rswitch") run_.rSwitch = t.real("rswitch");
    else if (v == "vfswitch") run_.vdw = t.boolean("vfswitch") ? BM_VSWITCH : BM_VFSWITCH;
    else if (v == "usepme") run_.elec = t.boolean("usepme") ? BM_PME : BM_FSWITCH;
    else if (v == "vdwmethod") {
      const std::string m = t.need("a vdW method");
      if (m == "vfswitch") run_.vdw = BM_VFSWITCH;
      else if (m == "vswitch") run_.vdw = BM_VSWITCH;
      else if (m == "vshift") run_.vdw = BM_VSHIFT;
      else fail("vdwmethod takes vfswitch, vswitch or vshift, not " + m);
    } else if (v == "elecmethod") {
      const std::string m = t.need("an electrostatics method");
      if (m == "fswitch") run_.elec = BM_FSWITCH;
      else if (m == "pme") run_.elec = BM_PME;
      else if (m == "fshift") run_.elec = BM_FSHIFT;
      else fail("elecmethod takes fswitch, pme or fshift, not " + m);
    } else if (v == "gridspace") {
      run_.gridSpace = t.real("gridspace");
      run_.grid[0] = run_.grid[1] = run_.grid[2] = -1;
    } else if (v == "grid") {
      for (int &g : run_.grid) g = (int)t.integer("a grid size");
      run_.gridSpace = -1;
    } else if (v == "orderewald") {
      run_.order = (int)t.integer("orderewald");
      if (run_.order != 4 && run_.order != 6 && run_.order != 8) fail("orderewald must be 4, 6 or 8");
    } else if (v == "shaketolerance") run_.shakeTol = t.real("shaketolerance");
    else if (v == "freqnpt") run_.freqNPT = (int)t.integer("freqnpt");
    else if (v == "volumefluctuation") run_.volumeFluctuation = t.real("volumefluctuation");
    else if (v == "pressure") run_.pressure_atm = t.real("pressure");
    else if (v == "domdecheuristic") (void)t.next();   // the reference ignores it too
    else if (v == "dxatommax" || v == "dxrmsinit" || v == "mintype" || v == "lbfgs_m" || v == "lbfgs_eps")
      (void)t.next();   // minimization settings (minimization is not supported)
    else fail("unrecognized run variable " + v);
  } else if (w == "dynamics" || w == "energy") {
    prepare_run();
    if (w == "energy") {
      blade_run_energy(s_);
      check_api("run energy");
      double e[BM_NTERMS];
      if (blade_get_energy(s_, e)) check_api("run energy");
      std::string line;
      char buf[32];
      for (double x : e) {
        snprintf(buf, sizeof buf, " %12.4f", x);
        line += buf;
      }
      log(line + "\n");
      return;
    }
    if (!outputsOpen_) {
      // output files open at the first dynamics call after they are named
      // (run/run.cu:572-590)
      if (blade_set_output_xtc(s_, run_.freqXTC > 0 ? run_.fnmXTC.c_str() : nullptr, std::max(run_.freqXTC, 0)) ||
          blade_set_output_nrg(s_, run_.freqNRG > 0 ? run_.fnmNRG.c_str() : nullptr, std::max(run_.freqNRG, 0)))
        check_api("run dynamics outputs");
      blade_set_lmd_format(s_, run_.hrLMD);
      if (blade_set_output_lmd(s_, run_.freqLMD > 0 ? run_.fnmLMD.c_str() : nullptr, std::max(run_.freqLMD, 0)))
        check_api("run dynamics outputs");
      outputsOpen_ = true;
    }
    const auto t0 = std::chrono::steady_clock::now();
    if (blade_run_dynamics(s_, (int)run_.nsteps, run_.freqNRG, 1)) check_api("run dynamics");
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    char buf[96];
    snprintf(buf, sizeof buf, "Elapsed dynamics time: %f\n", sec);
    log(buf);
    if (!run_.fnmCPO.empty() && blade_write_checkpoint(s_, run_.fnmCPO.c_str())) check_api("fnmcpo");
  } else if (w == "reset") {
    run_ = RunVars();
    outputsOpen_ = false;
  } else if (w == "print") {
    char buf[256];
    snprintf(buf, sizeof buf, "run print> nsteps %ld dt %g ps T %g K gamma %g /ps rcut %g rswitch %g\n", run_.nsteps,
             run_.dt_ps, run_.T, run_.gamma_ps, run_.rCut, run_.rSwitch);
    log(buf);
  } else {
    fail("run " + w + " is not supported by this interpreter");
  }
}

}  // namespace

extern "C" void blade_interpretter(const char *fnm, System *system) {
  if (!system || !fnm) return;
  void *&slot = blade::api_script_state(system, [](void *p) { delete static_cast<ScriptState *>(p); });
  if (!slot) slot = new ScriptState();
  Interp in(system, *static_cast<ScriptState *>(slot));
  try {
    in.run_file(fnm);
  } catch (const ScriptError &e) {
    blade::api_report_error(system, std::string("script ") + fnm + ": " + e.what);
  }
}
