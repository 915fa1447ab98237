// files.cpp - CHARMM input files through the blade_* API: parameter files
// (PRM), protein structure files (PSF) and coordinates (CRD, PDB).
//
// Each reader feeds what it parses to the same blade_* setters a library
// caller uses, so a system read from files and one built by API calls are
// the same system.  Formats and unit handling follow the reference's
// readers: system/parameters.cxx:129-484 (PRM), system/structure.cxx:566-826
// (PSF), system/coordinates.cxx:152-274 (PDB, CRD).
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/blade.h"

namespace blade {
void api_report_error(System *s, const std::string &m);
bool api_has_structure(System *s);
}  // namespace blade

namespace {

using blade::api_report_error;

std::string upper(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// whitespace tokens of one line up to a '!' comment
std::vector<std::string> tokens(const std::string &line) {
  std::vector<std::string> out;
  std::istringstream is(line);
  std::string t;
  while (is >> t) {
    if (t[0] == '!') break;
    out.push_back(t);
  }
  return out;
}

bool to_double(const std::string &s, double &v) {
  char *end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == '\0' && end != s.c_str();
}

bool to_int(const std::string &s, int &v) {
  char *end = nullptr;
  long x = std::strtol(s.c_str(), &end, 10);
  v = (int)x;
  return end && *end == '\0' && end != s.c_str();
}

struct ParseError {
  std::string what;
};

double need_double(const std::vector<std::string> &t, size_t k, const std::string &where) {
  double v;
  if (k >= t.size() || !to_double(t[k], v)) throw ParseError{"expected a number in " + where};
  return v;
}

int need_int(const std::vector<std::string> &t, size_t k, const std::string &where) {
  int v;
  if (k >= t.size() || !to_int(t[k], v)) throw ParseError{"expected an integer in " + where};
  return v;
}

double opt_double(const std::vector<std::string> &t, size_t k, double dflt) {
  double v;
  return (k < t.size() && to_double(t[k], v)) ? v : dflt;
}

// section keywords of a parameter file, by their first four letters
enum class Sec { None, Atoms, Bonds, Angles, Dihes, Imprs, Cmap, Nonb, Nbfix, Hbond };

bool section_of(const std::string &tok, Sec &s) {
  const std::string u = upper(tok).substr(0, 4);
  static const std::map<std::string, Sec> m = {
      {"ATOM", Sec::Atoms}, {"BOND", Sec::Bonds}, {"ANGL", Sec::Angles}, {"THET", Sec::Angles},
      {"DIHE", Sec::Dihes}, {"PHI", Sec::Dihes},  {"IMPR", Sec::Imprs},  {"IMPH", Sec::Imprs},
      {"CMAP", Sec::Cmap},  {"NONB", Sec::Nonb},  {"NBON", Sec::Nonb},   {"NBFI", Sec::Nbfix},
      {"HBON", Sec::Hbond}};
  auto it = m.find(u);
  if (it == m.end()) return false;
  s = it->second;
  return true;
}

// one line, without a trailing carriage return (option lines continued with
// '-' need no joining: the continuation starts with an option keyword)
bool next_line(std::istream &in, std::string &line) {
  if (!std::getline(in, line)) return false;
  if (!line.empty() && line.back() == '\r') line.pop_back();
  return true;
}

}  // namespace

extern "C" {

int blade_read_parameter_file(System *s, const char *path) {
  if (!s || !path) return -1;
  std::ifstream in(path);
  if (!in) {
    api_report_error(s, std::string("cannot open parameter file ") + path);
    return -1;
  }
  Sec sec = Sec::None;
  double e14fac = 1.0;
  int combine = 0;
  std::string line;
  int lineno = 0;
  // a CMAP grid being read: its eight types, its size and the values read
  // so far (ngrid*ngrid values, phi-major, over as many lines as they take)
  std::vector<std::string> cmapTypes;
  int cmapGrid = 0, cmapRead = 0;
  try {
    while (next_line(in, line)) {
      lineno++;
      if (!line.empty() && line[0] == '*') continue;   // title
      auto t = tokens(line);
      if (t.empty()) continue;
      if (cmapRead < cmapGrid * cmapGrid) {
        const std::string where = std::string(path) + ":" + std::to_string(lineno);
        for (size_t k = 0; k < t.size(); k++) {
          if (cmapRead == cmapGrid * cmapGrid) throw ParseError{"extra values after a CMAP grid in " + where};
          const double v = need_double(t, k, where);
          const auto &c = cmapTypes;
          blade_add_parameter_cmaps_fill(s, c[0].c_str(), c[1].c_str(), c[2].c_str(), c[3].c_str(), c[4].c_str(),
                                         c[5].c_str(), c[6].c_str(), c[7].c_str(), cmapRead / cmapGrid + 1,
                                         cmapRead % cmapGrid + 1, v);
          cmapRead++;
        }
        continue;
      }
      const std::string U = upper(t[0]);
      if (U == "END") break;
      Sec ns;
      if (section_of(t[0], ns)) {
        sec = ns;
        if (sec == Sec::Nonb) {
          // header options: e14fac and geometric combination are used
          e14fac = 1.0;
          combine = 0;
          for (size_t k = 1; k < t.size(); k++) {
            const std::string o = upper(t[k]);
            if (o == "E14FAC") e14fac = need_double(t, k + 1, "NONBONDED e14fac");
            if (o == "GEOM") combine = 1;
          }
        }
        continue;
      }
      const std::string where = std::string(path) + ":" + std::to_string(lineno);
      switch (sec) {
        case Sec::Atoms:
          // MASS index type mass [element]
          if (U == "MASS") blade_add_parameter_atoms(s, t.at(2).c_str(), need_double(t, 3, where));
          break;
        case Sec::Bonds:
          blade_add_parameter_bonds(s, t.at(0).c_str(), t.at(1).c_str(), need_double(t, 2, where),
                                    need_double(t, 3, where));
          break;
        case Sec::Angles:
          blade_add_parameter_angles(s, t.at(0).c_str(), t.at(1).c_str(), t.at(2).c_str(), need_double(t, 3, where),
                                     need_double(t, 4, where), opt_double(t, 5, 0.0), opt_double(t, 6, 0.0));
          break;
        case Sec::Dihes:
          blade_add_parameter_dihes(s, t.at(0).c_str(), t.at(1).c_str(), t.at(2).c_str(), t.at(3).c_str(),
                                    need_double(t, 4, where), need_int(t, 5, where), need_double(t, 6, where));
          break;
        case Sec::Imprs:
          blade_add_parameter_imprs(s, t.at(0).c_str(), t.at(1).c_str(), t.at(2).c_str(), t.at(3).c_str(),
                                    need_double(t, 4, where), need_int(t, 5, where), need_double(t, 6, where));
          break;
        case Sec::Nonb: {
          // type ignored eps Rmin/2 [ignored eps14 Rmin/2_14]; header
          // continuation lines carry options only
          const std::string o = upper(t[0]);
          if (o == "CUTNB" || o == "CTOFNB" || o == "CTONNB" || o == "EPS" || o == "E14FAC" || o == "WMIN" ||
              o == "GEOM" || o == "NBXMOD" || o == "ATOM" || o == "CDIEL" || o == "SHIFT" || o == "VATOM" ||
              o == "VDISTANCE" || o == "VSWITCH" || o == "SWITCH" || o == "VFSWITCH" || o == "FSWITCH") {
            for (size_t k = 0; k < t.size(); k++) {
              const std::string oo = upper(t[k]);
              if (oo == "E14FAC") e14fac = need_double(t, k + 1, where);
              if (oo == "GEOM") combine = 1;
            }
            break;
          }
          const double eps = need_double(t, 2, where), sig = need_double(t, 3, where);
          const double eps14 = opt_double(t, 5, eps), sig14 = opt_double(t, 6, sig);
          blade_add_parameter_nbonds(s, t[0].c_str(), eps, sig, eps14, sig14, e14fac, combine);
          break;
        }
        case Sec::Nbfix: {
          const double eps = need_double(t, 2, where), sig = need_double(t, 3, where);
          blade_add_parameter_nbfixs(s, t.at(0).c_str(), t.at(1).c_str(), eps, sig, opt_double(t, 4, eps),
                                     opt_double(t, 5, sig));
          break;
        }
        case Sec::Cmap: {
          // eight atom types (two dihedrals), then the grid size; the grid's
          // values follow on the next lines
          const int ng = need_int(t, 8, where);
          if (ng < 4 || ng > 60) throw ParseError{"CMAP grid must have 4 to 60 points per 360 degrees in " + where};
          cmapTypes.assign(t.begin(), t.begin() + 8);
          blade_add_parameter_cmaps(s, t[0].c_str(), t[1].c_str(), t[2].c_str(), t[3].c_str(), t[4].c_str(),
                                    t[5].c_str(), t[6].c_str(), t[7].c_str(), ng);
          cmapGrid = ng;
          cmapRead = 0;
          break;
        }
        case Sec::Hbond:
        case Sec::None:
          break;
      }
    }
  } catch (const ParseError &e) {
    api_report_error(s, e.what);
    return -1;
  } catch (const std::out_of_range &) {
    api_report_error(s, std::string("short parameter line in ") + path + ":" + std::to_string(lineno));
    return -1;
  }
  return 0;
}

int blade_read_psf(System *s, const char *path) {
  if (!s || !path) return -1;
  std::ifstream in(path);
  if (!in) {
    api_report_error(s, std::string("cannot open PSF ") + path);
    return -1;
  }
  std::string line;
  try {
    if (!next_line(in, line) || tokens(line).empty() || tokens(line)[0] != "PSF")
      throw ParseError{"the first line of a PSF starts with PSF"};
    bool cheq = false;
    for (auto &h : tokens(line)) cheq = cheq || h == "CHEQ";
    // section header: a count as the first token ("  648 !NATOM")
    auto header = [&](const char *what) {
      while (next_line(in, line)) {
        auto t = tokens(line);
        if (!t.empty()) return need_int(t, 0, std::string("PSF ") + what);
      }
      throw ParseError{std::string("PSF ends before ") + what};
    };
    // the integers following a section header, across lines
    std::vector<std::string> pend;
    size_t pos = 0;
    auto next_int = [&](const char *what) {
      while (pos >= pend.size()) {
        if (!next_line(in, line)) throw ParseError{std::string("PSF ends inside ") + what};
        pend = tokens(line);
        pos = 0;
      }
      return need_int(pend, pos++, std::string("PSF ") + what);
    };
    auto section_ints = [&](const char *what, int count) {
      std::vector<int> v(count);
      pend.clear();
      pos = 0;
      for (int k = 0; k < count; k++) v[k] = next_int(what);
      return v;
    };
    const int ntitle = header("!NTITLE");
    for (int k = 0; k < ntitle; k++) next_line(in, line);
    const int natom = header("!NATOM");
    blade_init_structure(s);
    for (int k = 0; k < natom; k++) {
      if (!next_line(in, line)) throw ParseError{"PSF ends inside the atom list"};
      auto t = tokens(line);
      const int idx = need_int(t, 0, "PSF atom");
      if (idx != k + 1) throw ParseError{"PSF atom " + std::to_string(idx) + " out of order"};
      blade_add_atom(s, idx, t.at(1).c_str(), t.at(2).c_str(), t.at(3).c_str(), t.at(4).c_str(), t.at(5).c_str(),
                     need_double(t, 6, "PSF atom charge"), need_double(t, 7, "PSF atom mass"));
    }
    auto check = [&](int a, const char *what) {
      if (a < 1 || a > natom) throw ParseError{std::string("PSF ") + what + " atom out of range"};
      return a;
    };
    const int nbond = header("!NBOND");
    auto b = section_ints("bonds", 2 * nbond);
    for (int k = 0; k < nbond; k++) blade_add_bond(s, check(b[2 * k], "bond"), check(b[2 * k + 1], "bond"));
    const int ntheta = header("!NTHETA");
    auto a = section_ints("angles", 3 * ntheta);
    for (int k = 0; k < ntheta; k++)
      blade_add_angle(s, check(a[3 * k], "angle"), check(a[3 * k + 1], "angle"), check(a[3 * k + 2], "angle"));
    const int nphi = header("!NPHI");
    auto p = section_ints("dihedrals", 4 * nphi);
    for (int k = 0; k < nphi; k++)
      blade_add_dihe(s, check(p[4 * k], "dihedral"), check(p[4 * k + 1], "dihedral"), check(p[4 * k + 2], "dihedral"),
                     check(p[4 * k + 3], "dihedral"));
    const int nimphi = header("!NIMPHI");
    auto q = section_ints("impropers", 4 * nimphi);
    for (int k = 0; k < nimphi; k++)
      blade_add_impr(s, check(q[4 * k], "improper"), check(q[4 * k + 1], "improper"), check(q[4 * k + 2], "improper"),
                     check(q[4 * k + 3], "improper"));
    // donors, acceptors, non-bonded exclusions, groups (and CHEQ molecule
    // numbers): read past
    section_ints("donors", 2 * header("!NDON"));
    section_ints("acceptors", 2 * header("!NACC"));
    header("!NNB");
    section_ints("nnb", natom);
    section_ints("groups", 3 * header("!NGRP"));
    if (cheq) {
      header("!MOLNT");
      section_ints("molnt", natom);
    }
    if (next_line(in, line)) {
      auto t = tokens(line);
      while (t.empty() && next_line(in, line)) t = tokens(line);
      const int nlp = t.empty() ? 0 : need_int(t, 0, "PSF lone pairs");
      const int nlph = t.size() > 1 ? need_int(t, 1, "PSF lone pair hosts") : 0;
      if (nlp > 0) {
        // lone pairs (system/structure.cxx:734-800): per pair its host
        // count, a pointer into the host list, a weighting flag and three
        // values; then the host list, each entry the site followed by its
        // hosts.  Colinear pairs (2 hosts, no weighting) are two-host
        // sites (distance, scale); 3 hosts with a nonzero distance are
        // three-host sites (distance, angle, dihedral in degrees).
        struct Lp { int nh, ptr; double a, b, c; };
        std::vector<Lp> lps;
        pend.clear();
        pos = 0;
        auto next_tok = [&](const char *what) {
          while (pos >= pend.size()) {
            if (!next_line(in, line)) throw ParseError{std::string("PSF ends inside ") + what};
            pend = tokens(line);
            pos = 0;
          }
          return pend[pos++];
        };
        for (int k = 0; k < nlp; k++) {
          Lp lp;
          std::vector<std::string> one = {next_tok("lone pairs")};
          lp.nh = need_int(one, 0, "PSF lone pair host count");
          one = {next_tok("lone pairs")};
          lp.ptr = need_int(one, 0, "PSF lone pair host pointer") - 1;
          const std::string w = next_tok("lone pairs");
          const bool weighted = !(w == "F" || w == "f" || w == "0" || w == "false" || w == "no" || w == "off");
          double v[3];
          for (double &x : v) {
            one = {next_tok("lone pairs")};
            x = need_double(one, 0, "PSF lone pair value");
          }
          lp.a = v[0]; lp.b = v[1]; lp.c = v[2];
          if (!((lp.nh == 2 && !weighted) || (lp.nh == 3 && lp.a != 0)))
            throw ParseError{"PSF lone pair of an unsupported kind (" + std::to_string(lp.nh) + " hosts)"};
          lps.push_back(lp);
        }
        auto hosts = section_ints("lone pair hosts", nlph);
        for (const Lp &lp : lps) {
          if (lp.ptr < 0 || lp.ptr + lp.nh >= nlph) throw ParseError{"PSF lone pair host pointer out of range"};
          int a[4];
          for (int k = 0; k <= lp.nh; k++) a[k] = check(hosts[lp.ptr + k], "lone pair");
          if (lp.nh == 2) blade_add_virt2(s, a[0], a[1], a[2], lp.a, lp.b);
          else blade_add_virt3(s, a[0], a[1], a[2], a[3], lp.a, lp.b, lp.c);
        }
      }
    }
    // CMAP cross terms (two dihedrals of eight atoms each)
    while (next_line(in, line)) {
      auto t = tokens(line);
      if (t.empty()) continue;
      int ncrt = 0;
      if (to_int(t[0], ncrt) && line.find("NCRTERM") != std::string::npos && ncrt > 0) {
        auto c = section_ints("cross terms", 8 * ncrt);
        for (int k = 0; k < ncrt; k++)
          blade_add_cmap(s, c[8 * k], c[8 * k + 1], c[8 * k + 2], c[8 * k + 3], c[8 * k + 4], c[8 * k + 5],
                         c[8 * k + 6], c[8 * k + 7]);
      }
      break;
    }
  } catch (const ParseError &e) {
    api_report_error(s, e.what);
    return -1;
  } catch (const std::out_of_range &) {
    api_report_error(s, "short atom line in PSF");
    return -1;
  }
  return 0;
}

// atoms are matched to the structure by (segment, residue id, atom name)
static int place_coordinates(System *s, const std::map<std::tuple<std::string, std::string, std::string>,
                                                     std::tuple<double, double, double>> &xyz,
                             const char *what) {
  if (!blade::api_has_structure(s)) {
    api_report_error(s, std::string(what) + ": read the structure first");
    return -1;
  }
  const int n = blade_structure_atom_count(s);
  blade_init_coordinates(s, n);
  for (int i = 0; i < n; i++) {
    char seg[64], res[64], name[64];
    blade_structure_atom_names(s, i + 1, seg, res, name, 64);
    auto it = xyz.find({seg, res, name});
    if (it == xyz.end()) {
      api_report_error(s, std::string(what) + ": atom " + seg + " " + res + " " + name + " missing");
      return -1;
    }
    blade_add_coordinates_position(s, i + 1, std::get<0>(it->second), std::get<1>(it->second),
                                   std::get<2>(it->second));
  }
  return 0;
}

int blade_read_crd(System *s, const char *path) {
  if (!s || !path) return -1;
  std::ifstream in(path);
  if (!in) {
    api_report_error(s, std::string("cannot open CRD ") + path);
    return -1;
  }
  std::map<std::tuple<std::string, std::string, std::string>, std::tuple<double, double, double>> xyz;
  std::string line;
  try {
    // title lines start with '*'; then NATOM [EXT]; then one atom a line:
    // ATOMNO RESNO RES TYPE X Y Z SEGID RESID [WEIGHT]
    int natom = -1;
    while (next_line(in, line)) {
      if (!line.empty() && line[0] == '*') continue;
      auto t = tokens(line);
      if (t.empty()) continue;
      natom = need_int(t, 0, "CRD atom count");
      break;
    }
    if (natom < 0) throw ParseError{"CRD without an atom count"};
    for (int k = 0; k < natom; k++) {
      if (!next_line(in, line)) throw ParseError{"CRD ends inside the atom list"};
      auto t = tokens(line);
      xyz[{t.at(7), t.at(8), t.at(3)}] = {need_double(t, 4, "CRD x"), need_double(t, 5, "CRD y"),
                                          need_double(t, 6, "CRD z")};
    }
  } catch (const ParseError &e) {
    api_report_error(s, e.what);
    return -1;
  } catch (const std::out_of_range &) {
    api_report_error(s, "short atom line in CRD");
    return -1;
  }
  return place_coordinates(s, xyz, "CRD");
}

int blade_read_pdb(System *s, const char *path) {
  if (!s || !path) return -1;
  std::ifstream in(path);
  if (!in) {
    api_report_error(s, std::string("cannot open PDB ") + path);
    return -1;
  }
  // fixed columns: atom name 13-16, residue number 23-27, x/y/z 31-54,
  // segment 73-76; the first record of each atom wins
  std::map<std::tuple<std::string, std::string, std::string>, std::tuple<double, double, double>> xyz;
  std::string line;
  auto field = [&](size_t a, size_t len) {
    std::string f = line.size() > a ? line.substr(a, len) : std::string();
    std::istringstream is(f);
    std::string w;
    is >> w;
    return w;
  };
  while (next_line(in, line)) {
    if (line.compare(0, 6, "ATOM  ") != 0 && line.compare(0, 6, "HETATM") != 0) continue;
    double x, y, z;
    if (!to_double(field(30, 8), x) || !to_double(field(38, 8), y) || !to_double(field(46, 8), z)) {
      api_report_error(s, "PDB coordinate field unreadable");
      return -1;
    }
    auto key = std::make_tuple(field(72, 4), field(22, 5), field(12, 4));
    if (!xyz.count(key)) xyz[key] = {x, y, z};
  }
  return place_coordinates(s, xyz, "PDB");
}

}  // extern "C"
