// api.cpp - the reference-compatible blade_* library API over the MI355X
// runtime (restates the library functions of system/system.cxx:309-329,
// system/parameters.cxx:585-786, system/structure.cxx:960-1092,
// system/coordinates.cxx:323-368, msld/msld.cu:1570-1735 and
// run/run.cu:633-719).  Indices arrive 1-based as from CHARMM.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/blade.h"
#include "setup.h"

using namespace blade;

namespace {
thread_local std::string g_apiErr;
const double kDegrees = 0.017453292519943;  // main/defines.h:32
}  // namespace

struct System {
  int device = 0;
  int verbose = 0;
  bool failed = false;
  std::string error;               // this system's first error (empty when none)
  std::unique_ptr<Parameters> parameters;
  std::unique_ptr<Structure> structure;
  std::unique_ptr<Coordinates> coordinates;
  std::unique_ptr<MsldSetup> msld;
  std::unique_ptr<RunSetup> run;
  bool unsupported = false;
  std::string unsupportedWhat;
  Topology topo;
  bm_runparams rp{};
  bool prepared = false;
  bool fromCheckpoint = false;   // the coordinates came from blade_read_checkpoint
  bm_system *gpu = nullptr;
  long step = 0;
  unsigned long long seed = 0xB1ADE000ull;
  int freqDomdec = 10;
  int fftMode = 0;
  int tileChunk = 0;
  // output files written by blade_run_dynamics (run/run.cu:572-590)
  FILE *fpNRG = nullptr, *fpLMD = nullptr, *fpXTC = nullptr;
  int freqNRG = 0, freqLMD = 0, freqXTC = 0;
  bool hrLMD = true;                    // text .lmd, else XDR frames
  long nptMoves = 0, nptAccepted = 0;   // barostat counters
  // the script interpreter's state (script.cpp), released with the system
  void *script = nullptr;
  void (*scriptRelease)(void *) = nullptr;
  ~System() {
    if (script && scriptRelease) scriptRelease(script);
    if (fpNRG) fclose(fpNRG);
    if (fpLMD) fclose(fpLMD);
    if (fpXTC) fclose(fpXTC);
  }
};

extern "C" int blade_xtc_frame(FILE *fp, int natoms, int step, float time, const float *box9, const float *xyz,
                               float precision);

extern "C" __attribute__((weak)) void blade_log(const char *message) { fprintf(stderr, "%s", message); }

static void api_fail(System *s, const std::string &m) {
  g_apiErr = m;
  if (s) {
    if (!s->failed) s->error = m;
    s->failed = true;
  }
  blade_log(("BLaDE-MI355X error: " + m + "\n").c_str());
}

namespace blade {
// error path shared with the file readers (files.cpp)
void api_report_error(System *s, const std::string &m) { api_fail(s, m); }
bool api_has_structure(System *s) { return s && s->structure; }
// what the script interpreter (script.cpp) reads of a system
const Structure *api_structure(System *s) { return s ? s->structure.get() : nullptr; }
bool api_has_parameters(System *s) { return s && s->parameters; }
bool api_has_coordinates(System *s) { return s && s->coordinates; }
bool api_has_msld(System *s) { return s && s->msld; }
int api_reset_device(System *s);
void *&api_script_state(System *s, void (*release)(void *)) {
  s->scriptRelease = release;
  return s->script;
}
}  // namespace blade

static void mark_unsupported(System *s, const char *what) {
  s->unsupported = true;
  s->unsupportedWhat = what;
}

extern "C" {

const char *blade_last_error(void) { return g_apiErr.c_str(); }
const char *blade_system_error(System *s) { return s ? s->error.c_str() : "null system"; }

System *blade_init_system(int ngpus, int *gpus) {
  System *s = new System();
  if (ngpus != 1) api_fail(s, "this build runs one system on one GPU (ngpus must be 1)");
  s->device = (gpus && ngpus > 0) ? gpus[0] : 0;
  s->msld.reset(new MsldSetup());
  return s;
}

void blade_dest_system(System *s) {
  if (!s) return;
  if (s->gpu) bm_system_destroy(s->gpu);
  delete s;
}

void blade_set_device(System *) {}
void blade_set_verbose(System *s, int v) { if (s) s->verbose = v; }

// ---- parameters (units exactly as the reference's setters) -------------
void blade_init_parameters(System *s) { s->parameters.reset(new Parameters()); }
void blade_dest_parameters(System *s) { s->parameters.reset(); }
void blade_add_parameter_atoms(System *s, const char *name, double mass) { s->parameters->atomMass[name] = mass; }
void blade_add_parameter_bonds(System *s, const char *t1, const char *t2, double kb, double b0) {
  s->parameters->bonds[canon({t1, t2})] = BondParam{2.0 * kb, b0};
}
void blade_add_parameter_angles(System *s, const char *t1, const char *t2, const char *t3, double kangle,
                                double angle0, double kureyb, double ureyb0) {
  s->parameters->angles[canon({t1, t2, t3})] = AngleParam{2.0 * kangle, kDegrees * angle0, 2.0 * kureyb, ureyb0};
}
void blade_add_parameter_dihes(System *s, const char *t1, const char *t2, const char *t3, const char *t4,
                               double kdih, int ndih, double dih0) {
  s->parameters->dihes[canon({t1, t2, t3, t4})].push_back(DiheParam{kdih, ndih, kDegrees * dih0});
}
void blade_add_parameter_imprs(System *s, const char *t1, const char *t2, const char *t3, const char *t4,
                               double kimp, int nimp, double imp0) {
  if (nimp < 0) { api_fail(s, "improper periodicity less than 0"); return; }
  s->parameters->imprs[canon({t1, t2, t3, t4})] = ImprParam{nimp == 0 ? 2.0 * kimp : kimp, nimp, kDegrees * imp0};
}
void blade_add_parameter_cmaps(System *s, const char *t1, const char *t2, const char *t3, const char *t4,
                               const char *t5, const char *t6, const char *t7, const char *t8, int ngrid) {
  if (ngrid < 4 || ngrid > 60) { api_fail(s, "CMAP grid must have 4 to 60 points per 360 degrees"); return; }
  s->parameters->cmaps[{t1, t2, t3, t4, t5, t6, t7, t8}] = {ngrid, std::vector<double>((size_t)ngrid * ngrid, 0.0)};
}
// grid value (i, j), 1-based: i steps phi and j psi from -180 degrees
void blade_add_parameter_cmaps_fill(System *s, const char *t1, const char *t2, const char *t3, const char *t4,
                                    const char *t5, const char *t6, const char *t7, const char *t8, int i, int j,
                                    double kcmapij) {
  auto it = s->parameters->cmaps.find({t1, t2, t3, t4, t5, t6, t7, t8});
  if (it == s->parameters->cmaps.end()) { api_fail(s, "CMAP fill before its grid was declared"); return; }
  const int n = it->second.first;
  if (i >= 1 && i <= n && j >= 1 && j <= n) it->second.second[(size_t)n * (i - 1) + (j - 1)] = kcmapij;
}
void blade_add_parameter_nbonds(System *s, const char *t1, double eps, double sig, double eps14, double sig14,
                                double e14fac, int combine) {
  s->parameters->nbonds[t1] = NbondParam{-eps, sig, -eps14, sig14, e14fac, combine};
}
void blade_add_parameter_nbfixs(System *s, const char *t1, const char *t2, double eps, double sig, double eps14,
                                double sig14) {
  s->parameters->nbfix[canon({t1, t2})] = NbondParam{-eps, sig, -eps14, sig14, 1.0, 0};
}

// ---- structure -----------------------------------------------------------
void blade_init_structure(System *s) { s->structure.reset(new Structure()); }
void blade_dest_structure(System *s) { s->structure.reset(); }
void blade_add_atom(System *s, int atomIdx, const char *segName, const char *resIdx, const char *resName,
                    const char *atomName, const char *atomTypeName, double charge, double mass) {
  auto &v = s->structure->atoms;
  if ((int)v.size() < atomIdx) v.resize(atomIdx);
  v[atomIdx - 1] = AtomRec{segName, resIdx, resName, atomName, atomTypeName, charge, mass};
}
void blade_add_bond(System *s, int i, int j) { s->structure->bonds.push_back({i - 1, j - 1}); }
void blade_add_angle(System *s, int i, int j, int k) { s->structure->angles.push_back({i - 1, j - 1, k - 1}); }
void blade_add_dihe(System *s, int i, int j, int k, int l) {
  s->structure->dihes.push_back({i - 1, j - 1, k - 1, l - 1});
}
void blade_add_impr(System *s, int i, int j, int k, int l) {
  s->structure->imprs.push_back({i - 1, j - 1, k - 1, l - 1});
}
void blade_add_cmap(System *s, int i1, int j1, int k1, int l1, int i2, int j2, int k2, int l2) {
  s->structure->cmaps.push_back({i1 - 1, j1 - 1, k1 - 1, l1 - 1, i2 - 1, j2 - 1, k2 - 1, l2 - 1});
}
void blade_add_shake(System *s, int shakeHbond) { s->structure->shakeHbond = shakeHbond != 0; }
// virtual sites (system/structure.cxx:1059-1085): 1-based atoms; angles in
// degrees, kept in radians
void blade_add_virt2(System *s, int v, int h1, int h2, double dist, double scale) {
  s->structure->virts.push_back(Structure::Virt{v - 1, {h1 - 1, h2 - 1, -1}, 2, dist, scale, 0.0});
}
void blade_add_virt3(System *s, int v, int h1, int h2, int h3, double dist, double theta, double phi) {
  s->structure->virts.push_back(
      Structure::Virt{v - 1, {h1 - 1, h2 - 1, h3 - 1}, 3, dist, kDegrees * theta, kDegrees * phi});
}
// restraints (system/structure.cxx:1093-1196): 1-based atoms and blocks,
// angles in degrees, AKMA units
void blade_add_noe(System *s, int i, int j, double rmin, double kmin, double rmax, double kmax, double rpeak,
                   double rswitch, double nswitch, double c0x, double c0y, double c0z, bool is_pnoe) {
  const int n = (int)s->structure->atoms.size();
  if (i < 1 || i > n || (!is_pnoe && (j < 1 || j > n))) { api_fail(s, "NOE: invalid atom index"); return; }
  s->structure->noes.push_back(NoeRec{i - 1, j - 1, rmin, kmin, rmax, kmax, rpeak, rswitch, nswitch, {c0x, c0y, c0z},
                                      is_pnoe});
}
void blade_add_harmonic(System *s, int i, double k, double x0, double y0, double z0, double n) {
  s->structure->harms.push_back(HarmRec{i - 1, k, {x0, y0, z0}, n});
}
void blade_add_borest(System *s, int i, int j, double kr, double r0, int lambdaBlock) {
  s->structure->boRests.push_back(BoRestRec{i - 1, j - 1, kr, r0, lambdaBlock - 1});
}
void blade_add_anrest(System *s, int i, int j, int k, double kt, double t0, int lambdaBlock) {
  s->structure->anRests.push_back(AnRestRec{i - 1, j - 1, k - 1, kt, kDegrees * t0, lambdaBlock - 1});
}
void blade_add_direst(System *s, int i, int j, int k, int l, double kphi, int nphi, double phi0, double width,
                      int lambdaBlock) {
  s->structure->diRests.push_back(
      DiRestRec{i - 1, j - 1, k - 1, l - 1, kphi, nphi, kDegrees * phi0, kDegrees * width, lambdaBlock - 1});
}
void blade_add_resd(System *s, int i1, int i2, int j1, int j2, double ci, double cj, double rdist, double kdist) {
  s->structure->resds.push_back(ResdRec{i1 - 1, i2 - 1, j1 - 1, j2 - 1, ci, cj, rdist, kdist});
}
// the TorchANI machine-learned potential (system/structure.cxx:1198-1221)
// needs LibTorch: the system is marked unsupported
void blade_tani_internal_setup(System *s, const int, const char *, const int, const int *, const int *, const int *,
                               const int) {
  mark_unsupported(s, "machine-learned potentials (blade_tani_internal_setup)");
}
int blade_structure_atom_count(System *s) { return (s && s->structure) ? (int)s->structure->atoms.size() : 0; }
int blade_structure_atom_names(System *s, int atomIdx, char *seg, char *res, char *name, int len) {
  if (!s || !s->structure || atomIdx < 1 || atomIdx > (int)s->structure->atoms.size() || len <= 0) return -1;
  const AtomRec &a = s->structure->atoms[atomIdx - 1];
  std::snprintf(seg, len, "%s", a.segName.c_str());
  std::snprintf(res, len, "%s", a.resIdx.c_str());
  std::snprintf(name, len, "%s", a.atomName.c_str());
  return 0;
}

// ---- coordinates -----------------------------------------------------------
void blade_init_coordinates(System *s, int n) {
  s->coordinates.reset(new Coordinates());
  s->coordinates->n = n;
  s->coordinates->pos.assign(3 * (size_t)n, 0.0);
  s->coordinates->vel.assign(3 * (size_t)n, 0.0);
}
void blade_dest_coordinates(System *s) { s->coordinates.reset(); }
void blade_add_coordinates_position(System *s, int i, double x, double y, double z) {
  double *p = &s->coordinates->pos[3 * (size_t)(i - 1)];
  p[0] = x; p[1] = y; p[2] = z;
}
void blade_add_coordinates_velocity(System *s, int i, double vx, double vy, double vz) {
  double *p = &s->coordinates->vel[3 * (size_t)(i - 1)];
  p[0] = vx; p[1] = vy; p[2] = vz;
}
void blade_add_coordinates_box(System *s, int name, double a, double b, double c, double alpha, double beta,
                               double gamma) {
  auto &C = *s->coordinates;
  C.boxName = name;
  C.box[0] = a; C.box[1] = b; C.box[2] = c;
  C.angles[0] = alpha; C.angles[1] = beta; C.angles[2] = gamma;
  if (alpha != 90 || beta != 90 || gamma != 90) mark_unsupported(s, "non-orthorhombic boxes");
}

// ---- msld ------------------------------------------------------------------
// nblocks substituent blocks plus the environment block 0
// (msld/msld.cu:1555-1563: blockCount = nblocks + 1)
void blade_init_msld(System *s, int nblocks) {
  s->msld.reset(new MsldSetup());
  int n = s->structure ? (int)s->structure->atoms.size() : 0;
  s->msld->resize(nblocks < 0 ? 1 : nblocks + 1, n);
}
void blade_dest_msld(System *s) { s->msld.reset(new MsldSetup()); }
void blade_add_msld_atomassignment(System *s, int atomIdx, int blockIdx) {
  auto &m = *s->msld;
  if ((int)m.atomBlock.size() < atomIdx) m.atomBlock.resize(atomIdx, 0);
  m.atomBlock[atomIdx - 1] = blockIdx - 1;
}
void blade_add_msld_initialconditions(System *s, int blockIdx, int siteIdx, double theta0, double thetaVelocity,
                                      double thetaMass, double fixBias, double blockCharge) {
  auto &m = *s->msld;
  int b = blockIdx - 1;
  m.lambdaSite[b] = siteIdx - 1;
  m.theta[b] = theta0;
  m.thetaVelocity[b] = thetaVelocity;
  m.thetaMass[b] = thetaMass;
  m.lambdaBias[b] = fixBias;
  m.lambdaCharge[b] = blockCharge;
}
void blade_add_msld_termscaling(System *s, int sb, int su, int sa, int sd, int si, int sc) {
  bool *t = s->msld->scaleTerms;
  t[0] = sb; t[1] = su; t[2] = sa; t[3] = sd; t[4] = si; t[5] = sc;
}
void blade_add_msld_flags(System *s, double gamma, double fnex, int useSoftCore, int useSoftCore14,
                          int msldEwaldType, double kRestraint, double softBondRadius, double softBondExponent,
                          double softNotBondExponent, int fix) {
  auto &m = *s->msld;
  m.gamma = gamma;
  m.softBondRadius = softBondRadius;
  m.softBondExponent = softBondExponent;
  m.softNotBondExponent = softNotBondExponent;
  m.kRestraint = kRestraint;
  m.fnex = fnex;
  m.useSoftCore = useSoftCore;
  m.useSoftCore14 = useSoftCore14;
  m.msldEwaldType = msldEwaldType;
  m.fix = fix;
}
// theta biases, per site (msld.cu:1675-1695): the first call sizes the
// table with `sites` entries, later calls fill entry i (1-based)
void blade_add_msld_thetacollbias(System *s, int sites, int i, double k, double n) {
  auto &m = *s->msld;
  if (m.thetaCollK.empty()) {
    m.thetaCollK.assign(std::max(sites, 0), 0.0);
    m.thetaCollN.assign(std::max(sites, 0), 0.0);
  }
  if (i < 1 || i > (int)m.thetaCollK.size()) { api_fail(s, "theta collective bias site out of range"); return; }
  m.thetaCollK[i - 1] = k;
  m.thetaCollN[i - 1] = n;
}
void blade_add_msld_thetaindebias(System *s, int sites, int i, double k) {
  auto &m = *s->msld;
  if (m.thetaIndeK.empty()) m.thetaIndeK.assign(std::max(sites, 0), 0.0);
  if (i < 1 || i > (int)m.thetaIndeK.size()) { api_fail(s, "theta independent bias site out of range"); return; }
  m.thetaIndeK[i - 1] = k;
}
// edge (target) bias of the piecewise constraint (msld.cu:1697-1704)
void blade_set_msld_thetaedgebias(System *s, double k, double N, double alpha, double phi) {
  auto &m = *s->msld;
  m.edgeK = k;
  m.edgeN = N;
  m.edgeAlpha = alpha;
  m.edgePhi = phi;
}
// piecewise lambda constraint (msld.cu:1706-1712)
void blade_set_msld_piecewise_constraint(System *s, int do_imp, double width, double k) {
  auto &m = *s->msld;
  m.piecewise = do_imp != 0;
  m.wellWidth = width;
  m.wellK = k;
}
// soft bonds (msld.cu:1714-1721) are recorded; a system that has any is
// rejected when its terms are derived
// msld/msld.cu:1714: 1-based atoms of a soft bond
void blade_add_msld_softbond(System *s, int i, int j) { s->msld->softBonds.push_back({i - 1, j - 1}); }
void blade_add_msld_bias(System *s, int i, int j, int type, double l0, double k, int n) {
  if (type <= 0 || type > 12) { api_fail(s, "type of variable bias is not a recognized type"); return; }
  s->msld->varBias.push_back(VarBias{i - 1, j - 1, type, l0, k, n});
}
void blade_add_msld_block_friction(System *s, int blockIdx, double friction) {
  auto &m = *s->msld;
  if (m.thetaFriction.empty()) m.thetaFriction.assign(m.blockCount, m.gamma);
  m.thetaFriction[blockIdx - 1] = friction;
}
void blade_set_msld_block_fixed(System *s, int blockIdx, int fixed) { s->msld->blockFixed[blockIdx - 1] = fixed != 0; }
void blade_add_msld_atomrestraint(System *s) { s->msld->atomRestraints.emplace_back(); }
void blade_add_msld_charges(System *s, double k1, double k2, double k3, double q0, double w) {
  auto &m = *s->msld;
  m.kCR1 = k1; m.kCR2 = k2; m.kCR3 = k3; m.q0CR3 = q0; m.wCR3 = w;
}
void blade_add_msld_atomrestraint_element(System *s, int i) {
  if (s->msld->atomRestraints.empty()) s->msld->atomRestraints.emplace_back();
  s->msld->atomRestraints.back().push_back(i - 1);
}

// ---- run -----------------------------------------------------------------
void blade_init_run(System *s) { s->run.reset(new RunSetup()); }
void blade_dest_run(System *s) { s->run.reset(); }
void blade_add_run_flags(System *s, double gamma, double betaEwald, double rCut, double rSwitch, int vdWfSwitch,
                         int elecPME, double gridSpace, int gridx, int gridy, int gridz, int orderEwald,
                         double shakeTolerance) {
  auto &r = *s->run;
  r.gamma = gamma;
  r.betaEwald = betaEwald;
  r.rCut = rCut;
  r.rSwitch = rSwitch;
  r.vdwMethod = vdWfSwitch;
  r.elecMethod = elecPME;
  r.gridSpace = gridSpace;
  r.grid[0] = gridx; r.grid[1] = gridy; r.grid[2] = gridz;
  r.orderEwald = orderEwald;
  r.shakeTolerance = shakeTolerance;
}
void blade_add_run_dynopts(System *s, int step, int step0, int nsteps, double dt, double T, int freqNPT,
                           double volumeFluctuation, double pressure) {
  auto &r = *s->run;
  r.step = step;
  r.step0 = step0;
  r.nsteps = nsteps;
  r.dt = dt;
  r.T = T;
  r.freqNPT = freqNPT;
  r.volumeFluctuation = volumeFluctuation;
  r.pressure = pressure;
}

void blade_set_seed(System *s, unsigned long long seed) { s->seed = seed; }
void blade_set_domdec_frequency(System *s, int steps) { s->freqDomdec = steps; }
void blade_set_fft_mode(System *s, int mode) { s->fftMode = mode; }
void blade_set_tile_chunk(System *s, int tiles) { s->tileChunk = tiles; }

int blade_prepare(System *s) {
  if (!s || s->failed) return BM_EINVAL;
  if (s->prepared) return BM_OK;
  if (s->unsupported) { api_fail(s, "unsupported feature: " + s->unsupportedWhat); return BM_ENOTSUP; }
  if (!s->parameters || !s->structure || !s->coordinates) { api_fail(s, "parameters, structure and coordinates are required"); return BM_EINVAL; }
  if (!s->run) s->run.reset(new RunSetup());
  auto &m = *s->msld;
  const int n = (int)s->structure->atoms.size();
  if ((int)m.atomBlock.size() != n) {
    if (m.blockCount <= 1) m.resize(1, n);
    else m.atomBlock.resize(n, 0);
  }
  s->run->seed = s->seed;
  s->run->freqDomdec = s->freqDomdec;
  s->run->fftMode = s->fftMode;
  s->run->tileChunk = s->tileChunk;
  try {
    build_topology(*s->parameters, *s->structure, m, *s->run, s->topo);
  } catch (const SetupError &e) {
    api_fail(s, e.msg);
    return BM_EINVAL;
  }
  build_runparams(*s->run, m, *s->coordinates, n, s->rp);
  s->topo.view();
  s->prepared = true;
  return BM_OK;
}

// the coordinates and velocities as supplied, SoA
static void soa_state(System *s, std::vector<double> &x, std::vector<double> &v) {
  const int n = s->topo.n;
  x.assign(3 * (size_t)n, 0.0);
  v.assign(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      x[k * (size_t)n + i] = s->coordinates->pos[3 * (size_t)i + k];
      v[k * (size_t)n + i] = s->coordinates->vel[3 * (size_t)i + k];
    }
}

// Run::dynamics_initialize (run/run.cu:572-622): derive terms, build the
// device system from the coordinates.
static int ensure_gpu(System *s) {
  if (s->gpu) return BM_OK;
  int rc = blade_prepare(s);
  if (rc) return rc;
  std::vector<double> x, v;
  soa_state(s, x, v);
  rc = bm_system_create(s->device, &s->topo.top, &s->rp, x.data(), v.data(), s->msld->theta.data(),
                        s->msld->thetaVelocity.data(), &s->gpu);
  // supplied coordinates have their constraint groups snapped to the exact
  // lengths (run/run.cu:606); a checkpoint's state is continued as written
  // (it is read after the rectification there)
  if (!rc && !s->fromCheckpoint) rc = bm_holonomic_rectify(s->gpu);
  if (!rc) rc = bm_sync(s->gpu);
  if (rc) api_fail(s, std::string("device system: ") + bm_last_error());
  s->step = s->run->step;
  return rc;
}

// the device system's state (positions, velocities, theta, box and step)
// back into the host setup, and the device system released, so that the
// next energy or dynamics call builds it again with the current run
// parameters (the reference builds its state at every dynamics call,
// run/run.cu:572-622)
static int reset_device(System *s) {
  if (!s || !s->gpu) return BM_OK;
  const int n = s->topo.n, nl = std::max(1, s->topo.top.n_blocks);
  std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), th(nl), tv(nl);
  double box[3];
  int rc = bm_get_positions(s->gpu, x.data(), th.data());
  if (!rc) rc = bm_get_velocities(s->gpu, v.data(), tv.data());
  if (!rc) rc = bm_get_box(s->gpu, box);
  if (rc) {
    api_fail(s, std::string("device state: ") + bm_last_error());
    return rc;
  }
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      s->coordinates->pos[3 * (size_t)i + k] = x[(size_t)k * n + i];
      s->coordinates->vel[3 * (size_t)i + k] = v[(size_t)k * n + i];
    }
  for (int k = 0; k < 3; k++) s->coordinates->box[k] = box[k];
  if (s->topo.top.n_blocks > 1) {
    s->msld->theta.assign(th.begin(), th.end());
    s->msld->thetaVelocity.assign(tv.begin(), tv.end());
  }
  bm_system_destroy(s->gpu);
  s->gpu = nullptr;
  s->run->step = s->step;
  s->prepared = false;
  s->fromCheckpoint = false;
  return BM_OK;
}

int blade_autotune(System *s, int steps, bm_tune_result *result) {
  if (!s || s->failed) return BM_EINVAL;
  if (s->gpu) {
    api_fail(s, "blade_autotune comes before the first energy or dynamics call");
    return BM_EINVAL;
  }
  int rc = blade_prepare(s);
  if (rc) return rc;
  std::vector<double> x, v;
  soa_state(s, x, v);
  bm_tune_result r{};
  rc = bm_autotune(s->device, &s->topo.top, &s->rp, x.data(), v.data(), s->msld->theta.data(),
                   s->msld->thetaVelocity.data(), steps, &r);
  if (rc) {
    api_fail(s, std::string("autotune: ") + bm_last_error());
    return rc;
  }
  s->fftMode = r.fft_mode;
  s->freqDomdec = r.freq_domdec;
  s->tileChunk = r.tile_chunk;
  s->prepared = false;   // run parameters are derived again with the chosen values
  if (result) *result = r;
  return BM_OK;
}

int blade_batch_autotune(int n, System *const *sys, int steps, bm_tune_result *result) {
  if (n < 1 || !sys) return BM_EINVAL;
  std::vector<std::vector<double>> x(n), v(n);
  std::vector<const bm_topology *> tops(n);
  std::vector<const bm_runparams *> rps(n);
  std::vector<const double *> px(n), pv(n), pt(n), ptv(n);
  for (int k = 0; k < n; k++) {
    System *s = sys[k];
    if (!s || s->failed) return BM_EINVAL;
    if (s->gpu) {
      api_fail(s, "blade_batch_autotune comes before the first energy or dynamics call");
      return BM_EINVAL;
    }
    if (int rc = blade_prepare(s)) return rc;
    soa_state(s, x[k], v[k]);
    tops[k] = &s->topo.top;
    rps[k] = &s->rp;
    px[k] = x[k].data();
    pv[k] = v[k].data();
    pt[k] = s->msld->theta.data();
    ptv[k] = s->msld->thetaVelocity.data();
  }
  bm_tune_result r{};
  int rc = bm_batch_autotune(sys[0]->device, n, tops.data(), rps.data(), px.data(), pv.data(), pt.data(),
                             ptv.data(), steps, &r);
  if (rc) {
    api_fail(sys[0], std::string("batch autotune: ") + bm_last_error());
    return rc;
  }
  for (int k = 0; k < n; k++) {
    sys[k]->freqDomdec = r.freq_domdec;
    sys[k]->tileChunk = r.tile_chunk;
    sys[k]->prepared = false;
  }
  if (result) *result = r;
  return BM_OK;
}

void blade_run_energy(System *s) {
  if (!s || s->failed) return;
  if (ensure_gpu(s)) return;
  int rc = bm_domdec_update(s->gpu, 1);
  if (!rc) rc = bm_calc_force(s->gpu, 0, BM_CALC_ENERGY);
  if (!rc) rc = bm_sync(s->gpu);
  if (rc) api_fail(s, std::string("energy: ") + bm_last_error());
}

// one .nrg line (io/io.cxx:376-387): the step, then every energy term
static void write_nrg(System *s, long step) {
  double e[BM_NTERMS];
  if (bm_get_energy(s->gpu, e)) return;
  fprintf(s->fpNRG, "%10ld", step);
  for (int i = 0; i < BM_NTERMS; i++) fprintf(s->fpNRG, " %12.4f", e[i]);
  fprintf(s->fpNRG, "\n");
}

// one text .lmd line (io/io.cxx:350-358): the step, then lambda of every
// block after the environment
// one XDR word: 4 bytes, most significant first
static void put_xdr(FILE *fp, uint32_t w) {
  const unsigned char b[4] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8),
                              (unsigned char)w};
  fwrite(b, 1, 4, fp);
}

// lambda frame (io/io.cxx:347-370): text "%10ld" step and " %8.6f" lambdas
// of the substituent blocks, or in XDR the block count then each lambda as
// an IEEE single
static void write_lmd(System *s, long step) {
  const int nl = s->topo.top.n_blocks;
  std::vector<double> l(nl > 0 ? nl : 1);
  if (bm_get_lambda(s->gpu, l.data())) return;
  if (!s->hrLMD) {
    put_xdr(s->fpLMD, (uint32_t)(nl - 1));
    for (int i = 1; i < nl; i++) {
      const float f = (float)l[i];
      uint32_t w;
      memcpy(&w, &f, 4);
      put_xdr(s->fpLMD, w);
    }
    return;
  }
  fprintf(s->fpLMD, "%10ld", step);
  for (int i = 1; i < nl; i++) fprintf(s->fpLMD, " %8.6f", l[i]);
  fprintf(s->fpLMD, "\n");
}

// coordinate frame (io/io.cxx:314-345): positions and box in nm, the time
// in ps, precision 1000 per nm
static void write_xtc(System *s, long step) {
  const int n = s->topo.n;
  std::vector<double> x(3 * (size_t)n);
  double box[3];
  if (bm_get_positions(s->gpu, x.data(), nullptr) || bm_get_box(s->gpu, box)) return;
  std::vector<float> a(3 * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) a[3 * (size_t)i + k] = (float)(x[(size_t)k * n + i] / 10.0);
  const float b9[9] = {(float)(box[0] / 10.0), 0, 0, 0, (float)(box[1] / 10.0), 0, 0, 0, (float)(box[2] / 10.0)};
  const float t = (float)(step * s->run->dt / 20.454828284385908);   // AKMA time -> ps
  blade_xtc_frame(s->fpXTC, n, (int)step, t, b9, a.data(), 1000.0f);
}

// the first step >= from that is a multiple of f (f > 0)
static long next_multiple(long from, int f) { return ((from + f - 1) / f) * f; }

// potential energy at the current positions (a list rebuild and one
// energy evaluation)
static int potential_now(bm_system *g, double &e) {
  double E[BM_NTERMS];
  int rc = bm_domdec_update(g, 1);
  if (!rc) rc = bm_calc_force(g, 0, BM_CALC_ENERGY);
  if (!rc) rc = bm_sync(g);
  if (!rc) rc = bm_get_energy(g, E);
  e = E[BM_EPOTENTIAL];
  return rc;
}

// uniforms of the barostat, keyed by (seed, step, draw); stream tag "NPT"
static double npt_uniform(unsigned long long seed, long step, unsigned k) {
  return bm_uniform(seed, step, 0x4E505400u | k);
}

// Monte Carlo volume move (update/pressure.cu:15-185): Gaussian volume
// draws (one for a cubic box, x=y and z for a tetragonal one, one per axis
// otherwise), every atom scaled with constraint groups kept rigid, and the
// move kept with probability min(1, exp(-dW/kT)),
//   dW = E_new - E_old + P (V_new - V_old) - N kT log(V_new / V_old),
// N = atoms - constraints; a rejected move restores positions and box.
static int pressure_coupling(System *s, long step) {
  bm_system *g = s->gpu;
  const RunSetup &r = *s->run;
  double box0[3];
  int rc = bm_get_box(g, box0);
  double e0 = 0.0, e1 = 0.0;
  if (!rc) rc = potential_now(g, e0);
  if (rc) return rc;
  const double V0 = box0[0] * box0[1] * box0[2];
  auto gauss = [&](unsigned k) {
    const double u1 = npt_uniform(s->seed, step, 2 * k + 1), u2 = npt_uniform(s->seed, step, 2 * k + 2);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  double scale[3], vn[3];
  const int name = s->coordinates->boxName;
  for (int k = 0; k < 3; k++) {
    const int draw = name == 0 ? 0 : (name == 1 ? (k < 2 ? 0 : 1) : k);
    vn[k] = V0 + r.volumeFluctuation * gauss((unsigned)draw);
  }
  for (int k = 0; k < 3; k++) {
    if (!(vn[k] > 0)) return BM_OK;   // no move
    scale[k] = std::exp(std::log(vn[k] / V0) / 3);
  }
  const double V1 = V0 * scale[0] * scale[1] * scale[2];
  rc = bm_volume_move(g, scale);
  if (rc == BM_EINVAL) return BM_OK;   // the box would pass rcut: no move
  if (!rc) rc = potential_now(g, e1);
  if (rc) return rc;
  const bm_topology &t = s->topo.top;
  const double N = t.n_atoms - (2.0 * t.n_tri + t.n_br1 + 2.0 * t.n_br2 + 3.0 * t.n_br3);
  const double kT = 0.001987191 * r.T;
  const double dW = e1 - e0 + r.pressure * (V1 - V0) - N * kT * std::log(V1 / V0);
  s->nptMoves++;
  if (npt_uniform(s->seed, step, 0) < std::exp(-dW / kT)) {
    s->nptAccepted++;
    double b[3];
    bm_get_box(g, b);
    for (int k = 0; k < 3; k++) s->rp.box[k] = s->coordinates->box[k] = b[k];
  } else {
    rc = bm_volume_restore(g, box0);
  }
  if (!rc) rc = bm_domdec_update(g, 1);
  if (!rc) rc = bm_sync(g);
  return rc;
}

int blade_run_dynamics(System *s, int nsteps, int freqNRG, int use_graph) {
  if (!s || s->failed) return BM_EINVAL;
  int rc = ensure_gpu(s);
  if (rc) return rc;
  const int fnpt = s->run->freqNPT;
  if (s->fpNRG || s->fpLMD || s->fpXTC || fnpt > 0) {
    // segments that end on output steps; energies on the .nrg steps
    // (run/run.cu:545-570 prints after each step's update); a barostat
    // move before every step that is a multiple of freqNPT
    const int fe = s->fpNRG ? s->freqNRG : freqNRG;
    const long end = s->step + nsteps;
    long t = s->step;
    while (!rc && t < end) {
      if (fnpt > 0 && t % fnpt == 0) {
        rc = pressure_coupling(s, t);
        if (rc) break;
      }
      long stop = end - 1;
      if (s->fpNRG) stop = std::min(stop, next_multiple(t, s->freqNRG));
      if (s->fpLMD) stop = std::min(stop, next_multiple(t, s->freqLMD));
      if (s->fpXTC) stop = std::min(stop, next_multiple(t, s->freqXTC));
      if (fnpt > 0) stop = std::min(stop, next_multiple(t + 1, fnpt) - 1);
      rc = bm_dynamics(s->gpu, t, stop - t + 1, fe, use_graph);
      if (!rc) rc = bm_sync(s->gpu);
      if (!rc && s->fpNRG && stop % s->freqNRG == 0) write_nrg(s, stop);
      if (!rc && s->fpLMD && stop % s->freqLMD == 0) write_lmd(s, stop);
      if (!rc && s->fpXTC && stop % s->freqXTC == 0) write_xtc(s, stop);
      t = stop + 1;
    }
    if (s->fpNRG) fflush(s->fpNRG);
    if (s->fpLMD) fflush(s->fpLMD);
    if (s->fpXTC) fflush(s->fpXTC);
  } else {
    rc = bm_dynamics(s->gpu, s->step, nsteps, freqNRG, use_graph);
    if (!rc) rc = bm_sync(s->gpu);
  }
  if (!rc) {
    int flag = -1;
    bm_get_nan_flag(s->gpu, &flag);
    if (flag != -1) {
      api_fail(s, "NaN/Inf detected at atom " + std::to_string(flag));
      return BM_ENAN;
    }
  }
  if (rc) api_fail(s, std::string("dynamics: ") + bm_last_error());
  s->step += nsteps;
  return rc;
}

int blade_set_output_nrg(System *s, const char *path, int freq) {
  if (!s || freq < 0) return BM_EINVAL;
  if (s->fpNRG) fclose(s->fpNRG);
  s->fpNRG = nullptr;
  s->freqNRG = 0;
  if (!path || !freq) return BM_OK;
  s->fpNRG = fopen(path, "w");
  if (!s->fpNRG) {
    api_fail(s, std::string("cannot open ") + path);
    return BM_EINVAL;
  }
  s->freqNRG = freq;
  return BM_OK;
}

int blade_set_output_xtc(System *s, const char *path, int freq) {
  if (!s || freq < 0) return BM_EINVAL;
  if (s->fpXTC) fclose(s->fpXTC);
  s->fpXTC = nullptr;
  s->freqXTC = 0;
  if (!path || !freq) return BM_OK;
  s->fpXTC = fopen(path, "wb");
  if (!s->fpXTC) {
    api_fail(s, std::string("cannot open ") + path);
    return BM_EINVAL;
  }
  s->freqXTC = freq;
  return BM_OK;
}
int blade_set_output_lmd(System *s, const char *path, int freq) {
  if (!s || freq < 0) return BM_EINVAL;
  if (s->fpLMD) fclose(s->fpLMD);
  s->fpLMD = nullptr;
  s->freqLMD = 0;
  if (!path || !freq) return BM_OK;
  s->fpLMD = fopen(path, "w");
  if (!s->fpLMD) {
    api_fail(s, std::string("cannot open ") + path);
    return BM_EINVAL;
  }
  s->freqLMD = freq;
  return BM_OK;
}

void blade_set_lmd_format(System *s, int humanReadable) { s->hrLMD = humanReadable != 0; }

// Binary checkpoint: the exact FP64 state, so a run continued from it is
// bit-identical to the uninterrupted one when it was written on a list
// rebuild step (step a multiple of the domdec cadence).  Layout: 8-byte
// magic "BLADEMCK", int64 step, int32 atoms, int32 blocks, f64 box[3],
// then f64 positions [3N] and velocities [3N] (SoA, x then y then z),
// theta [nL] and theta velocities [nL].
static const char kCkMagic[8] = {'B', 'L', 'A', 'D', 'E', 'M', 'C', 'K'};

int blade_write_checkpoint(System *s, const char *path) {
  if (!s || !path || !s->gpu) return BM_EINVAL;
  const int n = s->topo.n, nl = std::max(1, s->topo.top.n_blocks);
  std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), th(nl), tv(nl);
  int rc = bm_get_positions(s->gpu, x.data(), th.data());
  if (!rc) rc = bm_get_velocities(s->gpu, v.data(), tv.data());
  if (rc) return rc;
  FILE *fp = fopen(path, "wb");
  if (!fp) {
    api_fail(s, std::string("cannot open ") + path);
    return BM_EINVAL;
  }
  const int64_t step = s->step;
  const int32_t hn[2] = {n, nl};
  bool ok = fwrite(kCkMagic, 1, 8, fp) == 8 && fwrite(&step, 8, 1, fp) == 1 && fwrite(hn, 4, 2, fp) == 2 &&
            fwrite(s->rp.box, 8, 3, fp) == 3 && fwrite(x.data(), 8, x.size(), fp) == x.size() &&
            fwrite(v.data(), 8, v.size(), fp) == v.size() && fwrite(th.data(), 8, nl, fp) == (size_t)nl &&
            fwrite(tv.data(), 8, nl, fp) == (size_t)nl;
  ok = (fclose(fp) == 0) && ok;
  if (!ok) api_fail(s, std::string("checkpoint write failed: ") + path);
  return ok ? BM_OK : BM_EINVAL;
}

int blade_read_checkpoint(System *s, const char *path) {
  if (!s || !path || s->failed) return BM_EINVAL;
  if (s->gpu) {
    api_fail(s, "blade_read_checkpoint comes before the first energy or dynamics call");
    return BM_EINVAL;
  }
  int rc = blade_prepare(s);
  if (rc) return rc;
  FILE *fp = fopen(path, "rb");
  if (!fp) {
    api_fail(s, std::string("cannot open ") + path);
    return BM_EINVAL;
  }
  char magic[8];
  int64_t step = 0;
  int32_t hn[2] = {0, 0};
  double box[3];
  const int n = s->topo.n, nl = std::max(1, s->topo.top.n_blocks);
  bool ok = fread(magic, 1, 8, fp) == 8 && memcmp(magic, kCkMagic, 8) == 0 && fread(&step, 8, 1, fp) == 1 &&
            fread(hn, 4, 2, fp) == 2 && fread(box, 8, 3, fp) == 3;
  if (ok && (hn[0] != n || hn[1] != nl)) {
    fclose(fp);
    api_fail(s, "checkpoint does not match the system (atoms or blocks)");
    return BM_EINVAL;
  }
  std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), th(nl), tv(nl);
  ok = ok && fread(x.data(), 8, x.size(), fp) == x.size() && fread(v.data(), 8, v.size(), fp) == v.size() &&
       fread(th.data(), 8, nl, fp) == (size_t)nl && fread(tv.data(), 8, nl, fp) == (size_t)nl;
  fclose(fp);
  if (!ok) {
    api_fail(s, std::string("not a checkpoint file: ") + path);
    return BM_EINVAL;
  }
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      s->coordinates->pos[3 * (size_t)i + k] = x[(size_t)k * n + i];
      s->coordinates->vel[3 * (size_t)i + k] = v[(size_t)k * n + i];
    }
  for (int k = 0; k < 3; k++) s->coordinates->box[k] = box[k];
  if (s->topo.top.n_blocks > 1) {
    s->msld->theta.assign(th.begin(), th.end());
    s->msld->thetaVelocity.assign(tv.begin(), tv.end());
  }
  s->run->step = step;
  s->fromCheckpoint = true;
  s->prepared = false;   // the box may differ from the setup's
  return BM_OK;
}

int blade_get_npt_stats(System *s, long *moves, long *accepted, double *box) {
  if (!s) return BM_EINVAL;
  if (moves) *moves = s->nptMoves;
  if (accepted) *accepted = s->nptAccepted;
  if (box) {
    if (s->gpu) return bm_get_box(s->gpu, box);
    for (int k = 0; k < 3; k++) box[k] = s->coordinates ? s->coordinates->box[k] : 0.0;
  }
  return BM_OK;
}

int blade_get_energy(System *s, double *e) { return s && s->gpu ? bm_get_energy(s->gpu, e) : BM_EINVAL; }
int blade_get_forces(System *s, double *f) { return s && s->gpu ? bm_get_forces(s->gpu, f, nullptr) : BM_EINVAL; }
// before the device system exists: the coordinates as supplied (SoA)
int blade_get_positions(System *s, double *p) {
  if (!s || !p) return BM_EINVAL;
  if (s->gpu) return bm_get_positions(s->gpu, p, nullptr);
  if (!s->coordinates) return BM_EINVAL;
  const int n = s->coordinates->n;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) p[(size_t)k * n + i] = s->coordinates->pos[3 * (size_t)i + k];
  return BM_OK;
}
int blade_set_positions(System *s, const double *p) {
  if (!s) return BM_EINVAL;
  int rc = ensure_gpu(s);
  return rc ? rc : bm_set_positions(s->gpu, p, nullptr);
}
int blade_get_lambda(System *s, double *l) { return s && s->gpu ? bm_get_lambda(s->gpu, l) : BM_EINVAL; }
int blade_get_dudl(System *s, double *d) { return s && s->gpu ? bm_get_dudl(s->gpu, d) : BM_EINVAL; }
const bm_topology *blade_get_topology(System *s) { return s && s->prepared ? &s->topo.top : nullptr; }
const bm_runparams *blade_get_runparams(System *s) { return s && s->prepared ? &s->rp : nullptr; }
bm_system *blade_get_device_system(System *s) { return s ? s->gpu : nullptr; }

}  // extern "C"

namespace blade {
int api_reset_device(System *s) { return reset_device(s); }
}  // namespace blade
