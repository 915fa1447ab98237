// setup.h - host-side system description and term derivation.
//
// Mirrors the reference's host model: Parameters (system/parameters.h),
// Structure (system/structure.h), Coordinates (system/coordinates.h), the
// setup half of Msld (msld/msld.h) and Run's options (run/run.h).  The
// derivation in Topology::build restates Potential::initialize
// (system/potential.cxx:525-1837): bonded term lookup, Urey-Bradley,
// exclusion sets, 1-4 pairs, PME exclusion corrections, vdW type table,
// constraint grouping and MSLD siteBlock codes.  No HIP in this file.
#pragma once

#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/blade_mi355x.h"

namespace blade {

struct BondParam { double kb, b0; };
struct AngleParam { double kangle, angle0, kureyb, ureyb0; };
struct DiheParam { double kdih; int ndih; double dih0; };
struct ImprParam { double kimp; int nimp; double imp0; };
struct NbondParam { double eps, sig, eps14, sig14, e14fac; int combine; };

// Type-name tuples match forwards or reversed (parameters.cxx:40-88).
std::string canon(std::vector<std::string> t);

struct Parameters {
  std::map<std::string, double> atomMass;
  std::map<std::string, BondParam> bonds;
  std::map<std::string, AngleParam> angles;
  std::map<std::string, std::vector<DiheParam>> dihes;
  std::map<std::string, ImprParam> imprs;
  std::map<std::string, NbondParam> nbonds;
  std::map<std::string, NbondParam> nbfix;  // key canon({t1,t2})
  std::map<std::vector<std::string>, std::pair<int, std::vector<double>>> cmaps;  // ordered 8 types -> (ngrid, grid)
};

struct AtomRec {
  std::string segName, resIdx, resName, atomName, atomTypeName;
  double charge, mass;
};

// restraint records as the setters store them (system/structure.cxx:1093-1196)
struct NoeRec { int i, j; double rmin, kmin, rmax, kmax, rpeak, rswitch, nswitch, c0[3]; bool pnoe; };
struct HarmRec { int i; double k, x0[3], n; };
struct BoRestRec { int i, j; double kr, r0; int block; };
struct AnRestRec { int i, j, k; double kt, t0; int block; };
struct DiRestRec { int i, j, k, l; double kphi; int nphi; double phi0, width; int block; };
struct ResdRec { int i1, i2, j1, j2; double ci, cj, rdist, kdist; };

struct Structure {
  std::vector<AtomRec> atoms;
  std::vector<std::array<int, 2>> bonds;
  std::vector<std::array<int, 3>> angles;
  std::vector<std::array<int, 4>> dihes, imprs;
  std::vector<std::array<int, 8>> cmaps;
  bool shakeHbond = false;
  std::vector<NoeRec> noes;
  std::vector<HarmRec> harms;
  std::vector<BoRestRec> boRests;
  std::vector<AnRestRec> anRests;
  std::vector<DiRestRec> diRests;
  std::vector<ResdRec> resds;
  // virtual sites (system/structure.cxx:1059-1085): 0-based site and hosts;
  // two hosts: dist (A) and scale; three hosts: dist (A, negative: the
  // bisector of hosts 1 and 2), theta and phi (radians)
  struct Virt { int v, h[3], nh; double dist, a, b; };
  std::vector<Virt> virts;
};

struct Coordinates {
  int n = 0;
  std::vector<double> pos, vel;  // AoS [n][3] as the reference stores them
  double box[3] = {0, 0, 0};
  double angles[3] = {90, 90, 90};
  int boxName = 0;
};

struct VarBias { int i, j, type; double l0, k; int n; };

struct MsldSetup {
  int blockCount = 1;
  std::vector<int> atomBlock;   // [N]
  std::vector<int> lambdaSite;  // [nL]
  std::vector<double> theta, thetaVelocity, thetaMass, lambdaBias, lambdaCharge;
  std::vector<double> thetaFriction;  // empty => msld gamma
  std::vector<int> blockFixed;
  bool scaleTerms[6] = {true, true, true, true, true, true};
  double gamma = 1.0 / 20.454828284385908;  // ps^-1 in AKMA (msld.cu:35)
  double fnex = 5.5;
  bool useSoftCore = false, useSoftCore14 = false;
  int msldEwaldType = 2;
  bool fix = false;
  std::vector<VarBias> varBias;
  std::vector<std::vector<int>> atomRestraints;   // 0-based atoms of each group
  double kRestraint = 59.2;                       // kcal/mol/A^2 (msld.cu:83)
  double kCR1 = 0, kCR2 = 0, kCR3 = 0, q0CR3 = 0, wCR3 = 0;   // charge restraints
  std::vector<std::array<int, 2>> softBonds;      // 0-based atom pairs
  double softBondRadius = 1.0, softBondExponent = 2.0, softNotBondExponent = 1.0;   // msld.cu:89-91
  // theta biases (msld.cu:1675-1695): per-site collective (k, n) and
  // independent (k) biases, sized by their first setter call
  std::vector<double> thetaCollK, thetaCollN, thetaIndeK;
  // piecewise constraint and its edge bias (msld.cu:1697-1712)
  bool piecewise = false;
  double wellWidth = 1.0, wellK = 100.0;
  double edgeK = 0.0, edgeN = 2.0, edgeAlpha = 10.0, edgePhi = 0.0;
  void resize(int nblocks, int natoms);
};

struct RunSetup {
  long step = 0, step0 = 0, nsteps = 5000;
  double dt = 0.001 * 20.454828284385908;
  double T = 300;
  double gamma = 1.0 / 20.454828284385908;
  double betaEwald = 1 / 3.2, rCut = 10, rSwitch = 8.5;
  int vdwMethod = BM_VFSWITCH, elecMethod = BM_PME;
  double gridSpace = 1.0;
  int grid[3] = {-1, -1, -1};
  int orderEwald = 6;
  double shakeTolerance = 2e-7;
  int freqNRG = 10;
  int freqNPT = 0;
  double volumeFluctuation = 100.0;   // A^3 (run.cu:67)
  double pressure = 0.000014584319218586;   // 1 atm in kcal/mol/A^3 (run.cu:68)
  unsigned long long seed = 0xB1ADE000ull;
  int freqDomdec = 10;
  int fftMode = 0;       // PME FFT form (bm_runparams.fft_mode)
  int tileChunk = 0;     // directed tiles per nonbonded work unit (bm_runparams.tile_chunk)
};

// Derived term tables owning their storage; view() exposes bm_topology.
struct Topology {
  int n = 0;
  std::vector<double> charge, mass;
  std::vector<int> typeIdx, siteBlock;
  std::vector<std::string> typeNames;
  int nTypes = 0;
  std::vector<double> c12, c6;
  int nBonds12 = 0;
  std::vector<int> bondIdx, bondSb; std::vector<double> bondKb, bondB0;
  std::vector<int> angleIdx, angleSb; std::vector<double> angleK, angleT0;
  std::vector<int> diheIdx, diheSb, diheN; std::vector<double> diheK, diheD0;
  std::vector<int> imprIdx, imprSb, imprN; std::vector<double> imprK, imprD0;
  std::vector<int> bondSoft, angleSoft, diheSoft, imprSoft; bool anySoft = false;
  std::vector<int> virt2Idx, virt3Idx; std::vector<double> virt2Par, virt3Par;
  std::vector<int> nb14Idx, nb14Sb; std::vector<double> nb14qq, nb14c12, nb14c6, nb14e14;
  std::vector<int> nbexIdx, nbexSb; std::vector<double> nbexqq;
  std::vector<int> exclIdx;
  std::vector<int> triIdx, br1Idx, br2Idx, br3Idx;
  std::vector<double> triB0, br1B0, br2B0, br3B0;
  int nBlocks = 1, nSites = 1;
  std::vector<int> blockSite, siteBound, blockFixed;
  std::vector<double> thetaMass, fixedBias, thetaFriction;
  std::vector<int> vbI, vbJ, vbType, vbN; std::vector<double> vbL0, vbK;
  std::vector<int> catsBound, catsIdx; double catsK = 0;
  std::vector<int> cmapIdx, cmapSb, cmapType, cmapNgrid, cmapOff; std::vector<double> cmapGrid;
  std::vector<double> blockCharge; double crK1 = 0, crK2 = 0, crK3 = 0, crQ0 = 0, crW = 0;
  std::vector<double> tcollK, tcollN, tindeK;
  int piecewise = 0; double pwWidth = 1.0, pwK = 100.0;
  double edgeK = 0, edgeN = 2, edgeAlpha = 10, edgePhi = 0;
  int fixLambda = 0;
  std::vector<int> noeIdx, harmIdx, borIdx, borBlock, anrIdx, anrBlock, dirIdx, dirBlock, dirN, resdIdx;
  std::vector<double> noePar, harmPar, borPar, anrPar, dirPar, resdPar;
  bm_topology top{};
  const bm_topology *view();
};

// Error raised by derivation with a message (the reference calls fatal()).
struct SetupError { std::string msg; };

// Build the derived topology (Potential::initialize restatement).
void build_topology(const Parameters &p, const Structure &s, const MsldSetup &m,
                    const RunSetup &r, Topology &t);

// Run parameters for the kernels; box from coordinates.
void build_runparams(const RunSetup &r, const MsldSetup &m, const Coordinates &c, int natoms,
                     bm_runparams &rp);

// PME grid size rule (system/potential.cxx:1205-1221).
int pme_grid_size(double L, double gridSpace, int explicitSize);

}  // namespace blade
