// setup.cpp - term derivation (restates system/potential.cxx:525-1837 and
// the scaling rules of msld/msld.cu:470-678).  Host code only.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <set>
#include <sstream>

namespace blade {

static const double kBoltz = 0.001987191;

std::string canon(std::vector<std::string> t) {
  std::vector<std::string> r(t.rbegin(), t.rend());
  const std::vector<std::string> &k = (r < t) ? r : t;
  std::string s;
  for (size_t i = 0; i < k.size(); i++) s += (i ? " " : "") + k[i];
  return s;
}

void MsldSetup::resize(int nblocks, int natoms) {
  blockCount = nblocks;
  atomBlock.assign(natoms, 0);
  lambdaSite.assign(nblocks, 0);
  theta.assign(nblocks, 0);
  thetaVelocity.assign(nblocks, 0);
  thetaMass.assign(nblocks, 0);
  lambdaBias.assign(nblocks, 0);
  lambdaCharge.assign(nblocks, 0);
  blockFixed.assign(nblocks, 0);
  thetaFriction.clear();
}

int pme_grid_size(double L, double gridSpace, int explicitSize) {
  static const int good[] = {32, 27, 24, 20, 18, 16};
  if (gridSpace <= 0) return explicitSize;
  double minDim = L / gridSpace;
  int j = 1;
  while (minDim >= 32 * j) j *= 2;
  int k = 0;
  while (k < 5 && minDim < j * good[k]) k++;
  return j * good[k - 1 < 0 ? 0 : k - 1];
}

namespace {

[[noreturn]] void fail(const std::string &m) { throw SetupError{m}; }

int merge_site_block(int site, int block) {
  if (site >= (1 << 16) || block >= (1 << 16)) fail("site or block cap of 2^16 exceeded");
  return (site << 16) | block;
}

struct Scaler {
  const MsldSetup &m;
  std::vector<char> restrained;   // [N] atom is in an atom restraint group
  explicit Scaler(const MsldSetup &ms) : m(ms), restrained(ms.atomBlock.size(), 0) {
    for (const auto &g : m.atomRestraints)
      for (int a : g)
        if (a >= 0 && a < (int)restrained.size()) restrained[a] = 1;
  }
  // Msld::check_soft (msld/msld.cu:478-495): the term holds both atoms of
  // a soft bond
  bool soft(const int *idx, int nat) const {
    for (const auto &b : m.softBonds) {
      bool f0 = false, f1 = false;
      for (int k = 0; k < nat; k++) {
        f0 = f0 || idx[k] == b[0];
        f1 = f1 || idx[k] == b[1];
      }
      if (f0 && f1) return true;
    }
    return false;
  }
  // Msld::bonded_scaling (msld/msld.cu:509-553): distinct blocks of the
  // atoms in descending order, at most nsc of them.  A type named by
  // removescaling keeps an atom's block only when every atom of the term in
  // that block is restrained, or when the term is soft.  Returns whether it
  // is soft.
  bool bonded(const int *idx, int nat, int type, int nsc, int *sb) const {
    int block[5] = {m.blockCount, 0, 0, 0, 0};
    const bool isSoft = soft(idx, nat);
    bool scale = m.scaleTerms[type] || isSoft;
    for (int i = 1; i < nsc + 2; i++)
      for (int j = 0; j < nat; j++) {
        int ab = m.atomBlock[idx[j]];
        if (!scale)
          for (int k = 0; k < nat; k++)
            if (m.atomBlock[idx[k]] == m.atomBlock[idx[j]] && !restrained[idx[k]]) ab = 0;
        if (ab > block[i] && ab < block[i - 1]) block[i] = ab;
      }
    for (int i = 1; i < nsc + 1; i++)
      for (int j = i + 1; j < nsc + 1; j++)
        if (block[i] > 0 && block[j] > 0 && block[i] != block[j] &&
            m.lambdaSite[block[i]] == m.lambdaSite[block[j]])
          fail("illegal MSLD scaling between two blocks of the same site");
    if (block[nsc + 1] != 0) fail("too many lambda scalings in one bonded term");
    for (int i = 0; i < nsc; i++) sb[i] = merge_site_block(m.lambdaSite[block[i + 1]], block[i + 1]);
    return isSoft;
  }
  // Msld::nonbonded_scaling (msld.cu:555-577)
  void nonbonded(const int *idx, int nat, int *sb) const {
    int block[5] = {m.blockCount, 0, 0, 0, 0};
    for (int i = 1; i < nat + 2; i++)
      for (int j = 0; j < nat; j++) {
        int ab = m.atomBlock[idx[j]];
        if (ab > block[i] && ab < block[i - 1]) block[i] = ab;
      }
    for (int i = 0; i < nat; i++) sb[i] = merge_site_block(m.lambdaSite[block[i + 1]], block[i + 1]);
  }
  // Msld::nbex_scaling (msld.cu:617-666)
  bool nbex(const int *idx, int *sb) const {
    int block[2] = {m.atomBlock[idx[0]], m.atomBlock[idx[1]]};
    if (block[0] < block[1]) std::swap(block[0], block[1]);
    bool include = true;
    if (m.msldEwaldType == 1) {
      if (block[0] == block[1]) block[1] = 0;
      if (block[0] != block[1] && m.lambdaSite[block[0]] == m.lambdaSite[block[1]]) include = false;
    } else if (m.msldEwaldType == 2) {
      if (block[0] != block[1] && m.lambdaSite[block[0]] == m.lambdaSite[block[1]]) include = false;
    } else if (m.msldEwaldType != 3) {
      fail("illegal msldEwaldType");
    }
    for (int i = 0; i < 2; i++) sb[i] = merge_site_block(m.lambdaSite[block[i]], block[i]);
    return include;
  }
  // Msld::interacting (msld.cu:675-678)
  bool interacting(int i, int j) const {
    return m.atomBlock[i] == m.atomBlock[j] || m.lambdaSite[m.atomBlock[i]] != m.lambdaSite[m.atomBlock[j]];
  }
};

}  // namespace

void build_topology(const Parameters &p, const Structure &s, const MsldSetup &m,
                    const RunSetup &r, Topology &t) {
  const int n = (int)s.atoms.size();
  t = Topology();
  t.n = n;
  if ((int)m.atomBlock.size() != n) fail("msld atom assignment size mismatch");
  for (const auto &b : m.softBonds)
    if (b[0] < 0 || b[0] >= n || b[1] < 0 || b[1] >= n) fail("soft bond atom out of range");
  Scaler sc{m};
  auto type = [&](int i) { return s.atoms[i].atomTypeName; };
  auto isH = [&](int i) { return !type(i).empty() && type(i)[0] == 'H'; };

  t.charge.resize(n);
  t.mass.resize(n);
  for (int i = 0; i < n; i++) {
    t.charge[i] = s.atoms[i].charge;
    t.mass[i] = s.atoms[i].mass;
  }

  // bonds; hydrogen bonds become constraints under shake hbond
  // (potential.cxx:550-595)
  std::map<int, std::set<int>> cons;
  for (auto &b : s.bonds) {
    int idx[2] = {b[0], b[1]};
    if (s.shakeHbond && (isH(idx[0]) || isH(idx[1]))) {
      cons[idx[0]].insert(idx[1]);
      cons[idx[1]].insert(idx[0]);
      continue;
    }
    auto it = p.bonds.find(canon({type(idx[0]), type(idx[1])}));
    if (it == p.bonds.end()) fail("no bond parameter for " + type(idx[0]) + " " + type(idx[1]));
    int sb[2];
    t.bondSoft.push_back(sc.bonded(idx, 2, 0, 2, sb));
    t.bondIdx.insert(t.bondIdx.end(), idx, idx + 2);
    t.bondSb.insert(t.bondSb.end(), sb, sb + 2);
    t.bondKb.push_back(it->second.kb);
    t.bondB0.push_back(it->second.b0);
  }
  t.nBonds12 = (int)t.bondKb.size();
  // angles and Urey-Bradley (potential.cxx:597-645)
  for (auto &a : s.angles) {
    int idx[3] = {a[0], a[1], a[2]};
    auto it = p.angles.find(canon({type(idx[0]), type(idx[1]), type(idx[2])}));
    if (it == p.angles.end()) fail("no angle parameter for " + type(idx[0]) + " " + type(idx[1]) + " " + type(idx[2]));
    int sb[2], sbu[2];
    t.angleSoft.push_back(sc.bonded(idx, 3, 2, 2, sb));
    const bool softU = sc.bonded(idx, 3, 1, 2, sbu);
    t.angleIdx.insert(t.angleIdx.end(), idx, idx + 3);
    t.angleSb.insert(t.angleSb.end(), sb, sb + 2);
    t.angleK.push_back(it->second.kangle);
    t.angleT0.push_back(it->second.angle0);
    if (it->second.kureyb != 0) {
      // appended after the 1-2 bonds below
      t.bondIdx.push_back(idx[0]);
      t.bondIdx.push_back(idx[2]);
      t.bondSb.insert(t.bondSb.end(), sbu, sbu + 2);
      t.bondKb.push_back(it->second.kureyb);
      t.bondB0.push_back(it->second.ureyb0);
      t.bondSoft.push_back(softU);
    }
  }
  // dihedrals with X-X wildcard fallback (potential.cxx:647-705)
  for (auto &d : s.dihes) {
    int idx[4] = {d[0], d[1], d[2], d[3]};
    std::vector<std::string> ty = {type(idx[0]), type(idx[1]), type(idx[2]), type(idx[3])};
    auto it = p.dihes.find(canon(ty));
    if (it == p.dihes.end()) it = p.dihes.find(canon({"X", ty[1], ty[2], "X"}));
    if (it == p.dihes.end()) fail("no dihedral parameter for " + canon(ty));
    int sb[2];
    const bool soft = sc.bonded(idx, 4, 3, 2, sb);
    for (auto &dp : it->second) {
      t.diheSoft.push_back(soft);
      t.diheIdx.insert(t.diheIdx.end(), idx, idx + 4);
      t.diheSb.insert(t.diheSb.end(), sb, sb + 2);
      t.diheK.push_back(dp.kdih);
      t.diheN.push_back(dp.ndih);
      t.diheD0.push_back(dp.dih0);
    }
  }
  // CMAP cross terms: ordered 8-type match (potential.cxx:770-800), up to
  // three lambda scalings (msld/msld.cu:610); one grid per distinct type
  {
    std::map<std::vector<std::string>, int> seen;
    for (auto &c : s.cmaps) {
      int idx[8];
      std::vector<std::string> ty;
      for (int k = 0; k < 8; k++) { idx[k] = c[k]; ty.push_back(type(idx[k])); }
      auto it = p.cmaps.find(ty);
      if (it == p.cmaps.end()) {
        std::string nm;
        for (auto &x : ty) nm += x + " ";
        fail("no CMAP parameter for " + nm);
      }
      auto sn = seen.find(ty);
      int tid;
      if (sn == seen.end()) {
        tid = (int)t.cmapNgrid.size();
        seen[ty] = tid;
        t.cmapNgrid.push_back(it->second.first);
        t.cmapOff.push_back((int)t.cmapGrid.size());
        t.cmapGrid.insert(t.cmapGrid.end(), it->second.second.begin(), it->second.second.end());
      } else {
        tid = sn->second;
      }
      int sb[3];
      sc.bonded(idx, 8, 5, 3, sb);
      t.cmapIdx.insert(t.cmapIdx.end(), idx, idx + 8);
      t.cmapSb.insert(t.cmapSb.end(), sb, sb + 3);
      t.cmapType.push_back(tid);
    }
  }
  // impropers with the reference's wildcard search order (potential.cxx:707-760)
  for (auto &d : s.imprs) {
    int idx[4] = {d[0], d[1], d[2], d[3]};
    std::vector<std::string> ty = {type(idx[0]), type(idx[1]), type(idx[2]), type(idx[3])};
    std::vector<std::vector<std::string>> tries = {
        ty, {ty[0], "X", "X", ty[3]}, {"X", ty[1], ty[2], ty[3]}, {"X", ty[1], ty[2], "X"}, {"X", "X", ty[2], ty[3]}};
    const ImprParam *ip = nullptr;
    for (auto &tr : tries) {
      auto it = p.imprs.find(canon(tr));
      if (it != p.imprs.end()) { ip = &it->second; break; }
    }
    if (!ip) fail("no improper parameter for " + canon(ty));
    int sb[2];
    t.imprSoft.push_back(sc.bonded(idx, 4, 4, 2, sb));
    t.imprIdx.insert(t.imprIdx.end(), idx, idx + 4);
    t.imprSb.insert(t.imprSb.end(), sb, sb + 2);
    t.imprK.push_back(ip->kimp);
    t.imprN.push_back(ip->nimp);
    t.imprD0.push_back(ip->imp0);
  }

  // exclusion sets (potential.cxx:929-1062)
  std::vector<std::set<int>> bondEx(n), angleEx(n), diheEx(n), allEx(n);
  auto add = [&](std::vector<std::set<int>> &v, int a, int b) {
    if (a == b || allEx[a].count(b) || !sc.interacting(a, b)) return;
    v[a].insert(b); v[b].insert(a); allEx[a].insert(b); allEx[b].insert(a);
  };
  // a virtual site is excluded from its first host like a bond partner,
  // and from every bond partner of that host (potential.cxx:947-995)
  std::vector<std::set<int>> virtEx(n);
  for (const auto &v : s.virts) {
    auto ok = [&](int a) { return a >= 0 && a < n; };
    if (!ok(v.v) || !ok(v.h[0]) || !ok(v.h[1]) || (v.nh == 3 && !ok(v.h[2]))) fail("virtual site atom out of range");
    const int a = v.v, b = v.h[0];
    if (a == b || allEx[a].count(b) || !sc.interacting(a, b)) continue;
    virtEx[a].insert(b); virtEx[b].insert(a);
    add(bondEx, a, b);
  }
  for (auto &b : s.bonds) {
    add(bondEx, b[0], b[1]);
    for (int j = 0; j < 2; j++) {
      std::vector<int> ks(virtEx[b[1 - j]].begin(), virtEx[b[1 - j]].end());
      for (int k : ks) add(bondEx, b[j], k);
    }
  }
  for (auto &b : s.bonds)
    for (int j = 0; j < 2; j++) {
      std::vector<int> ks(bondEx[b[1 - j]].begin(), bondEx[b[1 - j]].end());
      for (int k : ks) add(angleEx, b[j], k);
    }
  for (auto &b : s.bonds)
    for (int j = 0; j < 2; j++) {
      std::vector<int> ks(angleEx[b[1 - j]].begin(), angleEx[b[1 - j]].end());
      for (int k : ks) add(diheEx, b[j], k);
    }
  // atoms of different blocks of one site never interact (potential.cxx:1048-1062)
  {
    std::vector<std::vector<int>> byBlock(m.blockCount);
    for (int i = 0; i < n; i++) byBlock[m.atomBlock[i]].push_back(i);
    for (int b1 = 1; b1 < m.blockCount; b1++)
      for (int b2 = b1 + 1; b2 < m.blockCount; b2++) {
        if (m.lambdaSite[b1] != m.lambdaSite[b2]) continue;
        for (int i : byBlock[b1])
          for (int j : byBlock[b2]) {
            allEx[i].insert(j);
            allEx[j].insert(i);
          }
      }
  }
  // vdW types sorted by count (potential.cxx:1309-1339)
  std::map<std::string, int> tcount;
  for (int i = 0; i < n; i++) tcount[type(i)]++;
  std::vector<std::pair<int, std::string>> order;
  for (auto &kv : tcount) order.push_back({-kv.second, kv.first});
  std::sort(order.begin(), order.end());
  std::map<std::string, int> lookup;
  for (auto &o : order) {
    lookup[o.second] = (int)t.typeNames.size();
    t.typeNames.push_back(o.second);
  }
  t.nTypes = (int)t.typeNames.size();
  t.typeIdx.resize(n);
  t.siteBlock.resize(n);
  for (int i = 0; i < n; i++) {
    t.typeIdx[i] = lookup[type(i)];
    sc.nonbonded(&i, 1, &t.siteBlock[i]);
  }
  auto nbpair = [&](const std::string &a, const std::string &b, bool is14, double &c12, double &c6) {
    double eps, sig;
    auto fx = p.nbfix.find(canon({a, b}));
    if (fx != p.nbfix.end()) {
      eps = is14 ? fx->second.eps14 : fx->second.eps;
      sig = is14 ? fx->second.sig14 : fx->second.sig;
    } else {
      auto ia = p.nbonds.find(a), ib = p.nbonds.find(b);
      if (ia == p.nbonds.end() || ib == p.nbonds.end()) fail("nonbonded parameter not found for " + a + "/" + b);
      const NbondParam &A = ia->second, &B = ib->second;
      eps = is14 ? std::sqrt(A.eps14 * B.eps14) : std::sqrt(A.eps * B.eps);
      double sa = is14 ? A.sig14 : A.sig, sbb = is14 ? B.sig14 : B.sig;
      sig = sa + sbb;
      if (A.combine == 1 && B.combine == 1) sig = 2 * std::sqrt(sa * sbb);
    }
    double s6 = sig * sig;
    s6 = s6 * s6 * s6;
    c12 = eps * s6 * s6;
    c6 = 2 * eps * s6;
  };
  t.c12.resize((size_t)t.nTypes * t.nTypes);
  t.c6.resize(t.c12.size());
  for (int a = 0; a < t.nTypes; a++)
    for (int b = 0; b < t.nTypes; b++)
      nbpair(t.typeNames[a], t.typeNames[b], false, t.c12[a * t.nTypes + b], t.c6[a * t.nTypes + b]);

  // 1-4 pairs (potential.cxx:1093-1142)
  for (int i = 0; i < n; i++)
    for (int j : diheEx[i]) {
      if (i >= j) continue;
      int idx[2] = {i, j}, sb[2];
      sc.nonbonded(idx, 2, sb);
      if (sb[0] != sb[1] && (sb[0] & 0xFFFF0000) == (sb[1] & 0xFFFF0000)) fail("illegal 1-4 interaction");
      auto ia = p.nbonds.find(type(i)), ib = p.nbonds.find(type(j));
      if (ia == p.nbonds.end() || ib == p.nbonds.end()) fail("nonbonded parameter not found");
      double c12, c6;
      nbpair(type(i), type(j), true, c12, c6);
      t.nb14Idx.push_back(i);
      t.nb14Idx.push_back(j);
      t.nb14Sb.insert(t.nb14Sb.end(), sb, sb + 2);
      t.nb14qq.push_back(t.charge[i] * t.charge[j]);
      t.nb14c12.push_back(c12);
      t.nb14c6.push_back(c6);
      t.nb14e14.push_back(std::sqrt(ia->second.e14fac) * std::sqrt(ib->second.e14fac));
    }
  // PME exclusion corrections (potential.cxx:1144-1163) and exclusions
  for (int i = 0; i < n; i++)
    for (int j : allEx[i]) {
      if (i >= j) continue;
      t.exclIdx.push_back(i);
      t.exclIdx.push_back(j);
      if (r.elecMethod == BM_PME && !diheEx[i].count(j)) {
        int idx[2] = {i, j}, sb[2];
        if (sc.nbex(idx, sb)) {
          t.nbexIdx.push_back(i);
          t.nbexIdx.push_back(j);
          t.nbexSb.insert(t.nbexSb.end(), sb, sb + 2);
          t.nbexqq.push_back(t.charge[i] * t.charge[j]);
        }
      }
    }

  // constraint groups (potential.cxx:1422-1621)
  auto b0of = [&](int a, int b) {
    auto it = p.bonds.find(canon({type(a), type(b)}));
    if (it == p.bonds.end()) fail("no bond parameter for constraint " + type(a) + " " + type(b));
    return it->second.b0;
  };
  for (auto &kv : cons) {
    std::vector<int> ci = {kv.first};
    ci.insert(ci.end(), kv.second.begin(), kv.second.end());
    std::vector<bool> h;
    for (int a : ci) h.push_back(isH(a));
    bool bad = false;
    if (ci.size() == 3) {
      if (cons[ci[1]].size() == 2 && cons[ci[2]].size() == 2) {
        if (cons[ci[1]].count(ci[2]) && cons[ci[2]].count(ci[1])) {
          bool first1 = h[0] < h[1] || (h[0] == h[1] && ci[0] < ci[1]);
          bool first2 = h[0] < h[2] || (h[0] == h[2] && ci[0] < ci[2]);
          if (first1 && first2) {
            t.triIdx.insert(t.triIdx.end(), ci.begin(), ci.end());
            t.triB0.push_back(b0of(ci[0], ci[1]));
            t.triB0.push_back(b0of(ci[0], ci[2]));
            t.triB0.push_back(b0of(ci[1], ci[2]));
          }
        } else {
          bad = true;
        }
      } else if (cons[ci[1]].size() == 1 && cons[ci[2]].size() == 1) {
        t.br2Idx.insert(t.br2Idx.end(), ci.begin(), ci.end());
        t.br2B0.push_back(b0of(ci[0], ci[1]));
        t.br2B0.push_back(b0of(ci[0], ci[2]));
      } else {
        bad = true;
      }
    } else if (ci.size() == 4) {
      if (cons[ci[1]].size() == 1 && cons[ci[2]].size() == 1 && cons[ci[3]].size() == 1) {
        t.br3Idx.insert(t.br3Idx.end(), ci.begin(), ci.end());
        for (int k = 1; k < 4; k++) t.br3B0.push_back(b0of(ci[0], ci[k]));
      } else {
        bad = true;
      }
    } else if (ci.size() == 2) {
      if (cons[ci[1]].size() == 1) {
        if (h[0] < h[1] || (h[0] == h[1] && ci[0] < ci[1])) {
          t.br1Idx.insert(t.br1Idx.end(), ci.begin(), ci.end());
          t.br1B0.push_back(b0of(ci[0], ci[1]));
        }
      }
    } else {
      bad = true;
    }
    if (bad) fail("unrecognized constraint group at atom " + std::to_string(ci[0]));
  }

  // MSLD blocks and sites (msld.cu:681-740)
  t.nBlocks = m.blockCount;
  t.blockSite = m.lambdaSite;
  int nsite = 1;
  for (int b = 0; b < m.blockCount; b++) {
    nsite = std::max(nsite, m.lambdaSite[b] + 1);
    if (b && m.lambdaSite[b] != m.lambdaSite[b - 1] && m.lambdaSite[b] != m.lambdaSite[b - 1] + 1)
      fail("blocks must be ordered by consecutive sites");
  }
  t.nSites = nsite;
  std::vector<int> per(nsite, 0);
  for (int b = 0; b < m.blockCount; b++) per[m.lambdaSite[b]]++;
  if (per[0] != 1) fail("only one block allowed in site 0");
  t.siteBound.assign(nsite + 1, 0);
  for (int s2 = 0; s2 < nsite; s2++) {
    if (s2 && per[s2] < 2 && !(per[s2] == 1 && m.blockFixed[t.siteBound[s2]]))
      fail("at least two blocks are required in each site");
    t.siteBound[s2 + 1] = t.siteBound[s2] + per[s2];
  }
  t.blockFixed = m.blockFixed;
  t.thetaMass = m.thetaMass;
  for (int b = 0; b < m.blockCount; b++)
    if (m.blockFixed[b]) t.thetaMass[b] = 0;
  t.fixedBias = m.lambdaBias;
  t.thetaFriction = m.thetaFriction;
  for (auto &v : m.varBias) {
    t.vbI.push_back(v.i); t.vbJ.push_back(v.j); t.vbType.push_back(v.type);
    t.vbN.push_back(v.n); t.vbL0.push_back(v.l0); t.vbK.push_back(v.k);
  }
  // atom restraint groups (msld.cu:755-776), empty groups dropped
  t.catsBound.assign(1, 0);
  t.catsIdx.clear();
  for (const auto &g : m.atomRestraints) {
    if (g.empty()) continue;
    for (int a : g) {
      if (a < 0 || a >= n) fail("atom restraint index out of range");
      t.catsIdx.push_back(a);
    }
    t.catsBound.push_back((int)t.catsIdx.size());
  }
  t.catsK = m.kRestraint;
  // each block's charge is the sum of its atoms' charges, whatever
  // blockCharge the initial conditions carried (Msld::initialize,
  // msld.cu:684-690)
  t.blockCharge.assign(m.blockCount, 0.0);
  for (int i = 0; i < n; i++) t.blockCharge[m.atomBlock[i]] += t.charge[i];
  if (m.kCR2 != 0 && m.kCR2 != 1) fail("only kChargeRestraint2 values of 0 or 1 are allowed");
  t.crK1 = m.kCR1; t.crK2 = m.kCR2; t.crK3 = m.kCR3; t.crQ0 = m.q0CR3; t.crW = m.wCR3;
  // theta biases: one value per site (msld.cu:1157-1160 and 1184 index
  // them by site)
  if (!m.thetaCollK.empty() && (int)m.thetaCollK.size() != nsite)
    fail("theta collective bias needs one entry per site (" + std::to_string(nsite) + ")");
  if (!m.thetaIndeK.empty() && (int)m.thetaIndeK.size() != nsite)
    fail("theta independent bias needs one entry per site (" + std::to_string(nsite) + ")");
  t.tcollK = m.thetaCollK;
  t.tcollN = m.thetaCollN;
  t.tindeK = m.thetaIndeK;
  t.piecewise = m.piecewise ? 1 : 0;
  if (t.piecewise)
    for (int b = 0; b < m.blockCount; b++)
      if (m.blockFixed[b]) fail("fixed blocks are not supported with the piecewise lambda constraint");
  t.pwWidth = m.wellWidth;
  t.pwK = m.wellK;
  t.edgeK = m.edgeK; t.edgeN = m.edgeN; t.edgeAlpha = m.edgeAlpha; t.edgePhi = m.edgePhi;
  t.fixLambda = m.fix ? 1 : 0;
  t.anySoft = false;
  for (const auto *v : {&t.bondSoft, &t.angleSoft, &t.diheSoft, &t.imprSoft})
    for (int f : *v) t.anySoft = t.anySoft || f;
  // virtual sites (indices checked with the exclusions)
  for (const auto &v : s.virts) {
    if (v.nh == 2) {
      const int idx[3] = {v.v, v.h[0], v.h[1]};
      t.virt2Idx.insert(t.virt2Idx.end(), idx, idx + 3);
      t.virt2Par.push_back(v.dist);
      t.virt2Par.push_back(v.a);
    } else {
      const int idx[4] = {v.v, v.h[0], v.h[1], v.h[2]};
      t.virt3Idx.insert(t.virt3Idx.end(), idx, idx + 4);
      t.virt3Par.push_back(v.dist);
      t.virt3Par.push_back(v.a);
      t.virt3Par.push_back(v.b);
    }
  }
  // restraints (Potential::initialize copies the structure's lists,
  // potential.cxx:1620-1660)
  auto atom_ok = [&](int a) {
    if (a < 0 || a >= n) fail("restraint atom index out of range");
    return a;
  };
  auto block_ok = [&](int b) {
    if (b < 0 || b >= m.blockCount) fail("restraint lambda block out of range");
    return b;
  };
  for (const NoeRec &r : s.noes) {
    t.noeIdx.push_back(atom_ok(r.i));
    t.noeIdx.push_back(r.pnoe ? -1 : atom_ok(r.j));
    const double p[10] = {r.rmin, r.kmin, r.rmax, r.kmax, r.rpeak, r.rswitch, r.nswitch, r.c0[0], r.c0[1], r.c0[2]};
    t.noePar.insert(t.noePar.end(), p, p + 10);
  }
  for (const HarmRec &r : s.harms) {
    t.harmIdx.push_back(atom_ok(r.i));
    const double p[5] = {r.k, r.x0[0], r.x0[1], r.x0[2], r.n};
    t.harmPar.insert(t.harmPar.end(), p, p + 5);
  }
  for (const BoRestRec &r : s.boRests) {
    t.borIdx.push_back(atom_ok(r.i));
    t.borIdx.push_back(atom_ok(r.j));
    t.borBlock.push_back(block_ok(r.block));
    t.borPar.push_back(r.kr);
    t.borPar.push_back(r.r0);
  }
  for (const AnRestRec &r : s.anRests) {
    for (int a : {r.i, r.j, r.k}) t.anrIdx.push_back(atom_ok(a));
    t.anrBlock.push_back(block_ok(r.block));
    t.anrPar.push_back(r.kt);
    t.anrPar.push_back(r.t0);
  }
  for (const DiRestRec &r : s.diRests) {
    for (int a : {r.i, r.j, r.k, r.l}) t.dirIdx.push_back(atom_ok(a));
    t.dirBlock.push_back(block_ok(r.block));
    t.dirN.push_back(r.nphi);
    t.dirPar.push_back(r.kphi);
    t.dirPar.push_back(r.phi0);
    t.dirPar.push_back(r.width);
  }
  for (const ResdRec &r : s.resds) {
    for (int a : {r.i1, r.i2, r.j1, r.j2}) t.resdIdx.push_back(atom_ok(a));
    const double p[4] = {r.ci, r.cj, r.rdist, r.kdist};
    t.resdPar.insert(t.resdPar.end(), p, p + 4);
  }
  (void)kBoltz;
}

const bm_topology *Topology::view() {
  bm_topology &v = top;
  v = bm_topology{};
  v.n_atoms = n;
  v.charge = charge.data(); v.mass = mass.data();
  v.type_idx = typeIdx.data(); v.site_block = siteBlock.data();
  v.n_types = nTypes; v.vdw_c12 = c12.data(); v.vdw_c6 = c6.data();
  v.n_bonds = (int)bondKb.size(); v.n_bonds12 = nBonds12;
  v.bond_idx = bondIdx.data(); v.bond_sb = bondSb.data(); v.bond_kb = bondKb.data(); v.bond_b0 = bondB0.data();
  v.n_angles = (int)angleK.size();
  v.angle_idx = angleIdx.data(); v.angle_sb = angleSb.data(); v.angle_k = angleK.data(); v.angle_t0 = angleT0.data();
  v.n_dihes = (int)diheK.size();
  v.dihe_idx = diheIdx.data(); v.dihe_sb = diheSb.data(); v.dihe_k = diheK.data(); v.dihe_d0 = diheD0.data(); v.dihe_n = diheN.data();
  v.n_imprs = (int)imprK.size();
  v.impr_idx = imprIdx.data(); v.impr_sb = imprSb.data(); v.impr_k = imprK.data(); v.impr_d0 = imprD0.data(); v.impr_n = imprN.data();
  if (anySoft) {
    v.bond_soft = bondSoft.data(); v.angle_soft = angleSoft.data();
    v.dihe_soft = diheSoft.data(); v.impr_soft = imprSoft.data();
  }
  v.n_virt2 = (int)virt2Par.size() / 2; v.virt2_idx = virt2Idx.data(); v.virt2_par = virt2Par.data();
  v.n_virt3 = (int)virt3Par.size() / 3; v.virt3_idx = virt3Idx.data(); v.virt3_par = virt3Par.data();
  v.n_nb14 = (int)nb14qq.size();
  v.nb14_idx = nb14Idx.data(); v.nb14_sb = nb14Sb.data(); v.nb14_qq = nb14qq.data();
  v.nb14_c12 = nb14c12.data(); v.nb14_c6 = nb14c6.data(); v.nb14_e14fac = nb14e14.data();
  v.n_nbex = (int)nbexqq.size();
  v.nbex_idx = nbexIdx.data(); v.nbex_sb = nbexSb.data(); v.nbex_qq = nbexqq.data();
  v.n_excl = (int)exclIdx.size() / 2; v.excl_idx = exclIdx.data();
  v.n_tri = (int)triIdx.size() / 3; v.n_br1 = (int)br1Idx.size() / 2;
  v.n_br2 = (int)br2Idx.size() / 3; v.n_br3 = (int)br3Idx.size() / 4;
  v.tri_idx = triIdx.data(); v.tri_b0 = triB0.data();
  v.br1_idx = br1Idx.data(); v.br1_b0 = br1B0.data();
  v.br2_idx = br2Idx.data(); v.br2_b0 = br2B0.data();
  v.br3_idx = br3Idx.data(); v.br3_b0 = br3B0.data();
  v.n_blocks = nBlocks; v.n_sites = nSites;
  v.block_site = blockSite.data(); v.site_bound = siteBound.data();
  v.theta_mass = thetaMass.data(); v.block_fixed = blockFixed.data();
  v.fixed_bias = fixedBias.data();
  v.theta_friction = thetaFriction.empty() ? nullptr : thetaFriction.data();
  v.n_varbias = (int)vbK.size();
  v.vb_i = vbI.data(); v.vb_j = vbJ.data(); v.vb_type = vbType.data(); v.vb_n = vbN.data();
  v.vb_l0 = vbL0.data(); v.vb_k = vbK.data();
  v.n_cats = catsBound.empty() ? 0 : (int)catsBound.size() - 1;
  v.cats_bound = catsBound.empty() ? nullptr : catsBound.data();
  v.cats_idx = catsIdx.data();
  v.cats_k = catsK;
  v.block_charge = blockCharge.data();
  v.n_cmaps = (int)cmapType.size();
  v.cmap_idx = cmapIdx.data(); v.cmap_sb = cmapSb.data(); v.cmap_type = cmapType.data();
  v.n_cmap_types = (int)cmapNgrid.size();
  v.cmap_ngrid = cmapNgrid.data(); v.cmap_off = cmapOff.data(); v.cmap_grid = cmapGrid.data();
  v.cr_k1 = crK1; v.cr_k2 = crK2; v.cr_k3 = crK3; v.cr_q0 = crQ0; v.cr_w = crW;
  v.tcoll_k = tcollK.empty() ? nullptr : tcollK.data();
  v.tcoll_n = tcollN.empty() ? nullptr : tcollN.data();
  v.tinde_k = tindeK.empty() ? nullptr : tindeK.data();
  v.piecewise = piecewise;
  v.pw_width = pwWidth; v.pw_k = pwK;
  v.edge_k = edgeK; v.edge_n = edgeN; v.edge_alpha = edgeAlpha; v.edge_phi = edgePhi;
  v.fix_lambda = fixLambda;
  v.n_noe = (int)noeIdx.size() / 2; v.noe_idx = noeIdx.data(); v.noe_par = noePar.data();
  v.n_harm = (int)harmIdx.size(); v.harm_idx = harmIdx.data(); v.harm_par = harmPar.data();
  v.n_borest = (int)borBlock.size(); v.borest_idx = borIdx.data(); v.borest_block = borBlock.data();
  v.borest_par = borPar.data();
  v.n_anrest = (int)anrBlock.size(); v.anrest_idx = anrIdx.data(); v.anrest_block = anrBlock.data();
  v.anrest_par = anrPar.data();
  v.n_direst = (int)dirBlock.size(); v.direst_idx = dirIdx.data(); v.direst_block = dirBlock.data();
  v.direst_n = dirN.data(); v.direst_par = dirPar.data();
  v.n_resd = (int)resdIdx.size() / 4; v.resd_idx = resdIdx.data(); v.resd_par = resdPar.data();
  return &top;
}

void build_runparams(const RunSetup &r, const MsldSetup &m, const Coordinates &c, int natoms,
                     bm_runparams &rp) {
  rp = bm_runparams{};
  rp.rcut = r.rCut;
  rp.rswitch = r.rSwitch;
  rp.beta_ewald = r.betaEwald;
  rp.vdw_method = r.vdwMethod;
  rp.elec_method = r.elecMethod;
  for (int d = 0; d < 3; d++) {
    rp.box[d] = c.box[d];
    rp.grid[d] = pme_grid_size(c.box[d], r.gridSpace, r.grid[d]);
  }
  rp.order = r.orderEwald;
  rp.dt = r.dt;
  rp.temperature = r.T;
  rp.gamma = r.gamma;
  rp.msld_gamma = m.gamma;
  rp.shake_tol = r.shakeTolerance;
  rp.fnex = m.fnex;
  rp.soft_core = m.useSoftCore;
  rp.soft_core14 = m.useSoftCore14;
  rp.soft_bond_exp = m.softNotBondExponent;
  rp.soft_bond_alpha = 1.0 / (m.softBondRadius * m.softBondRadius);
  rp.soft_bond_power = m.softBondExponent;
  rp.freq_domdec = r.freqDomdec > 0 ? r.freqDomdec : 10;
  rp.fft_mode = r.fftMode;
  rp.tile_chunk = r.tileChunk;
  // cullPad = 2 sqrt(30 kB T) freqDomdec dt (domdec/domdec.cu:138)
  rp.cull_pad = 2 * std::sqrt(30 * kBoltz * r.T) * rp.freq_domdec * r.dt;
  rp.seed = r.seed;
  (void)natoms;
}

}  // namespace blade
