// xtc.cpp - GROMACS XTC trajectory frames (the format the reference writes
// through its vendored xdrfile library, io/io.cxx:314-345 and src/xdr/).
//
// A frame is XDR (big-endian) data: magic 1995, atom count, step, time,
// the 3x3 box in nm, then the coordinates.  Up to 9 atoms are stored as
// plain floats; larger frames use the XTC integer compression: coordinates
// rounded to integers at `precision` per nm, the bounding box of those
// integers, and a bit stream in which each atom is either a full triple
// (mixed-radix packed against the box extents) or, for atoms close to the
// previous one (waters), a run of small differences packed against one of a
// table of size classes whose index adapts from atom to atom.  Readers
// decode with the same size-class table, so it is part of the format.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// size classes of the small-difference runs (about 2^(k/3)); index 9 is
// the first one used, the last is the table's end
const int kSizeClass[] = {0,       0,       0,       0,       0,       0,       0,       0,       0,
                          8,       10,      12,      16,      20,      25,      32,      40,      50,
                          64,      80,      101,     128,     161,     203,     256,     322,     406,
                          512,     645,     812,     1024,    1290,    1625,    2048,    2580,    3250,
                          4096,    5060,    6501,    8192,    10321,   13003,   16384,   20642,   26007,
                          32768,   41285,   52015,   65536,   82570,   104031,  131072,  165140,  208063,
                          262144,  330280,  416127,  524287,  660561,  832255,  1048576, 1321122, 1664510,
                          2097152, 2642245, 3329021, 4194304, 5284491, 6658042, 8388607, 10568983, 13316085,
                          16777216};
constexpr int kFirstClass = 9;
constexpr int kLastClass = (int)(sizeof(kSizeClass) / sizeof(kSizeClass[0]));
constexpr int kMaxAbs = INT_MAX - 2;

// bits appended most significant first; the last byte is left-aligned
struct BitStream {
  std::vector<unsigned char> bytes;
  unsigned acc = 0;   // pending bits (low `nacc` bits of acc)
  int nacc = 0;
  void put(int nbits, unsigned v) {
    for (int b = nbits - 1; b >= 0; b--) {
      acc = (acc << 1) | (b < 32 ? (v >> b) & 1u : 0u);
      if (++nacc == 8) {
        bytes.push_back((unsigned char)acc);
        acc = 0;
        nacc = 0;
      }
    }
  }
  std::vector<unsigned char> finish() const {
    std::vector<unsigned char> out = bytes;
    if (nacc) out.push_back((unsigned char)(acc << (8 - nacc)));
    return out;
  }
};

// bits needed to hold the value `size` itself
int bits_of(unsigned size) {
  int n = 0;
  unsigned long long v = 1;
  while (size >= v && n < 32) {
    n++;
    v <<= 1;
  }
  return n;
}

// little-endian bytes of the mixed-radix number v0*s1*s2 + v1*s2 + v2
// (v0 first, then each next value multiplied in)
std::vector<unsigned> radix_bytes(int n, const unsigned *sizes, const unsigned *vals) {
  std::vector<unsigned> b;
  unsigned t = vals[0];
  do {
    b.push_back(t & 0xffu);
    t >>= 8;
  } while (t);
  for (int i = 1; i < n; i++) {
    unsigned long long carry = vals[i];
    for (auto &x : b) {
      carry += (unsigned long long)x * sizes[i];
      x = (unsigned)(carry & 0xffu);
      carry >>= 8;
    }
    while (carry) {
      b.push_back((unsigned)(carry & 0xffu));
      carry >>= 8;
    }
  }
  return b;
}

// bits of the largest product of the sizes (the bit length of the product)
int bits_of_product(int n, const unsigned *sizes) {
  std::vector<unsigned> b(1, 1u);
  for (int i = 0; i < n; i++) {
    unsigned long long carry = 0;
    for (auto &x : b) {
      carry += (unsigned long long)x * sizes[i];
      x = (unsigned)(carry & 0xffu);
      carry >>= 8;
    }
    while (carry) {
      b.push_back((unsigned)(carry & 0xffu));
      carry >>= 8;
    }
  }
  int top = 0;
  unsigned v = 1;
  while (b.back() >= v) {
    top++;
    v <<= 1;
  }
  return top + 8 * ((int)b.size() - 1);
}

// a packed triple of `nbits` bits: its bytes least significant first, the
// last one cut to the bits left
void put_triple(BitStream &bs, int nbits, const unsigned *sizes, const unsigned *vals) {
  const std::vector<unsigned> b = radix_bytes(3, sizes, vals);
  const int nb = (int)b.size();
  if (nbits >= 8 * nb) {
    for (int i = 0; i < nb; i++) bs.put(8, b[i]);
    bs.put(nbits - 8 * nb, 0);
  } else {
    for (int i = 0; i < nb - 1; i++) bs.put(8, b[i]);
    bs.put(nbits - 8 * (nb - 1), b[nb - 1]);
  }
}

struct Xdr {
  std::vector<unsigned char> out;
  void i32(int32_t v) {
    const uint32_t u = (uint32_t)v;
    for (int s = 24; s >= 0; s -= 8) out.push_back((unsigned char)(u >> s));
  }
  void f32(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    i32((int32_t)u);
  }
  void opaque(const std::vector<unsigned char> &b) {
    out.insert(out.end(), b.begin(), b.end());
    while (out.size() % 4) out.push_back(0);
  }
};

int quantize(float x, float prec, int &q) {
  const float s = x * prec;
  const float lf = (float)((double)s + (s >= 0.0f ? 0.5 : -0.5));
  if (std::fabs(lf) > (float)kMaxAbs) return -1;
  q = (int)lf;
  return 0;
}

// the compressed coordinate block of one frame
int put_coords(Xdr &x, int n, const float *xyz, float prec) {
  x.i32(n);
  if (n <= 9) {
    for (int i = 0; i < 3 * n; i++) x.f32(xyz[i]);
    return 0;
  }
  x.f32(prec);
  std::vector<int> q(3 * (size_t)n);
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  long long mindiff = LLONG_MAX;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      if (quantize(xyz[3 * i + k], prec, q[3 * i + k])) return -1;
      lo[k] = std::min(lo[k], q[3 * i + k]);
      hi[k] = std::max(hi[k], q[3 * i + k]);
    }
    if (i > 0) {
      long long d = 0;
      for (int k = 0; k < 3; k++) d += std::llabs((long long)q[3 * i + k] - q[3 * (i - 1) + k]);
      mindiff = std::min(mindiff, d);
    }
  }
  for (int k = 0; k < 3; k++) x.i32(lo[k]);
  for (int k = 0; k < 3; k++) x.i32(hi[k]);
  unsigned ext[3];
  int extbits[3] = {0, 0, 0};
  bool wide = false;
  for (int k = 0; k < 3; k++) {
    if ((long long)hi[k] - lo[k] >= kMaxAbs) return -1;
    ext[k] = (unsigned)(hi[k] - lo[k] + 1);
    if (ext[k] > 0xffffffu) wide = true;
  }
  int tripbits = 0;
  if (wide)
    for (int k = 0; k < 3; k++) extbits[k] = bits_of(ext[k]);
  else
    tripbits = bits_of_product(3, ext);
  int cls = kFirstClass;
  while (cls < kLastClass - 1 && kSizeClass[cls] < mindiff) cls++;
  x.i32(cls);
  const int maxcls = std::min(kLastClass, cls + 8), mincls = maxcls - 8;
  int smaller = kSizeClass[std::max(kFirstClass, cls - 1)] / 2;
  int smallnum = kSizeClass[cls] / 2;
  unsigned small3[3] = {(unsigned)kSizeClass[cls], (unsigned)kSizeClass[cls], (unsigned)kSizeClass[cls]};
  const int larger = kSizeClass[std::min(maxcls, kLastClass - 1)] / 2;
  BitStream bs;
  int prev[3] = {0, 0, 0};
  int prevrun = -1;
  int i = 0;
  unsigned diffs[8 * 3];
  auto near = [](const int *a, const int *b, int lim) {
    return std::abs(a[0] - b[0]) < lim && std::abs(a[1] - b[1]) < lim && std::abs(a[2] - b[2]) < lim;
  };
  while (i < n) {
    int *cur = q.data() + 3 * (size_t)i;
    int step_class;
    if (cls < maxcls && i >= 1 && near(cur, prev, larger)) step_class = 1;
    else if (cls > mincls) step_class = -1;
    else step_class = 0;
    bool run_next = false;
    if (i + 1 < n && near(cur, cur + 3, smallnum)) {
      // the first two atoms of a close pair are stored in swapped order
      // (a water's oxygen after its first hydrogen)
      for (int k = 0; k < 3; k++) std::swap(cur[k], cur[3 + k]);
      run_next = true;
    }
    unsigned full[3];
    for (int k = 0; k < 3; k++) full[k] = (unsigned)(cur[k] - lo[k]);
    if (wide)
      for (int k = 0; k < 3; k++) bs.put(extbits[k], full[k]);
    else
      put_triple(bs, tripbits, ext, full);
    for (int k = 0; k < 3; k++) prev[k] = cur[k];
    i++;
    cur += 3;
    int run = 0;
    if (!run_next && step_class == -1) step_class = 0;
    while (run_next && run < 8 * 3) {
      if (step_class == -1) {
        // the squared distance and bound in 32-bit two's-complement
        // arithmetic, as readers and writers of the format compute them
        // (they wrap for the coarsest size classes)
        uint32_t d2 = 0;
        for (int k = 0; k < 3; k++) {
          const uint32_t d = (uint32_t)(cur[k] - prev[k]);
          d2 += d * d;
        }
        const uint32_t b2 = (uint32_t)smaller * (uint32_t)smaller;
        if ((int32_t)d2 >= (int32_t)b2) step_class = 0;
      }
      for (int k = 0; k < 3; k++) diffs[run++] = (unsigned)(cur[k] - prev[k] + smallnum);
      for (int k = 0; k < 3; k++) prev[k] = cur[k];
      i++;
      cur += 3;
      run_next = i < n && near(cur, prev, smallnum);
    }
    if (run != prevrun || step_class != 0) {
      prevrun = run;
      bs.put(1, 1);
      bs.put(5, (unsigned)(run + step_class + 1));
    } else {
      bs.put(1, 0);
    }
    for (int k = 0; k < run; k += 3) put_triple(bs, cls, small3, diffs + k);
    if (step_class != 0) {
      cls += step_class;
      if (step_class < 0) {
        smallnum = smaller;
        smaller = cls > kFirstClass ? kSizeClass[cls - 1] / 2 : 0;
      } else {
        smaller = smallnum;
        smallnum = kSizeClass[cls] / 2;
      }
      small3[0] = small3[1] = small3[2] = (unsigned)kSizeClass[cls];
    }
  }
  const std::vector<unsigned char> packed = bs.finish();
  x.i32((int32_t)packed.size());
  x.opaque(packed);
  return 0;
}

}  // namespace

extern "C" {

// one XTC frame appended to fp: positions AoS in nm, box diagonal-or-full
// 3x3 in nm (row-major), time in ps; 0 on success
int blade_xtc_frame(FILE *fp, int natoms, int step, float time, const float *box9, const float *xyz,
                    float precision) {
  if (!fp || natoms < 0 || (natoms && !xyz)) return -1;
  Xdr x;
  x.i32(1995);
  x.i32(natoms);
  x.i32(step);
  x.f32(time);
  for (int k = 0; k < 9; k++) x.f32(box9 ? box9[k] : 0.0f);
  if (put_coords(x, natoms, xyz, precision)) return -1;
  return fwrite(x.out.data(), 1, x.out.size(), fp) == x.out.size() ? 0 : -1;
}

// the same to a file by name (append nonzero: added at its end)
int blade_write_xtc_frame(const char *path, int append, int natoms, int step, float time, const float *box9,
                          const float *xyz, float precision) {
  FILE *fp = fopen(path, append ? "ab" : "wb");
  if (!fp) return -1;
  const int rc = blade_xtc_frame(fp, natoms, step, time, box9, xyz, precision);
  return (fclose(fp) == 0 && rc == 0) ? 0 : -1;
}

}  // extern "C"
