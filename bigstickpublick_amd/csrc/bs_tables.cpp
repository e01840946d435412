// Host-side table construction (see bs_tables.hpp for the reference map).
//
// The floating-point steps that produce the FP32 matrix elements are kept in
// the reference's order (single-precision Racah formula for the
// Clebsch-Gordan coefficients, left-to-right products, J-ascending sums) so
// that the matrix elements come out bit-identical to the reference's hmat*.
#include "bs_tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <utility>

namespace bs {

namespace {

int pm1(int k) { return (k % 2 == 0) ? 1 : -1; }

struct FacLog {
  double v[501];
  FacLog() {
    double fn = 1.0;
    v[1] = 0.0;
    v[2] = 0.0;
    for (int i = 3; i <= 500; i++) {
      fn = fn + 1.0;
      v[i] = v[i - 1] + std::log(fn);
    }
  }
};
const FacLog& faclog() {
  static FacLog f;
  return f;
}

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

std::vector<std::string> data_tokens(std::istream& is, bool first_line_only_check = false) {
  (void)first_line_only_check;
  std::vector<std::string> toks;
  std::string line;
  while (std::getline(is, line)) {
    size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos) continue;
    if (line[p] == '!' || line[p] == '#') continue;
    std::istringstream ss(line);
    std::string t;
    while (ss >> t) {
      if (t[0] == '!') break;
      toks.push_back(t);
    }
  }
  return toks;
}

// Fortran list-directed reads over the data lines of a file (comment lines
// starting with ! or # skipped): each take(n) starts on a new line, continues
// over the following lines until n values are read, and drops the rest of its
// last line, as one READ(unit, *) statement does.
struct RecordReader {
  std::vector<std::vector<std::string>> lines;
  size_t next = 0;
  explicit RecordReader(std::istream& is) {
    std::string line;
    while (std::getline(is, line)) {
      size_t p = line.find_first_not_of(" \t\r");
      if (p == std::string::npos || line[p] == '!' || line[p] == '#') continue;
      std::istringstream ss(line);
      std::vector<std::string> toks;
      std::string t;
      while (ss >> t) {
        if (t[0] == '!') break;
        toks.push_back(t);
      }
      if (!toks.empty()) lines.push_back(toks);
    }
  }
  bool take(size_t n, std::vector<std::string>& out) {
    out.clear();
    while (out.size() < n && next < lines.size()) {
      for (const std::string& t : lines[next])
        if (out.size() < n) out.push_back(t);
      next++;
    }
    return out.size() == n;
  }
};

}  // namespace

// Clebsch-Gordan <j1 m1 j2 m2 | j3 m3> with 2j, 2m arguments: the Racah sum in
// single precision over double log-factorials, with the argument reordering
// of src/libra.f:419-521.
float cleb(int ia, int id, int ib, int ie, int ic, int iff) {
  const double* fl = faclog().v;
  if (id + ie - iff != 0) return 0.0f;
  if (pm1(ia + ib + ic) < 0) return 0.0f;
  int k1 = ia + ib - ic, k2 = ic + ia - ib, k3 = ib + ic - ia;
  int k4 = ia - std::abs(ib - ic), k5 = ib - std::abs(ic - ia), k6 = ic - std::abs(ia - ib);
  if (std::min({k1, k2, k3, k4, k5, k6}) < 0) return 0.0f;
  if (pm1(ia + id) <= 0 || pm1(ib + ie) <= 0 || pm1(ic + iff) <= 0) return 0.0f;
  if (ia < std::abs(id) || ib < std::abs(ie) || ic < std::abs(iff)) return 0.0f;
  const int iam = ia, ibm = ib, icm = ic, idm = id;
  float signfc = 1.0f;
  bool reorder_c = true;
  if (ia < ib) {
    if (ia < ic) {
      std::swap(ia, ib);
      std::swap(id, ie);
      signfc = (float)pm1((ia + ib - ic) / 2);
      reorder_c = false;
    }
  } else if (ic >= ib) {
    reorder_c = false;
  }
  if (reorder_c) {
    std::swap(ic, ib);
    int t = iff;
    iff = -ie;
    ie = -t;
    float fibm = (float)(ibm + 1), ficm = (float)(icm + 1);
    signfc = (float)pm1((iam - idm) / 2) * std::sqrt(ficm / fibm);
  }
  if (ib == 0) return signfc;
  if (ie > 0) {
    signfc = signfc * (float)pm1((ia + ib - ic) / 2);
    id = -id;
    ie = -ie;
    iff = -iff;
  }
  float fc2 = (float)(ic + 1);
  int iabcp = (ia + ib + ic) / 2 + 1;
  int iabc = iabcp - ic, icab = iabcp - ib, ibca = iabcp - ia;
  int iapd = (ia + id) / 2 + 1, iamd = iapd - id;
  int ibpe = (ib + ie) / 2 + 1, ibme = ibpe - ie;
  int icpf = (ic + iff) / 2 + 1, icmf = icpf - iff;
  double s = (double)std::log(fc2) - fl[iabcp + 1];
  s = s + fl[iabc];
  s = s + fl[icab];
  s = s + fl[ibca];
  s = s + fl[iapd];
  s = s + fl[iamd];
  s = s + fl[ibpe];
  s = s + fl[ibme];
  s = s + fl[icpf];
  s = s + fl[icmf];
  float sqfclg = (float)(0.5 * s);
  int nzmic2 = (ib - ic - id) / 2, nzmic3 = (ia - ic + ie) / 2;
  int nzmi = std::max({0, nzmic2, nzmic3}) + 1;
  int nzmx = std::min({iabc, iamd, ibpe});
  if (nzmi > nzmx) return 0.0f;
  float ss = 0.0f;
  float s1 = (float)pm1(nzmi - 1);
  for (int nz = nzmi; nz <= nzmx; nz++) {
    int nzm1 = nz - 1;
    double tl = (double)sqfclg - fl[nz];
    tl = tl - fl[iabc - nzm1];
    tl = tl - fl[iamd - nzm1];
    tl = tl - fl[ibpe - nzm1];
    tl = tl - fl[nz - nzmic2];
    tl = tl - fl[nz - nzmic3];
    float termlg = (float)tl;
    ss = ss + s1 * std::exp(termlg);
    s1 = -s1;
  }
  return signfc * ss;
}

void read_sps(const std::string& path, ModelInput& in) {
  std::ifstream f(path);
  if (!f) fail("cannot open .sps file " + path);
  auto t = data_tokens(f);
  if (t.size() < 2 || t[0] != "iso") fail(".sps: only the 'iso' format is supported: " + path);
  int norb = std::atoi(t[1].c_str());
  if (norb < 1 || norb > kMaxOrb) fail(".sps: bad number of orbits");
  if ((int)t.size() < 2 + 4 * norb) fail(".sps: truncated orbit list");
  in.orbits.clear();
  for (int i = 0; i < norb; i++) {
    Orbit o;
    o.n = (int)std::lround(std::atof(t[2 + 4 * i].c_str()));
    o.l = (int)std::lround(std::atof(t[3 + 4 * i].c_str()));
    o.j2 = (int)std::lround(2.0 * std::atof(t[4 + 4 * i].c_str()));
    o.w = (int)std::lround(std::atof(t[5 + 4 * i].c_str()));
    if (o.j2 % 2 == 0 || o.j2 < 1) fail(".sps: j must be half-integer");
    in.orbits.push_back(o);
  }
}

namespace {

struct Builder {
  Tables& t;
  int norb;
  explicit Builder(Tables& tt) : t(tt), norb((int)tt.in.orbits.size()) {}

  const Orbit& orb(int i1) const { return t.in.orbits[i1 - 1]; }  // 1-based
  static int xxpair(int a, int b) { return a * (a - 1) / 2 + b; }
  int pnpair(int a, int b) const { return norb * (a - 1) + b; }
  size_t key_xx(int p1, int p2) const {
    int np = norb * (norb + 1) / 2 + 1;
    if (p1 < p2) std::swap(p1, p2);
    return ((size_t)p1 * np + p2) * 32;
  }
  size_t key_pn(int p1, int p2) const {
    int np = norb * norb + 1;
    if (p1 < p2) std::swap(p1, p2);
    return ((size_t)p1 * np + p2) * 32;
  }

  // src/binput_tbme.f90:898-1990 (iso format, no autoscaling)
  void read_int(const std::string& path) {
    std::ifstream f(path);
    if (!f) fail("cannot open .int file " + path);
    read_int_stream(f);
  }
  void read_int_stream(std::istream& f) {
    // one record: the count and the SPEs (src/binput_tbme.f90:1086); one per
    // element: a b c d J T V (:1171); values past those on a line are ignored
    RecordReader rr(f);
    std::vector<std::string> head, rec;
    if (rr.lines.empty()) fail(".int: empty file");
    const long nme0 = std::strtol(rr.lines[0][0].c_str(), nullptr, 10);
    // a negative count marks an autoscaled file (src/binput_tbme.f90:1005-1011,
    // 1075-1080, 1119-1140): three more values on the SPE record (core A, A0,
    // x), SPEs unscaled, TBMEs scaled by (A0 / (core + Z + N))^x, or by 1
    // when A0 or x is 0; the scaling line of the menu is not used
    const bool autoscale = nme0 < 0;
    if (!rr.take(1 + norb + (autoscale ? 3 : 0), head)) fail(".int: truncated SPE line");
    const long nme = autoscale ? -nme0 : nme0;
    if (autoscale) {
      const float core = std::strtof(head[1 + norb].c_str(), nullptr);
      const float ax = std::strtof(head[2 + norb].c_str(), nullptr);
      const float x = std::strtof(head[3 + norb].c_str(), nullptr);
      t.in.spscale = 1.0f;
      if (ax == 0.0f || x == 0.0f) {
        t.in.vscale = 1.0f;
      } else {
        const float bx = core + (float)t.in.Z + (float)t.in.N;
        t.in.vscale = std::pow(ax / bx, x);
      }
      head.resize(1 + norb);
    }
    std::vector<std::string> tok(head);
    for (long n = 0; n < nme; n++) {
      if (!rr.take(7, rec)) fail(".int: truncated");
      tok.insert(tok.end(), rec.begin(), rec.end());
    }
    const float spscale = t.in.spscale, vscale = t.in.vscale;
    t.pspe.assign(norb, 0.0f);
    t.nspe.assign(norb, 0.0f);
    t.spe_raw.assign(norb, 0.0f);
    t.lines.clear();
    for (int i = 0; i < norb; i++) {
      float e = std::strtof(tok[1 + i].c_str(), nullptr);
      t.spe_raw[i] = e;
      t.pspe[i] = t.pspe[i] + spscale * e;
      t.nspe[i] = t.nspe[i] + spscale * e;
    }
    size_t nxx = (size_t)(norb * (norb + 1) / 2 + 1), npn = (size_t)(norb * norb + 1);
    t.cpp.assign(nxx * nxx * 32, 0.0f);
    t.cnn.assign(nxx * nxx * 32, 0.0f);
    t.cpn.assign(npn * npn * 32, 0.0f);
    const int np1 = t.in.Z, np2 = t.in.N;
    size_t p = 1 + norb;
    for (long n = 0; n < nme; n++, p += 7) {
      int ia = std::atoi(tok[p].c_str()), ib = std::atoi(tok[p + 1].c_str());
      int ic = std::atoi(tok[p + 2].c_str()), id = std::atoi(tok[p + 3].c_str());
      int J = std::atoi(tok[p + 4].c_str()), T = std::atoi(tok[p + 5].c_str());
      float vv = std::strtof(tok[p + 6].c_str(), nullptr);
      t.lines.push_back({ia, ib, ic, id, J, T, vv});
      if (T != 0 && T != 1) fail(".int: bad T value");
      if (std::min({ia, ib, ic, id}) < 1 || std::max({ia, ib, ic, id}) > norb) fail(".int: orbit label out of range");
      if ((ia == ib || ic == id) && (J + T) % 2 == 0 && vv != 0.0f) fail(".int: illegal matrix element for identical orbits");
      if ((orb(ia).l + orb(ib).l + orb(ic).l + orb(id).l) % 2) fail(".int: parity violated");
      for (int it = 0; it < 2 && T == 1; it++) {
        if ((it == 0 ? np1 : np2) < 2) continue;
        int a = ia, b = ib, c = ic, d = id, phase = 1;
        if (a < b) { std::swap(a, b); phase = pm1(J + T + (orb(a).j2 + orb(b).j2) / 2); }
        if (c < d) { std::swap(c, d); phase = phase * pm1(J + T + (orb(c).j2 + orb(d).j2) / 2); }
        if (a < c || (a == c && b < d)) { std::swap(a, c); std::swap(b, d); }
        size_t k = key_xx(xxpair(a, b), xxpair(c, d));
        std::vector<float>& v = it == 0 ? t.cpp : t.cnn;
        v[k + J] = v[k + J] + ((vv * vscale) * (float)phase) * 1.0f;
      }
      if (np1 * np2 == 0) continue;
      float factor = 0.5f;
      if (ia == ib) factor = factor * std::sqrt(2.0f);
      if (ic == id) factor = factor * std::sqrt(2.0f);
      for (int part = 1; part <= 4; part++) {
        int a, b, c, d, phase = 1;
        if (part == 1) { a = ia; b = ib; c = ic; d = id; }
        else if (part == 2) { b = ia; a = ib; c = ic; d = id; if (a == b) continue; }
        else if (part == 3) { a = ia; b = ib; d = ic; c = id; if (c == d || (c == b && d == a)) continue; }
        else { b = ia; a = ib; d = ic; c = id; if (a == b || c == d) continue; }
        if (a < c || (a == c && b < d)) { std::swap(a, c); std::swap(b, d); }
        if (part == 2) phase = pm1(J + T + (orb(ia).j2 + orb(ib).j2) / 2);
        if (part == 3) phase = pm1(J + T + (orb(ic).j2 + orb(id).j2) / 2);
        if (part == 4) phase = pm1((orb(a).j2 + orb(b).j2 + orb(c).j2 + orb(d).j2) / 2);
        size_t k = key_pn(pnpair(a, b), pnpair(c, d));
        t.cpn[k + J] = t.cpn[k + J] + (((vv * vscale) * (float)phase) * factor) * 1.0f;
      }
    }
  }

  static float zeta(int i, int j) { return i == j ? std::sqrt(2.0f) : 1.0f; }

  // src/bsppot.f90:20-240
  void subsume() {
    const int np1 = t.in.Z, np2 = t.in.N;
    t.call_spe = true;
    if (np1 + np2 <= 1) return;
    t.call_spe = false;
    const float afactor = 1.0f / (float)(np1 + np2 - 1);
    std::vector<float> psp((norb + 1) * (norb + 1), 0.0f), nsp((norb + 1) * (norb + 1), 0.0f);
    auto P = [&](int a, int b) -> float& { return psp[a * (norb + 1) + b]; };
    auto Nn = [&](int a, int b) -> float& { return nsp[a * (norb + 1) + b]; };
    if (np2 > 0)
      for (int a = 1; a <= norb; a++) Nn(a, a) = Nn(a, a) + t.nspe[a - 1];
    for (int pass = 0; pass < 2; pass++) {
      const int nump = pass == 0 ? np2 : np1;
      if (pass == 1) {
        if (np1 <= 0) break;
        for (int a = 1; a <= norb; a++) P(a, a) = P(a, a) + t.pspe[a - 1];
      }
      auto SP = [&](int a, int b) -> float { return pass == 0 ? Nn(a, b) : P(a, b); };
      std::vector<float>& cv = pass == 0 ? t.cnn : t.cpp;
      for (int a = 1; a <= norb; a++) {
        int ja = (orb(a).j2 + 1) / 2;
        for (int c = 1; c <= a; c++) {
          int jc = (orb(c).j2 + 1) / 2;
          if (nump > 1) {
            for (int b = 1; b <= a; b++) {
              int jb = (orb(b).j2 + 1) / 2;
              int dfin = (a == c) ? b : c;
              for (int d = 1; d <= dfin; d++) {
                int jd = (orb(d).j2 + 1) / 2;
                if (b != d && a != c && b != c && a != d) continue;
                if ((orb(a).l + orb(b).l + orb(c).l + orb(d).l) % 2) continue;
                size_t k = key_xx(xxpair(a, b), xxpair(c, d));
                int jjmin = std::max(std::abs(orb(a).j2 - orb(b).j2), std::abs(orb(c).j2 - orb(d).j2)) / 2;
                int jjmax = std::min(orb(a).j2 + orb(b).j2, orb(c).j2 + orb(d).j2) / 2;
                for (int jj = jjmin; jj <= jjmax; jj++) {
                  int phab = pm1((ja - jb) + jj), phcd = pm1((jc - jd) + jj);
                  float vvv = SP(a, c) * (float)(b == d);
                  vvv = vvv + (SP(a, d) * (float)(b == c)) * (float)phcd;
                  vvv = vvv + (SP(b, c) * (float)(a == d)) * (float)phab;
                  vvv = vvv + SP(b, d) * (float)(a == c);
                  if (pass == 1 && vvv == 0.0f) continue;
                  vvv = ((afactor / zeta(a, b)) / zeta(d, c)) * vvv;
                  cv[k + jj] = cv[k + jj] + vvv;
                }
              }
            }
          }
          if (pass == 0 || np2 < 1) continue;
          for (int b = 1; b <= norb; b++) {
            int dfin = (a == c) ? b : norb;
            for (int d = 1; d <= dfin; d++) {
              if ((orb(a).l + orb(b).l + orb(c).l + orb(d).l) % 2) continue;
              size_t k = key_pn(pnpair(a, b), pnpair(c, d));
              float vvv = P(a, c) * (float)(b == d) + Nn(b, d) * (float)(a == c);
              vvv = afactor * vvv;
              int jjmin = std::max(std::abs(orb(a).j2 - orb(b).j2), std::abs(orb(c).j2 - orb(d).j2)) / 2;
              int jjmax = std::min(orb(a).j2 + orb(b).j2, orb(c).j2 + orb(d).j2) / 2;
              for (int jj = jjmin; jj <= jjmax; jj++) t.cpn[k + jj] = t.cpn[k + jj] + vvv;
            }
          }
        }
      }
    }
  }

  // src/bspstates.f90:47-330 for one W and one parity: states orbit by
  // orbit, m ascending; selection sort with swaps on |m|; then split into
  // left (m<0) and right (m>=0) halves keeping order.
  void make_sps() {
    std::vector<int> o, m;
    for (int i = 0; i < norb; i++)
      for (int mm = -t.in.orbits[i].j2; mm <= t.in.orbits[i].j2; mm += 2) { o.push_back(i); m.push_back(mm); }
    int n = (int)o.size();
    if (n > kMaxSps) fail("too many single-particle states");
    for (int is = 0; is < n; is++) {
      int k = is, mm = m[is], so = o[is], sm = m[is];
      for (int j = is + 1; j < n; j++)
        if (std::abs(m[j]) < std::abs(mm)) { k = j; mm = m[j]; so = o[j]; sm = m[j]; }
      if (k != is) { o[k] = o[is]; m[k] = m[is]; o[is] = so; m[is] = sm; }
    }
    t.ssorb = o;
    t.ssm2 = m;
    std::vector<int> lo, lm, ro, rm;
    for (int i = 0; i < n; i++) {
      if (m[i] >= 0) { ro.push_back(o[i]); rm.push_back(m[i]); }
      else { lo.push_back(o[i]); lm.push_back(m[i]); }
    }
    t.nL = (int)lo.size();
    t.nR = (int)ro.size();
    t.nsps = n;
    t.sorb.clear(); t.sm2.clear(); t.sj2.clear();
    for (int i = 0; i < t.nL; i++) { t.sorb.push_back(lo[i]); t.sm2.push_back(lm[i]); t.sj2.push_back(t.in.orbits[lo[i]].j2); }
    for (int i = 0; i < t.nR; i++) { t.sorb.push_back(ro[i]); t.sm2.push_back(rm[i]); t.sj2.push_back(t.in.orbits[ro[i]].j2); }
  }

  struct Haiku { uint64_t bits; int jz; };

  // Haikus of nh particles in a half space (src/bbasis_haikus.f90:327-411):
  // generation order == lexicographic descending order of the ascending
  // index tuples.
  static std::vector<Haiku> haikus(int nstates, const int* m2, int nh) {
    std::vector<Haiku> h;
    if (nh < 0 || nh > nstates) return h;
    if (nh == 0) { h.push_back({0, 0}); return h; }
    uint64_t s = (1ULL << nh) - 1, lim = 1ULL << nstates;
    while (s < lim) {
      int jz = 0;
      for (uint64_t q = s; q; q &= q - 1) jz += m2[__builtin_ctzll(q)];
      h.push_back({s, jz});
      uint64_t u = s & (~s + 1), v = s + u;
      s = v + (((v ^ s) / u) >> 2);
    }
    std::sort(h.begin(), h.end(), [](const Haiku& x, const Haiku& y) {
      uint64_t a = x.bits, b = y.bits;
      while (a && b) {
        int ia = __builtin_ctzll(a), ib = __builtin_ctzll(b);
        if (ia != ib) return ia > ib;
        a &= a - 1;
        b &= b - 1;
      }
      return false;
    });
    return h;
  }

  void build_basis() {
    const int nL = t.nL, nR = t.nR;
    std::vector<int> lm2(t.sm2.begin(), t.sm2.begin() + nL), rm2(t.sm2.begin() + nL, t.sm2.end());
    const int np[2] = {t.in.Z, t.in.N};
    std::vector<std::vector<Haiku>> HR(nR + 1), HL(nL + 1);
    for (int k = 0; k <= nR; k++) HR[k] = haikus(nR, rm2.data(), k);
    for (int k = 0; k <= nL; k++) HL[k] = haikus(nL, lm2.data(), k);
    int jzmin[2], jzmax[2];
    for (int it = 0; it < 2; it++) {
      int Np = np[it];
      int maxR = std::min(Np, nR), maxL = std::min(Np, nL);
      int mn = 1 << 20, mx = -(1 << 20);
      for (int nh = Np - maxL; nh <= maxR; nh++) {
        if (nh < 0) continue;
        for (auto& r : HR[nh])
          for (auto& l : HL[Np - nh]) { mn = std::min(mn, r.jz + l.jz); mx = std::max(mx, r.jz + l.jz); }
      }
      if (mn > mx) fail("no Slater determinants for a species");
      jzmin[it] = mn;
      jzmax[it] = mx;
    }
    const int jz = t.in.jz2;
    jzmax[0] = std::min(jzmax[0], jz - jzmin[1]);
    jzmax[1] = std::min(jzmax[1], jz - jzmin[0]);
    jzmin[0] = std::max(jzmin[0], jz - jzmax[1]);
    jzmin[1] = std::max(jzmin[1], jz - jzmax[0]);
    for (int it = 0; it < 2; it++) {
      Species& S = t.sp[it];
      S = Species();
      int Np = np[it];
      S.np = Np;
      int maxR = std::min(Np, nR), maxL = std::min(Np, nL);
      int start = it == 0 ? jzmin[0] : jzmax[1], stop = it == 0 ? jzmax[0] : jzmin[1];
      int step = it == 0 ? 2 : -2;
      for (int jzx = start; it == 0 ? jzx <= stop : jzx >= stop; jzx += step) {
        int64_t first = (int64_t)S.mask.size();
        for (int nh = Np - maxL; nh <= maxR; nh++) {
          if (nh < 0) continue;
          const auto& hr = HR[nh];
          const auto& hl = HL[Np - nh];
          int rjmin = 1 << 20, rjmax = -(1 << 20);
          for (auto& r : hr) { rjmin = std::min(rjmin, r.jz); rjmax = std::max(rjmax, r.jz); }
          for (int jzh = rjmin; jzh <= rjmax; jzh += 2) {
            int jzhc = jzx - jzh;
            for (auto& l : hl) {
              if (l.jz != jzhc) continue;
              for (auto& r : hr) {
                if (r.jz != jzh) continue;
                S.mask.push_back(l.bits | (r.bits << nL));
              }
            }
          }
        }
        int64_t cnt = (int64_t)S.mask.size() - first;
        if (cnt == 0) continue;
        S.secjz.push_back(jzx);
        S.secfirst.push_back(first);
        S.seccount.push_back(cnt);
        for (int64_t q = 0; q < cnt; q++) S.sdsec.push_back(S.nsec);
        S.nsec++;
      }
      S.start.assign(S.mask.size(), -1);
    }
    for (int it = 0; it < 2; it++) {
      Species& S = t.sp[it];
      const Species& C = t.sp[1 - it];
      S.conj.assign(S.nsec, -1);
      for (int s = 0; s < S.nsec; s++)
        for (int c = 0; c < C.nsec; c++)
          if (C.secjz[c] == jz - S.secjz[s]) S.conj[s] = c;
      for (int s = 0; s < S.nsec; s++)
        if (S.conj[s] < 0) fail("sector without conjugate sector");
    }
    int64_t dim = 0, ncount = 0, pcount = 0;
    Species& P = t.sp[0];
    Species& Nn = t.sp[1];
    for (int s = 0; s < P.nsec; s++) {
      int cs = P.conj[s];
      int64_t ndim = Nn.seccount[cs];
      if (Nn.conj[cs] == s)
        for (int64_t in = 0; in < ndim; in++) Nn.start[ncount++] = in;
      for (int64_t ip = 0; ip < P.seccount[s]; ip++) {
        P.start[pcount++] = dim;
        dim += ndim;
      }
    }
    t.dim = dim;
  }

  struct PairQn { int m, par, indx, ia, ib; };

  // src/btbme.f90:548-703 (heapsort on M, then on parity within equal M)
  static void sortpair(std::vector<PairQn>& v) {
    int npair = (int)v.size();
    PairQn* pair = v.data() - 1;
    PairQn tmp;
    int i, j, l = npair / 2 + 1, n = npair;
    while (n > 0) {
      if (l > 1) { l = l - 1; tmp = pair[l]; }
      else {
        tmp = pair[n];
        pair[n] = pair[1];
        n = n - 1;
        if (n == 1) { pair[1] = tmp; break; }
      }
      i = l;
      j = l + l;
      while (j <= n) {
        if (j < n && pair[j].m > pair[j + 1].m) j = j + 1;
        if (tmp.m > pair[j].m) { pair[i] = pair[j]; i = j; j = j + j; }
        else j = n + 1;
      }
      pair[i] = tmp;
    }
    int istart = 1, istop = 1;
    while (istart <= npair) {
      int m = pair[istart].m;
      if (istart == npair) istop = istart;
      else {
        while (pair[istop + 1].m == m) {
          istop = istop + 1;
          if (istop == npair) break;
        }
      }
      l = (istop - istart + 1) / 2 + istart;
      n = istop;
      while (n >= istart) {
        if (l > istart) { l = l - 1; tmp = pair[l]; }
        else {
          tmp = pair[n];
          pair[n] = pair[istart];
          n = n - 1;
          if (n == istart) { pair[istart] = tmp; break; }
        }
        i = l;
        j = l + l - (istart - 1);
        while (j <= n) {
          if (j < n && pair[j].par < pair[j + 1].par) j = j + 1;
          if (tmp.par < pair[j].par) { pair[i] = pair[j]; i = j; j = j + j - (istart - 1); }
          else j = n + 1;
        }
        pair[i] = tmp;
      }
      istart = istop + 1;
      istop = istart;
    }
  }

  // src/btbme.f90:1209-1558
  void uncouple_xx(int it) {
    const int ns = t.nsps, np = ns * (ns - 1) / 2;
    std::vector<float> v((size_t)np * np, 0.0f);
    std::vector<PairQn> pl;
    for (int a = 1; a < ns; a++)
      for (int b = 0; b < a; b++) pl.push_back({(t.sm2[a] + t.sm2[b]) / 2, 1, a * (a - 1) / 2 + b, a, b});
    sortpair(pl);
    const std::vector<float>& cv = it == 0 ? t.cpp : t.cnn;
    for (int dp = 0; dp < np; dp++) {
      const int m = pl[dp].m;
      int k = pl[dp].ia, l = pl[dp].ib;
      int ia = t.sorb[k] + 1, ib = t.sorb[l] + 1;
      int jk = t.sj2[k], mk = t.sm2[k], jl = t.sj2[l], ml = t.sm2[l];
      int phasekl = 1;
      if (ia < ib) { std::swap(ia, ib); std::swap(jk, jl); std::swap(mk, ml); phasekl = -1; }
      for (int cp = 0; cp <= dp; cp++) {
        if (pl[cp].m != m || pl[cp].par != pl[dp].par) continue;
        int i = pl[cp].ia, j = pl[cp].ib;
        int ic = t.sorb[i] + 1, id = t.sorb[j] + 1;
        int ji = t.sj2[i], mi = t.sm2[i], jj = t.sj2[j], mj = t.sm2[j];
        int phaseij = 1;
        if (ic < id) { std::swap(ic, id); std::swap(ji, jj); std::swap(mi, mj); phaseij = -1; }
        size_t key = key_xx(xxpair(ia, ib), xxpair(ic, id));
        int jmax = std::min(jk + jl, ji + jj) / 2;
        int jmin = std::max(std::abs(jk - jl), std::abs(ji - jj)) / 2;
        jmin = std::max(jmin, std::abs(m));
        float vtmp = 0.0f;
        for (int J = jmin; J <= jmax; J++) {
          float q = cv[key + J] * zeta(ia, ib);
          q = q * zeta(ic, id);
          q = q * cleb(jk, mk, jl, ml, 2 * J, 2 * m);
          q = q * cleb(ji, mi, jj, mj, 2 * J, 2 * m);
          vtmp = vtmp + q;
        }
        float val = (vtmp * (float)phasekl) * (float)phaseij;
        v[(size_t)pl[cp].indx * np + pl[dp].indx] = val;
        v[(size_t)pl[dp].indx * np + pl[cp].indx] = val;
      }
    }
    (it == 0 ? t.vpp : t.vnn).swap(v);
    t.npairxx = np;
  }

  // src/btbmepn.f90:177-594
  void uncouple_pn() {
    const int ns = t.nsps, np = ns * ns;
    std::vector<float> v((size_t)np * np, 0.0f);
    std::vector<PairQn> pl;
    for (int a = 0; a < ns; a++)
      for (int b = 0; b < ns; b++) pl.push_back({(t.sm2[a] + t.sm2[b]) / 2, 1, a * ns + b, a, b});
    sortpair(pl);
    int s0 = 0;
    while (s0 < np) {
      int s1 = s0;
      while (s1 + 1 < np && pl[s1 + 1].m == pl[s0].m && pl[s1 + 1].par == pl[s0].par) s1++;
      for (int dp = s0; dp <= s1; dp++) {
        int k = pl[dp].ia, l = pl[dp].ib;
        int ia = t.sorb[k] + 1, ib = t.sorb[l] + 1;
        int jk = t.sj2[k], mk = t.sm2[k], jl = t.sj2[l], ml = t.sm2[l];
        const int m = pl[dp].m;
        for (int cp = s0; cp <= dp; cp++) {
          int i = pl[cp].ia, j = pl[cp].ib;
          int ic = t.sorb[i] + 1, id = t.sorb[j] + 1;
          int ji = t.sj2[i], mi = t.sm2[i], jj = t.sj2[j], mj = t.sm2[j];
          size_t key = key_pn(pnpair(ia, ib), pnpair(ic, id));
          int jmax = std::min(jk + jl, ji + jj) / 2;
          int jmin = std::max(std::abs(jk - jl), std::abs(ji - jj)) / 2;
          jmin = std::max(jmin, std::abs(m));
          float vtmp = 0.0f;
          for (int J = jmin; J <= jmax; J++) {
            float q = t.cpn[key + J] * cleb(jk, mk, jl, ml, 2 * J, mk + ml);
            q = q * cleb(ji, mi, jj, mj, 2 * J, mi + mj);
            vtmp = vtmp + q;
          }
          v[(size_t)pl[cp].indx * np + pl[dp].indx] = vtmp;
          v[(size_t)pl[dp].indx * np + pl[cp].indx] = vtmp;
        }
      }
      s0 = s1 + 1;
    }
    t.vpn.swap(v);
    t.npairpn = np;
  }

  static int fsign(uint64_t mask, int c) {
    uint64_t below = c == 0 ? 0 : (mask & ((1ULL << c) - 1));
    return (__builtin_popcountll(below) & 1) ? -1 : 1;
  }

  // Delta classes of one-body operators a+_a a_c (Delta = m_a - m_c) and the
  // PN element blocks W[class](pp_local, nn_local) = <a b|V|c d>, where the
  // neutron operator (b,d) lies in the class of -Delta.
  void build_classes() {
    const int ns = t.nsps;
    std::vector<int> deltas;
    for (int a = 0; a < ns; a++)
      for (int c = 0; c < ns; c++) deltas.push_back(t.sm2[a] - t.sm2[c]);
    std::sort(deltas.begin(), deltas.end());
    deltas.erase(std::unique(deltas.begin(), deltas.end()), deltas.end());
    t.ndelta = (int)deltas.size();
    t.delta_of_class = deltas;
    t.class_of_op.assign(ns * ns, -1);
    t.local_of_op.assign(ns * ns, -1);
    t.class_size.assign(t.ndelta, 0);
    for (int a = 0; a < ns; a++)
      for (int c = 0; c < ns; c++) {
        int d = t.sm2[a] - t.sm2[c];
        int k = (int)(std::lower_bound(deltas.begin(), deltas.end(), d) - deltas.begin());
        t.class_of_op[a * ns + c] = k;
        t.local_of_op[a * ns + c] = t.class_size[k]++;
      }
    t.max_class = *std::max_element(t.class_size.begin(), t.class_size.end());
    t.class_offset.assign(t.ndelta + 1, 0);
    for (int k = 0; k < t.ndelta; k++) {
      int kn = (int)(std::lower_bound(deltas.begin(), deltas.end(), -deltas[k]) - deltas.begin());
      t.class_offset[k + 1] = t.class_offset[k] + t.class_size[k] * t.class_size[kn];
    }
    t.wpn.assign(t.class_offset[t.ndelta], 0.0f);
    const int npn = ns * ns;
    for (int a = 0; a < ns; a++)
      for (int c = 0; c < ns; c++) {
        int k = t.class_of_op[a * ns + c], pl = t.local_of_op[a * ns + c];
        int kn = (int)(std::lower_bound(deltas.begin(), deltas.end(), -deltas[k]) - deltas.begin());
        for (int b = 0; b < ns; b++)
          for (int d = 0; d < ns; d++) {
            if (t.class_of_op[b * ns + d] != kn) continue;
            int nl = t.local_of_op[b * ns + d];
            float w = t.vpn[(size_t)(a * ns + b) * npn + (c * ns + d)];
            t.wpn[t.class_offset[k] + pl * t.class_size[kn] + nl] = w;
          }
      }
  }

  // Runs body(f0, f1, chunk) over contiguous chunks of [0, n) on up to 16
  // host threads; chunk c covers the c-th range, so results gathered per
  // chunk and concatenated in chunk order equal a serial loop's.
  template <typename F>
  static int parallel_chunks(int64_t n, F body) {
    int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (int c = 1; c < nt; c++) th.emplace_back(body, n * c / nt, n * (c + 1) / nt, c);
    body(0, n / nt, 0);
    for (auto& x : th) x.join();
    return nt;
  }

  void build_jumps(int it) {
    Species& S = t.sp[it];
    const int ns = t.nsps;
    std::unordered_map<uint64_t, int32_t> idx;
    idx.reserve(S.mask.size() * 2);
    for (size_t i = 0; i < S.mask.size(); i++) idx[S.mask[i]] = (int32_t)i;
    const int64_t nsd = (int64_t)S.mask.size();
    // one-body: rows of final SD f, entries sorted by (class, source, op)
    OneBody& O = t.ob[it];
    O = OneBody();
    O.rowptr.assign(nsd + 1, 0);
    struct E1 { int cls; int32_t src; int16_t op; int8_t ph; };
    struct Chunk1 { std::vector<int32_t> src; std::vector<int16_t> op; std::vector<int8_t> ph; };
    std::vector<Chunk1> c1(16);
    const int nt1 = parallel_chunks(nsd, [&](int64_t f0, int64_t f1, int c) {
      std::vector<E1> row;
      Chunk1& out = c1[c];
      for (int64_t f = f0; f < f1; f++) {
        row.clear();
        uint64_t fm = S.mask[f];
        for (uint64_t ta = fm; ta; ta &= ta - 1) {
          int a = __builtin_ctzll(ta);
          uint64_t m1 = fm & ~(1ULL << a);
          for (int cc = 0; cc < ns; cc++) {
            if (cc != a && ((m1 >> cc) & 1)) continue;
            uint64_t im = m1 | (1ULL << cc);
            auto itr = idx.find(im);
            if (itr == idx.end()) continue;  // outside the basis (cannot happen for full spaces)
            int ph = fsign(im, cc) * fsign(im & ~(1ULL << cc), a);
            row.push_back({t.class_of_op[a * ns + cc], itr->second, (int16_t)(a * ns + cc), (int8_t)ph});
          }
        }
        std::sort(row.begin(), row.end(), [](const E1& x, const E1& y) {
          return x.cls != y.cls ? x.cls < y.cls : (x.src != y.src ? x.src < y.src : x.op < y.op);
        });
        for (auto& e : row) { out.src.push_back(e.src); out.op.push_back(e.op); out.ph.push_back(e.ph); }
        O.rowptr[f + 1] = (int64_t)row.size();
      }
    });
    for (int64_t f = 0; f < nsd; f++) O.rowptr[f + 1] += O.rowptr[f];
    O.src.reserve(O.rowptr[nsd]);
    for (int c = 0; c < nt1; c++) {
      O.src.insert(O.src.end(), c1[c].src.begin(), c1[c].src.end());
      O.op.insert(O.op.end(), c1[c].op.begin(), c1[c].op.end());
      O.phase.insert(O.phase.end(), c1[c].ph.begin(), c1[c].ph.end());
    }
    // two-body, merged per source
    TwoBody& T = t.tb[it];
    T = TwoBody();
    T.rowptr.assign(nsd + 1, 0);
    t.tb_unmerged[it].assign(nsd, 0);
    if (S.np >= 2) {
      const std::vector<float>& v = it == 0 ? t.vpp : t.vnn;
      const int np = t.npairxx;
      struct Chunk2 { std::vector<int32_t> src; std::vector<double> val; };
      std::vector<Chunk2> c2(16);
      const int nt2 = parallel_chunks(nsd, [&](int64_t f0, int64_t f1, int c) {
        std::vector<std::pair<int32_t, double>> acc;
        Chunk2& out = c2[c];
        for (int64_t f = f0; f < f1; f++) {
          acc.clear();
          const size_t before = out.src.size();
          uint64_t fm = S.mask[f];
          for (uint64_t ti = fm; ti; ti &= ti - 1) {
            int i = __builtin_ctzll(ti);
            for (uint64_t tj = fm & ((1ULL << i) - 1); tj; tj &= tj - 1) {
              int j = __builtin_ctzll(tj);
              uint64_t m2 = fm & ~(1ULL << i) & ~(1ULL << j);
              int mij = t.sm2[i] + t.sm2[j];
              int cidx = i * (i - 1) / 2 + j;
              for (int k = 1; k < ns; k++) {
                if ((m2 >> k) & 1) continue;
                for (int l = 0; l < k; l++) {
                  if ((m2 >> l) & 1) continue;
                  if (t.sm2[k] + t.sm2[l] != mij) continue;
                  float me = v[(size_t)cidx * np + (k * (k - 1) / 2 + l)];
                  uint64_t sm = m2 | (1ULL << k) | (1ULL << l);
                  // the reference stores and counts jumps whatever their element value
                  t.tb_unmerged[it][f] += (sm == fm) ? 2 : 1;
                  if (me == 0.0f && !t.keep_structural) continue;
                  // sign of a+_i a+_j a_l a_k acting on |s>
                  int sg = fsign(sm, k);
                  uint64_t s1 = sm & ~(1ULL << k);
                  sg *= fsign(s1, l);
                  uint64_t s2 = s1 & ~(1ULL << l);
                  sg *= fsign(s2, j);
                  uint64_t s3 = s2 | (1ULL << j);
                  sg *= fsign(s3, i);
                  auto itr = idx.find(sm);
                  if (itr == idx.end()) continue;
                  acc.push_back({itr->second, (double)me * (double)sg});
                }
              }
            }
          }
          std::sort(acc.begin(), acc.end(), [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) {
            return x.first < y.first;
          });
          for (size_t q = 0; q < acc.size();) {
            size_t r = q;
            double sum = 0.0;
            while (r < acc.size() && acc[r].first == acc[q].first) { sum += acc[r].second; r++; }
            if (sum != 0.0 || t.keep_structural) { out.src.push_back(acc[q].first); out.val.push_back(sum); }
            q = r;
          }
          T.rowptr[f + 1] = (int64_t)(out.src.size() - before);
        }
      });
      for (int64_t f = 0; f < nsd; f++) T.rowptr[f + 1] += T.rowptr[f];
      T.src.reserve(T.rowptr[nsd]);
      T.val.reserve(T.rowptr[nsd]);
      for (int c = 0; c < nt2; c++) {
        T.src.insert(T.src.end(), c2[c].src.begin(), c2[c].src.end());
        T.val.insert(T.val.end(), c2[c].val.begin(), c2[c].val.end());
      }
    }
  }

  void count_ops() {
    const Species& P = t.sp[0];
    const Species& Nn = t.sp[1];
    t.nops_pp = t.nops_nn = t.nops_pn = t.nops_spe = 0;
    if (t.call_spe) t.nops_spe = t.dim;
    // PP: merged entries into each proton SD x its row length
    for (int64_t f = 0; f < (int64_t)P.mask.size(); f++) {
      int64_t L = Nn.seccount[P.conj[P.sdsec[f]]];
      t.nops_pp += (t.tb[0].rowptr[f + 1] - t.tb[0].rowptr[f]) * L;
    }
    for (int64_t f = 0; f < (int64_t)P.mask.size(); f++)
      t.ref_pp += t.tb_unmerged[0][f] * Nn.seccount[P.conj[P.sdsec[f]]];
    for (int s = 0; s < P.nsec; s++) {
      int c = P.conj[s];
      int64_t u = 0;
      for (int64_t n = Nn.secfirst[c]; n < Nn.secfirst[c] + Nn.seccount[c]; n++) u += t.tb_unmerged[1][n];
      t.ref_nn += u * P.seccount[s];
    }
    for (int s = 0; s < P.nsec; s++) {
      int c = P.conj[s];
      int64_t e = 0;
      for (int64_t n = Nn.secfirst[c]; n < Nn.secfirst[c] + Nn.seccount[c]; n++) e += t.tb[1].rowptr[n + 1] - t.tb[1].rowptr[n];
      t.nops_nn += e * P.seccount[s];
    }
    // PN: for each proton jump i->f (class k), neutron jumps from conj(i) to conj(f) in class -k
    const int ns = t.nsps;
    std::vector<int> negk(t.ndelta);
    for (int k = 0; k < t.ndelta; k++)
      negk[k] = (int)(std::lower_bound(t.delta_of_class.begin(), t.delta_of_class.end(), -t.delta_of_class[k]) - t.delta_of_class.begin());
    // neutron class counts per target sector: cnt[c_target][class]
    std::vector<int64_t> ncnt((size_t)Nn.nsec * t.ndelta, 0);
    for (int64_t f = 0; f < (int64_t)Nn.mask.size(); f++)
      for (int64_t q = t.ob[1].rowptr[f]; q < t.ob[1].rowptr[f + 1]; q++)
        ncnt[(size_t)Nn.sdsec[f] * t.ndelta + t.class_of_op[t.ob[1].op[q]]]++;
    for (int64_t f = 0; f < (int64_t)P.mask.size(); f++) {
      int cf = P.conj[P.sdsec[f]];
      for (int64_t q = t.ob[0].rowptr[f]; q < t.ob[0].rowptr[f + 1]; q++) {
        int k = t.class_of_op[t.ob[0].op[q]];
        t.nops_pn += ncnt[(size_t)cf * t.ndelta + negk[k]];
      }
    }
    (void)ns;
  }
};

}  // namespace

namespace {
Tables build_any(const std::vector<Orbit>& orbits, const std::string* intfile, const std::string* inttext, int Z,
                 int N, int jz2, float spscale, float a0, float a, float x, bool with_jumps, bool structural = false);
}

Tables build_tables_structural(const std::vector<Orbit>& orbits, const std::string& intfile, int Z, int N,
                               int jz2, float spscale, float a0, float a, float x) {
  return build_any(orbits, &intfile, nullptr, Z, N, jz2, spscale, a0, a, x, true, true);
}

Tables build_tables(const std::vector<Orbit>& orbits, const std::string& intfile, int Z, int N, int jz2,
                    float spscale, float a0, float a, float x, bool with_jumps) {
  return build_any(orbits, &intfile, nullptr, Z, N, jz2, spscale, a0, a, x, with_jumps);
}

Tables build_tables_text(const std::vector<Orbit>& orbits, const std::string& inttext, int Z, int N, int jz2,
                         bool with_jumps) {
  return build_any(orbits, nullptr, &inttext, Z, N, jz2, 1.0f, 1.0f, 1.0f, 1.0f, with_jumps);
}

// J^2 = sum_i j_i^2 + sum_{i<j} (J_ij^2 - j_i^2 - j_j^2) and the same for T^2,
// written as an iso interaction on the orbits of `orbits`.
std::string observable_int_text(const std::vector<Orbit>& orbits, int which) {
  const int no = (int)orbits.size();
  std::ostringstream os;
  std::vector<std::string> lines;
  char buf[160];
  for (int a = 1; a <= no; a++)
    for (int b = 1; b <= a; b++) {
      const double ja = orbits[a - 1].j2 / 2.0, jb = orbits[b - 1].j2 / 2.0;
      const int jmin = std::abs(orbits[a - 1].j2 - orbits[b - 1].j2) / 2;
      const int jmax = (orbits[a - 1].j2 + orbits[b - 1].j2) / 2;
      for (int J = jmin; J <= jmax; J++)
        for (int T = 0; T <= 1; T++) {
          if (a == b && (J + T) % 2 == 0) continue;
          const double v = which == 0 ? J * (J + 1.0) - ja * (ja + 1.0) - jb * (jb + 1.0) : T * (T + 1.0) - 1.5;
          std::snprintf(buf, sizeof buf, "%d %d %d %d %d %d %.9g", a, b, a, b, J, T, v);
          lines.push_back(buf);
        }
    }
  os << lines.size();
  for (int a = 0; a < no; a++) {
    const double ja = orbits[a].j2 / 2.0;
    std::snprintf(buf, sizeof buf, " %.9g", which == 0 ? ja * (ja + 1.0) : 0.75);
    os << buf;
  }
  os << "\n";
  for (auto& l : lines) os << l << "\n";
  return os.str();
}

namespace {
Tables build_any(const std::vector<Orbit>& orbits, const std::string* intfile, const std::string* inttext, int Z,
                 int N, int jz2, float spscale, float a0, float a, float x, bool with_jumps, bool structural) {
  Tables t;
  t.keep_structural = structural;
  if (orbits.empty() || (int)orbits.size() > kMaxOrb) fail("bad number of orbits");
  for (const Orbit& o : orbits)
    if (o.j2 < 1 || o.j2 % 2 == 0 || o.l < 0 || o.n < 0) fail("bad orbit quantum numbers");
  t.in.orbits = orbits;
  for (auto& o : t.in.orbits) {
    if (o.w != t.in.orbits[0].w) fail("W truncation / mixed-W spaces are not supported");
    if (o.l % 2 != t.in.orbits[0].l % 2) fail("mixed-parity spaces are not supported");
  }
  if (Z < 0 || N < 0 || Z + N < 1) fail("bad particle numbers");
  if (((Z + N) % 2 == 0) != (jz2 % 2 == 0)) fail("2Jz parity does not match the particle number");
  t.in.Z = Z;
  t.in.N = N;
  t.in.jz2 = jz2;
  t.in.spscale = spscale;
  // src/binput_tbme.f90:1024-1031: TBMEs scaled by (A0/A)^x, or by A0 when A or x is 0
  t.in.vscale = (a == 0.0f || x == 0.0f) ? a0 : std::pow(a0 / a, x);
  Builder b(t);
  if (intfile) {
    b.read_int(*intfile);
  } else {
    std::istringstream is(*inttext);
    b.read_int_stream(is);
  }
  b.subsume();
  b.make_sps();
  int nmax = 0;
  for (auto& o : t.in.orbits) nmax += o.j2 + 1;
  if (Z > nmax || N > nmax) fail("more particles than single-particle states");
  b.build_basis();
  b.uncouple_xx(0);
  b.uncouple_xx(1);
  b.uncouple_pn();
  b.build_classes();
  if (with_jumps) {
    b.build_jumps(0);
    b.build_jumps(1);
    b.count_ops();
  }
  return t;
}

}  // namespace

MeTables build_me_tables(const Tables& base, const std::string& intfile) {
  Tables t;
  t.in = base.in;
  Builder b(t);
  b.read_int(intfile);
  if ((int)t.pspe.size() != (int)base.in.orbits.size()) fail("interaction has a different number of orbits");
  b.subsume();
  b.make_sps();
  b.uncouple_xx(0);
  b.uncouple_xx(1);
  b.uncouple_pn();
  MeTables m;
  m.vpp.swap(t.vpp);
  m.vnn.swap(t.vnn);
  m.vpn.swap(t.vpn);
  m.pspe = t.pspe;
  m.nspe = t.nspe;
  m.call_spe = t.call_spe;
  m.same_lines = t.lines.size() == base.lines.size();
  for (size_t i = 0; m.same_lines && i < t.lines.size(); i++) {
    const TbmeLine &u = t.lines[i], &w = base.lines[i];
    m.same_lines = u.a == w.a && u.b == w.b && u.c == w.c && u.d == w.d && u.J == w.J && u.T == w.T;
  }
  // pn elements in base's class layout
  const int ns = base.nsps, npn = ns * ns;
  m.wpn.assign(base.wpn.size(), 0.0f);
  for (int a = 0; a < ns; a++)
    for (int c = 0; c < ns; c++) {
      const int k = base.class_of_op[a * ns + c], pl = base.local_of_op[a * ns + c];
      const int kn = (int)(std::lower_bound(base.delta_of_class.begin(), base.delta_of_class.end(),
                                            -base.delta_of_class[k]) - base.delta_of_class.begin());
      for (int bb = 0; bb < ns; bb++)
        for (int d = 0; d < ns; d++) {
          if (base.class_of_op[bb * ns + d] != kn) continue;
          const int nl = base.local_of_op[bb * ns + d];
          m.wpn[base.class_offset[k] + pl * base.class_size[kn] + nl] = m.vpn[(size_t)(a * ns + bb) * npn + (c * ns + d)];
        }
    }
  return m;
}

std::vector<double> merged_values(const Tables& t, int it, const std::vector<float>& v) {
  if (!t.keep_structural) fail("merged_values needs structural tables");
  const Species& S = t.sp[it];
  const TwoBody& T = t.tb[it];
  std::vector<double> out(T.src.size(), 0.0);
  if (S.np < 2) return out;
  std::unordered_map<uint64_t, int32_t> idx;
  for (size_t i = 0; i < S.mask.size(); i++) idx[S.mask[i]] = (int32_t)i;
  const int ns = t.nsps, np = t.npairxx;
  for (int64_t f = 0; f < (int64_t)S.mask.size(); f++) {
    const uint64_t fm = S.mask[f];
    const int64_t r0 = T.rowptr[f], r1 = T.rowptr[f + 1];
    for (uint64_t ti = fm; ti; ti &= ti - 1) {
      const int i = __builtin_ctzll(ti);
      for (uint64_t tj = fm & ((1ULL << i) - 1); tj; tj &= tj - 1) {
        const int j = __builtin_ctzll(tj);
        const uint64_t m2 = fm & ~(1ULL << i) & ~(1ULL << j);
        const int mij = t.sm2[i] + t.sm2[j];
        const int cidx = i * (i - 1) / 2 + j;
        for (int k = 1; k < ns; k++) {
          if ((m2 >> k) & 1) continue;
          for (int l = 0; l < k; l++) {
            if ((m2 >> l) & 1) continue;
            if (t.sm2[k] + t.sm2[l] != mij) continue;
            const uint64_t sm = m2 | (1ULL << k) | (1ULL << l);
            auto itr = idx.find(sm);
            if (itr == idx.end()) continue;
            int sg = Builder::fsign(sm, k);
            const uint64_t s1 = sm & ~(1ULL << k);
            sg *= Builder::fsign(s1, l);
            const uint64_t s2 = s1 & ~(1ULL << l);
            sg *= Builder::fsign(s2, j);
            sg *= Builder::fsign(s2 | (1ULL << j), i);
            const int32_t src = itr->second;
            const auto pos = std::lower_bound(T.src.begin() + r0, T.src.begin() + r1, src);
            if (pos == T.src.begin() + r1 || *pos != src) fail("structural two-body entry missing");
            out[pos - T.src.begin()] += (double)v[(size_t)cidx * np + (k * (k - 1) / 2 + l)] * (double)sg;
          }
        }
      }
    }
  }
  return out;
}

std::vector<ParamColumn> param_jacobian(const Tables& t0) {
  const int norb = (int)t0.in.orbits.size();
  const int np = norb + (int)t0.lines.size();
  std::vector<ParamColumn> cols(np);
  for (int p = 0; p < np; p++) {
    std::ostringstream os;
    os << t0.lines.size();
    for (int a = 0; a < norb; a++) os << (p == a ? " 1" : " 0");
    os << "\n";
    for (size_t q = 0; q < t0.lines.size(); q++) {
      const TbmeLine& L = t0.lines[q];
      os << L.a << " " << L.b << " " << L.c << " " << L.d << " " << L.J << " " << L.T << " "
         << ((int)q + norb == p ? "1" : "0") << "\n";
    }
    Tables t;
    t.in = t0.in;
    Builder b(t);
    std::istringstream is(os.str());
    b.read_int_stream(is);
    b.subsume();
    b.make_sps();
    b.uncouple_xx(0);
    b.uncouple_xx(1);
    b.uncouple_pn();
    const std::vector<float>* tab[3] = {&t.vpp, &t.vnn, &t.vpn};
    for (int k = 0; k < 3; k++)
      for (size_t i = 0; i < tab[k]->size(); i++)
        if ((*tab[k])[i] != 0.0f) {
          cols[p].idx[k].push_back((int32_t)i);
          cols[p].coef[k].push_back((*tab[k])[i]);
        }
  }
  return cols;
}

void build_twobody_ops(const Tables& t, int it, TwoBodyOps& out) {
  const Species& S = t.sp[it];
  const int ns = t.nsps;
  std::unordered_map<uint64_t, int32_t> idx;
  for (size_t i = 0; i < S.mask.size(); i++) idx[S.mask[i]] = (int32_t)i;
  const int64_t nsd = (int64_t)S.mask.size();
  out = TwoBodyOps();
  out.rowptr.assign(nsd + 1, 0);
  if (S.np >= 2) {
    for (int64_t f = 0; f < nsd; f++) {
      uint64_t fm = S.mask[f];
      for (uint64_t ti = fm; ti; ti &= ti - 1) {
        int i = __builtin_ctzll(ti);
        for (uint64_t tj = fm & ((1ULL << i) - 1); tj; tj &= tj - 1) {
          int j = __builtin_ctzll(tj);
          uint64_t m2 = fm & ~(1ULL << i) & ~(1ULL << j);
          int mij = t.sm2[i] + t.sm2[j];
          int cidx = i * (i - 1) / 2 + j;
          for (int k = 1; k < ns; k++) {
            if ((m2 >> k) & 1) continue;
            for (int l = 0; l < k; l++) {
              if ((m2 >> l) & 1) continue;
              if (t.sm2[k] + t.sm2[l] != mij) continue;
              uint64_t sm = m2 | (1ULL << k) | (1ULL << l);
              auto fs = [](uint64_t mask, int c) {
                uint64_t below = c == 0 ? 0 : (mask & ((1ULL << c) - 1));
                return (__builtin_popcountll(below) & 1) ? -1 : 1;
              };
              int sg = fs(sm, k);
              uint64_t s1 = sm & ~(1ULL << k);
              sg *= fs(s1, l);
              uint64_t s2 = s1 & ~(1ULL << l);
              sg *= fs(s2, j);
              sg *= fs(s2 | (1ULL << j), i);
              auto itr = idx.find(sm);
              if (itr == idx.end()) continue;
              out.src.push_back(itr->second);
              out.me.push_back(cidx * t.npairxx + (k * (k - 1) / 2 + l));
              out.phase.push_back((int8_t)sg);
            }
          }
        }
      }
      out.rowptr[f + 1] = (int64_t)out.src.size();
    }
  }
}

std::string wfn_header(const Tables& t) {
  // Layout (little-endian stream, no record markers), as written by
  // write_wfn_header: magic, version, 0, offset of the first vector, 4 zeros;
  // Z, N; isoflag, numorb(1), numorb(2); per species and orbit
  // (nr, 2j, l, par, w); nsps(1), nsps(2); per species and s.p. state in the
  // |m|-sorted order (nr, 2j, 2m, l, w, par, orbit, group); 2Jz, parity
  // character, maxWtot, allsameparity, allsameW, spinless; dimension (int64).
  // Single-parity, single-W spaces (the only ones this engine builds): parity
  // flags are written as +1 and '+', W as 0, maxWtot 0.
  std::string h;
  auto i4 = [&](int32_t v) { h.append(reinterpret_cast<const char*>(&v), 4); };
  const std::vector<Orbit>& orb = t.in.orbits;
  const int norb = (int)orb.size();
  i4(31415926);
  i4(1);
  i4(0);
  i4(0);  // vector offset, patched below
  for (int i = 0; i < 4; i++) i4(0);
  i4(t.in.Z);
  i4(t.in.N);
  i4(1);
  i4(norb);
  i4(norb);
  for (int it = 0; it < 2; it++)
    for (const Orbit& o : orb) {
      i4(o.n);
      i4(o.j2);
      i4(o.l);
      i4(1);
      i4(0);
    }
  const int ns = (int)t.ssorb.size();
  i4(ns);
  i4(ns);
  for (int it = 0; it < 2; it++)
    for (int i = 0; i < ns; i++) {
      const Orbit& o = orb[t.ssorb[i]];
      i4(o.n);
      i4(o.j2);
      i4(t.ssm2[i]);
      i4(o.l);
      i4(0);
      i4(1);
      i4(t.ssorb[i] + 1);
      i4(0);
    }
  i4(t.in.jz2);
  i4('+');
  i4(0);
  i4(1);
  i4(1);
  i4(0);
  const int64_t d = t.dim;
  h.append(reinterpret_cast<const char*>(&d), 8);
  const int32_t off = (int32_t)h.size() + 4;
  std::memcpy(&h[12], &off, 4);
  return h;
}

OneBodyOp read_opme(const std::string& path, const Tables& t) {
  std::ifstream f(path);
  if (!f) fail("cannot open .opme file " + path);
  std::string title;
  std::getline(f, title);
  OneBodyOp op;
  bool xpn = false;
  const std::string kind = title.substr(0, 3);
  if (kind == "iso") op.pn = false;
  else if (kind == "pns") op.pn = true;
  else if (kind == "xpn") op.pn = xpn = true;
  else fail(".opme: something wrong with the first line (iso, pns or xpn expected): " + path);
  const std::vector<Orbit>& orb = t.in.orbits;
  const int no = (int)orb.size();
  int n = -1;
  if (!(f >> n) || n != no) fail(".opme: mismatch in single particle orbits: " + path);
  for (int i = 1; i <= no; i++) {
    int i1 = 0, i2 = 0, nr = 0, l = 0;
    double xj = 0.0;
    if (xpn) {
      if (!(f >> i1 >> i2 >> nr >> l >> xj) || i2 != i + no) fail(".opme: bad orbit line: " + path);
    } else if (!(f >> i1 >> nr >> l >> xj)) {
      fail(".opme: bad orbit line: " + path);
    }
    if (i1 != i || nr != orb[i - 1].n || l != orb[i - 1].l || (int)std::lround(2.0 * xj) != orb[i - 1].j2)
      fail(".opme: mismatch in orbits at orbit " + std::to_string(i) + ": " + path);
  }
  op.top = 2;
  if (!(f >> op.jop)) fail(".opme: missing J of the operator: " + path);
  if (!op.pn && !(f >> op.top)) fail(".opme: missing T of the operator: " + path);
  if (op.jop < 0 || op.top < 0 || (!op.pn && op.top > 1)) fail(".opme: bad J or T of the operator: " + path);
  std::vector<float> iso(no * no, 0.0f), pv(no * no, 0.0f), nv(no * no, 0.0f);
  for (int k = 0; k < 10000; k++) {
    int a = 0, b = 0;
    float x = 0.0f, y = 0.0f;
    if (op.pn && !xpn) {
      if (!(f >> a >> b >> x >> y)) break;
    } else if (!(f >> a >> b >> x)) {
      break;
    }
    bool hasp = true, hasn = true;
    if (xpn) {
      if ((a > no) != (b > no)) continue;  // charge-changing: skipped, as the reference does
      if (a > no) {
        a -= no;
        b -= no;
        y = x;
        x = 0.0f;
        hasp = false;
      } else {
        y = 0.0f;
        hasn = false;
      }
    }
    if (a < 1 || a > no || b < 1 || b > no) fail(".opme: orbit index out of range: " + path);
    const int ja = orb[a - 1].j2, jb = orb[b - 1].j2;
    if (ja + jb < 2 * op.jop || std::abs(ja - jb) > 2 * op.jop)
      fail(".opme: mismatch in angular momentum of operator for orbits " + std::to_string(a) + " " +
           std::to_string(b) + ": " + path);
    if (op.pn) {
      if (t.in.Z > 0 && hasp) pv[(a - 1) * no + b - 1] = x;
      if (t.in.N > 0 && hasn) nv[(a - 1) * no + b - 1] = y;
    } else {
      iso[(a - 1) * no + b - 1] = x;
    }
  }
  // decouple, in float arithmetic in the reference's order of operations
  const int ns = t.nsps;
  op.p.assign((size_t)ns * ns, 0.0f);
  op.n.assign((size_t)ns * ns, 0.0f);
  for (int it = 0; it < 2; it++) {
    std::vector<float>& out = it == 0 ? op.p : op.n;
    if (op.pn && (it == 0 ? t.in.Z : t.in.N) == 0) continue;
    for (int sa = 0; sa < ns; sa++)
      for (int sb = 0; sb < ns; sb++) {
        const int a = t.sorb[sa], b = t.sorb[sb];
        const int ja = t.sj2[sa], ma = t.sm2[sa], jb = t.sj2[sb], mb = t.sm2[sb];
        const float v = op.pn ? (it == 0 ? pv : nv)[a * no + b] : iso[a * no + b];
        if (v == 0.0f || ma != mb) continue;
        if (op.jop > (ja + jb) / 2 || op.jop < std::abs(ja - jb) / 2) continue;
        int iphase = ((jb - mb) / 2) % 2 == 0 ? 1 : -1;
        float r;
        if (op.pn) {
          r = (float)iphase * cleb(ja, ma, jb, -mb, 2 * op.jop, ma - mb) * v / std::sqrt(2.0f * op.jop + 1.0f);
        } else {
          if (it == 1) iphase = -iphase;  // neutrons: T_z = -1/2
          const float ct = it == 0 ? cleb(1, 1, 1, -1, 2 * op.top, 0) : cleb(1, -1, 1, 1, 2 * op.top, 0);
          r = (float)iphase * cleb(ja, ma, jb, -mb, 2 * op.jop, ma - mb) * v / std::sqrt(2.0f * op.top + 1.0f) /
              std::sqrt(2.0f * op.jop + 1.0f) * ct;
        }
        out[(size_t)sa * ns + sb] = r;
      }
  }
  return op;
}

Tables onebody_tables(const Tables& base, const OneBodyOp& op) {
  Tables t;
  t.in = base.in;
  t.nsps = base.nsps;
  t.nL = base.nL;
  t.nR = base.nR;
  t.sorb = base.sorb;
  t.sj2 = base.sj2;
  t.sm2 = base.sm2;
  t.ssorb = base.ssorb;
  t.ssm2 = base.ssm2;
  t.sp[0] = base.sp[0];
  t.sp[1] = base.sp[1];
  t.dim = base.dim;
  t.call_spe = false;
  t.pspe.assign(base.in.orbits.size(), 0.0f);
  t.nspe.assign(base.in.orbits.size(), 0.0f);
  t.ndelta = base.ndelta;
  t.delta_of_class = base.delta_of_class;
  t.class_of_op = base.class_of_op;
  t.local_of_op = base.local_of_op;
  t.class_size = base.class_size;
  t.class_offset = base.class_offset;
  t.wpn = base.wpn;
  t.max_class = base.max_class;
  const int ns = base.nsps;
  for (int it = 0; it < 2; it++) {
    const OneBody& O = base.ob[it];
    const std::vector<float>& pot = it == 0 ? op.p : op.n;
    const size_t nsd = base.sp[it].mask.size();
    if (O.rowptr.size() != nsd + 1) fail("one-body jumps missing");
    TwoBody& T = t.tb[it];
    T.rowptr.assign(1, 0);
    std::vector<std::pair<int32_t, double>> row;
    for (size_t f = 0; f < nsd; f++) {
      row.clear();
      for (int64_t q = O.rowptr[f]; q < O.rowptr[f + 1]; q++) {
        const int a = O.op[q] / ns, c = O.op[q] % ns;
        if (t.sm2[a] != t.sm2[c]) continue;
        const float v = pot[(size_t)a * ns + c];
        if (v == 0.0f) continue;
        row.emplace_back(O.src[q], (double)v * O.phase[q]);
      }
      std::sort(row.begin(), row.end(),
                [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) { return x.first < y.first; });
      for (size_t e = 0; e < row.size(); e++) {
        if (e > 0 && row[e].first == T.src.back() && T.src.size() > (size_t)T.rowptr.back()) {
          T.val.back() += row[e].second;
        } else {
          T.src.push_back(row[e].first);
          T.val.push_back(row[e].second);
        }
      }
      T.rowptr.push_back((int64_t)T.src.size());
    }
    t.ob[it].rowptr.assign(nsd + 1, 0);
  }
  return t;
}

WfnInfo read_wfn_info(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail("cannot open .wfn file " + path);
  // the header of the largest space this engine builds (16 orbits, 64
  // states per species) is under 5 KB
  std::string b(1 << 16, '\0');
  f.read(&b[0], (std::streamsize)b.size());
  b.resize((size_t)f.gcount());
  size_t pos = 0;
  auto i4 = [&]() {
    if (pos + 4 > b.size()) fail(".wfn: truncated header: " + path);
    int32_t v;
    std::memcpy(&v, &b[pos], 4);
    pos += 4;
    return v;
  };
  WfnInfo w;
  if (i4() != 31415926) fail(".wfn: bad magic number (not a BIGSTICK .wfn file): " + path);
  if (i4() != 1) fail(".wfn: unsupported file version: " + path);
  i4();
  w.offset = i4();
  for (int i = 0; i < 4; i++) i4();
  w.Z = i4();
  w.N = i4();
  if (i4() != 1) fail(".wfn: only isospin-format (isoflag 1) files are supported: " + path);
  const int no1 = i4(), no2 = i4();
  if (no1 != no2 || no1 < 1 || no1 > kMaxOrb) fail(".wfn: bad orbit counts: " + path);
  for (int it = 0; it < 2; it++)
    for (int i = 0; i < no1; i++) {
      Orbit o;
      o.n = i4();
      o.j2 = i4();
      o.l = i4();
      i4();  // parity flag
      o.w = i4();
      if (it == 0) w.orbits.push_back(o);
    }
  const int ns1 = i4(), ns2 = i4();
  if (ns1 < 1 || ns1 > kMaxSps || ns2 < 1 || ns2 > kMaxSps) fail(".wfn: bad s.p. state counts: " + path);
  for (int i = 0; i < 8 * (ns1 + ns2); i++) i4();
  w.jz2 = i4();
  for (int i = 0; i < 5; i++) i4();  // parity character, maxWtot, same-parity, same-W, spinless
  if (pos + 8 > b.size()) fail(".wfn: truncated header: " + path);
  std::memcpy(&w.dim, &b[pos], 8);
  pos += 8;
  w.header = b.substr(0, pos);
  if (w.offset != (int64_t)pos + 4) fail(".wfn: vector offset does not follow the header: " + path);
  w.nkeep = i4();
  if (w.nkeep < 0 || w.dim < 1) fail(".wfn: bad dimension or vector count: " + path);
  f.clear();
  f.seekg(0, std::ios::end);
  const int64_t body = (int64_t)f.tellg() - w.offset;
  if (w.nkeep > 0) {
    if (body == (int64_t)w.nkeep * (20 + 4 * w.dim)) w.elem = 4;
    else if (body == (int64_t)w.nkeep * (20 + 8 * w.dim)) w.elem = 8;
    else fail(".wfn: file size does not match " + std::to_string(w.nkeep) + " vectors of dimension " +
              std::to_string(w.dim) + ": " + path);
  }
  return w;
}

}  // namespace bs
