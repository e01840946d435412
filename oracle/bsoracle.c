/*
 * bsoracle.c -- CPU test oracle for the M-scheme Hamiltonian of BIGSTICK.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * file; only tests/, the smoke check and bench.py's CPU-baseline leg use it, as
 * the checker.
 *
 * It restates, in plain C, the parts of the reference that fix the numbers the
 * accelerated path must reproduce:
 *   - single-particle order          src/bspstates.f90:108-218 (sort_spsqn),
 *                                    :224-330 (make_hspsqn)
 *   - basis order (haikus, blocks,   src/bbasis_haikus.f90:327-411,
 *     sectors, pstart/nstart)        src/bbasis_blocks.f90:183-336,458-560,
 *                                    src/bbasis_limits.f90:64-291,
 *                                    src/bbasislib.f90:189-677
 *   - coupled TBMEs from .int        src/binput_tbme.f90:898-1990
 *   - SPE subsumption                src/bsppot.f90:39-240
 *   - Clebsch-Gordan (float)         src/libra.f:320-521 (TH_FACINIT, CLEB)
 *   - pair lists and their sort      src/btbme.f90:208-308,324-423,548-703
 *   - FP32 m-scheme MEs              src/btbme.f90:1209-1558 (uncoupleXXtbme),
 *                                    src/btbmepn.f90:177-594 (count_uncouplepn)
 *   - one-body operators (.opme)     src/bdenslib1.f90:1823-2071 (readin1bodyop),
 *                                    :2444-2597 (decouple1bodyop)
 * and then applies H (or a one-body operator) to a vector by brute-force
 * second quantization over the
 * basis (a different algorithm from the reference's factorized jumps,
 * src/bapplyhlibOMP.f90:979-1450), accumulating in FP64 with the FP32 matrix
 * elements promoted, which is what the reference computes with lanc_prec=8.
 *
 * Scope: full (untruncated) valence spaces whose orbits all share one W and one
 * parity (sd, pf, ...), iso-format .sps/.int files.  Anything else is refused.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXORB 16
#define MAXSPS 64
#define MAXJ 32

typedef struct {
  int norb;
  int on[MAXORB], ol[MAXORB], oj2[MAXORB], ow[MAXORB];
  float spe[MAXORB];
  int Z, N, jz2;
  /* single-particle states in combined label order: left (m<0) then right */
  int nL, nR, nsps;
  int sorb[MAXSPS], sj2[MAXSPS], sm2[MAXSPS];
  /* basis */
  int64_t nsd[2];
  uint64_t *mask[2];
  int *sdsec[2];
  int64_t *start[2];
  int nsec[2];
  int secjz[2][128];
  int64_t secfirst[2][128], seccount[2][128];
  int conj[2][128]; /* for the single-W, single-parity case each sector has one conjugate */
  int64_t dim;
  /* coupled TBMEs (float), symmetric storage keyed by canonical pair-of-pairs */
  float *cpp, *cnn, *cpn;
  int call_spe;
  float pspe[MAXORB], nspe[MAXORB];
  /* uncoupled FP32 m-scheme MEs, dense over pair indices */
  int npairxx, npairpn;
  float *vpp, *vnn, *vpn;
} bso_t;

/* ------------------------------------------------------------------ CLEB -- */
/* src/libra.f:332-347 (TH_FACINIT) and :386-521 (CLEB), float arithmetic with
 * double log-factorials exactly as typed there (implicit REAL for S*, T*, R*). */
static double faclog[501];
static int fac_init = 0;
static void th_facinit(void) {
  double fn = 1.0;
  faclog[1] = 0.0;
  faclog[2] = 0.0;
  for (int i = 3; i <= 500; i++) {
    fn = fn + 1.0;
    faclog[i] = faclog[i - 1] + log(fn);
  }
  fac_init = 1;
}
static int ipow_m1(int k) { return (k % 2 == 0) ? 1 : -1; }

float bso_cleb(int ia, int id, int ib, int ie, int ic, int iff) {
  if (!fac_init) th_facinit();
  float rac = 0.0f;
  if (id + ie - iff != 0) return rac;
  int k1 = ia + ib + ic;
  if (ipow_m1(k1) < 0) return rac;
  k1 = ia + ib - ic;
  int k2 = ic + ia - ib, k3 = ib + ic - ia;
  int k4 = ia - abs(ib - ic), k5 = ib - abs(ic - ia), k6 = ic - abs(ia - ib);
  int k7 = k1;
  if (k2 < k7) k7 = k2;
  if (k3 < k7) k7 = k3;
  if (k4 < k7) k7 = k4;
  if (k5 < k7) k7 = k5;
  if (k6 < k7) k7 = k6;
  if (k7 < 0) return rac;
  if (ipow_m1(ia + id) <= 0) return rac;
  if (ipow_m1(ib + ie) <= 0) return rac;
  if (ipow_m1(ic + iff) <= 0) return rac;
  if (ia - abs(id) < 0) return rac;
  if (ib - abs(ie) < 0) return rac;
  if (ic - abs(iff) < 0) return rac;
  float signfc = 1.0f;
  int iam = ia, ibm = ib, icm = ic, idm = id;
  int it;
  (void)iam;
  int goto235 = 0;
  if (ia - ib < 0) {
    if (ia - ic < 0) { /* 215 */
      it = ia; ia = ib; ib = it;
      it = id; id = ie; ie = it;
      signfc = (float)ipow_m1((ia + ib - ic) / 2);
      goto235 = 1;
    }
  } else {
    if (ic - ib >= 0) goto235 = 1; /* 220: IF(IC-IB) 225,235,235 */
  }
  if (!goto235) { /* 225 */
    it = ic; ic = ib; ib = it;
    it = iff; iff = -ie; ie = -it;
    float fibm = (float)(ibm + 1);
    float ficm = (float)(icm + 1);
    signfc = (float)ipow_m1((iam - idm) / 2) * sqrtf(ficm / fibm);
  }
  /* 235 */
  if (ib == 0) return signfc;
  if (ie > 0) { /* 240 */
    signfc = signfc * (float)ipow_m1((ia + ib - ic) / 2);
    id = -id; ie = -ie; iff = -iff;
  }
  float fc2 = (float)(ic + 1);
  int iabcp = (ia + ib + ic) / 2 + 1;
  int iabc = iabcp - ic, icab = iabcp - ib, ibca = iabcp - ia;
  int iapd = (ia + id) / 2 + 1, iamd = iapd - id;
  int ibpe = (ib + ie) / 2 + 1, ibme = ibpe - ie;
  int icpf = (ic + iff) / 2 + 1, icmf = icpf - iff;
  double t = (double)logf(fc2) - faclog[iabcp + 1];
  t = t + faclog[iabc];
  t = t + faclog[icab];
  t = t + faclog[ibca];
  t = t + faclog[iapd];
  t = t + faclog[iamd];
  t = t + faclog[ibpe];
  t = t + faclog[ibme];
  t = t + faclog[icpf];
  t = t + faclog[icmf];
  float sqfclg = (float)(0.5 * t);
  int nzmic2 = (ib - ic - id) / 2, nzmic3 = (ia - ic + ie) / 2;
  int nzmi = 0;
  if (nzmic2 > nzmi) nzmi = nzmic2;
  if (nzmic3 > nzmi) nzmi = nzmic3;
  nzmi += 1;
  int nzmx = iabc;
  if (iamd < nzmx) nzmx = iamd;
  if (ibpe < nzmx) nzmx = ibpe;
  if (nzmi > nzmx) return 0.0f;
  float ss = 0.0f;
  float s1 = (float)ipow_m1(nzmi - 1);
  for (int nz = nzmi; nz <= nzmx; nz++) {
    int nzm1 = nz - 1;
    int nzt1 = iabc - nzm1, nzt2 = iamd - nzm1, nzt3 = ibpe - nzm1;
    int nzt4 = nz - nzmic2, nzt5 = nz - nzmic3;
    double tl = (double)sqfclg - faclog[nz];
    tl = tl - faclog[nzt1];
    tl = tl - faclog[nzt2];
    tl = tl - faclog[nzt3];
    tl = tl - faclog[nzt4];
    tl = tl - faclog[nzt5];
    float termlg = (float)tl;
    float ssterm = s1 * expf(termlg);
    ss = ss + ssterm;
    s1 = -s1;
  }
  rac = signfc * ss;
  return rac;
}

/* ------------------------------------------------------------- parsing -- */
static int read_data_line(FILE *f, char *buf, int n) {
  while (fgets(buf, n, f)) {
    char *p = buf;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '!' || *p == '#' || *p == '\n' || *p == 0) continue;
    return 1;
  }
  return 0;
}

static int parse_sps(bso_t *o, const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char buf[512];
  if (!read_data_line(f, buf, sizeof buf)) { fclose(f); return -2; }
  if (strncmp(buf, "iso", 3) != 0) { fclose(f); return -3; }
  if (!read_data_line(f, buf, sizeof buf)) { fclose(f); return -2; }
  o->norb = atoi(buf);
  if (o->norb < 1 || o->norb > MAXORB) { fclose(f); return -4; }
  for (int i = 0; i < o->norb; i++) {
    double xn, xl, xj, xw;
    if (!read_data_line(f, buf, sizeof buf)) { fclose(f); return -2; }
    if (sscanf(buf, "%lf %lf %lf %lf", &xn, &xl, &xj, &xw) != 4) { fclose(f); return -5; }
    o->on[i] = (int)lround(xn);
    o->ol[i] = (int)lround(xl);
    o->oj2[i] = (int)lround(2.0 * xj);
    o->ow[i] = (int)lround(xw);
  }
  fclose(f);
  return 0;
}

/* key for XX coupled MEs: pairs (a>=b), (c>=d), 1-based orbits */
static int xxpair(int a, int b) { return a * (a - 1) / 2 + b; }
static size_t ckey_xx(const bso_t *o, int p1, int p2) {
  int np = o->norb * (o->norb + 1) / 2 + 1;
  if (p1 < p2) { int t = p1; p1 = p2; p2 = t; }
  return ((size_t)p1 * np + p2) * MAXJ;
}
static int pnpair(const bso_t *o, int a, int b) { return o->norb * (a - 1) + b; }
static size_t ckey_pn(const bso_t *o, int p1, int p2) {
  int np = o->norb * o->norb + 1;
  if (p1 < p2) { int t = p1; p1 = p2; p2 = t; }
  return ((size_t)p1 * np + p2) * MAXJ;
}

static int parse_int(bso_t *o, const char *path, float spscale, float vscale) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char buf[1024];
  if (!read_data_line(f, buf, sizeof buf)) { fclose(f); return -2; }
  char *p = buf;
  long nme = strtol(p, &p, 10);
  if (nme < 0) { fclose(f); return -6; } /* autoscaled files: not supported */
  for (int i = 0; i < o->norb; i++) {
    float e = strtof(p, &p);
    o->spe[i] = e;
    o->pspe[i] = 0.0f + spscale * e; /* src/binput_tbme.f90:1119-1120 */
    o->nspe[i] = 0.0f + spscale * e;
  }
  int np1 = o->Z, np2 = o->N;
  for (long n = 0; n < nme; n++) {
    if (!read_data_line(f, buf, sizeof buf)) { fclose(f); return -2; }
    int ia, ib, ic, id, J, T;
    float vv;
    if (sscanf(buf, "%d %d %d %d %d %d %f", &ia, &ib, &ic, &id, &J, &T, &vv) != 7) { fclose(f); return -7; }
    if (T != 0 && T != 1) { fclose(f); return -8; }
    if ((ia == ib || ic == id) && ((J + T) % 2 == 0) && vv != 0.0f) { fclose(f); return -9; }
    /* PP and NN parts, src/binput_tbme.f90:1234-1480 (T=1 only) */
    for (int it = 0; it < 2 && T == 1; it++) {
      if ((it == 0 ? np1 : np2) < 2) continue;
      int a = ia, b = ib, c = ic, d = id, phase = 1;
      if (a < b) { int t = a; a = b; b = t; phase = ipow_m1(J + T + (o->oj2[a - 1] + o->oj2[b - 1]) / 2); }
      if (c < d) { int t = c; c = d; d = t; phase = phase * ipow_m1(J + T + (o->oj2[c - 1] + o->oj2[d - 1]) / 2); }
      if (a < c || (a == c && b < d)) { int na = c, nb = d; c = a; d = b; a = na; b = nb; }
      size_t k = ckey_xx(o, xxpair(a, b), xxpair(c, d));
      float *v = (it == 0 ? o->cpp : o->cnn);
      v[k + J] = v[k + J] + ((vv * vscale) * (float)phase) * 1.0f;
    }
    if (np1 * np2 == 0) continue;
    /* PN parts 1-4, src/binput_tbme.f90:1516-1990 */
    float factor = 0.5f;
    if (ia == ib) factor = factor * sqrtf(2.0f);
    if (ic == id) factor = factor * sqrtf(2.0f);
    for (int part = 1; part <= 4; part++) {
      int a, b, c, d, phase = 1;
      if (part == 1) { a = ia; b = ib; c = ic; d = id; }
      else if (part == 2) { b = ia; a = ib; c = ic; d = id; if (a == b) continue; }
      else if (part == 3) { a = ia; b = ib; d = ic; c = id; if (c == d) continue; if (c == b && d == a) continue; }
      else { b = ia; a = ib; d = ic; c = id; if (a == b || c == d) continue; }
      if (a < c || (a == c && b < d)) { int t = a; a = c; c = t; t = b; b = d; d = t; }
      if (part == 2) phase = ipow_m1(J + T + (o->oj2[ia - 1] + o->oj2[ib - 1]) / 2);
      if (part == 3) phase = ipow_m1(J + T + (o->oj2[ic - 1] + o->oj2[id - 1]) / 2);
      if (part == 4) phase = ipow_m1((o->oj2[a - 1] + o->oj2[b - 1] + o->oj2[c - 1] + o->oj2[d - 1]) / 2);
      size_t k = ckey_pn(o, pnpair(o, a, b), pnpair(o, c, d));
      o->cpn[k + J] = o->cpn[k + J] + (((vv * vscale) * (float)phase) * factor) * 1.0f;
    }
  }
  fclose(f);
  return 0;
}

/* src/bsppot.f90:39-240 */
static float zeta(int i, int j) { return (i == j) ? sqrtf(2.0f) : 1.0f; }
static int kd(int i, int j) { return i == j; }
static void subsume_spe(bso_t *o) {
  int np1 = o->Z, np2 = o->N;
  o->call_spe = 1;
  if (np1 + np2 <= 1) return; /* src/bsppot.f90:20-34 */
  o->call_spe = 0;
  float afactor = 1.0f / (float)(np1 + np2 - 1);
  float psp[MAXORB + 1][MAXORB + 1], nsp[MAXORB + 1][MAXORB + 1];
  memset(psp, 0, sizeof psp);
  memset(nsp, 0, sizeof nsp);
  int no = o->norb;
  if (np2 > 0)
    for (int a = 1; a <= no; a++) nsp[a][a] = nsp[a][a] + o->nspe[a - 1];
  for (int pass = 0; pass < 2; pass++) {
    /* pass 0: NN (if np2>1); pass 1: PP (if np1>1) */
    int nump = pass == 0 ? np2 : np1;
    if (pass == 1) {
      if (np1 <= 0) break;
      for (int a = 1; a <= no; a++) psp[a][a] = psp[a][a] + o->pspe[a - 1];
    }
    float (*sp)[MAXORB + 1] = pass == 0 ? nsp : psp;
    float *cv = pass == 0 ? o->cnn : o->cpp;
    for (int a = 1; a <= no; a++) {
      int ja = (o->oj2[a - 1] + 1) / 2;
      for (int c = 1; c <= a; c++) {
        int jc = (o->oj2[c - 1] + 1) / 2;
        if (nump > 1) {
          for (int b = 1; b <= a; b++) {
            int jb = (o->oj2[b - 1] + 1) / 2;
            int dfinish = (a == c) ? b : c;
            for (int d = 1; d <= dfinish; d++) {
              int jd = (o->oj2[d - 1] + 1) / 2;
              if (b != d && a != c && b != c && a != d) continue;
              if ((o->ol[a - 1] + o->ol[b - 1] + o->ol[c - 1] + o->ol[d - 1]) % 2) continue;
              size_t k = ckey_xx(o, xxpair(a, b), xxpair(c, d));
              int jjmin = abs(o->oj2[a - 1] - o->oj2[b - 1]);
              int t2 = abs(o->oj2[c - 1] - o->oj2[d - 1]);
              if (t2 > jjmin) jjmin = t2;
              jjmin /= 2;
              int jjmax = o->oj2[a - 1] + o->oj2[b - 1];
              t2 = o->oj2[c - 1] + o->oj2[d - 1];
              if (t2 < jjmax) jjmax = t2;
              jjmax /= 2;
              for (int jj = jjmin; jj <= jjmax; jj++) {
                int phaseab = ipow_m1((ja - jb) + jj);
                int phasecd = ipow_m1((jc - jd) + jj);
                float vvv = sp[a][c] * (float)kd(b, d);
                vvv = vvv + (sp[a][d] * (float)kd(b, c)) * (float)phasecd;
                vvv = vvv + (sp[b][c] * (float)kd(a, d)) * (float)phaseab;
                vvv = vvv + sp[b][d] * (float)kd(a, c);
                if (pass == 1 && vvv == 0.0f) continue;
                vvv = ((afactor / zeta(a, b)) / zeta(d, c)) * vvv;
                cv[k + jj] = cv[k + jj] + vvv;
              }
            }
          }
        }
        if (pass == 0) continue;
        if (np2 < 1) continue;
        for (int b = 1; b <= no; b++) {
          int dfinish = (a == c) ? b : no;
          for (int d = 1; d <= dfinish; d++) {
            if ((o->ol[a - 1] + o->ol[b - 1] + o->ol[c - 1] + o->ol[d - 1]) % 2) continue;
            size_t k = ckey_pn(o, pnpair(o, a, b), pnpair(o, c, d));
            float vvv = psp[a][c] * (float)kd(b, d) + nsp[b][d] * (float)kd(a, c);
            vvv = afactor * vvv;
            int jjmin = abs(o->oj2[a - 1] - o->oj2[b - 1]);
            int t2 = abs(o->oj2[c - 1] - o->oj2[d - 1]);
            if (t2 > jjmin) jjmin = t2;
            jjmin /= 2;
            int jjmax = o->oj2[a - 1] + o->oj2[b - 1];
            t2 = o->oj2[c - 1] + o->oj2[d - 1];
            if (t2 < jjmax) jjmax = t2;
            jjmax /= 2;
            for (int jj = jjmin; jj <= jjmax; jj++) o->cpn[k + jj] = o->cpn[k + jj] + vvv;
          }
        }
      }
    }
  }
}

/* ---------------------------------------------------- single particles -- */
/* src/bspstates.f90:47-90 (extract_sps), :108-218 (sort_spsqn; W and parity
 * sorts are no-ops for a single-W, single-parity space), :224-330 */
static void make_sps(bso_t *o) {
  int orb[MAXSPS], m2[MAXSPS], n = 0;
  for (int i = 0; i < o->norb; i++)
    for (int m = -o->oj2[i]; m <= o->oj2[i]; m += 2) { orb[n] = i; m2[n] = m; n++; }
  /* selection sort on |m| with swap (not stable) */
  for (int is = 0; is < n; is++) {
    int k = is, m = m2[is], so = orb[is], sm = m2[is];
    for (int j = is + 1; j < n; j++)
      if (abs(m2[j]) < abs(m)) { k = j; m = m2[j]; so = orb[j]; sm = m2[j]; }
    if (k != is) {
      orb[k] = orb[is]; m2[k] = m2[is];
      orb[is] = so; m2[is] = sm;
    }
  }
  int nl = 0, nr = 0;
  int lo[MAXSPS], lm[MAXSPS], ro[MAXSPS], rm[MAXSPS];
  for (int i = 0; i < n; i++) {
    if (m2[i] >= 0) { ro[nr] = orb[i]; rm[nr] = m2[i]; nr++; }
    else { lo[nl] = orb[i]; lm[nl] = m2[i]; nl++; }
  }
  o->nL = nl; o->nR = nr; o->nsps = n;
  for (int i = 0; i < nl; i++) { o->sorb[i] = lo[i]; o->sm2[i] = lm[i]; o->sj2[i] = o->oj2[lo[i]]; }
  for (int i = 0; i < nr; i++) { o->sorb[nl + i] = ro[i]; o->sm2[nl + i] = rm[i]; o->sj2[nl + i] = o->oj2[ro[i]]; }
}

/* ------------------------------------------------------------- haikus -- */
/* A haiku is a subset of one half-space (bits = half-space index, 0-based).
 * src/bbasis_haikus.f90:327-411 builds them for one W by adding, to each
 * parent in order, every state above the parent's highest occupied one, from
 * the top down.  That order is the lexicographic *descending* order of the
 * ascending index tuples; the oracle enumerates subsets and sorts by it. */
typedef struct { uint64_t bits; int jz; } haiku_t;

static int cmp_haiku_desc(const void *pa, const void *pb) {
  uint64_t a = ((const haiku_t *)pa)->bits, b = ((const haiku_t *)pb)->bits;
  /* compare ascending tuples lexicographically; larger first */
  while (a && b) {
    int ia = __builtin_ctzll(a), ib = __builtin_ctzll(b);
    if (ia != ib) return ia > ib ? -1 : 1;
    a &= a - 1;
    b &= b - 1;
  }
  if (!a && !b) return 0;
  return a ? -1 : 1; /* not reached for equal particle numbers */
}

static haiku_t *list_haikus(int nstates, const int *m2, int nh, int *count) {
  int64_t cap = 1;
  for (int i = 0; i < nh; i++) cap = cap * (nstates - i) / (i + 1);
  haiku_t *h = (haiku_t *)malloc(sizeof(haiku_t) * (cap > 0 ? cap : 1));
  int c = 0;
  if (nh == 0) { h[0].bits = 0; h[0].jz = 0; *count = 1; return h; }
  /* Gosper enumeration of nh-subsets */
  uint64_t s = (nh >= 64) ? ~0ULL : ((1ULL << nh) - 1);
  uint64_t lim = (nstates >= 64) ? 0 : (1ULL << nstates);
  while (nstates >= nh && (lim == 0 || s < lim)) {
    int jz = 0;
    for (uint64_t t = s; t; t &= t - 1) jz += m2[__builtin_ctzll(t)];
    h[c].bits = s; h[c].jz = jz; c++;
    uint64_t u = s & -s, v = s + u;
    if (v == 0) break;
    s = v + (((v ^ s) / u) >> 2);
  }
  qsort(h, c, sizeof(haiku_t), cmp_haiku_desc);
  *count = c;
  return h;
}

/* --------------------------------------------------------------- basis -- */
typedef struct { int nh; int jz; int n; uint64_t *bits; } hblock_t;

static int build_basis(bso_t *o) {
  int nL = o->nL, nR = o->nR;
  int lm2[MAXSPS], rm2[MAXSPS];
  for (int i = 0; i < nL; i++) lm2[i] = o->sm2[i];
  for (int i = 0; i < nR; i++) rm2[i] = o->sm2[nL + i];
  int np[2] = {o->Z, o->N};
  /* per species: blocks for right (h=0) and left (h=1) haikus, by nh then jz */
  int jzmin[2], jzmax[2];
  for (int it = 0; it < 2; it++) {
    int Np = np[it];
    int maxR = Np < nR ? Np : nR, maxL = Np < nL ? Np : nL;
    int mn = 100000, mx = -100000;
    for (int nh = Np - maxL; nh <= maxR; nh++) {
      int nhc = Np - nh;
      int cr, cl;
      haiku_t *hr = list_haikus(nR, rm2, nh, &cr);
      haiku_t *hl = list_haikus(nL, lm2, nhc, &cl);
      for (int i = 0; i < cr; i++)
        for (int j = 0; j < cl; j++) {
          int z = hr[i].jz + hl[j].jz;
          if (z < mn) mn = z;
          if (z > mx) mx = z;
        }
      free(hr);
      free(hl);
    }
    jzmin[it] = mn; jzmax[it] = mx;
  }
  /* src/bbasis_limits.f90:269-291 (reconcile_sdjz), sequential updates */
  int jz = o->jz2;
  if (jz - jzmin[1] < jzmax[0]) jzmax[0] = jz - jzmin[1];
  if (jz - jzmin[0] < jzmax[1]) jzmax[1] = jz - jzmin[0];
  if (jz - jzmax[1] > jzmin[0]) jzmin[0] = jz - jzmax[1];
  if (jz - jzmax[0] > jzmin[1]) jzmin[1] = jz - jzmax[0];
  for (int it = 0; it < 2; it++) {
    int Np = np[it];
    int maxR = Np < nR ? Np : nR, maxL = Np < nL ? Np : nL;
    /* count SDs first */
    int64_t total = 0;
    int ns = 0;
    int start = it == 0 ? jzmin[0] : jzmax[1];
    int stop = it == 0 ? jzmax[0] : jzmin[1];
    int step = it == 0 ? 2 : -2;
    for (int jzx = start; it == 0 ? jzx <= stop : jzx >= stop; jzx += step) {
      int64_t cnt = 0;
      for (int nh = Np - maxL; nh <= maxR; nh++) {
        int cr, cl;
        haiku_t *hr = list_haikus(nR, rm2, nh, &cr);
        haiku_t *hl = list_haikus(nL, lm2, Np - nh, &cl);
        for (int i = 0; i < cr; i++)
          for (int j = 0; j < cl; j++)
            if (hr[i].jz + hl[j].jz == jzx) cnt++;
        free(hr);
        free(hl);
      }
      if (cnt == 0) continue;
      o->secjz[it][ns] = jzx;
      o->secfirst[it][ns] = total;
      o->seccount[it][ns] = cnt;
      total += cnt;
      ns++;
    }
    o->nsec[it] = ns;
    o->nsd[it] = total;
    o->mask[it] = (uint64_t *)malloc(sizeof(uint64_t) * (total ? total : 1));
    o->sdsec[it] = (int *)malloc(sizeof(int) * (total ? total : 1));
    o->start[it] = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
    /* fill: within a sector, blocks by nh (right count) then right jz
     * ascending; inside a block, left haiku outer and right haiku inner
     * (src/bbasislib.f90:380-430; ix = start + radd + (ladd-1)*nradd) */
    for (int s = 0; s < ns; s++) {
      int jzx = o->secjz[it][s];
      int64_t pos = o->secfirst[it][s];
      for (int nh = Np - maxL; nh <= maxR; nh++) {
        int cr, cl;
        haiku_t *hr = list_haikus(nR, rm2, nh, &cr);
        haiku_t *hl = list_haikus(nL, lm2, Np - nh, &cl);
        int rjmin = 100000, rjmax = -100000;
        for (int i = 0; i < cr; i++) {
          if (hr[i].jz < rjmin) rjmin = hr[i].jz;
          if (hr[i].jz > rjmax) rjmax = hr[i].jz;
        }
        for (int jzh = rjmin; jzh <= rjmax; jzh += 2) {
          int jzhc = jzx - jzh;
          for (int j = 0; j < cl; j++) {
            if (hl[j].jz != jzhc) continue;
            for (int i = 0; i < cr; i++) {
              if (hr[i].jz != jzh) continue;
              o->mask[it][pos] = hl[j].bits | (hr[i].bits << nL);
              o->sdsec[it][pos] = s;
              pos++;
            }
          }
        }
        free(hr);
        free(hl);
      }
      if (pos != o->secfirst[it][s] + o->seccount[it][s]) return -20;
    }
  }
  /* conjugate sectors: single W/parity => neutron jz = Jz - proton jz */
  for (int it = 0; it < 2; it++)
    for (int s = 0; s < o->nsec[it]; s++) {
      int want = o->jz2 - o->secjz[it][s];
      o->conj[it][s] = -1;
      for (int t = 0; t < o->nsec[1 - it]; t++)
        if (o->secjz[1 - it][t] == want) o->conj[it][s] = t;
      if (o->conj[it][s] < 0) return -21;
    }
  /* src/bbasislib.f90:539-677 (combinebasis), 0-based states */
  int64_t dim = 0, ncount = 0, pcount = 0;
  for (int s = 0; s < o->nsec[0]; s++) {
    int cs = o->conj[0][s];
    int64_t ndim = o->seccount[1][cs];
    if (o->conj[1][cs] == s)
      for (int64_t in = 0; in < ndim; in++) o->start[1][ncount++] = in;
    for (int64_t ip = 0; ip < o->seccount[0][s]; ip++) {
      o->start[0][pcount++] = dim;
      dim += ndim;
    }
  }
  o->dim = dim;
  return 0;
}

/* ---------------------------------------------------- m-scheme MEs ------ */
typedef struct { int m, par, indx, ia, ib; } pairqn_t; /* ia, ib: combined labels (0-based) */

/* src/btbme.f90:548-703: heapsort on M (descending), then heapsort on parity
 * inside each run of equal M; 1-based loops restated as written. */
static void sortpair(int npair, pairqn_t *pair1based) {
  pairqn_t *pair = pair1based - 1; /* pair[1..npair] */
  pairqn_t tmp;
  int i, j, l = npair / 2 + 1, n = npair;
  while (n > 0) {
    if (l > 1) {
      l = l - 1;
      tmp = pair[l];
    } else {
      tmp = pair[n];
      pair[n] = pair[1];
      n = n - 1;
      if (n == 1) { pair[1] = tmp; break; }
    }
    i = l;
    j = l + l;
    while (j <= n) {
      if (j < n && pair[j].m > pair[j + 1].m) j = j + 1;
      if (tmp.m > pair[j].m) {
        pair[i] = pair[j];
        i = j;
        j = j + j;
      } else j = n + 1;
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
      if (l > istart) {
        l = l - 1;
        tmp = pair[l];
      } else {
        tmp = pair[n];
        pair[n] = pair[istart];
        n = n - 1;
        if (n == istart) { pair[istart] = tmp; break; }
      }
      i = l;
      j = l + l - (istart - 1);
      while (j <= n) {
        if (j < n && pair[j].par < pair[j + 1].par) j = j + 1;
        if (tmp.par < pair[j].par) {
          pair[i] = pair[j];
          i = j;
          j = j + j - (istart - 1);
        } else j = n + 1;
      }
      pair[i] = tmp;
    }
    istart = istop + 1;
    istop = istart;
  }
}

static void uncouple_xx(bso_t *o, int it) {
  int ns = o->nsps;
  int np = ns * (ns - 1) / 2;
  float *v = (float *)calloc((size_t)np * np, sizeof(float));
  pairqn_t *pl = (pairqn_t *)malloc(sizeof(pairqn_t) * (np ? np : 1));
  int ip = 0;
  for (int a = 1; a < ns; a++)
    for (int b = 0; b < a; b++) {
      pl[ip].m = (o->sm2[a] + o->sm2[b]) / 2;
      pl[ip].par = 1;
      pl[ip].indx = a * (a - 1) / 2 + b;
      pl[ip].ia = a;
      pl[ip].ib = b;
      ip++;
    }
  sortpair(np, pl);
  const float *cv = it == 0 ? o->cpp : o->cnn;
  for (int dp = 0; dp < np; dp++) {
    int m = pl[dp].m;
    int k = pl[dp].ia, l = pl[dp].ib;
    int ia = o->sorb[k] + 1, ib = o->sorb[l] + 1;
    int jk = o->sj2[k], mk = o->sm2[k], jl = o->sj2[l], ml = o->sm2[l];
    int phasekl = 1;
    if (ia < ib) {
      int t = ia; ia = ib; ib = t;
      t = jk; jk = jl; jl = t;
      t = mk; mk = ml; ml = t;
      phasekl = -1;
    }
    for (int cp = 0; cp <= dp; cp++) {
      if (pl[cp].m != m || pl[cp].par != pl[dp].par) continue;
      int i = pl[cp].ia, j = pl[cp].ib;
      int ic = o->sorb[i] + 1, id = o->sorb[j] + 1;
      int ji = o->sj2[i], mi = o->sm2[i], jj = o->sj2[j], mj = o->sm2[j];
      int phaseij = 1;
      if (ic < id) {
        int t = ic; ic = id; id = t;
        t = ji; ji = jj; jj = t;
        t = mi; mi = mj; mj = t;
        phaseij = -1;
      }
      size_t key = ckey_xx(o, xxpair(ia, ib), xxpair(ic, id));
      int jmax = (jk + jl < ji + jj ? jk + jl : ji + jj) / 2;
      int jmin = (abs(jk - jl) > abs(ji - jj) ? abs(jk - jl) : abs(ji - jj)) / 2;
      if (abs(m) > jmin) jmin = abs(m);
      float vtmp = 0.0f;
      for (int jt = jmin; jt <= jmax; jt++) {
        float t = cv[key + jt] * zeta(ia, ib);
        t = t * zeta(ic, id);
        t = t * bso_cleb(jk, mk, jl, ml, 2 * jt, 2 * m);
        t = t * bso_cleb(ji, mi, jj, mj, 2 * jt, 2 * m);
        vtmp = vtmp + t;
      }
      float val = (vtmp * (float)phasekl) * (float)phaseij;
      v[(size_t)pl[cp].indx * np + pl[dp].indx] = val;
      v[(size_t)pl[dp].indx * np + pl[cp].indx] = val;
    }
  }
  free(pl);
  if (it == 0) o->vpp = v; else o->vnn = v;
  o->npairxx = np;
}

static void uncouple_pn(bso_t *o) {
  int ns = o->nsps;
  int np = ns * ns;
  float *v = (float *)calloc((size_t)np * np, sizeof(float));
  pairqn_t *pl = (pairqn_t *)malloc(sizeof(pairqn_t) * np);
  int ip = 0;
  for (int a = 0; a < ns; a++)
    for (int b = 0; b < ns; b++) {
      pl[ip].m = (o->sm2[a] + o->sm2[b]) / 2;
      pl[ip].par = 1;
      pl[ip].indx = a * ns + b;
      pl[ip].ia = a;
      pl[ip].ib = b;
      ip++;
    }
  sortpair(np, pl);
  int s0 = 0;
  while (s0 < np) {
    int s1 = s0;
    while (s1 + 1 < np && pl[s1 + 1].m == pl[s0].m && pl[s1 + 1].par == pl[s0].par) s1++;
    for (int dp = s0; dp <= s1; dp++) {
      int k = pl[dp].ia, l = pl[dp].ib;
      int ia = o->sorb[k] + 1, ib = o->sorb[l] + 1;
      int jk = o->sj2[k], mk = o->sm2[k], jl = o->sj2[l], ml = o->sm2[l];
      int m = pl[dp].m;
      for (int cp = s0; cp <= dp; cp++) {
        int i = pl[cp].ia, j = pl[cp].ib;
        int ic = o->sorb[i] + 1, id = o->sorb[j] + 1;
        int ji = o->sj2[i], mi = o->sm2[i], jj = o->sj2[j], mj = o->sm2[j];
        size_t key = ckey_pn(o, pnpair(o, ia, ib), pnpair(o, ic, id));
        int jmax = (jk + jl < ji + jj ? jk + jl : ji + jj) / 2;
        int jmin = (abs(jk - jl) > abs(ji - jj) ? abs(jk - jl) : abs(ji - jj)) / 2;
        if (abs(m) > jmin) jmin = abs(m);
        float vtmp = 0.0f;
        for (int jt = jmin; jt <= jmax; jt++) {
          float t = o->cpn[key + jt] * bso_cleb(jk, mk, jl, ml, 2 * jt, mk + ml);
          t = t * bso_cleb(ji, mi, jj, mj, 2 * jt, mi + mj);
          vtmp = vtmp + t;
        }
        v[(size_t)pl[cp].indx * np + pl[dp].indx] = vtmp;
        v[(size_t)pl[dp].indx * np + pl[cp].indx] = vtmp;
      }
    }
    s0 = s1 + 1;
  }
  free(pl);
  o->vpn = v;
  o->npairpn = np;
}

/* ------------------------------------------------------------ public API -- */
void bso_free(bso_t *o);

bso_t *bso_create(const char *sps, const char *intf, int Z, int N, int jz2,
                  float spscale, float vscale, int *err) {
  bso_t *o = (bso_t *)calloc(1, sizeof(bso_t));
  int e = parse_sps(o, sps);
  if (e) { *err = e; free(o); return NULL; }
  for (int i = 1; i < o->norb; i++)
    if (o->ow[i] != o->ow[0] || (o->ol[i] % 2) != (o->ol[0] % 2)) { *err = -30; free(o); return NULL; }
  o->Z = Z; o->N = N; o->jz2 = jz2;
  if (((Z + N) % 2 == 0) != (jz2 % 2 == 0)) { *err = -31; free(o); return NULL; }
  size_t nxx = (size_t)(o->norb * (o->norb + 1) / 2 + 1);
  size_t npn = (size_t)(o->norb * o->norb + 1);
  o->cpp = (float *)calloc(nxx * nxx * MAXJ, sizeof(float));
  o->cnn = (float *)calloc(nxx * nxx * MAXJ, sizeof(float));
  o->cpn = (float *)calloc(npn * npn * MAXJ, sizeof(float));
  e = parse_int(o, intf, spscale, vscale);
  if (e) { *err = e; bso_free(o); return NULL; }
  subsume_spe(o);
  make_sps(o);
  e = build_basis(o);
  if (e) { *err = e; bso_free(o); return NULL; }
  uncouple_xx(o, 0);
  uncouple_xx(o, 1);
  uncouple_pn(o);
  *err = 0;
  return o;
}

void bso_free(bso_t *o) {
  if (!o) return;
  for (int it = 0; it < 2; it++) { free(o->mask[it]); free(o->sdsec[it]); free(o->start[it]); }
  free(o->cpp); free(o->cnn); free(o->cpn);
  free(o->vpp); free(o->vnn); free(o->vpn);
  free(o);
}

int64_t bso_dim(const bso_t *o) { return o->dim; }
int64_t bso_nsd(const bso_t *o, int it) { return o->nsd[it]; }
int bso_nsps(const bso_t *o) { return o->nsps; }
int bso_nleft(const bso_t *o) { return o->nL; }
int bso_call_spe(const bso_t *o) { return o->call_spe; }

/* SD bit masks over combined labels (left states first, then right) and the
 * pstart/nstart offsets (0-based state = pstart + nstart). */
void bso_sds(const bso_t *o, int it, uint64_t *mask, int64_t *start, int *sector) {
  memcpy(mask, o->mask[it], sizeof(uint64_t) * o->nsd[it]);
  memcpy(start, o->start[it], sizeof(int64_t) * o->nsd[it]);
  memcpy(sector, o->sdsec[it], sizeof(int) * o->nsd[it]);
}
/* s.p. states in combined label order: orbit (0-based), 2j, 2m */
void bso_sps(const bso_t *o, int *orb, int *j2, int *m2) {
  for (int i = 0; i < o->nsps; i++) { orb[i] = o->sorb[i]; j2[i] = o->sj2[i]; m2[i] = o->sm2[i]; }
}
/* Uncoupled FP32 MEs.  kind 0: pp, 1: nn  (pair index c1*(c1-1)/2+c2, c1>c2);
 * kind 2: pn (pair index cp*nsps+cn).  Returns a pointer to the dense
 * npair x npair row-major table. */
const float *bso_me_table(const bso_t *o, int kind, int *npair) {
  if (kind == 2) { *npair = o->npairpn; return o->vpn; }
  *npair = o->npairxx;
  return kind == 0 ? o->vpp : o->vnn;
}
/* Coupled FP32 values after reading and SPE subsumption: kind 0 pp, 1 nn
 * (orbits a>=b, c>=d), 2 pn (proton a,c; neutron b,d); 1-based orbits. */
float bso_coupled(const bso_t *o, int kind, int a, int b, int c, int d, int J) {
  if (kind == 2) return o->cpn[ckey_pn(o, pnpair(o, a, b), pnpair(o, c, d)) + J];
  const float *cv = kind == 0 ? o->cpp : o->cnn;
  return cv[ckey_xx(o, xxpair(a, b), xxpair(c, d)) + J];
}

static int64_t find_sd(const bso_t *o, int it, uint64_t m, int sec) {
  int64_t lo = o->secfirst[it][sec], hi = lo + o->seccount[it][sec];
  /* within a sector SDs are not sorted by mask: linear scan is too slow, so
   * the caller uses the lookup table built in bso_apply */
  for (int64_t i = lo; i < hi; i++)
    if (o->mask[it][i] == m) return i;
  return -1;
}

/* open-addressing map mask -> SD index */
typedef struct { uint64_t *keys; int64_t *vals; size_t cap; } sdmap_t;
static void map_build(sdmap_t *mp, const uint64_t *masks, int64_t n) {
  size_t cap = 16;
  while (cap < (size_t)n * 2 + 16) cap <<= 1;
  mp->cap = cap;
  mp->keys = (uint64_t *)malloc(sizeof(uint64_t) * cap);
  mp->vals = (int64_t *)malloc(sizeof(int64_t) * cap);
  for (size_t i = 0; i < cap; i++) mp->keys[i] = ~0ULL;
  for (int64_t i = 0; i < n; i++) {
    size_t h = (size_t)((masks[i] * 0x9E3779B97F4A7C15ULL) >> 17) & (cap - 1);
    while (mp->keys[h] != ~0ULL) h = (h + 1) & (cap - 1);
    mp->keys[h] = masks[i];
    mp->vals[h] = i;
  }
}
static int64_t map_get(const sdmap_t *mp, uint64_t key) {
  size_t h = (size_t)((key * 0x9E3779B97F4A7C15ULL) >> 17) & (mp->cap - 1);
  while (mp->keys[h] != ~0ULL) {
    if (mp->keys[h] == key) return mp->vals[h];
    h = (h + 1) & (mp->cap - 1);
  }
  return -1;
}

/* fermion sign of removing/adding label c on mask: (-1)^(# occupied below c) */
static int fsign(uint64_t mask, int c) {
  uint64_t below = (c == 0) ? 0 : (mask & ((1ULL << c) - 1));
  return (__builtin_popcountll(below) & 1) ? -1 : 1;
}

/* y += H x with FP64 accumulation (x,y of length dim).  Basis states are
 * 0-based.  Slater determinants are a^+ products in ascending combined label
 * order, protons to the left of neutrons.  H is real symmetric, so each row
 * y[s] is gathered from the states the operators reach from s:
 * y[s] = sum_t <t|H|s> x[t]; rows are independent (OpenMP over proton SDs). */
/* The states of the listed proton SDs, one OpenMP work item per state (the
 * rows' states flattened, so that a few rows still spread over every
 * thread). */
static void apply_rows(const bso_t *o, const double *x, double *y, const int64_t *rows, int64_t nrows) {
  int ns = o->nsps;
  sdmap_t mp[2];
  for (int it = 0; it < 2; it++) map_build(&mp[it], o->mask[it], o->nsd[it]);
  int npx = ns * (ns - 1) / 2;
  int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nrows + 1));
  off[0] = 0;
  for (int64_t r = 0; r < nrows; r++) off[r + 1] = off[r] + o->seccount[1][o->conj[0][o->sdsec[0][rows[r]]]];
  const int64_t total = off[nrows];
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t q = 0; q < total; q++) {
    int64_t lo = 0, hi = nrows - 1;  /* the row r with off[r] <= q < off[r + 1] */
    while (lo < hi) {
      int64_t mid = (lo + hi + 1) / 2;
      if (off[mid] <= q) lo = mid; else hi = mid - 1;
    }
    const int64_t ip = rows[lo];
    int sp = o->sdsec[0][ip];
    int sn = o->conj[0][sp];
    uint64_t pm = o->mask[0][ip];
    {
      const int64_t in = o->secfirst[1][sn] + (q - off[lo]);
      uint64_t nm = o->mask[1][in];
      int64_t s = o->start[0][ip] + o->start[1][in];
      double acc = 0.0;
      if (o->call_spe) {
        double e = 0.0;
        for (uint64_t t = pm; t; t &= t - 1) e += o->pspe[o->sorb[__builtin_ctzll(t)]];
        for (uint64_t t = nm; t; t &= t - 1) e += o->nspe[o->sorb[__builtin_ctzll(t)]];
        acc += e * x[s];
      }
      /* PP and NN: a+_i a+_j a_l a_k, c(k)>c(l), c(i)>c(j) */
      for (int it = 0; it < 2; it++) {
        uint64_t m0 = it == 0 ? pm : nm;
        const float *v = it == 0 ? o->vpp : o->vnn;
        if (!v || (it == 0 ? o->Z : o->N) < 2) continue;
        for (uint64_t tk = m0; tk; tk &= tk - 1) {
          int k = __builtin_ctzll(tk);
          for (uint64_t tl = m0 & ((1ULL << k) - 1); tl; tl &= tl - 1) {
            int l = __builtin_ctzll(tl);
            int sg = fsign(m0, k);
            uint64_t m1 = m0 & ~(1ULL << k);
            sg *= fsign(m1, l);
            uint64_t m2 = m1 & ~(1ULL << l);
            int dkl = k * (k - 1) / 2 + l;
            int mkl = o->sm2[k] + o->sm2[l];
            for (int i = 1; i < ns; i++) {
              if (m2 >> i & 1) continue;
              for (int j = 0; j < i; j++) {
                if (m2 >> j & 1) continue;
                if (o->sm2[i] + o->sm2[j] != mkl) continue;
                float me = v[(size_t)(i * (i - 1) / 2 + j) * npx + dkl];
                if (me == 0.0f) continue;
                int s2 = sg * fsign(m2, j);
                uint64_t m3 = m2 | (1ULL << j);
                s2 *= fsign(m3, i);
                uint64_t m4 = m3 | (1ULL << i);
                int64_t t;
                if (it == 0) {
                  int64_t ipf = map_get(&mp[0], m4);
                  t = o->start[0][ipf] + o->start[1][in];
                } else {
                  int64_t inf = map_get(&mp[1], m4);
                  t = o->start[0][ip] + o->start[1][inf];
                }
                acc += (double)me * (double)s2 * x[t];
              }
            }
          }
        }
      }
      /* PN: a+_i b+_j b_l a_k */
      if (o->Z > 0 && o->N > 0) {
        for (uint64_t tk = pm; tk; tk &= tk - 1) {
          int k = __builtin_ctzll(tk);
          int sk = fsign(pm, k);
          uint64_t pm1 = pm & ~(1ULL << k);
          for (uint64_t tl = nm; tl; tl &= tl - 1) {
            int l = __builtin_ctzll(tl);
            int sl = fsign(nm, l);
            uint64_t nm1 = nm & ~(1ULL << l);
            int mkl = o->sm2[k] + o->sm2[l];
            for (int i = 0; i < ns; i++) {
              if (pm1 >> i & 1) continue;
              int si = fsign(pm1, i);
              uint64_t pm2 = pm1 | (1ULL << i);
              int64_t ipf = -1;
              for (int j = 0; j < ns; j++) {
                if (nm1 >> j & 1) continue;
                if (o->sm2[i] + o->sm2[j] != mkl) continue;
                float me = o->vpn[(size_t)(i * ns + j) * (ns * ns) + (k * ns + l)];
                if (me == 0.0f) continue;
                int sj = fsign(nm1, j);
                uint64_t nm2 = nm1 | (1ULL << j);
                if (ipf < 0) ipf = map_get(&mp[0], pm2);
                int64_t inf = map_get(&mp[1], nm2);
                int64_t t = o->start[0][ipf] + o->start[1][inf];
                acc += (double)me * (double)(sk * sl * si * sj) * x[t];
              }
            }
          }
        }
      }
      y[s] += acc;
    }
  }
  free(off);
  for (int it = 0; it < 2; it++) { free(mp[it].keys); free(mp[it].vals); }
  (void)find_sd;
}

void bso_apply(const bso_t *o, const double *x, double *y) {
  int64_t n = o->nsd[0];
  int64_t *rows = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) rows[i] = i;
  apply_rows(o, x, y, rows, n);
  free(rows);
}

/* y += H x on the states of the listed proton SDs only (the states
 * pstart[ip] + nstart[in], in over the conjugate neutron sector); the other
 * entries of y are left alone.  x must hold the whole vector.  Returns -1
 * for a row index out of range. */
int bso_apply_rows(const bso_t *o, const double *x, double *y, const int64_t *rows, int64_t nrows) {
  for (int64_t r = 0; r < nrows; r++)
    if (rows[r] < 0 || rows[r] >= o->nsd[0]) return -1;
  if (nrows > 0) apply_rows(o, x, y, rows, nrows);
  return 0;
}

/* ------------------------------------------------ factorized CPU port -- */
/* A CPU restatement of the reference's factorized apply,
 * applyHbundled_omp (src/bapplyhlibOMP.f90:979-1450): precomputed proton and
 * neutron one- and two-body jump lists, FP32 vectors and FP32 matrix elements,
 * OpenMP threads owning disjoint output rows (the reference splits output
 * ranges between threads, :34-868).  It is used only as the CPU timing
 * baseline (bench.py) and as a second, independent check of the factorized
 * data flow against bso_apply. */
typedef struct {
  int64_t *p1ptr[2]; int32_t *p1src[2]; int16_t *p1a[2], *p1c[2]; int8_t *p1ph[2];
  int64_t *p1dptr[2];  /* per target SD: (ndm+1) offsets of its jumps grouped by m_a - m_c */
  int dmmin, ndm;
  int64_t *p2ptr[2]; int32_t *p2src[2]; float *p2val[2];
  int built;
} bso_jumps_t;

static bso_jumps_t g_jumps_dummy;

typedef struct bso_port {
  const bso_t *o;
  bso_jumps_t j;
} bso_port_t;

bso_port_t *bso_port_create(const bso_t *o) {
  bso_port_t *p = (bso_port_t *)calloc(1, sizeof(bso_port_t));
  p->o = o;
  int ns = o->nsps, npx = ns * (ns - 1) / 2;
  for (int it = 0; it < 2; it++) {
    sdmap_t mp;
    map_build(&mp, o->mask[it], o->nsd[it]);
    int64_t n = o->nsd[it];
    /* one-body a+_a a_c (all a, c incl. a == c), listed by target */
    int np = it == 0 ? o->Z : o->N;
    int64_t cap1 = n * (int64_t)(np * (ns - np + 1) + 1);
    p->j.p1ptr[it] = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    p->j.p1src[it] = (int32_t *)malloc(sizeof(int32_t) * (cap1 ? cap1 : 1));
    p->j.p1a[it] = (int16_t *)malloc(sizeof(int16_t) * (cap1 ? cap1 : 1));
    p->j.p1c[it] = (int16_t *)malloc(sizeof(int16_t) * (cap1 ? cap1 : 1));
    p->j.p1ph[it] = (int8_t *)malloc(cap1 ? cap1 : 1);
    int64_t e = 0;
    p->j.p1ptr[it][0] = 0;
    for (int64_t f = 0; f < n; f++) {
      uint64_t fm = o->mask[it][f];
      for (uint64_t ta = fm; ta; ta &= ta - 1) {
        int a = __builtin_ctzll(ta);
        uint64_t m1 = fm & ~(1ULL << a);
        for (int c = 0; c < ns; c++) {
          if (c != a && ((m1 >> c) & 1)) continue;
          uint64_t im = m1 | (1ULL << c);
          int64_t i = map_get(&mp, im);
          if (i < 0) continue;
          p->j.p1src[it][e] = (int32_t)i;
          p->j.p1a[it][e] = (int16_t)a;
          p->j.p1c[it][e] = (int16_t)c;
          p->j.p1ph[it][e] = (int8_t)(fsign(im, c) * fsign(im & ~(1ULL << c), a));
          e++;
        }
      }
      p->j.p1ptr[it][f + 1] = e;
    }
    /* group each target's jumps by Delta = m_a - m_c (insertion sort) */
    {
      int dmax = 0;
      for (int q = 0; q < ns; q++) if (o->sm2[q] > dmax) dmax = o->sm2[q];
      p->j.dmmin = -2 * dmax;
      p->j.ndm = 2 * dmax + 1;
      const int ndm = p->j.ndm;
      p->j.p1dptr[it] = (int64_t *)malloc(sizeof(int64_t) * (n * (ndm + 1) + 1));
      for (int64_t f = 0; f < n; f++) {
        int64_t b = p->j.p1ptr[it][f], en = p->j.p1ptr[it][f + 1];
        for (int64_t q = b + 1; q < en; q++) {
          int32_t s0 = p->j.p1src[it][q]; int16_t a0 = p->j.p1a[it][q], c0 = p->j.p1c[it][q]; int8_t h0 = p->j.p1ph[it][q];
          int d0 = o->sm2[a0] - o->sm2[c0];
          int64_t r = q - 1;
          while (r >= b && o->sm2[p->j.p1a[it][r]] - o->sm2[p->j.p1c[it][r]] > d0) {
            p->j.p1src[it][r + 1] = p->j.p1src[it][r]; p->j.p1a[it][r + 1] = p->j.p1a[it][r];
            p->j.p1c[it][r + 1] = p->j.p1c[it][r]; p->j.p1ph[it][r + 1] = p->j.p1ph[it][r];
            r--;
          }
          p->j.p1src[it][r + 1] = s0; p->j.p1a[it][r + 1] = a0; p->j.p1c[it][r + 1] = c0; p->j.p1ph[it][r + 1] = h0;
        }
        int64_t q = b;
        for (int k = 0; k <= ndm; k++) {
          int dk = p->j.dmmin + 2 * k;  /* group k holds Delta == dk (odd k slots stay empty) */
          while (q < en && o->sm2[p->j.p1a[it][q]] - o->sm2[p->j.p1c[it][q]] < dk) q++;
          p->j.p1dptr[it][f * (ndm + 1) + k] = q;
        }
        p->j.p1dptr[it][f * (ndm + 1) + ndm] = en;
      }
    }
    /* two-body a+_i a+_j a_l a_k, listed by target with value*phase */
    const float *v = it == 0 ? o->vpp : o->vnn;
    int64_t cap2 = 0;
    for (int64_t f = 0; f < n; f++) cap2 += (int64_t)(np * (np - 1) / 2) * (npx > 0 ? npx : 1);
    if (cap2 > 400000000LL) cap2 = 400000000LL;
    p->j.p2ptr[it] = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    int64_t cap = 1024;
    p->j.p2src[it] = (int32_t *)malloc(sizeof(int32_t) * cap);
    p->j.p2val[it] = (float *)malloc(sizeof(float) * cap);
    e = 0;
    p->j.p2ptr[it][0] = 0;
    for (int64_t f = 0; f < n; f++) {
      uint64_t fm = o->mask[it][f];
      if (np >= 2)
        for (uint64_t ti = fm; ti; ti &= ti - 1) {
          int i = __builtin_ctzll(ti);
          for (uint64_t tj = fm & ((1ULL << i) - 1); tj; tj &= tj - 1) {
            int j = __builtin_ctzll(tj);
            uint64_t m2 = fm & ~(1ULL << i) & ~(1ULL << j);
            int mij = o->sm2[i] + o->sm2[j];
            int cidx = i * (i - 1) / 2 + j;
            for (int k = 1; k < ns; k++) {
              if ((m2 >> k) & 1) continue;
              for (int l = 0; l < k; l++) {
                if ((m2 >> l) & 1) continue;
                if (o->sm2[k] + o->sm2[l] != mij) continue;
                float me = v[(size_t)cidx * npx + (k * (k - 1) / 2 + l)];
                if (me == 0.0f) continue;
                uint64_t sm = m2 | (1ULL << k) | (1ULL << l);
                int sg = fsign(sm, k);
                uint64_t s1 = sm & ~(1ULL << k);
                sg *= fsign(s1, l);
                uint64_t s2 = s1 & ~(1ULL << l);
                sg *= fsign(s2, j);
                sg *= fsign(s2 | (1ULL << j), i);
                int64_t src = map_get(&mp, sm);
                if (src < 0) continue;
                if (e == cap) {
                  cap *= 2;
                  p->j.p2src[it] = (int32_t *)realloc(p->j.p2src[it], sizeof(int32_t) * cap);
                  p->j.p2val[it] = (float *)realloc(p->j.p2val[it], sizeof(float) * cap);
                }
                p->j.p2src[it][e] = (int32_t)src;
                p->j.p2val[it][e] = me * (float)sg;
                e++;
              }
            }
          }
        }
      p->j.p2ptr[it][f + 1] = e;
    }
    free(mp.keys);
    free(mp.vals);
  }
  p->j.built = 1;
  (void)g_jumps_dummy;
  return p;
}

void bso_port_free(bso_port_t *p) {
  if (!p) return;
  for (int it = 0; it < 2; it++) {
    free(p->j.p1ptr[it]); free(p->j.p1src[it]); free(p->j.p1a[it]); free(p->j.p1c[it]); free(p->j.p1ph[it]);
    free(p->j.p1dptr[it]);
    free(p->j.p2ptr[it]); free(p->j.p2src[it]); free(p->j.p2val[it]);
  }
  free(p);
}

/* Number of updates bso_port_apply performs for proton rows lo..hi. */
int64_t bso_port_count(const bso_port_t *p, int64_t lo, int64_t hi) {
  const bso_t *o = p->o;
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
  for (int64_t f = lo; f < hi; f++) {
    int cn = o->conj[0][o->sdsec[0][f]];
    int64_t L = o->seccount[1][cn], n0 = o->secfirst[1][cn];
    int64_t u = o->call_spe ? L : 0;
    u += (p->j.p2ptr[0][f + 1] - p->j.p2ptr[0][f]) * L;
    u += p->j.p2ptr[1][n0 + L] - p->j.p2ptr[1][n0];
    if (o->Z > 0 && o->N > 0) {
      const int ndm = p->j.ndm;
      for (int64_t e = p->j.p1ptr[0][f]; e < p->j.p1ptr[0][f + 1]; e++) {
        const int dm = o->sm2[p->j.p1a[0][e]] - o->sm2[p->j.p1c[0][e]];
        const int k = (-dm - p->j.dmmin) / 2;
        for (int64_t c = 0; c < L; c++)
          u += p->j.p1dptr[1][(n0 + c) * (ndm + 1) + k + 1] - p->j.p1dptr[1][(n0 + c) * (ndm + 1) + k];
      }
    }
    total += u;
  }
  return total;
}

/* y[rows of proton SDs lo..hi) = H x, FP32 throughout; returns the number of
 * updates performed.  Threads take whole proton rows. */
int64_t bso_port_apply(const bso_port_t *p, const float *x, float *y, int64_t lo, int64_t hi) {
  const bso_t *o = p->o;
  const int ns = o->nsps;
  const int npn = ns * ns;
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 2) reduction(+ : total)
  for (int64_t f = lo; f < hi; f++) {
    int sp = o->sdsec[0][f];
    int cn = o->conj[0][sp];
    int64_t L = o->seccount[1][cn], n0 = o->secfirst[1][cn];
    float *yr = y + o->start[0][f];
    int64_t upd = 0;
    for (int64_t c = 0; c < L; c++) yr[c] = 0.0f;
    if (o->call_spe) {
      float ep = 0.0f;
      for (uint64_t t = o->mask[0][f]; t; t &= t - 1) ep += o->pspe[o->sorb[__builtin_ctzll(t)]];
      const float *xr = x + o->start[0][f];
      for (int64_t c = 0; c < L; c++) {
        float en = 0.0f;
        for (uint64_t t = o->mask[1][n0 + c]; t; t &= t - 1) en += o->nspe[o->sorb[__builtin_ctzll(t)]];
        yr[c] += (ep + en) * xr[c];
      }
      upd += L;
    }
    /* PP: conjugate neutron runs are contiguous (bapplyhlibOMP.f90:1090-1105) */
    for (int64_t e = p->j.p2ptr[0][f]; e < p->j.p2ptr[0][f + 1]; e++) {
      const float v = p->j.p2val[0][e];
      const float *xr = x + o->start[0][p->j.p2src[0][e]];
      for (int64_t c = 0; c < L; c++) yr[c] += v * xr[c];
      upd += L;
    }
    /* NN within the row (:1032-1042) */
    {
      const float *xr = x + o->start[0][f];
      for (int64_t c = 0; c < L; c++) {
        int64_t n = n0 + c;
        float s = 0.0f;
        for (int64_t e = p->j.p2ptr[1][n]; e < p->j.p2ptr[1][n + 1]; e++)
          s += p->j.p2val[1][e] * xr[p->j.p2src[1][e] - n0];
        yr[c] += s;
        upd += p->j.p2ptr[1][n + 1] - p->j.p2ptr[1][n];
      }
    }
    /* PN: proton jump x neutron jump (:1253-1277) */
    if (o->Z > 0 && o->N > 0) {
      for (int64_t e = p->j.p1ptr[0][f]; e < p->j.p1ptr[0][f + 1]; e++) {
        const int64_t i = p->j.p1src[0][e];
        const int a = p->j.p1a[0][e], cc = p->j.p1c[0][e];
        const float php = (float)p->j.p1ph[0][e];
        const int dm = o->sm2[a] - o->sm2[cc];
        const int ci = o->conj[0][o->sdsec[0][i]];
        const int64_t n0i = o->secfirst[1][ci];
        const float *xr = x + o->start[0][i];
        const int k = (-dm - p->j.dmmin) / 2;
        const int ndm = p->j.ndm;
        for (int64_t c = 0; c < L; c++) {
          int64_t n = n0 + c;
          float s = 0.0f;
          const int64_t q0 = p->j.p1dptr[1][n * (ndm + 1) + k], q1 = p->j.p1dptr[1][n * (ndm + 1) + k + 1];
          for (int64_t q = q0; q < q1; q++) {
            const int b = p->j.p1a[1][q], d = p->j.p1c[1][q];
            const float me = o->vpn[(size_t)(a * ns + b) * npn + (cc * ns + d)];
            s += me * (float)p->j.p1ph[1][q] * xr[p->j.p1src[1][q] - n0i];
            upd++;
          }
          yr[c] += php * s;
        }
      }
    }
    total += upd;
  }
  return total;
}

/* ------------------------------------------------------- one-body operators */
/* readin1bodyop (src/bdenslib1.f90:1823-2071): .opme file, first line iso /
 * pns / xpn, number of orbits, the orbit list (checked against the space),
 * "J T" (iso) or "J" (pns, xpn), then "a b value" (iso, xpn) or "a b vp vn"
 * (pns) lines to the end of the file (at most 10000); xpn lines that connect
 * a proton and a neutron orbit are skipped.  decouple1bodyop
 * (src/bdenslib1.f90:2444-2597): the M = 0 component in the m-scheme, in
 * float arithmetic in the order written there.
 * op_p, op_n: nsps x nsps, [a * nsps + b] is the coefficient of a+_a a_b.
 * Returns 0, or a negative code: -1 file, -2 format, -3 orbit mismatch,
 * -4 angular momentum out of range. */
int bso_opme(const bso_t *o, const char *path, double *op_p, double *op_n, int *jop_out, int *top_out) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[512];
  if (!fgets(line, sizeof line, f)) { fclose(f); return -2; }
  int pn = 0, xpn = 0;
  if (!strncmp(line, "iso", 3)) pn = 0;
  else if (!strncmp(line, "pns", 3)) pn = 1;
  else if (!strncmp(line, "xpn", 3)) { pn = 1; xpn = 1; }
  else { fclose(f); return -2; }
  int n;
  if (fscanf(f, "%d", &n) != 1 || n != o->norb) { fclose(f); return -3; }
  for (int i = 1; i <= n; i++) {
    int i1, i2, nr, l;
    float xj;
    if (xpn) {
      if (fscanf(f, "%d %d %d %d %f", &i1, &i2, &nr, &l, &xj) != 5 || i2 != i + n) { fclose(f); return -3; }
    } else if (fscanf(f, "%d %d %d %f", &i1, &nr, &l, &xj) != 4) { fclose(f); return -3; }
    if (i1 != i || nr != o->on[i - 1] || l != o->ol[i - 1] || (int)lrintf(2.0f * xj) != o->oj2[i - 1]) {
      fclose(f);
      return -3;
    }
  }
  int jop, top = 2;
  if (fscanf(f, "%d", &jop) != 1) { fclose(f); return -2; }
  if (!pn && fscanf(f, "%d", &top) != 1) { fclose(f); return -2; }
  float op[MAXORB][MAXORB], pop[MAXORB][MAXORB], nop[MAXORB][MAXORB];
  memset(op, 0, sizeof op);
  memset(pop, 0, sizeof pop);
  memset(nop, 0, sizeof nop);
  for (int k = 0; k < 10000; k++) {
    int a, b;
    float x = 0.0f, y = 0.0f;
    if (pn && !xpn) {
      if (fscanf(f, "%d %d %f %f", &a, &b, &x, &y) != 4) break;
    } else if (fscanf(f, "%d %d %f", &a, &b, &x) != 3) {
      break;
    }
    int pv = 1, nv = 1;
    if (xpn) {
      if ((a > n) != (b > n)) continue; /* charge changing: skipped */
      if (a > n) { a -= n; b -= n; y = x; x = 0.0f; pv = 0; }
      else { y = 0.0f; nv = 0; }
    }
    if (a < 1 || a > n || b < 1 || b > n) { fclose(f); return -2; }
    int ja = o->oj2[a - 1], jb = o->oj2[b - 1];
    if (ja + jb < 2 * jop || abs(ja - jb) > 2 * jop) { fclose(f); return -4; }
    if (pn) {
      if (o->Z > 0 && pv) pop[a - 1][b - 1] = x;
      if (o->N > 0 && nv) nop[a - 1][b - 1] = y;
    } else {
      op[a - 1][b - 1] = x;
    }
  }
  fclose(f);
  const int ns = o->nsps;
  for (int i = 0; i < ns * ns; i++) op_p[i] = op_n[i] = 0.0;
  for (int it = 0; it < 2; it++) {
    double *out = it == 0 ? op_p : op_n;
    for (int sa = 0; sa < ns; sa++)
      for (int sb = 0; sb < ns; sb++) {
        const int a = o->sorb[sa], b = o->sorb[sb];
        const int ja = o->sj2[sa], ma = o->sm2[sa], jb = o->sj2[sb], mb = o->sm2[sb];
        const float v = pn ? (it == 0 ? pop[a][b] : nop[a][b]) : op[a][b];
        if (pn && (it == 0 ? o->Z : o->N) == 0) continue;
        if (v == 0.0f || ma != mb) continue;
        if (jop > (ja + jb) / 2 || jop < abs(ja - jb) / 2) continue;
        int iphase = ipow_m1((jb - mb) / 2);
        float r;
        if (pn) {
          r = (float)iphase * bso_cleb(ja, ma, jb, -mb, 2 * jop, ma - mb) * v / sqrtf(2.0f * jop + 1.0f);
        } else {
          if (it == 1) iphase = -iphase;
          const float ct = it == 0 ? bso_cleb(1, 1, 1, -1, 2 * top, 0) : bso_cleb(1, -1, 1, 1, 2 * top, 0);
          r = (float)iphase * bso_cleb(ja, ma, jb, -mb, 2 * jop, ma - mb) * v / sqrtf(2.0f * top + 1.0f) /
              sqrtf(2.0f * jop + 1.0f) * ct;
        }
        out[sa * ns + sb] = r;
      }
  }
  *jop_out = jop;
  *top_out = top;
  return 0;
}

/* y += O x for the one-body operator sum_ab op_p[a,b] a+_a a_b (protons) +
 * op_n (neutrons), by brute force over the basis: for every source state and
 * every a+_a a_b that does not annihilate it (src/bdenslib3.f90:319-427 and
 * its neutron twin apply the same operator through jumps).  FP64. */
void bso_apply_onebody(const bso_t *o, const double *op_p, const double *op_n, const double *x, double *y) {
  const int ns = o->nsps;
  sdmap_t mp[2];
  for (int it = 0; it < 2; it++) map_build(&mp[it], o->mask[it], o->nsd[it]);
  for (int64_t ip = 0; ip < o->nsd[0]; ip++) {
    const int sn = o->conj[0][o->sdsec[0][ip]];
    for (int64_t in = o->secfirst[1][sn]; in < o->secfirst[1][sn] + o->seccount[1][sn]; in++) {
      const int64_t s = o->start[0][ip] + o->start[1][in];
      if (x[s] == 0.0) continue;
      for (int it = 0; it < 2; it++) {
        const uint64_t m0 = it == 0 ? o->mask[0][ip] : o->mask[1][in];
        const double *op = it == 0 ? op_p : op_n;
        for (uint64_t tb = m0; tb; tb &= tb - 1) {
          const int b = __builtin_ctzll(tb);
          const int s1 = fsign(m0, b);
          const uint64_t m1 = m0 & ~(1ULL << b);
          for (int a = 0; a < ns; a++) {
            if (m1 >> a & 1) continue;
            const double v = op[a * ns + b];
            if (v == 0.0) continue;
            const uint64_t m2 = m1 | (1ULL << a);
            const int sg = s1 * fsign(m1, a);
            int64_t t;
            if (it == 0) t = o->start[0][map_get(&mp[0], m2)] + o->start[1][in];
            else t = o->start[0][ip] + o->start[1][map_get(&mp[1], m2)];
            y[t] += v * sg * x[s];
          }
        }
      }
    }
  }
  for (int it = 0; it < 2; it++) {
    free(mp[it].keys);
    free(mp[it].vals);
  }
}
