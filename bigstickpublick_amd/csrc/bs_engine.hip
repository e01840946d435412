// bs_engine.hip -- MI355X (gfx950) engine for the M-scheme Hamiltonian and
// its Lanczos driver, behind the C ABI of include/bigstick_amd.h.
//
// Data layout in HBM.  The Lanczos vector is the reference's basis order
// (src/bbasislib.f90:539-677): one dense block per proton sector s,
// X_s[nP_s][nN_c(s)] row-major, neutrons innermost.  Every kernel that
// applies H is output-stationary: one workgroup owns one proton-SD row of the
// output, accumulates SPE + PP + NN + PN for that row in registers and writes
// it once -- no atomics, no pre-zeroed output, no per-thread scratch copies
// (the reference instead splits output ranges between threads,
// src/bapplyhlibOMP.f90:34-868).
//
//   PP  (proton two-body, sector-diagonal): y_f += v * x_i row AXPYs,
//       coalesced along the neutron index.
//   NN  (neutron two-body): gathers inside row f through the neutron CSR.
//   PN  (one-body x one-body): for each proton jump i->f of Delta class k,
//       y_f[n] += ph_p * sum_{neutron jumps n'->n of class -k}
//                 ph_n * W_k[pp][nn] * x_i[n'], with W_k the dense block of
//       pn elements for that class (<= 60x60 for the pf shell).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bigstick_amd.h"
#include "bs_tables.hpp"
#include <rocsolver/rocsolver.h>
#include <atomic>
#include <mutex>
#include <thread>

// Device scratch of bs_lanczos_noalloc: one array per GPU, held from the
// library's load until the process ends.  In host code its name is only a
// stand-in for the symbol; the device address comes from
// hipGetSymbolAddress (scratch_ptr below) and is the only thing passed to
// kernels or copies.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_err;

int set_err(const std::string& m, int code = -1) {
  g_err = m;
  return code;
}

#define HIP_OK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kBlock = 256;

struct DevTables {
  int64_t dim;
  int nP, nN, ndelta;
  int call_spe;
  int lmax;                  // longest neutron row
  int wcols;                 // LDS row stride of the pn element block (max class size + 1)
  const int64_t* pstart;     // [nP] row offset of proton SD
  const float* pspe_sd;      // [nP]
  const float* nspe_sd;      // [nN]
  // tiles: up to R consecutive proton rows of one sector; tile_map sends
  // workgroup b to a tile (-1: idle slot)
  const int32_t* tile_f0;
  const int32_t* tile_nr;
  const int32_t* tile_map;
  // PP per tile: the union of the tile rows' source rows (sorted), each entry
  // the source SD | (mask of the tile rows it feeds) << 27 and R values
  // (zero where it does not feed a row)
  const int64_t* tpp_ptr;     // [ntiles + 1]
  const uint32_t* tpp_src;
  const float* tpp_val32;     // [entry * R + r]
  const double* tpp_val64;
  const int32_t* psec;       // [nP] proton sector of each SD
  const int32_t* pconj;      // [nsecP] conjugate neutron sector
  const int32_t* nlen;       // [nsecN] SDs in neutron sector
  const int32_t* nfirst;     // [nsecN] first SD of neutron sector
  const int32_t* srcsec;     // [nsecP * ndelta] proton source sector for class k (-1: none)
  // PP merged CSR by target proton SD
  const int64_t* pp_ptr;
  const int32_t* pp_src;
  const float* pp_val32;
  const double* pp_val64;
  // NN merged, ELL per neutron sector: slot e of column n at nn_base[c] + e*L_c + n
  const int64_t* nn_base;
  const int32_t* nn_emax;
  const int32_t* nn_src;     // local source column
  const float* nn_val32;
  const double* nn_val64;
  // proton one-body: class pointers per row [nP][ndelta+1] (absolute), entries
  const int32_t* pcls;
  const int32_t* p1_src;     // proton source SD (global)
  const int16_t* p1_meta;    // pp_local | (phase<0) << 8
  // neutron one-body, ELL per (target neutron sector c, class k):
  // slot e of column n at n1_base[c*ndelta+k] + e*L_c + n; packed
  // src_local (16b) | nn_local (8b) << 16 | (phase<0) << 24; padding slots
  // point at the zero column nn_local = size of the class
  const int64_t* n1_base;
  const int32_t* n1_emax;
  const uint32_t* n1_pk;      // src | nl << 16 | sign << 24
  const uint32_t* n1_pkv[3][2];  // per G (1, 2, 4) and T (f32, f64): LDS byte offsets, see k_apply_tiles;
                                 // quad layout: slot e of column n at n1q_base[c*ndelta+k] + ((e/4)*L_c + n)*4 + e%4
  const int64_t* n1q_base;
  const uint8_t* n1_chunk;       // [(c*ndelta + k)*64 + j]: entries used by columns 64j..64j+63
  const uint32_t* n1_pkr;        // src | (2*op + sign) << 16 (multi-vector kernels)
  // pn element blocks: class k block is [size_k][size_-k] at woff[k]
  const float* W;
  const int32_t* woff;
  const int32_t* wrows;      // size_k
  const int32_t* wstride;    // size_-k
  const int32_t* negk;
};

template <typename T> __device__ __forceinline__ T vload(const float* a, const double* b, int64_t i);
template <> __device__ __forceinline__ float vload<float>(const float* a, const double*, int64_t i) { return a[i]; }
template <> __device__ __forceinline__ double vload<double>(const float*, const double* b, int64_t i) { return b[i]; }

// Raw buffer loads: a uniform (scalar) descriptor and byte offset plus a
// per-lane 32-bit byte offset, so table walks cost no vector address math.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ uint32_t buf_u32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
__device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 buf_u32x4(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
}

// The neutron one-body entries of one PN batch for the thread's CPT columns
// (column q: entry count Eq[q], wave-uniform), four entries to a 16-byte
// load ([e/4][column][4], qstep bytes per quad row).  The next quad (the
// rest of this column, or the next column's first) is in flight while the
// current one is applied; load(pk) reads one entry's operands, apply(q, e) adds
// its products.  PAIRS issues a quad as two pairs (half the registers for
// operands in flight).
template <int CPT, bool PAIRS = false, typename Load, typename Apply>
__device__ __forceinline__ void walk_quads(__amdgpu_buffer_rsrc_t rpk, const uint32_t (&cl)[CPT], const int (&Eq)[CPT],
                                           uint32_t qstep, Load&& load, Apply&& apply) {
  u32x4 pq = {0u, 0u, 0u, 0u};
  if (Eq[0] > 0) pq = buf_u32x4(rpk, cl[0] * 16u, 0u);
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int Ee = Eq[q];
    u32x4 pn = {0u, 0u, 0u, 0u};
    if (q + 1 < CPT) {
      if (Eq[q + 1] > 0) pn = buf_u32x4(rpk, cl[q + 1] * 16u, 0u);
    }
    for (int e = 0; e < Ee; e += 4) {
      const u32x4 cur = pq;
      if (e + 4 < Ee) pq = buf_u32x4(rpk, cl[q] * 16u, (uint32_t)((e >> 2) + 1) * qstep);
      // the reads of a quad (or of a pair, for a tail of one or two) are
      // issued together, then its products; slots past Ee hold padding
      // entries (a zero coefficient row), so a whole pair or quad is applied
      if (Ee - e > 2) {
        if constexpr (PAIRS) {
          {
            const auto l0 = load(cur.x), l1 = load(cur.y);
            apply(q, l0);
            apply(q, l1);
          }
          const auto l2 = load(cur.z), l3 = load(cur.w);
          apply(q, l2);
          apply(q, l3);
        } else {
          const auto l0 = load(cur.x), l1 = load(cur.y), l2 = load(cur.z), l3 = load(cur.w);
          apply(q, l0);
          apply(q, l1);
          apply(q, l2);
          apply(q, l3);
        }
      } else {
        const auto l0 = load(cur.x), l1 = load(cur.y);
        apply(q, l0);
        apply(q, l1);
      }
    }
    pq = pn;
  }
}

// 16-byte vector of T for streaming passes
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; static constexpr int n = 4; };
template <> struct Vec16<double> { using type = double2; static constexpr int n = 2; };

// G values of one LDS column, read with one wide LDS load.
template <typename T, int G>
struct alignas(sizeof(T) * G) VecG {
  T v[G];
};

// One neutron entry of a PN batch, as read from LDS: the G staged source
// values at its source column and the G signed coefficients of its operator.
template <typename T, int G>
struct PnEntry {
  VecG<T, G> x;
  VecG<float, G> w;
};

// One workgroup per tile of up to R proton rows of one sector; BLOCK threads,
// CPT columns per thread (column = tid + q*BLOCK, clamped to the row for
// loads; only columns < L are stored).  Static LDS, in bytes:
//   [0, kWbBytes*G)    Wb[2*nl + s][g] = (-1)^s * phase_g * W_k[pp_g][nl],
//                      the batch's coefficients with the neutron jump's sign
//                      folded in (rows 2*wsd, 2*wsd + 1: zero, for padding)
//   [kWbBytes*G, ...)  staged x rows, column-major: [col][R] for the NN part,
//                      [col][G] for each PN batch
// The neutron one-body entries of the PN part are pre-scaled LDS byte
// offsets for this G and T: x offset | (Wb offset << 17), so an entry costs
// one AND, one shift, two LDS reads and G FMAs.
constexpr int kWbRows = 128;                    // 2 * (max class size + 1) <= 128
constexpr int kWbBytes = kWbRows * (int)sizeof(float);
constexpr int kMaxCols = 4096;                  // longest neutron sector (staged row length)

// PF: the PN source rows of the next batch are loaded into registers while
// the current batch is computed from LDS (register double buffering).
template <typename T, int BLOCK, int CPT, int R, int G, bool PF>
__global__ __launch_bounds__(BLOCK, PF ? 4 : 1) void k_apply_tiles(DevTables tb, const T* __restrict__ x,
                                                       T* __restrict__ y, int accumulate) {
  constexpr int S = R > G ? R : G;
  constexpr int XS = kWbBytes * G;  // byte offset of the staged rows
  // static, so that every LDS address is a register offset plus a constant
  __shared__ __align__(16) unsigned char smem[XS + S * kMaxCols * sizeof(T)];
  T* xs = reinterpret_cast<T*>(smem + XS);
  const int tid = threadIdx.x;
  const int tile = tb.tile_map[blockIdx.x];
  if (tile < 0) return;  // padding slot of the XCD map (uniform per workgroup)
  const int f0 = tb.tile_f0[tile];
  const int nr = tb.tile_nr[tile];
  const int sp = tb.psec[f0];
  const int cn = tb.pconj[sp];
  const int L = tb.nlen[cn];
  const int nf0 = tb.nfirst[cn];
  uint32_t cl[CPT];  // clamped columns: loads never leave the row
#pragma unroll
  for (int q = 0; q < CPT; q++) cl[q] = (uint32_t)min(tid + q * BLOCK, L - 1);
  T acc[R][CPT];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int q = 0; q < CPT; q++) acc[r][q] = T(0);

  // ---- SPE (only when the single-particle energies are not subsumed)
  if (tb.call_spe) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r < nr) {
        const T ep = (T)tb.pspe_sd[f0 + r];
        const int64_t bf = tb.pstart[f0 + r];
#pragma unroll
        for (int q = 0; q < CPT; q++) acc[r][q] = fma(ep + (T)tb.nspe_sd[nf0 + cl[q]], x[bf + cl[q]], acc[r][q]);
      }
    }
  }
  // ---- PP: row AXPYs over the union of the tile rows' source rows (sorted
  // by source SD): each source row is read once, with coalesced buffer loads
  // through a scalar descriptor, and applied to every tile row its mask
  // names (wave-uniform branches); four source rows' loads in flight
  {
    const int64_t e0 = tb.tpp_ptr[tile], e1 = tb.tpp_ptr[tile + 1];
    int64_t e = e0;
    for (; e + 4 <= e1; e += 4) {
      __amdgpu_buffer_rsrc_t rr[4];
      uint32_t mk[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t sm = tb.tpp_src[e + u];
        mk[u] = sm >> 27;
        rr[u] = buf_rsrc(x + tb.pstart[sm & 0x7ffffffu]);
      }
      T xx[4][CPT];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int q = 0; q < CPT; q++) {
          if constexpr (sizeof(T) == 4) xx[u][q] = __uint_as_float(buf_u32(rr[u], cl[q] * 4u, 0u));
          else xx[u][q] = buf_f64(rr[u], cl[q] * 8u, 0u);
        }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int r = 0; r < R; r++)
          if ((mk[u] >> r) & 1u) {
            const T v = vload<T>(tb.tpp_val32, tb.tpp_val64, (e + u) * R + r);
#pragma unroll
            for (int q = 0; q < CPT; q++) acc[r][q] = fma(v, xx[u][q], acc[r][q]);
          }
    }
    for (; e < e1; e++) {
      const uint32_t sm = tb.tpp_src[e];
      const __amdgpu_buffer_rsrc_t ra = buf_rsrc(x + tb.pstart[sm & 0x7ffffffu]);
      T xa[CPT];
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        if constexpr (sizeof(T) == 4) xa[q] = __uint_as_float(buf_u32(ra, cl[q] * 4u, 0u));
        else xa[q] = buf_f64(ra, cl[q] * 8u, 0u);
      }
#pragma unroll
      for (int r = 0; r < R; r++)
        if (((sm >> 27) >> r) & 1u) {
          const T v = vload<T>(tb.tpp_val32, tb.tpp_val64, e * R + r);
#pragma unroll
          for (int q = 0; q < CPT; q++) acc[r][q] = fma(v, xa[q], acc[r][q]);
        }
    }
  }
  // ---- NN: the tile's rows staged transposed ([col][R]); one wide LDS read
  // per entry serves every row of the tile
  {
    VecG<T, R>* xr = reinterpret_cast<VecG<T, R>*>(xs);
    {
      // one column of all R rows per thread, one wide conflict-free LDS store
      int64_t bfr[R];
#pragma unroll
      for (int r = 0; r < R; r++) bfr[r] = r < nr ? tb.pstart[f0 + r] : -1;
      for (int c = tid; c < L; c += BLOCK) {
        VecG<T, R> v;
#pragma unroll
        for (int r = 0; r < R; r++) v.v[r] = bfr[r] >= 0 ? x[bfr[r] + c] : T(0);
        xr[c] = v;
      }
    }
    __syncthreads();
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(tb.nn_src + tb.nn_base[cn]);
    const __amdgpu_buffer_rsrc_t rv =
        sizeof(T) == 4 ? buf_rsrc(tb.nn_val32 + tb.nn_base[cn]) : buf_rsrc(tb.nn_val64 + tb.nn_base[cn]);
    const int E = tb.nn_emax[cn];
    // entry-major over the thread's columns: CPT independent loads in flight
    for (int e = 0; e < E; e++) {
      const uint32_t so = (uint32_t)e * (uint32_t)L * 4u;
      int sl[CPT];
      T v[CPT];
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        sl[q] = (int)buf_u32(rs, cl[q] * 4u, so);
        if constexpr (sizeof(T) == 4) v[q] = __uint_as_float(buf_u32(rv, cl[q] * 4u, so));
        else v[q] = buf_f64(rv, cl[q] * 8u, so * 2u);
      }
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        const VecG<T, R> xv = xr[sl[q]];
#pragma unroll
        for (int r = 0; r < R; r++) acc[r][q] = fma(v[q], xv.v[r], acc[r][q]);
      }
    }
  }
  // ---- PN: per Delta class, batches of G source rows staged transposed
  const int gsel = G == 1 ? 0 : (G == 2 ? 1 : 2);
  const uint32_t* pkt = sizeof(T) == 4 ? tb.n1_pkv[gsel][0] : tb.n1_pkv[gsel][1];
  const int nd = tb.ndelta, nd1 = nd + 1;
  float* wb = reinterpret_cast<float*>(smem);
  constexpr int kWbPer = (kWbRows * G + BLOCK - 1) / BLOCK;
  if constexpr (R == G) {
    // staged columns wider than 16 bytes (FP64, G = 4) are split in two
    // halves in two LDS arrays, so that consecutive lanes' reads stride 16
    // bytes over all 64 banks; xo is the table's offset (column * G *
    // sizeof(T))
    constexpr bool SPLIT = G * sizeof(T) > 16;
    constexpr int HB = G * (int)sizeof(T) / 2;          // bytes per half column
    constexpr size_t HALF = (size_t)kMaxCols * HB;      // offset of the second array
    using HV = VecG<T, G / 2>;
    auto store_col = [&](int c, const VecG<T, G>& v) {
      if constexpr (SPLIT) {
        HV lo, hi;
#pragma unroll
        for (int g = 0; g < G / 2; g++) {
          lo.v[g] = v.v[g];
          hi.v[g] = v.v[g + G / 2];
        }
        *reinterpret_cast<HV*>(smem + XS + (size_t)c * HB) = lo;
        *reinterpret_cast<HV*>(smem + XS + HALF + (size_t)c * HB) = hi;
      } else {
        *reinterpret_cast<VecG<T, G>*>(xs + (size_t)c * G) = v;
      }
    };
    auto load_col = [&](uint32_t xo) {
      VecG<T, G> v;
      if constexpr (SPLIT) {
        const HV lo = *reinterpret_cast<const HV*>(smem + XS + xo / 2);
        const HV hi = *reinterpret_cast<const HV*>(smem + XS + HALF + xo / 2);
#pragma unroll
        for (int g = 0; g < G / 2; g++) {
          v.v[g] = lo.v[g];
          v.v[g + G / 2] = hi.v[g];
        }
      } else {
        v = *reinterpret_cast<const VecG<T, G>*>(smem + XS + xo);
      }
      return v;
    };
    // one staged row per tile row: batch b of class k holds jump b of every
    // tile row (slot g = row g; rows with fewer jumps leave their slot empty,
    // zero coefficients), so one walk of the neutron entries serves all R
    // rows and each entry's G products go to G independent accumulators
    for (int k = 0; k < nd; k++) {
      int p0[R], nj[R];
      int nbmax = 0;
#pragma unroll
      for (int r = 0; r < R; r++) {
        p0[r] = nj[r] = 0;
        if (r < nr) {
          p0[r] = tb.pcls[(int64_t)(f0 + r) * nd1 + k];
          nj[r] = tb.pcls[(int64_t)(f0 + r) * nd1 + k + 1] - p0[r];
          nbmax = max(nbmax, nj[r]);
        }
      }
      if (nbmax == 0) continue;
      const int kn = tb.negk[k];
      const int si = tb.srcsec[sp * nd + k];
      const int Li = tb.nlen[tb.pconj[si]];
      const int wsd = tb.wstride[k];
      const float* Wk = tb.W + tb.woff[k];
      const __amdgpu_buffer_rsrc_t rpk = buf_rsrc(pkt + tb.n1q_base[(int64_t)cn * nd + kn]);
      const uint8_t* ch = tb.n1_chunk + ((int64_t)cn * nd + kn) * 64 + __builtin_amdgcn_readfirstlane(tid >> 6);
      int Eq[CPT];
#pragma unroll
      for (int q = 0; q < CPT; q++) Eq[q] = ch[q * (BLOCK / 64)];
      VecG<T, G> pre[PF ? CPT : 1];
      if constexpr (PF) {
        int64_t bi[G];
#pragma unroll
        for (int g = 0; g < G; g++) bi[g] = nj[g] > 0 ? tb.pstart[tb.p1_src[p0[g]]] : -1;
#pragma unroll
        for (int q = 0; q < CPT; q++) {
          const int c = tid + q * BLOCK;
#pragma unroll
          for (int g = 0; g < G; g++) pre[q].v[g] = (c < Li && bi[g] >= 0) ? x[bi[g] + c] : T(0);
        }
      }
      for (int b = 0; b < nbmax; b++) {
        float wreg[kWbPer];
#pragma unroll
        for (int u = 0; u < kWbPer; u++) {
          const int i = tid + u * BLOCK;
          const int row = i / G, g = i - row * G, nl = row >> 1;
          float w = 0.0f;
          int njg = 0, p0g = 0;
#pragma unroll
          for (int gg = 0; gg < G; gg++)
            if (gg == g) {
              njg = nj[gg];
              p0g = p0[gg];
            }
          if (i < 2 * (wsd + 1) * G && b < njg && nl < wsd) {
            const int meta = tb.p1_meta[p0g + b];
            w = Wk[(meta & 0xff) * wsd + nl];
            if ((meta >> 8) ^ (row & 1)) w = -w;
          }
          wreg[u] = w;
        }
        __syncthreads();
        if constexpr (PF) {
#pragma unroll
          for (int q = 0; q < CPT; q++) {
            const int c = tid + q * BLOCK;
            if (c < Li) store_col(c, pre[q]);
          }
        } else {
          int64_t bi[G];
#pragma unroll
          for (int g = 0; g < G; g++) bi[g] = b < nj[g] ? tb.pstart[tb.p1_src[p0[g] + b]] : -1;
          for (int c = tid; c < Li; c += BLOCK) {
            VecG<T, G> v;
#pragma unroll
            for (int g = 0; g < G; g++) v.v[g] = bi[g] >= 0 ? x[bi[g] + c] : T(0);
            store_col(c, v);
          }
        }
#pragma unroll
        for (int u = 0; u < kWbPer; u++) {
          const int i = tid + u * BLOCK;
          if (i < 2 * (wsd + 1) * G) wb[i] = wreg[u];
        }
        __syncthreads();
        if constexpr (PF) {
          if (b + 1 < nbmax) {
            int64_t bi[G];
#pragma unroll
            for (int g = 0; g < G; g++) bi[g] = b + 1 < nj[g] ? tb.pstart[tb.p1_src[p0[g] + b + 1]] : -1;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
              const int c = tid + q * BLOCK;
#pragma unroll
              for (int g = 0; g < G; g++) pre[q].v[g] = (c < Li && bi[g] >= 0) ? x[bi[g] + c] : T(0);
            }
          }
        }
        // the PF forms issue pairs: the batch loaded ahead holds the registers
        // a quad's operands would take (FP64 at 1024 threads spills with quads)
        walk_quads<CPT, PF>(
            rpk, cl, Eq, (uint32_t)L * 16u,
            [&](uint32_t pk) {
              return PnEntry<T, G>{load_col(pk & 0x1ffffu),
                                   *reinterpret_cast<const VecG<float, G>*>(smem + (pk >> 17))};
            },
            [&](int q, const PnEntry<T, G>& en) {
#pragma unroll
              for (int g = 0; g < G; g++) acc[g][q] = fma((T)en.w.v[g], en.x.v[g], acc[g][q]);
            });
      }
    }
  } else
  for (int k = 0; k < nd; k++) {
    int p0[R], p1[R];
    int any = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      p0[r] = p1[r] = 0;
      if (r < nr) {
        p0[r] = tb.pcls[(int64_t)(f0 + r) * nd1 + k];
        p1[r] = tb.pcls[(int64_t)(f0 + r) * nd1 + k + 1];
        any |= (p1[r] > p0[r]);
      }
    }
    if (!any) continue;
    const int kn = tb.negk[k];
    const int si = tb.srcsec[sp * nd + k];
    const int Li = tb.nlen[tb.pconj[si]];
    const int wsd = tb.wstride[k];
    const float* Wk = tb.W + tb.woff[k];
    const __amdgpu_buffer_rsrc_t rpk = buf_rsrc(pkt + tb.n1q_base[(int64_t)cn * nd + kn]);
    // per 64-column chunk entry counts: the chunk of column tid + q*BLOCK is
    // q*(BLOCK/64) + wave, uniform across the wave
    const uint8_t* ch = tb.n1_chunk + ((int64_t)cn * nd + kn) * 64 + __builtin_amdgcn_readfirstlane(tid >> 6);
    int Eq[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) Eq[q] = ch[q * (BLOCK / 64)];
    // PF: columns tid + q*BLOCK (< Li <= CPT*BLOCK) of the G source rows of a
    // batch, loaded one batch ahead
    VecG<T, G> pre[PF ? CPT : 1];
    if constexpr (PF) {
      int fb = 0, fe = 0;
#pragma unroll
      for (int r = R - 1; r >= 0; r--)
        if (p1[r] > p0[r]) {
          fb = p0[r];
          fe = p1[r];
        }
      int64_t bi[G];
#pragma unroll
      for (int g = 0; g < G; g++) bi[g] = g < fe - fb ? tb.pstart[tb.p1_src[fb + g]] : -1;
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        const int c = tid + q * BLOCK;
#pragma unroll
        for (int g = 0; g < G; g++) pre[q].v[g] = (c < Li && bi[g] >= 0) ? x[bi[g] + c] : T(0);
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      for (int b0 = p0[r]; b0 < p1[r]; b0 += G) {
        const int nb = min(G, p1[r] - b0);
        // this batch's signed coefficients, loaded first so that their
        // latency overlaps the source-row staging (one or two per thread)
        float wreg[kWbPer];
#pragma unroll
        for (int u = 0; u < kWbPer; u++) {
          const int i = tid + u * BLOCK;
          const int row = i / G, g = i - row * G, nl = row >> 1;
          float w = 0.0f;
          if (i < 2 * (wsd + 1) * G && g < nb && nl < wsd) {
            const int meta = tb.p1_meta[b0 + g];
            w = Wk[(meta & 0xff) * wsd + nl];
            if ((meta >> 8) ^ (row & 1)) w = -w;
          }
          wreg[u] = w;
        }
        __syncthreads();
        if constexpr (PF) {
#pragma unroll
          for (int q = 0; q < CPT; q++) {
            const int c = tid + q * BLOCK;
            if (c < Li) *reinterpret_cast<VecG<T, G>*>(xs + (size_t)c * G) = pre[q];
          }
        } else {
          // each thread gathers one column of all G source rows (G coalesced
          // loads) and writes it with one wide LDS store: consecutive lanes
          // write consecutive LDS words, so the stores do not conflict
          int64_t bi[G];
#pragma unroll
          for (int g = 0; g < G; g++) bi[g] = g < nb ? tb.pstart[tb.p1_src[b0 + g]] : -1;
          for (int c = tid; c < Li; c += BLOCK) {
            VecG<T, G> v;
#pragma unroll
            for (int g = 0; g < G; g++) v.v[g] = bi[g] >= 0 ? x[bi[g] + c] : T(0);
            *reinterpret_cast<VecG<T, G>*>(xs + (size_t)c * G) = v;
          }
        }
#pragma unroll
        for (int u = 0; u < kWbPer; u++) {
          const int i = tid + u * BLOCK;
          if (i < 2 * (wsd + 1) * G) wb[i] = wreg[u];
        }
        __syncthreads();
        if constexpr (PF) {
          // the next batch: the rest of this row, else the first later row
          // with entries in this class; its loads overlap this batch's work
          int nb0 = 0, ne = 0;
          if (b0 + G < p1[r]) {
            nb0 = b0 + G;
            ne = p1[r];
          } else {
#pragma unroll
            for (int rr = R - 1; rr > r; rr--)
              if (p1[rr] > p0[rr]) {
                nb0 = p0[rr];
                ne = p1[rr];
              }
          }
          if (ne > nb0) {
            int64_t bi[G];
#pragma unroll
            for (int g = 0; g < G; g++) bi[g] = g < ne - nb0 ? tb.pstart[tb.p1_src[nb0 + g]] : -1;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
              const int c = tid + q * BLOCK;
#pragma unroll
              for (int g = 0; g < G; g++) pre[q].v[g] = (c < Li && bi[g] >= 0) ? x[bi[g] + c] : T(0);
            }
          }
        }
        // column-major over the thread's columns: each (wave, q) runs to its
        // own chunk's entry count (wave-uniform), four entries to a load
        T pa[CPT];
#pragma unroll
        for (int q = 0; q < CPT; q++) pa[q] = T(0);
        walk_quads<CPT>(
            rpk, cl, Eq, (uint32_t)L * 16u,
            [&](uint32_t pk) {
              return PnEntry<T, G>{*reinterpret_cast<const VecG<T, G>*>(smem + XS + (pk & 0x1ffffu)),
                                   *reinterpret_cast<const VecG<float, G>*>(smem + (pk >> 17))};
            },
            [&](int q, const PnEntry<T, G>& en) {
#pragma unroll
              for (int g = 0; g < G; g++) pa[q] = fma((T)en.w.v[g], en.x.v[g], pa[q]);
            });
#pragma unroll
        for (int q = 0; q < CPT; q++) acc[r][q] += pa[q];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (r < nr) {
      const int64_t bf = tb.pstart[f0 + r];
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        const int c = tid + q * BLOCK;
        if (c < L) {
          if (accumulate) y[bf + c] += acc[r][q];
          else y[bf + c] = acc[r][q];
        }
      }
    }
  }
}

// ----------------------------------------------------------- vector kernels
__device__ __forceinline__ double dot16(float4 a, float4 b) {
  return (double)a.x * (double)b.x + (double)a.y * (double)b.y + (double)a.z * (double)b.z + (double)a.w * (double)b.w;
}
__device__ __forceinline__ double dot16(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }

// Partial sums of <V_j, w> for j in [0, m): block b writes part[b*m + j].
// 16-byte loads over the first D - D % n elements of each vector (V_j starts
// 16-byte aligned: ld is a multiple of 64 elements), the tail by block 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_multidot(const T* __restrict__ V, int64_t ld, int m,
                                                     const T* __restrict__ w, int64_t D,
                                                     double* __restrict__ part) {
  using V16 = typename Vec16<T>::type;
  constexpr int VW = Vec16<T>::n;
  __shared__ double red[8][kBlock / 64];
  // 16-byte path only when every vector starts 16-byte aligned (uniform test)
  const bool vec = ((((uintptr_t)V | (uintptr_t)w) & 15) == 0) && (ld % VW == 0);
  const int64_t nv = vec ? D / VW : 0;
  const int64_t chunk = (nv + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < nv ? lo + chunk : nv;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const V16* w16 = reinterpret_cast<const V16*>(w);
  for (int j0 = 0; j0 < m; j0 += 8) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
      const V16 wi = w16[i];
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (j0 + u < m) a[u] += dot16(reinterpret_cast<const V16*>(V + (int64_t)(j0 + u) * ld)[i], wi);
    }
    // scalar part: the tail, or everything when not aligned (block-strided)
    {
      const int64_t s0 = nv * VW;
      const int64_t sc = (D - s0 + gridDim.x - 1) / gridDim.x;
      const int64_t slo = s0 + (int64_t)blockIdx.x * sc, shi = min(slo + sc, D);
      for (int64_t i = slo + threadIdx.x; i < shi; i += kBlock) {
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (j0 + u < m) a[u] += (double)V[(int64_t)(j0 + u) * ld + i] * (double)w[i];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      double v = a[u];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      if (lane == 0) red[u][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 8 && j0 + (int)threadIdx.x < m) {
      double s = 0.0;
      for (int q = 0; q < kBlock / 64; q++) s += red[threadIdx.x][q];
      part[(int64_t)blockIdx.x * m + j0 + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// out[j] = sum_b part[b*m + j]  (mode 0), or sqrt of it (mode 1)
__global__ __launch_bounds__(kBlock) void k_reduce(const double* __restrict__ part, int nb, int m,
                                                   double* __restrict__ out, int mode) {
  __shared__ double red[kBlock / 64];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) s += part[(int64_t)b * m + j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kBlock / 64; q++) t += red[q];
    out[j] = mode == 1 ? sqrt(t) : t;
  }
}

// w -= T(a * v)   (src/breorthog.f90:704-708)
template <typename T>
__global__ void k_sub_scaled(T* __restrict__ w, const T* __restrict__ v, const double* __restrict__ a,
                             int64_t D) {
  const double s = *a;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = w[i] - (T)(s * (double)v[i]);
}

// w -= T(sum_j d_j V_j)   (src/breorthog.f90:880-889); 16-byte loads, V_j
// 16-byte aligned (ld a multiple of 64 elements), scalar tail
template <typename T>
__global__ void k_multiaxpy(T* __restrict__ w, const T* __restrict__ V, int64_t ld, int m,
                            const double* __restrict__ d, int64_t D) {
  using V16 = typename Vec16<T>::type;
  constexpr int VW = Vec16<T>::n;
  const bool vec = ((((uintptr_t)V | (uintptr_t)w) & 15) == 0) && (ld % VW == 0);
  const int64_t nv = vec ? D / VW : 0;
  V16* w16 = reinterpret_cast<V16*>(w);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
    double s[VW];
#pragma unroll
    for (int e = 0; e < VW; e++) s[e] = 0.0;
    for (int j = 0; j < m; j++) {
      const V16 v = reinterpret_cast<const V16*>(V + (int64_t)j * ld)[i];
      const T* pv = reinterpret_cast<const T*>(&v);
#pragma unroll
      for (int e = 0; e < VW; e++) s[e] += d[j] * (double)pv[e];
    }
    V16 o = w16[i];
    T* po = reinterpret_cast<T*>(&o);
#pragma unroll
    for (int e = 0; e < VW; e++) po[e] = po[e] - (T)s[e];
    w16[i] = o;
  }
  for (int64_t i = nv * VW + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < m; j++) s += d[j] * (double)V[(int64_t)j * ld + i];
    w[i] = w[i] - (T)s;
  }
}

// v = T(double(w) * (1/beta))   (src/breorthog.f90:775-780)
template <typename T>
__global__ void k_normalize(T* __restrict__ v, const T* __restrict__ w, const double* __restrict__ beta,
                            int64_t D) {
  const double m = 1.0 / *beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (T)((double)w[i] * m);
}

// seeded values in [-1, 1): SplitMix64 (Steele, Lea, Flood 2014) of the index
template <typename T>
__global__ void k_pivot_random(T* __restrict__ v, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)i * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    v[i] = (T)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
  }
}

template <typename T>
__global__ void k_fill(T* __restrict__ v, T val, int64_t D) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = val;
}

// out_k = sum_j Y[j*ky + k] V_j (FP64 accumulation), k < nk
template <typename T, typename U>
__global__ void k_combine(U* __restrict__ out, int64_t ldo, const T* __restrict__ V, int64_t ld, int m,
                          const double* __restrict__ Y, int ky, int nk, int64_t D) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
    double acc[16];
    for (int k = 0; k < nk; k++) acc[k] = 0.0;
    for (int j = 0; j < m; j++) {
      const double v = (double)V[(int64_t)j * ld + i];
      for (int k = 0; k < nk; k++) acc[k] += Y[j * ky + k] * v;
    }
    for (int k = 0; k < nk; k++) out[(int64_t)k * ldo + i] = (U)acc[k];
  }
}

// X[:, 0:ncol] = V[:, 0:m] Y[0:m, 0:ncol] (thick-restart rotation and Ritz
// vectors, src/breorthog.f90:1211-1243) on the FP64 matrix cores, ncol <= 16.
// One wave per strip of 64 rows: lane (li, lk) = (lane % 16, lane / 16) loads
// rows 4 li .. 4 li + 3 of vector k + lk as one contiguous chunk (16 lanes
// cover the strip's 64 rows of a vector: whole 256-byte segments for FP32),
// converts them to FP64 and feeds four v_mfma_f64_16x16x4f64 tiles, tile u
// holding rows 4 i + u (i = 0..15).  Operand layouts of the instruction:
// A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane
// j + 16 (i % 4), register i / 4.
typedef double dbl4 __attribute__((ext_vector_type(4)));
template <typename T, typename U>
__global__ __launch_bounds__(256) void k_combine_mfma(U* __restrict__ X, int64_t ldo, const T* __restrict__ V,
                                                      int64_t ld, int m, const double* __restrict__ Y, int ys,
                                                      int ncol, int64_t D) {
  __shared__ double tile[4][16 * 64];
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t nstrip = (D + 63) / 64;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t st = wave; st < nstrip; st += nwave) {
    const int64_t i0 = st * 64;
    const int64_t r0 = i0 + 4 * li;
    const bool full = i0 + 64 <= D;
    dbl4 acc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) acc[u] = dbl4{0.0, 0.0, 0.0, 0.0};
    // four k-steps per pass: their loads (kept in the vectors' own type) are
    // in flight together
    constexpr int KS = 4;
    for (int k0 = 0; k0 < m; k0 += 4 * KS) {
      T a[KS][4];
      double b[KS];
#pragma unroll
      for (int s4 = 0; s4 < KS; s4++) {
        const int k = k0 + 4 * s4 + lk;
        b[s4] = (k < m && li < ncol) ? Y[(int64_t)k * ys + li] : 0.0;
        const T* src = V + (int64_t)(k < m ? k : 0) * ld + r0;
        if (k < m && full) {
          if constexpr (sizeof(T) == 4) {
            const float4 q = *reinterpret_cast<const float4*>(src);
            a[s4][0] = q.x; a[s4][1] = q.y; a[s4][2] = q.z; a[s4][3] = q.w;
          } else {
            const double2 q0 = *reinterpret_cast<const double2*>(src);
            const double2 q1 = *reinterpret_cast<const double2*>(src + 2);
            a[s4][0] = q0.x; a[s4][1] = q0.y; a[s4][2] = q1.x; a[s4][3] = q1.y;
          }
        } else {
#pragma unroll
          for (int u = 0; u < 4; u++) a[s4][u] = (k < m && r0 + u < D) ? src[u] : T(0);
        }
      }
#pragma unroll
      for (int s4 = 0; s4 < KS; s4++)
#pragma unroll
        for (int u = 0; u < 4; u++)
          acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[s4][u], b[s4], acc[u], 0, 0, 0);
    }
    // through this wave's LDS tile [column][64 rows], so that each column's
    // 64 rows leave as one contiguous store per wave
    double* tl = tile[threadIdx.x >> 6];
    if (li < ncol) {
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int r = 0; r < 4; r++) tl[li * 64 + 4 * (lk + 4 * r) + u] = acc[u][r];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (i0 + lane < D)
      for (int j = 0; j < ncol; j++) X[(int64_t)j * ldo + i0 + lane] = (U)tl[j * 64 + lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

template <typename T, typename U>
__global__ void k_convert(U* __restrict__ out, const T* __restrict__ in, int64_t D) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (U)in[i];
}

// ------------------------------------------------------------ block Lanczos
// Block vectors are interleaved like the multi-vector apply: element k of
// basis state s of block i at V[i*D*K + s*K + k].

// part[b * (m*K*K) + (i*K + k1)*K + k2] = workgroup b's share of
// sum_s V_i[s, k1] * W[s, k2]  (blockIdx.y = i), FP64.
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_block_dots(const T* __restrict__ V, int64_t D, int m,
                                                        const T* __restrict__ W, double* __restrict__ part) {
  const int i = blockIdx.y;
  const T* Vi = V + (int64_t)i * D * K;
  double acc[K][K];
#pragma unroll
  for (int a = 0; a < K; a++)
#pragma unroll
    for (int c = 0; c < K; c++) acc[a][c] = 0.0;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < D; s += (int64_t)gridDim.x * blockDim.x) {
    double v[K], w[K];
#pragma unroll
    for (int a = 0; a < K; a++) {
      v[a] = (double)Vi[s * K + a];
      w[a] = (double)W[s * K + a];
    }
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
      for (int c = 0; c < K; c++) acc[a][c] += v[a] * w[c];
  }
  __shared__ double red[kBlock / 64];
  double* out = part + (int64_t)blockIdx.x * m * K * K + (int64_t)i * K * K;
#pragma unroll
  for (int a = 0; a < K; a++)
#pragma unroll
    for (int c = 0; c < K; c++) {
      double x = acc[a][c];
      for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kBlock / 64; q++) t += red[q];
        out[a * K + c] = t;
      }
      __syncthreads();
    }
}

// W[s, k2] -= T(sum_{i, k1} V_i[s, k1] * C[(i*K + k1)*K + k2])  (block CGS)
template <typename T, int K>
__global__ void k_block_update(T* __restrict__ W, const T* __restrict__ V, int64_t D, int m,
                               const double* __restrict__ C) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < D; s += (int64_t)gridDim.x * blockDim.x) {
    double acc[K];
#pragma unroll
    for (int c = 0; c < K; c++) acc[c] = 0.0;
    for (int i = 0; i < m; i++) {
      const T* Vi = V + (int64_t)i * D * K + s * K;
#pragma unroll
      for (int a = 0; a < K; a++) {
        const double v = (double)Vi[a];
        const double* Cr = C + ((int64_t)i * K + a) * K;
#pragma unroll
        for (int c = 0; c < K; c++) acc[c] += v * Cr[c];
      }
    }
#pragma unroll
    for (int c = 0; c < K; c++) W[s * K + c] = W[s * K + c] - (T)acc[c];
  }
}

// dst[s, c] = T(sum_a src[s, a] * M[a*K + c])  (block orthonormalization)
template <typename T, int K>
__global__ void k_block_transform(T* __restrict__ dst, const T* __restrict__ src, int64_t D,
                                  const double* __restrict__ M) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < D; s += (int64_t)gridDim.x * blockDim.x) {
    double v[K];
#pragma unroll
    for (int a = 0; a < K; a++) v[a] = (double)src[s * K + a];
#pragma unroll
    for (int c = 0; c < K; c++) {
      double t = 0.0;
#pragma unroll
      for (int a = 0; a < K; a++) t += v[a] * M[a * K + c];
      dst[s * K + c] = (T)t;
    }
  }
}

// out_q[s] = sum_{c < n} col_c[s] * Y[c*ys + q] (q < nq <= 16), col_{iK+k} =
// V_i[., k]; FP64 output vectors at stride D (Ritz vectors of the block basis)
template <typename T, int K>
__global__ void k_block_combine(double* __restrict__ out, const T* __restrict__ V, int64_t D, int n,
                                const double* __restrict__ Y, int ys, int nq) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < D; s += (int64_t)gridDim.x * blockDim.x) {
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = 0.0;
    for (int c = 0; c < n; c++) {
      const double v = (double)V[(int64_t)(c / K) * D * K + s * K + c % K];
#pragma unroll
      for (int q = 0; q < 16; q++)
        if (q < nq) acc[q] += v * Y[(int64_t)c * ys + q];
    }
#pragma unroll
    for (int q = 0; q < 16; q++)
      if (q < nq) out[(int64_t)q * D + s] = acc[q];
  }
}

// X[s*64 + k] = 1 if s == j0 + k (k < kb) else 0: a block of unit vectors
__global__ void k_unit_block(double* __restrict__ X, int64_t D, int64_t j0, int kb) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D * 64; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i / 64;
    const int k = (int)(i % 64);
    X[i] = (k < kb && s == j0 + k) ? 1.0 : 0.0;
  }
}

// column j0 + k of the dense (column-major) H from the interleaved block H X
__global__ void k_scatter_cols(double* __restrict__ H, int64_t D, const double* __restrict__ Y, int64_t j0, int kb) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D * 64; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i / 64;
    const int k = (int)(i % 64);
    if (k < kb) H[(j0 + k) * D + s] = Y[i];
  }
}

// ------------------------------------------------------- adjoint (densities)
// rho[me] = <psi| O_me |psi> for every uncoupled operator, walking the same
// jump tables as the apply (the analogue of applyh{PP,NN,PN}bundled_den,
// src/bdenslib4.f90:129-725).  FP64 accumulation.
struct DensTables {
  // PP: unmerged operators by target proton row
  const int64_t* pp_ptr;
  const int32_t* pp_src;
  const int32_t* pp_me;
  const int8_t* pp_ph;
  // NN: per neutron sector, flat lists of (target col | source col << 24 | sign << 48) and compact element id
  const int64_t* nn_off;     // [nsecN + 1]
  const uint64_t* nn_pk;
  const int32_t* nn_mc;
  int nn_ncomp;
  // PN: lists per (target neutron sector, class, neutron operator)
  const int64_t* pn_off;     // [(c*nd + k)*(maxcls+1) + nl]
  const uint64_t* pn_pk;     // target col | source col << 24 | sign << 48
  int maxcls1;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// The density sums accumulate in 64-bit fixed point (value * scale, rounded,
// two's complement): integer addition is associative, so the result does not
// depend on the order in which workgroups and lanes add -- every run gives
// the same bits.  The host picks scale = 2^k so that no element can leave the
// int64 range (|element| <= |psi|^2: every operator here has norm <= 1).
__device__ __forceinline__ unsigned long long dfix(double v, double scale) {
  return (unsigned long long)__double2ll_rn(v * scale);
}

template <typename T>
__global__ __launch_bounds__(256) void k_density_pp(DevTables tb, DensTables dn, const T* __restrict__ x,
                                                    unsigned long long* __restrict__ rho, double scale) {
  const int f = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = tb.nlen[tb.pconj[tb.psec[f]]];
  const int64_t bf = tb.pstart[f];
  for (int64_t e = dn.pp_ptr[f] + wave; e < dn.pp_ptr[f + 1]; e += 4) {
    const int64_t bs = tb.pstart[dn.pp_src[e]];
    double a = 0.0;
    for (int c = lane; c < L; c += 64) a += (double)x[bf + c] * (double)x[bs + c];
    a = wave_sum(a);
    if (lane == 0) atomicAdd(&rho[dn.pp_me[e]], dfix((double)dn.pp_ph[e] * a, scale));
  }
}

// NN: LDS-staged tiles of R rows with per-workgroup element sums in LDS
// (needs R rows and the element accumulators in LDS)
template <typename T, int R>
__global__ __launch_bounds__(256) void k_density_nn(DevTables tb, DensTables dn, const T* __restrict__ x,
                                                    unsigned long long* __restrict__ rho_comp, double scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned long long* rl = reinterpret_cast<unsigned long long*>(smem);
  T* xs = reinterpret_cast<T*>(smem + (size_t)dn.nn_ncomp * sizeof(unsigned long long));
  const int tile = blockIdx.x;
  const int f0 = tb.tile_f0[tile], nr = tb.tile_nr[tile];
  const int cn = tb.pconj[tb.psec[f0]];
  const int L = tb.nlen[cn];
  for (int i = threadIdx.x; i < dn.nn_ncomp; i += 256) rl[i] = 0ull;
  for (int r = 0; r < R; r++) {
    const int64_t bf = r < nr ? tb.pstart[f0 + r] : 0;
    for (int c = threadIdx.x; c < L; c += 256) xs[c * R + r] = r < nr ? x[bf + c] : T(0);
  }
  __syncthreads();
  const VecG<T, R>* xv = reinterpret_cast<const VecG<T, R>*>(xs);
  for (int64_t q = dn.nn_off[cn] + threadIdx.x; q < dn.nn_off[cn + 1]; q += 256) {
    const uint64_t pk = dn.nn_pk[q];
    const VecG<T, R> a = xv[pk & 0xffffffu];
    const VecG<T, R> b = xv[(pk >> 24) & 0xffffffu];
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) v += (double)a.v[r] * (double)b.v[r];
    atomicAdd(&rl[dn.nn_mc[q]], dfix(((pk >> 48) & 1) ? -v : v, scale));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < dn.nn_ncomp; i += 256)
    if (rl[i] != 0ull) atomicAdd(&rho_comp[i], rl[i]);
}

// NN without LDS (rows or element sets too large for it): a workgroup per
// proton row, the row's values gathered from global memory
template <typename T>
__global__ __launch_bounds__(256) void k_density_nn_g(DevTables tb, DensTables dn, const T* __restrict__ x,
                                                      unsigned long long* __restrict__ rho_comp, double scale) {
  const int f = blockIdx.x;
  const int cn = tb.pconj[tb.psec[f]];
  const int64_t bf = tb.pstart[f];
  for (int64_t q = dn.nn_off[cn] + threadIdx.x; q < dn.nn_off[cn + 1]; q += 256) {
    const uint64_t pk = dn.nn_pk[q];
    const double v = (double)x[bf + (int64_t)(pk & 0xffffffu)] * (double)x[bf + (int64_t)((pk >> 24) & 0xffffffu)];
    atomicAdd(&rho_comp[dn.nn_mc[q]], dfix(((pk >> 48) & 1) ? -v : v, scale));
  }
}

// PN: per proton row, batches of G source rows of one class; GLB = 0 stages
// them and the row itself in LDS, GLB = 1 gathers from global memory
template <typename T, int G, bool GLB>
__global__ __launch_bounds__(256) void k_density_pn(DevTables tb, DensTables dn, const T* __restrict__ x,
                                                    unsigned long long* __restrict__ rho_w, double scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* xs = reinterpret_cast<T*>(smem);                                      // [lmax][G] staged sources
  T* yf = reinterpret_cast<T*>(smem + (GLB ? 0 : (size_t)G * tb.lmax * sizeof(T)));  // own row
  const int f = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sp = tb.psec[f];
  const int cn = tb.pconj[sp];
  const int L = tb.nlen[cn];
  const int64_t bf = tb.pstart[f];
  if constexpr (!GLB)
    for (int c = threadIdx.x; c < L; c += 256) yf[c] = x[bf + c];
  const VecG<T, G>* xg = reinterpret_cast<const VecG<T, G>*>(xs);
  const int nd = tb.ndelta, nd1 = nd + 1;
  for (int k = 0; k < nd; k++) {
    const int p0 = tb.pcls[(int64_t)f * nd1 + k], p1 = tb.pcls[(int64_t)f * nd1 + k + 1];
    if (p0 == p1) continue;
    const int kn = tb.negk[k];
    const int si = tb.srcsec[sp * nd + k];
    const int Li = tb.nlen[tb.pconj[si]];
    const int ws = tb.wstride[k];
    const int64_t lbase = ((int64_t)cn * nd + kn) * dn.maxcls1;
    for (int b0 = p0; b0 < p1; b0 += G) {
      const int nb = min(G, p1 - b0);
      int64_t bi[G];
#pragma unroll
      for (int g = 0; g < G; g++) bi[g] = g < nb ? tb.pstart[tb.p1_src[b0 + g]] : -1;
      if constexpr (!GLB) {
        __syncthreads();
        for (int g = 0; g < G; g++)
          for (int c = threadIdx.x; c < Li; c += 256) xs[c * G + g] = bi[g] >= 0 ? x[bi[g] + c] : T(0);
        __syncthreads();
      }
      for (int nl = wave; nl < ws; nl += 4) {
        double a[G];
#pragma unroll
        for (int g = 0; g < G; g++) a[g] = 0.0;
        for (int64_t q = dn.pn_off[lbase + nl] + lane; q < dn.pn_off[lbase + nl + 1]; q += 64) {
          const uint64_t pk = dn.pn_pk[q];
          const int tc = (int)(pk & 0xffffffu), sc = (int)((pk >> 24) & 0xffffffu);
          double yv;
          if constexpr (GLB) yv = (double)x[bf + tc];
          else yv = (double)yf[tc];
          if ((pk >> 48) & 1) yv = -yv;
          if constexpr (GLB) {
#pragma unroll
            for (int g = 0; g < G; g++)
              if (bi[g] >= 0) a[g] += yv * (double)x[bi[g] + sc];
          } else {
            const VecG<T, G> xv = xg[sc];
#pragma unroll
            for (int g = 0; g < G; g++) a[g] += yv * (double)xv.v[g];
          }
        }
#pragma unroll
        for (int g = 0; g < G; g++) a[g] = wave_sum(a[g]);
        if (lane == 0)
          for (int g = 0; g < nb; g++) {
            const int meta = tb.p1_meta[b0 + g];
            const double sgn = (meta >> 8) ? -a[g] : a[g];
            atomicAdd(&rho_w[tb.woff[k] + (meta & 0xff) * ws + nl], dfix(sgn, scale));
          }
      }
    }
  }
}

// ------------------------------------------------------ batched Hamiltonians
// K Hamiltonians on one basis; vectors interleaved X[state*K + k].  One wave
// per output basis state, lane = Hamiltonian: every jump-table read (uniform
// across the wave, so scalar) serves all K vectors, and the x and element
// loads are K-contiguous.
struct BatchVals {
  int K;       // vectors per state (interleaved)
  int vs;      // value stride: K (one Hamiltonian per vector) or 1 (one shared Hamiltonian)
  int vk;      // 1: value index adds the lane, 0: shared values
  const float* pp32;   // [merged PP entry][vs]
  const double* pp64;
  const float* nn32;   // [NN ELL slot][vs]
  const double* nn64;
  const float* W;      // [pn block element][vs]
  const float* pspe;   // [proton SD][vs]
  const float* nspe;   // [neutron SD][vs]
  const float* kt32;   // one Hamiltonian per vector: [two-row PP union entry][row][K]
  const double* kt64;
};

template <typename T> __device__ __forceinline__ T bval(const float* a, const double* b, int64_t i);
template <> __device__ __forceinline__ float bval<float>(const float* a, const double*, int64_t i) { return a[i]; }
template <> __device__ __forceinline__ double bval<double>(const float*, const double* b, int64_t i) { return b[i]; }

template <typename T>
__global__ __launch_bounds__(256) void k_apply_batched(DevTables tb, BatchVals bv, const T* __restrict__ x,
                                                       T* __restrict__ y) {
  const int K = bv.K;
  const int k = threadIdx.x & 63;
  const int f = blockIdx.x;
  const int c = __builtin_amdgcn_readfirstlane((int)(blockIdx.y * 4 + (threadIdx.x >> 6)));
  const int sp = tb.psec[f];
  const int cn = tb.pconj[sp];
  const int L = tb.nlen[cn];
  if (c >= L) return;
  const bool on = k < K;
  const int kk = on ? k : 0;
  const int vs = bv.vs, vk = bv.vk ? kk : 0;
  const int64_t bf = tb.pstart[f];
  const int nf0 = tb.nfirst[cn];
  T acc = T(0);
  if (tb.call_spe)
    acc += ((T)bv.pspe[(int64_t)f * vs + vk] + (T)bv.nspe[(int64_t)(nf0 + c) * vs + vk]) * x[(bf + c) * K + kk];
  for (int64_t e = tb.pp_ptr[f]; e < tb.pp_ptr[f + 1]; e++) {
    const int64_t bi = tb.pstart[tb.pp_src[e]];
    acc += bval<T>(bv.pp32, bv.pp64, e * vs + vk) * x[(bi + c) * K + kk];
  }
  {
    const int64_t base = tb.nn_base[cn];
    const int E = tb.nn_emax[cn];
    for (int e = 0; e < E; e++) {
      const int64_t o = base + (int64_t)e * L + c;
      acc += bval<T>(bv.nn32, bv.nn64, o * vs + vk) * x[(bf + tb.nn_src[o]) * K + kk];
    }
  }
  const int nd = tb.ndelta, nd1 = nd + 1;
  for (int cl = 0; cl < nd; cl++) {
    const int p0 = tb.pcls[(int64_t)f * nd1 + cl], p1 = tb.pcls[(int64_t)f * nd1 + cl + 1];
    if (p0 == p1) continue;
    const int kn = tb.negk[cl];
    const int ws = tb.wstride[cl];
    const int64_t ebase = tb.n1_base[(int64_t)cn * nd + kn];
    const int E = tb.n1_emax[(int64_t)cn * nd + kn];
    for (int pj = p0; pj < p1; pj++) {
      const int64_t bi = tb.pstart[tb.p1_src[pj]];
      const int meta = tb.p1_meta[pj];
      const int64_t wrow = tb.woff[cl] + (int64_t)(meta & 0xff) * ws;
      T a = T(0);
      for (int e = 0; e < E; e++) {
        const uint32_t pk = tb.n1_pk[ebase + (int64_t)e * L + c];
        const int nl = (pk >> 16) & 0xff;
        if (nl >= ws) continue;  // padding slot
        const T t = (T)bv.W[(wrow + nl) * vs + vk] * x[(bi + (pk & 0xffff)) * K + kk];
        a += ((pk >> 24) & 1) ? -t : t;
      }
      acc += (meta >> 8) ? -a : a;
    }
  }
  if (on) y[(bf + c) * K + k] = acc;
}

// Row kernel: y = H x with every gather from global memory and no LDS, one
// thread per output state, a workgroup per (proton row, run of 256 columns)
// from the rows_f / rows_c map.  It carries the spaces the tile kernel's LDS
// cannot: neutron sectors longer than its staged row (4,096 SDs) and Delta
// classes larger than its signed-coefficient table (63 operators); the
// neutron rows it gathers from stay in L2.  It is also one more variant on
// every system.  Same sums as k_apply_tiles: SPE, PP row AXPYs, NN ELL
// gathers within the row, PN per proton jump over the neutron entries of the
// conjugate class (padding slots carry the class size as operator index).
template <typename T>
__global__ __launch_bounds__(256) void k_apply_rows(DevTables tb, const int32_t* __restrict__ rows_f,
                                                    const int32_t* __restrict__ rows_c, const T* __restrict__ x,
                                                    T* __restrict__ y, int accumulate) {
  const int f = rows_f[blockIdx.x];
  const int c = rows_c[blockIdx.x] + (int)threadIdx.x;
  const int sp = tb.psec[f];
  const int cn = tb.pconj[sp];
  const int L = tb.nlen[cn];
  if (c >= L) return;  // no barriers below
  const int64_t bf = tb.pstart[f];
  T acc = T(0);
  if (tb.call_spe) acc = ((T)tb.pspe_sd[f] + (T)tb.nspe_sd[tb.nfirst[cn] + c]) * x[bf + c];
  for (int64_t e = tb.pp_ptr[f]; e < tb.pp_ptr[f + 1]; e++)
    acc = fma(vload<T>(tb.pp_val32, tb.pp_val64, e), x[tb.pstart[tb.pp_src[e]] + c], acc);
  {
    const int64_t base = tb.nn_base[cn];
    const int E = tb.nn_emax[cn];
    for (int e = 0; e < E; e++) {
      const int64_t o = base + (int64_t)e * L + c;
      acc = fma(vload<T>(tb.nn_val32, tb.nn_val64, o), x[bf + tb.nn_src[o]], acc);
    }
  }
  const int nd = tb.ndelta, nd1 = nd + 1;
  for (int k = 0; k < nd; k++) {
    const int p0 = tb.pcls[(int64_t)f * nd1 + k], p1 = tb.pcls[(int64_t)f * nd1 + k + 1];
    if (p0 == p1) continue;
    const int kn = tb.negk[k];
    const int ws = tb.wstride[k];
    const int64_t ebase = tb.n1_base[(int64_t)cn * nd + kn];
    const int E = tb.n1_emax[(int64_t)cn * nd + kn];
    const float* Wk = tb.W + tb.woff[k];
    for (int pj = p0; pj < p1; pj++) {
      const int64_t bi = tb.pstart[tb.p1_src[pj]];
      const int meta = tb.p1_meta[pj];
      const float* wr = Wk + (int64_t)(meta & 0xff) * ws;
      T a = T(0);
      for (int e = 0; e < E; e++) {
        const uint32_t pk = tb.n1_pk[ebase + (int64_t)e * L + c];
        const int nl = (pk >> 16) & 0xff;
        if (nl >= ws) continue;  // padding slot
        const T t = (T)wr[nl] * x[bi + (pk & 0xffffu)];
        a += ((pk >> 24) & 1u) ? -t : t;
      }
      acc += (meta >> 8) ? -a : a;
    }
  }
  y[bf + c] = accumulate ? y[bf + c] + acc : acc;
}

// One neutron entry of a K-vector PN batch, as read from LDS: the G*K staged
// source values at its source column and its GW signed coefficients.
template <typename T, int GK, int GW>
struct MultiEntry {
  VecG<T, GK> x;
  VecG<float, GW> w;
};

// K interleaved vectors per basis state on large bases: the tile kernel's
// design (one workgroup per proton row, CPT columns per thread, per-chunk
// entry bounds, signed coefficient tables, buffer-load table walks) with every
// jump-table read, index decode and barrier serving K vectors.  Dynamic LDS:
//   [0, kWbRows*G*WK*4)  Wb[2*op + s][g][k], the batch's signed coefficients
//                        of each Hamiltonian k (WK = K), or of the one
//                        Hamiltonian (SH, WK = 1)
//   [kWbRows*G*WK*4, ..) staged rows [col][K] (NN) or [col][G][K] (PN)
// Neutron entries come as src | (2*op + sign) << 16 and are scaled to LDS
// byte offsets with two shifts.  Values per Hamiltonian (vs = K, vk = 1) or
// shared (vs = 1, vk = 0; launched as SH when K > 1).
template <typename T, int BLOCK, int CPT, int K, int G, bool SH>
__global__ __launch_bounds__(BLOCK) void k_apply_multi(DevTables tb, BatchVals bv, const T* __restrict__ x,
                                                       T* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int GK = G * K;
  // SH: one Hamiltonian for all K vectors (bv.vk == 0), so each coefficient
  // row holds G values, not G*K copies: an entry reads 4*G bytes of
  // coefficients beside its G*K staged source values
  constexpr int WK = SH ? 1 : K;
  constexpr int GW = G * WK;
  constexpr int XS = kWbRows * GW * (int)sizeof(float);
  T* xs = reinterpret_cast<T*>(smem + XS);
  float* wb = reinterpret_cast<float*>(smem);
  const int tid = threadIdx.x;
  const int tile = tb.tile_map[blockIdx.x];
  if (tile < 0) return;
  const int f = tb.tile_f0[tile];
  const int sp = tb.psec[f];
  const int cn = tb.pconj[sp];
  const int L = tb.nlen[cn];
  const int nf0 = tb.nfirst[cn];
  const int vs = bv.vs, vk = bv.vk;
  const int64_t bf = tb.pstart[f];
  using XV = VecG<T, K>;
  uint32_t cl[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) cl[q] = (uint32_t)min(tid + q * BLOCK, L - 1);
  T acc[CPT][K];
#pragma unroll
  for (int q = 0; q < CPT; q++)
#pragma unroll
    for (int k = 0; k < K; k++) acc[q][k] = T(0);
  // ---- SPE
  if (tb.call_spe) {
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const XV xv = *reinterpret_cast<const XV*>(x + (bf + cl[q]) * K);
#pragma unroll
      for (int k = 0; k < K; k++)
        acc[q][k] = fma((T)bv.pspe[(int64_t)f * vs + k * vk] + (T)bv.nspe[(int64_t)(nf0 + cl[q]) * vs + k * vk],
                        xv.v[k], acc[q][k]);
    }
  }
  // ---- PP: row AXPYs, K-wide coalesced reads
  for (int64_t e = tb.pp_ptr[f]; e < tb.pp_ptr[f + 1]; e++) {
    const T* xi = x + tb.pstart[tb.pp_src[e]] * K;
    T v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = bval<T>(bv.pp32, bv.pp64, e * vs + k * vk);
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const XV xv = *reinterpret_cast<const XV*>(xi + (int64_t)cl[q] * K);
#pragma unroll
      for (int k = 0; k < K; k++) acc[q][k] = fma(v[k], xv.v[k], acc[q][k]);
    }
  }
  // ---- NN: the row staged [col][K]
  {
    for (int c = tid; c < L; c += BLOCK)
      *reinterpret_cast<XV*>(xs + (size_t)c * K) = *reinterpret_cast<const XV*>(x + (bf + c) * K);
    __syncthreads();
    const XV* xr = reinterpret_cast<const XV*>(xs);
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(tb.nn_src + tb.nn_base[cn]);
    const int64_t nb0 = tb.nn_base[cn];
    const int E = tb.nn_emax[cn];
    for (int e = 0; e < E; e++) {
      const uint32_t so = (uint32_t)e * (uint32_t)L * 4u;
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        const int sl = (int)buf_u32(rs, cl[q] * 4u, so);
        const int64_t o = nb0 + (int64_t)e * L + cl[q];
        const XV xv = xr[sl];
        T v[K];
        if (vk && sizeof(T) == 4) {  // K contiguous FP32 values of this slot
          const VecG<float, K> vv = *reinterpret_cast<const VecG<float, K>*>(bv.nn32 + o * K);
#pragma unroll
          for (int k = 0; k < K; k++) v[k] = (T)vv.v[k];
        } else {
#pragma unroll
          for (int k = 0; k < K; k++) v[k] = bval<T>(bv.nn32, bv.nn64, o * vs + k * vk);
        }
#pragma unroll
        for (int k = 0; k < K; k++) acc[q][k] = fma(v[k], xv.v[k], acc[q][k]);
      }
    }
  }
  // ---- PN: per Delta class, batches of G source rows staged [col][G][K]
  using GV = VecG<T, GK>;
  using WV = VecG<float, GW>;
  constexpr int XSH = __builtin_ctz(GK * (int)sizeof(T));  // log2 bytes per staged column
  constexpr int WSH = __builtin_ctz(GW * (int)sizeof(float));
  const int nd = tb.ndelta, nd1 = nd + 1;
  for (int k0 = 0; k0 < nd; k0++) {
    const int p0 = tb.pcls[(int64_t)f * nd1 + k0], p1 = tb.pcls[(int64_t)f * nd1 + k0 + 1];
    if (p0 == p1) continue;
    const int kn = tb.negk[k0];
    const int si = tb.srcsec[sp * nd + k0];
    const int Li = tb.nlen[tb.pconj[si]];
    const int wsd = tb.wstride[k0];
    const int64_t woff = tb.woff[k0];
    // the single-vector kernel's quad table for 4 staged FP32 rows: x offset
    // src * 16 bytes (bits 0-16), coefficient offset (2*op + sign) * 16 bytes
    // (bits 17-31), rescaled here to this kernel's column and row widths
    const __amdgpu_buffer_rsrc_t rpk = buf_rsrc(tb.n1_pkv[2][0] + tb.n1q_base[(int64_t)cn * nd + kn]);
    const uint8_t* ch = tb.n1_chunk + ((int64_t)cn * nd + kn) * 64 + __builtin_amdgcn_readfirstlane(tid >> 6);
    int Eq[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) Eq[q] = ch[q * (BLOCK / 64)];
    for (int b0 = p0; b0 < p1; b0 += G) {
      const int nb = min(G, p1 - b0);
      __syncthreads();
      {
        // one column of all G source rows per thread, one wide LDS store
        int64_t bi[G];
#pragma unroll
        for (int g = 0; g < G; g++) bi[g] = g < nb ? tb.pstart[tb.p1_src[b0 + g]] : -1;
        for (int c = tid; c < Li; c += BLOCK) {
          GV v;
#pragma unroll
          for (int g = 0; g < G; g++) {
            XV u;
            if (bi[g] >= 0) u = *reinterpret_cast<const XV*>(x + (bi[g] + c) * K);
            else
#pragma unroll
              for (int k = 0; k < K; k++) u.v[k] = T(0);
#pragma unroll
            for (int k = 0; k < K; k++) v.v[g * K + k] = u.v[k];
          }
          *reinterpret_cast<GV*>(xs + (size_t)c * GK) = v;
        }
      }
      for (int i = tid; i < 2 * (wsd + 1) * GW; i += BLOCK) {
        const int row = i / GW, r2 = i - row * GW, g = r2 / WK, k = r2 - g * WK, nl = row >> 1;
        float w = 0.0f;
        if (g < nb && nl < wsd) {
          const int meta = tb.p1_meta[b0 + g];
          w = bv.W[(woff + (int64_t)(meta & 0xff) * wsd + nl) * vs + k * vk];
          if ((meta >> 8) ^ (row & 1)) w = -w;
        }
        wb[i] = w;
      }
      __syncthreads();
      walk_quads<CPT>(
          rpk, cl, Eq, (uint32_t)L * 16u,
          [&](uint32_t pk) {
            const uint32_t xo = XSH >= 4 ? (pk & 0x1ffffu) << (XSH - 4) : (pk & 0x1ffffu) >> (4 - XSH);
            const uint32_t wo = WSH >= 4 ? (pk >> 17) << (WSH - 4) : (pk >> 17) >> (4 - WSH);
            return MultiEntry<T, GK, GW>{*reinterpret_cast<const GV*>(smem + XS + xo),
                                         *reinterpret_cast<const WV*>(smem + wo)};
          },
          [&](int q, const MultiEntry<T, GK, GW>& en) {
#pragma unroll
            for (int k = 0; k < K; k++)
#pragma unroll
              for (int g = 0; g < G; g++)
                acc[q][k] = fma((T)en.w.v[g * WK + (SH ? 0 : k)], en.x.v[g * K + k], acc[q][k]);
          });
    }
  }
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int c = tid + q * BLOCK;
    if (c < L) {
      XV o;
#pragma unroll
      for (int k = 0; k < K; k++) o.v[k] = acc[q][k];
      *reinterpret_cast<XV*>(y + (bf + c) * K) = o;
    }
  }
}

// K interleaved vectors with the single-vector kernel's row tiles: a
// workgroup of 1024 threads takes two proton rows of one sector (4 columns per
// thread, rows up to 4096 columns), accumulators [row][column][k].  PP: the
// union of the two rows' source rows, each read once with K-wide loads, two in
// flight.  NN: both rows staged.  PN: each batch stages one jump of each row
// (slot g) and the signed
// coefficients [2*op + s][slot] (SH: once for all K vectors of the one
// Hamiltonian; else [2*op + s][slot][k], one per vector's Hamiltonian); the
// neutron entries come by quads from the four-row FP32 quad table (offsets
// rescaled), so an entry reads 2*K staged values and 2 (or 2*K) coefficients
// for 2*K FMAs.  With 32-byte columns (FP32 K = 4, FP64 K = 2) one array per
// slot keeps consecutive lanes' reads at a 16-byte stride (a [col][slot][k]
// layout would put 32 bytes between them and halve the LDS banks a pass can
// use); 16-byte columns stay [col][slot][k], one read per entry.  Static LDS:
// kWbRows * 2 (* K) floats, then 2 * K * sizeof(T) bytes per column.
template <typename T, int K, bool SH>
__global__ __launch_bounds__(1024) void k_apply_ktiles(DevTables tb, BatchVals bv, const T* __restrict__ x,
                                                       T* __restrict__ y) {
  constexpr int R = 2, G = 2, BLOCK = 1024, CPT = 4;
  constexpr int GK = G * K;
  constexpr int WK = SH ? 1 : K;  // coefficients per slot: one Hamiltonian, or one per vector
  constexpr int GW = G * WK;
  constexpr int XS = kWbRows * GW * (int)sizeof(float);
  // staged rows: slot (or tile row) g of column c at XS + g * SLOTB + c * COLB.
  // Columns wider than 16 bytes are split into one array per slot, so that
  // consecutive lanes' reads stride 16 bytes; narrower ones stay [col][slot][k]
  // (one 16-byte read per entry)
  constexpr int KB = K * (int)sizeof(T);
  constexpr bool SPLIT = G * KB > 16;
  constexpr int SLOTB = SPLIT ? kMaxCols * KB : KB;
  constexpr int COLB = SPLIT ? KB : G * KB;
  __shared__ __align__(16) unsigned char smem[XS + G * kMaxCols * KB];
  float* wb = reinterpret_cast<float*>(smem);
  using XV = VecG<T, K>;
  using SV = VecG<T, GK>;
  const int tid = threadIdx.x;
  const int tile = tb.tile_map[blockIdx.x];
  if (tile < 0) return;  // padding slot of the XCD map (uniform per workgroup)
  const int f0 = tb.tile_f0[tile];
  const int nr = tb.tile_nr[tile];
  const int sp = tb.psec[f0];
  const int cn = tb.pconj[sp];
  const int L = tb.nlen[cn];
  const int nf0 = tb.nfirst[cn];
  uint32_t cl[CPT];  // clamped columns: loads never leave the row
#pragma unroll
  for (int q = 0; q < CPT; q++) cl[q] = (uint32_t)min(tid + q * BLOCK, L - 1);
  int64_t bfr[R];
#pragma unroll
  for (int r = 0; r < R; r++) bfr[r] = r < nr ? tb.pstart[f0 + r] : -1;
  T acc[R][CPT][K];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int q = 0; q < CPT; q++)
#pragma unroll
      for (int k = 0; k < K; k++) acc[r][q][k] = T(0);
  // ---- SPE
  if (tb.call_spe) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r < nr) {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
          const XV xv = *reinterpret_cast<const XV*>(x + (bfr[r] + cl[q]) * K);
#pragma unroll
          for (int k = 0; k < K; k++) {
            const T e = SH ? (T)tb.pspe_sd[f0 + r] + (T)tb.nspe_sd[nf0 + cl[q]]
                           : (T)bv.pspe[(int64_t)(f0 + r) * K + k] + (T)bv.nspe[(int64_t)(nf0 + cl[q]) * K + k];
            acc[r][q][k] = fma(e, xv.v[k], acc[r][q][k]);
          }
        }
      }
    }
  }
  // ---- PP: the union of the tile rows' source rows, two in flight
  {
    const int64_t e0 = tb.tpp_ptr[tile], e1 = tb.tpp_ptr[tile + 1];
    auto apply_src = [&](int64_t e, uint32_t sm, const XV (&xx)[CPT]) {
#pragma unroll
      for (int r = 0; r < R; r++)
        if (((sm >> 27) >> r) & 1u) {
          T v[K];
#pragma unroll
          for (int k = 0; k < K; k++)
            v[k] = SH ? vload<T>(tb.tpp_val32, tb.tpp_val64, e * R + r)
                      : bval<T>(bv.kt32, bv.kt64, (e * R + r) * K + k);
#pragma unroll
          for (int q = 0; q < CPT; q++)
#pragma unroll
            for (int k = 0; k < K; k++) acc[r][q][k] = fma(v[k], xx[q].v[k], acc[r][q][k]);
        }
    };
    int64_t e = e0;
    for (; e + 2 <= e1; e += 2) {
      const uint32_t sa = tb.tpp_src[e], sb = tb.tpp_src[e + 1];
      const T* xa = x + tb.pstart[sa & 0x7ffffffu] * K;
      const T* xb = x + tb.pstart[sb & 0x7ffffffu] * K;
      XV ua[CPT], ub[CPT];
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        ua[q] = *reinterpret_cast<const XV*>(xa + (int64_t)cl[q] * K);
        ub[q] = *reinterpret_cast<const XV*>(xb + (int64_t)cl[q] * K);
      }
      apply_src(e, sa, ua);
      apply_src(e + 1, sb, ub);
    }
    if (e < e1) {
      const uint32_t sa = tb.tpp_src[e];
      const T* xa = x + tb.pstart[sa & 0x7ffffffu] * K;
      XV ua[CPT];
#pragma unroll
      for (int q = 0; q < CPT; q++) ua[q] = *reinterpret_cast<const XV*>(xa + (int64_t)cl[q] * K);
      apply_src(e, sa, ua);
    }
  }
  // ---- NN: both rows staged [col][row][k]; one entry's value serves both
  // rows and all K vectors
  {
    for (int c = tid; c < L; c += BLOCK) {
      SV v;
#pragma unroll
      for (int r = 0; r < R; r++) {
        XV u;
        if (bfr[r] >= 0) u = *reinterpret_cast<const XV*>(x + (bfr[r] + c) * K);
        else
#pragma unroll
          for (int k = 0; k < K; k++) u.v[k] = T(0);
        if constexpr (SPLIT) *reinterpret_cast<XV*>(smem + XS + r * SLOTB + c * COLB) = u;
        else
#pragma unroll
          for (int k = 0; k < K; k++) v.v[r * K + k] = u.v[k];
      }
      if constexpr (!SPLIT) *reinterpret_cast<SV*>(smem + XS + c * COLB) = v;  // one wide store
    }
    __syncthreads();
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(tb.nn_src + tb.nn_base[cn]);
    const __amdgpu_buffer_rsrc_t rv =
        sizeof(T) == 4 ? buf_rsrc(tb.nn_val32 + tb.nn_base[cn]) : buf_rsrc(tb.nn_val64 + tb.nn_base[cn]);
    const int E = tb.nn_emax[cn];
    const int64_t nb0 = tb.nn_base[cn];
    for (int e = 0; e < E; e++) {
      const uint32_t so = (uint32_t)e * (uint32_t)L * 4u;
      int sl[CPT];
      T v[CPT][WK];
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        sl[q] = (int)buf_u32(rs, cl[q] * 4u, so);
        if constexpr (SH) {
          if constexpr (sizeof(T) == 4) v[q][0] = __uint_as_float(buf_u32(rv, cl[q] * 4u, so));
          else v[q][0] = buf_f64(rv, cl[q] * 8u, so * 2u);
        } else {
          // this slot's K values, one per Hamiltonian
          const int64_t o = nb0 + (int64_t)e * L + cl[q];
#pragma unroll
          for (int k = 0; k < WK; k++) v[q][k] = bval<T>(bv.nn32, bv.nn64, o * K + k);
        }
      }
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        SV xv;
        if constexpr (SPLIT) {
#pragma unroll
          for (int r = 0; r < R; r++) {
            const XV u = *reinterpret_cast<const XV*>(smem + XS + r * SLOTB + sl[q] * COLB);
#pragma unroll
            for (int k = 0; k < K; k++) xv.v[r * K + k] = u.v[k];
          }
        } else {
          xv = *reinterpret_cast<const SV*>(smem + XS + sl[q] * COLB);
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
          for (int k = 0; k < K; k++) acc[r][q][k] = fma(v[q][SH ? 0 : k], xv.v[r * K + k], acc[r][q][k]);
      }
    }
  }
  // ---- PN: per Delta class, batch b holds jump b of each tile row (slot g =
  // row g; a row with fewer jumps leaves its slot's coefficients zero)
  constexpr int XSH = __builtin_ctz(COLB);  // log2 bytes between staged columns
  constexpr int WSH = __builtin_ctz(GW * (int)sizeof(float));  // log2 bytes per coefficient row
  constexpr int kWbPer = (kWbRows * GW + BLOCK - 1) / BLOCK;
  const int nd = tb.ndelta, nd1 = nd + 1;
  for (int k0 = 0; k0 < nd; k0++) {
    int p0[R], nj[R];
    int nbmax = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      p0[r] = nj[r] = 0;
      if (r < nr) {
        p0[r] = tb.pcls[(int64_t)(f0 + r) * nd1 + k0];
        nj[r] = tb.pcls[(int64_t)(f0 + r) * nd1 + k0 + 1] - p0[r];
        nbmax = max(nbmax, nj[r]);
      }
    }
    if (nbmax == 0) continue;
    const int kn = tb.negk[k0];
    const int si = tb.srcsec[sp * nd + k0];
    const int Li = tb.nlen[tb.pconj[si]];
    const int wsd = tb.wstride[k0];
    const float* Wk = tb.W + tb.woff[k0];
    const float* Wb = bv.W + (int64_t)tb.woff[k0] * K;  // per Hamiltonian: [element][K]
    // the four-row FP32 quad table: x offset src * 16 bytes (bits 0-16),
    // coefficient offset (2*op + sign) * 16 bytes (bits 17-31)
    const __amdgpu_buffer_rsrc_t rpk = buf_rsrc(tb.n1_pkv[2][0] + tb.n1q_base[(int64_t)cn * nd + kn]);
    const uint8_t* ch = tb.n1_chunk + ((int64_t)cn * nd + kn) * 64 + __builtin_amdgcn_readfirstlane(tid >> 6);
    int Eq[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) Eq[q] = ch[q * (BLOCK / 64)];
    for (int b = 0; b < nbmax; b++) {
      float wreg[kWbPer];
#pragma unroll
      for (int u = 0; u < kWbPer; u++) {
        const int i = tid + u * BLOCK;
        const int row = i / GW, r2 = i - row * GW, g = r2 / WK, k = r2 - g * WK, nl = row >> 1;
        float w = 0.0f;
        const int njg = g == 0 ? nj[0] : nj[1], p0g = g == 0 ? p0[0] : p0[1];
        if (i < 2 * (wsd + 1) * GW && b < njg && nl < wsd) {
          const int meta = tb.p1_meta[p0g + b];
          w = SH ? Wk[(meta & 0xff) * wsd + nl] : Wb[(int64_t)((meta & 0xff) * wsd + nl) * K + k];
          if ((meta >> 8) ^ (row & 1)) w = -w;
        }
        wreg[u] = w;
      }
      __syncthreads();
      {
        int64_t bi[G];
#pragma unroll
        for (int g = 0; g < G; g++) bi[g] = b < nj[g] ? tb.pstart[tb.p1_src[p0[g] + b]] : -1;
        for (int c = tid; c < Li; c += BLOCK) {
          SV v;
#pragma unroll
          for (int g = 0; g < G; g++) {
            XV u;
            if (bi[g] >= 0) u = *reinterpret_cast<const XV*>(x + (bi[g] + c) * K);
            else
#pragma unroll
              for (int k = 0; k < K; k++) u.v[k] = T(0);
            if constexpr (SPLIT) *reinterpret_cast<XV*>(smem + XS + g * SLOTB + c * COLB) = u;
            else
#pragma unroll
              for (int k = 0; k < K; k++) v.v[g * K + k] = u.v[k];
          }
          if constexpr (!SPLIT) *reinterpret_cast<SV*>(smem + XS + c * COLB) = v;  // one wide store
        }
      }
#pragma unroll
      for (int u = 0; u < kWbPer; u++) {
        const int i = tid + u * BLOCK;
        if (i < 2 * (wsd + 1) * GW) wb[i] = wreg[u];
      }
      __syncthreads();
      walk_quads<CPT>(
          rpk, cl, Eq, (uint32_t)L * 16u,
          [&](uint32_t pk) {
            const uint32_t xo = XSH >= 4 ? (pk & 0x1ffffu) << (XSH - 4) : (pk & 0x1ffffu) >> (4 - XSH);
            const uint32_t wo = WSH >= 4 ? (pk >> 17) << (WSH - 4) : (pk >> 17) >> (4 - WSH);
            MultiEntry<T, GK, GW> en;
            if constexpr (SPLIT) {
#pragma unroll
              for (int g = 0; g < G; g++) {
                const XV u = *reinterpret_cast<const XV*>(smem + XS + g * SLOTB + xo);
#pragma unroll
                for (int k = 0; k < K; k++) en.x.v[g * K + k] = u.v[k];
              }
            } else {
              en.x = *reinterpret_cast<const SV*>(smem + XS + xo);
            }
            en.w = *reinterpret_cast<const VecG<float, GW>*>(smem + wo);
            return en;
          },
          [&](int q, const MultiEntry<T, GK, GW>& en) {
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
              for (int k = 0; k < K; k++)
                acc[g][q][k] = fma((T)en.w.v[g * WK + (SH ? 0 : k)], en.x.v[g * K + k], acc[g][q][k]);
          });
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (r < nr) {
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        const int c = tid + q * BLOCK;
        if (c < L) {
          XV o;
#pragma unroll
          for (int k = 0; k < K; k++) o.v[k] = acc[r][q][k];
          *reinterpret_cast<XV*>(y + (bfr[r] + c) * K) = o;
        }
      }
    }
  }
}

// part[b*m*K + j*K + k] = partial sum over this block's states of V_j[s][k] * w[s][k]
template <typename T>
__global__ __launch_bounds__(256) void k_bmultidot(const T* __restrict__ V, int64_t ldv, int m, const T* __restrict__ w,
                                                   int64_t D, int K, double* __restrict__ part) {
  __shared__ double red[4][64];
  const int k = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t chunk = (D + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(lo + chunk, D);
  for (int j = 0; j < m; j++) {
    double a = 0.0;
    if (k < K)
      for (int64_t st = lo + wv; st < hi; st += 4) a += (double)V[(int64_t)j * ldv + st * K + k] * (double)w[st * K + k];
    red[wv][k] = a;
    __syncthreads();
    if (wv == 0 && k < K)
      part[((int64_t)blockIdx.x * m + j) * K + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
    __syncthreads();
  }
}

// out[j*K + k] = sum_b part[...]; mode 1: sqrt
__global__ void k_breduce(const double* __restrict__ part, int nb, int m, int K, double* __restrict__ out, int mode) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * K) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[(int64_t)b * m * K + idx];
  out[idx] = mode == 1 ? sqrt(s) : s;
}

template <typename T>
__global__ void k_bsub_scaled(T* __restrict__ w, const T* __restrict__ v, const double* __restrict__ a, int64_t D, int K) {
  const int64_t n = D * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = w[i] - (T)(a[i % K] * (double)v[i]);
}

template <typename T>
__global__ void k_bmultiaxpy(T* __restrict__ w, const T* __restrict__ V, int64_t ldv, int m, const double* __restrict__ d,
                             int64_t D, int K) {
  const int64_t n = D * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % K);
    double s = 0.0;
    for (int j = 0; j < m; j++) s += d[j * K + k] * (double)V[(int64_t)j * ldv + i];
    w[i] = w[i] - (T)s;
  }
}

template <typename T>
__global__ void k_bnormalize(T* __restrict__ v, const T* __restrict__ w, const double* __restrict__ beta, int64_t D, int K) {
  const int64_t n = D * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (T)((double)w[i] * (1.0 / beta[i % K]));
}

// Breakdown handling of the batched recurrence, per Hamiltonian k, on the
// device so that whole cycles stay captured (the rule of lanczos_p,
// src/blanczosmain.f90:1634-1658): flags[k] = 2 when beta_i < 1e-4 of the
// previous beta (|H v_i| on a cycle's first step i == kstart: sqrt(alpha^2 +
// beta^2 + sum arrow^2)), 1 when beta_i < 0.1 of it, else 0; flags[K] = the
// largest, so that the gated kernels below exit at once when nothing broke.
__global__ void k_bbreak(const double* __restrict__ alpha, const double* __restrict__ beta,
                         const double* __restrict__ arrow, int i, int kstart, int K, int* __restrict__ flags) {
  const int k = threadIdx.x;  // one wave, K <= 64
  int f = 0;
  if (k < K) {
    const double a = alpha[(int64_t)i * K + k], b = beta[(int64_t)i * K + k];
    double scale;
    if (i == kstart) {
      double s2 = a * a + b * b;
      for (int j = 0; j < kstart; j++) s2 += arrow[(int64_t)j * K + k] * arrow[(int64_t)j * K + k];
      scale = sqrt(s2);
    } else {
      scale = beta[(int64_t)(i - 1) * K + k];
    }
    f = !(b >= 1e-4 * scale) || b == 0.0 ? 2 : (b < 0.1 * scale ? 1 : 0);
    flags[k] = f;
  }
  for (int o = 32; o > 0; o >>= 1) f = max(f, __shfl_down(f, o, 64));
  if (k == 0) flags[K] = f;
}

// v_{i+1}: w / beta for members without breakdown, a seeded random vector
// (SplitMix64 of the element index) for members with flag 2, zero when the
// Krylov space already spans the basis (last = 1)
template <typename T>
__global__ void k_bnext(T* __restrict__ v, const T* __restrict__ w, const double* __restrict__ beta, int64_t D, int K,
                        const int* __restrict__ flags, uint64_t seed, int last) {
  const int64_t n = D * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % K);
    T out;
    if (last) {
      out = T(0);
    } else if (flags[k] == 2) {
      uint64_t z = seed + (uint64_t)i * 0x9E3779B97F4A7C15ULL;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      z ^= z >> 31;
      out = (T)(((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0) / sqrt((double)D));
    } else {
      out = (T)((double)w[i] * (1.0 / beta[k]));
    }
    v[i] = out;
  }
}

// gated k_bmultidot: nothing unless some member has flags >= minflag
template <typename T>
__global__ __launch_bounds__(256) void k_bmultidot_g(const T* __restrict__ V, int64_t ldv, int m,
                                                     const T* __restrict__ w, int64_t D, int K,
                                                     double* __restrict__ part, const int* __restrict__ flags,
                                                     int minflag) {
  if (flags[K] < minflag) return;
  __shared__ double red[4][64];
  const int k = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t chunk = (D + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(lo + chunk, D);
  for (int j = 0; j < m; j++) {
    double a = 0.0;
    if (k < K && flags[k] >= minflag)
      for (int64_t st = lo + wv; st < hi; st += 4) a += (double)V[(int64_t)j * ldv + st * K + k] * (double)w[st * K + k];
    red[wv][k] = a;
    __syncthreads();
    if (wv == 0 && k < K)
      part[((int64_t)blockIdx.x * m + j) * K + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
    __syncthreads();
  }
}

__global__ void k_breduce_g(const double* __restrict__ part, int nb, int m, int K, double* __restrict__ out, int mode,
                            const int* __restrict__ flags, int minflag) {
  if (flags[K] < minflag) return;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * K) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[(int64_t)b * m * K + idx];
  out[idx] = mode == 1 ? sqrt(s) : s;
}

// members with flag 1 keep their vector unless it overlaps a stored one by
// more than 1e-5 (br_check_orthogonal, src/breorthog.f90:977-1045): their
// flag drops to 0 when it does not
__global__ void k_bloss(const double* __restrict__ dots, int m, int K, int* __restrict__ flags) {
  const int k = threadIdx.x;
  if (k >= K || flags[k] != 1) return;
  double dmax = 0.0;
  for (int j = 0; j < m; j++) dmax = fmax(dmax, fabs(dots[(int64_t)j * K + k]));
  if (dmax <= 1e-5) flags[k] = 0;
}

// single Lanczos run: step i needs the host's decision when beta_i is below
// 0.1 of its scale (beta_{i-1}, or |H v_i| on a cycle's first step, with the
// arrow couplings beta[0:kst]) or not finite; the first such step is kept
__global__ void k_bflag(const double* __restrict__ alpha, const double* __restrict__ beta, int i, int kst,
                        int* __restrict__ first) {
  if (threadIdx.x != 0) return;
  const double a = alpha[i], b = beta[i];
  double scale;
  if (i == kst) {
    double s2 = a * a + b * b;
    for (int j = 0; j < kst; j++) s2 += beta[j] * beta[j];
    scale = sqrt(s2);
  } else {
    scale = beta[i - 1];
  }
  if (!(b >= 0.1 * scale) || b == 0.0 || !isfinite(a) || !isfinite(b)) atomicMin(first, i);
}

// w -= T(sum_j d_j V_j) and then w /= norm, for members with flags >= minflag
template <typename T>
__global__ void k_bmultiaxpy_m(T* __restrict__ w, const T* __restrict__ V, int64_t ldv, int m,
                               const double* __restrict__ d, int64_t D, int K, const int* __restrict__ flags,
                               int minflag) {
  if (flags[K] < minflag) return;
  const int64_t n = D * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % K);
    if (flags[k] < minflag) continue;
    double s = 0.0;
    for (int j = 0; j < m; j++) s += d[j * K + k] * (double)V[(int64_t)j * ldv + i];
    w[i] = w[i] - (T)s;
  }
}

template <typename T>
__global__ void k_bnormalize_m(T* __restrict__ v, const double* __restrict__ nrm, int64_t D, int K,
                               const int* __restrict__ flags, int minflag) {
  if (flags[K] < minflag) return;
  const int64_t n = D * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % K);
    if (flags[k] >= minflag) v[i] = (T)((double)v[i] * (1.0 / nrm[k]));
  }
}

// Lowest nkeep eigenpairs of K symmetric tridiagonal matrices (alpha[i*K+k],
// beta[i*K+k] = coupling of i and i+1), FP64, one thread per matrix: Sturm
// bisection for the eigenvalues, inverse iteration (tridiagonal LU with
// partial pivoting) for the vectors.  Y[(k*m + i)*nkeep + n].
__global__ void k_tridiag_ritz(const double* __restrict__ alpha, const double* __restrict__ beta, int m, int K,
                               int nkeep, double* __restrict__ evals, double* __restrict__ Y, double* __restrict__ work) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  auto a = [&](int i) { return alpha[(int64_t)i * K + k]; };
  auto b = [&](int i) { return beta[(int64_t)i * K + k]; };
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; i++) {
    const double r = (i > 0 ? fabs(b(i - 1)) : 0.0) + (i + 1 < m ? fabs(b(i)) : 0.0);
    lo = fmin(lo, a(i) - r);
    hi = fmax(hi, a(i) + r);
  }
  double* wk = work + (int64_t)k * 6 * m;
  double* dd = wk;         // LU diagonal
  double* du = wk + m;     // first superdiagonal
  double* du2 = wk + 2 * m;  // second superdiagonal (pivoting)
  double* dl = wk + 3 * m;   // multipliers
  int* piv = (int*)(wk + 4 * m);
  double* v = wk + 5 * m;
  for (int n = 0; n < nkeep; n++) {
    // count of eigenvalues below x (Sturm sequence)
    double l = lo, h = hi;
    for (int it = 0; it < 200 && h - l > 1e-15 * fmax(1.0, fabs(l) + fabs(h)); it++) {
      const double x = 0.5 * (l + h);
      int cnt = 0;
      double q = 1.0;
      for (int i = 0; i < m; i++) {
        const double bb = i > 0 ? b(i - 1) : 0.0;
        q = a(i) - x - (i > 0 ? bb * bb / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0.0) cnt++;
      }
      if (cnt > n) h = x;
      else l = x;
    }
    const double lam = 0.5 * (l + h);
    evals[(int64_t)k * nkeep + n] = lam;
    // LU of (T - lam) with partial pivoting (tridiagonal -> bandwidth 2)
    const double shift = lam + 1e-13 * fmax(1.0, fabs(lam));
    for (int i = 0; i < m; i++) {
      dd[i] = a(i) - shift;
      du[i] = i + 1 < m ? b(i) : 0.0;
      du2[i] = 0.0;
    }
    double sub = m > 1 ? b(0) : 0.0;
    for (int i = 0; i < m - 1; i++) {
      const double nxt_sub = i + 2 < m ? b(i + 1) : 0.0;
      if (fabs(dd[i]) >= fabs(sub)) {
        piv[i] = 0;
        const double mult = sub / (dd[i] == 0.0 ? 1e-300 : dd[i]);
        dl[i] = mult;
        dd[i + 1] -= mult * du[i];
      } else {
        piv[i] = 1;
        const double mult = dd[i] / sub;
        dl[i] = mult;
        dd[i] = sub;
        double t0 = du[i];
        du[i] = dd[i + 1];
        dd[i + 1] = t0 - mult * du[i];
        du2[i] = du[i + 1];
        du[i + 1] = -mult * du[i + 1];
      }
      sub = nxt_sub;
    }
    for (int i = 0; i < m; i++) v[i] = 1.0 / sqrt((double)m) * (1.0 + 0.01 * (double)((i * 7919) % 13));
    for (int rep = 0; rep < 4; rep++) {
      // forward: apply the row operations
      for (int i = 0; i < m - 1; i++) {
        if (piv[i]) {
          const double t0 = v[i];
          v[i] = v[i + 1];
          v[i + 1] = t0 - dl[i] * v[i];
        } else {
          v[i + 1] -= dl[i] * v[i];
        }
      }
      // back substitution
      for (int i = m - 1; i >= 0; i--) {
        double t0 = v[i];
        if (i + 1 < m) t0 -= du[i] * v[i + 1];
        if (i + 2 < m) t0 -= du2[i] * v[i + 2];
        v[i] = t0 / (dd[i] == 0.0 ? 1e-300 : dd[i]);
      }
      // orthogonalize against the previous vectors of this matrix, normalize
      for (int p = 0; p < n; p++) {
        double dot = 0.0;
        for (int i = 0; i < m; i++) dot += v[i] * Y[((int64_t)k * m + i) * nkeep + p];
        for (int i = 0; i < m; i++) v[i] -= dot * Y[((int64_t)k * m + i) * nkeep + p];
      }
      double nrm = 0.0;
      for (int i = 0; i < m; i++) nrm += v[i] * v[i];
      nrm = 1.0 / sqrt(nrm);
      for (int i = 0; i < m; i++) v[i] *= nrm;
    }
    for (int i = 0; i < m; i++) Y[((int64_t)k * m + i) * nkeep + n] = v[i];
  }
}

// X_n[s][k] = sum_j Y[(k*m + j)*ys + n] V_j[s][k] for n < ncol (<= 16)
template <typename T, typename U>
__global__ void k_bcombine(U* __restrict__ X, const T* __restrict__ V, int64_t ldv, int m, const double* __restrict__ Y,
                           int ys, int ncol, int64_t D, int K) {
  const int64_t n = D * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % K);
    double acc[16];
    for (int q = 0; q < ncol; q++) acc[q] = 0.0;
    for (int j = 0; j < m; j++) {
      const double vj = (double)V[(int64_t)j * ldv + i];
      const double* yk = Y + ((int64_t)k * m + j) * ys;
      for (int q = 0; q < ncol; q++) acc[q] += yk[q] * vj;
    }
    for (int q = 0; q < ncol; q++) X[(int64_t)q * n + i] = (U)acc[q];
  }
}

// ---------------------------------------------------------------- helpers
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  bool owned = true;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIP_OK(hipMalloc(&p, count * sizeof(T)));
  }
  // use memory owned elsewhere (not freed by release)
  void wrap(void* ptr, size_t count) {
    release();
    p = static_cast<T*>(ptr);
    n = count;
    owned = false;
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = true;
  }
  ~DevBuf() { release(); }
};

// Symmetric eigensolver for the small Lanczos / arrowhead matrices (FP64,
// cyclic Jacobi).  Eigenvalues ascending; Y column-major m x m (Y[j*m+k]
// = component j of eigenvector k).
void jacobi_eigen(int m, std::vector<double> A, std::vector<double>& e, std::vector<double>& Y) {
  Y.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++) Y[(size_t)i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) {
        tot += A[(size_t)i * m + j] * A[(size_t)i * m + j];
        if (i != j) off += A[(size_t)i * m + j] * A[(size_t)i * m + j];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; p++)
      for (int q = p + 1; q < m; q++) {
        double apq = A[(size_t)p * m + q];
        if (std::fabs(apq) < 1e-300) continue;
        double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; k++) {
          double akp = A[(size_t)k * m + p], akq = A[(size_t)k * m + q];
          A[(size_t)k * m + p] = c * akp - s * akq;
          A[(size_t)k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; k++) {
          double apk = A[(size_t)p * m + k], aqk = A[(size_t)q * m + k];
          A[(size_t)p * m + k] = c * apk - s * aqk;
          A[(size_t)q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; k++) {
          double ykp = Y[(size_t)k * m + p], ykq = Y[(size_t)k * m + q];
          Y[(size_t)k * m + p] = c * ykp - s * ykq;
          Y[(size_t)k * m + q] = s * ykp + c * ykq;
        }
      }
  }
  std::vector<int> ord(m);
  for (int i = 0; i < m; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return A[(size_t)a * m + a] < A[(size_t)b * m + b]; });
  e.resize(m);
  std::vector<double> Ys((size_t)m * m);
  for (int k = 0; k < m; k++) {
    e[k] = A[(size_t)ord[k] * m + ord[k]];
    for (int j = 0; j < m; j++) Ys[(size_t)j * m + k] = Y[(size_t)j * m + ord[k]];
  }
  Y.swap(Ys);
}

// Ritz pairs ordered as the reference's option tx orders them
// (find_lanczos_eigenvalues, src/blanczoslib.f90:277-327, txchoice 1): by
// distance to the target energy (selection sort), then the nkeep nearest by
// energy.  e ascending on entry; Y column-major m x m, column k = pair k.
void target_order(int m, int nkeep, double et, std::vector<double>& e, std::vector<double>& Y) {
  auto swap_pair = [&](int a, int b) {
    std::swap(e[a], e[b]);
    for (int j = 0; j < m; j++) std::swap(Y[(size_t)j * m + a], Y[(size_t)j * m + b]);
  };
  std::vector<double> d(m);
  for (int i = 0; i < m; i++) d[i] = std::fabs(e[i] - et);
  for (int i = 0; i < m - 1; i++) {
    int k = i;
    for (int j = i + 1; j < m; j++)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      std::swap(d[k], d[i]);
      swap_pair(k, i);
    }
  }
  for (int i = 0; i < std::min(nkeep, m) - 1; i++) {
    int k = i;
    for (int j = i + 1; j < std::min(nkeep, m); j++)
      if (e[j] < e[k]) k = j;
    if (k != i) swap_pair(k, i);
  }
}

// The nk lowest eigenvalues of the symmetric tridiagonal matrix with diagonal
// a[0..n) and off-diagonal b[0..n-1) (b[j] couples j and j+1), by Sturm
// sequence bisection (the count of negative pivots of T - xI is the number of
// eigenvalues below x) inside the Gershgorin interval.
std::vector<double> tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, int nk) {
  const int n = (int)a.size();
  double lo = a[0], hi = a[0];
  for (int j = 0; j < n; j++) {
    const double r = (j > 0 ? std::fabs(b[j - 1]) : 0.0) + (j + 1 < n ? std::fabs(b[j]) : 0.0);
    lo = std::min(lo, a[j] - r);
    hi = std::max(hi, a[j] + r);
  }
  auto below = [&](double x) {
    int c = 0;
    double q = a[0] - x;
    for (int j = 0;; j++) {
      if (q == 0.0) q = -1e-300;
      if (q < 0.0) c++;
      if (j + 1 == n) break;
      q = a[j + 1] - x - b[j] * b[j] / q;
    }
    return c;
  };
  std::vector<double> e(std::min(nk, n));
  for (int k = 0; k < (int)e.size(); k++) {
    double l = lo, h = hi;
    for (int it = 0; it < 200 && h - l > 1e-14 * std::max(1.0, std::fabs(l) + std::fabs(h)); it++) {
      const double m = 0.5 * (l + h);
      if (below(m) > k) h = m;
      else l = m;
    }
    e[k] = 0.5 * (l + h);
  }
  return e;
}

// one wave per 16-row strip, at most 2048 workgroups of 4 waves
int mfma_grid(int64_t D) {
  const int64_t waves = (D + 63) / 64;
  return (int)std::min<int64_t>(std::max<int64_t>((waves + 3) / 4, 1), 2048);
}

int grid_for(int64_t D) {
  int64_t g = (D + kBlock - 1) / kBlock;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 4096);
}

}  // namespace

struct bs_system {
  bs::Tables t;
  std::unique_ptr<bs_system> obs[2];   // J^2 and T^2 on the same basis (built on demand)
  DevBuf<unsigned char> hostx, hosty;  // staging for bs_apply_h_host
  std::vector<int64_t> nn_slot;        // ELL slot of each merged NN entry (batched mode)
  std::vector<int32_t> tpp_csr2;       // two-row PP union: [entry][row] merged PP entry, or -1
  bool kt_ready = false;               // a batch on this system holds two-row union values
  std::map<int, int> multi_choice;     // multi-vector kernel variant per (dtype, shared values, K)
  std::map<int, std::array<float, 8>> multi_ms;  // its autotuner timings
  // adjoint tables (built on first bs_density)
  bool dens_ready = false;
  DensTables dn{};
  DevBuf<int64_t> d_pp_ptr, d_nn_off, d_pn_off;
  DevBuf<int32_t> d_pp_src, d_pp_me, d_nn_mc;
  DevBuf<int8_t> d_pp_ph;
  DevBuf<uint64_t> d_nn_pk, d_pn_pk;
  std::vector<int32_t> nn_comp_to_me;
  bool jac_ready = false;
  std::vector<bs::ParamColumn> jac;
  bool uploaded = false;
  int last_random_restarts = 0;  // breakdown handling of the last Lanczos run
  int last_second_passes = 0;
  int last_steps = 0;            // Lanczos steps of its last Krylov sequence
  int last_thick_restarts = 0;   // thick restarts it took
  int cpt = 1;
  bool tiles_ok = true;        // neutron rows and Delta classes fit the tile kernels' LDS
  DevBuf<int32_t> rows_f, rows_c;  // k_apply_rows: (proton row, first column) per workgroup
  int nrows_wg = 0;
  int variant[2] = {-1, -1};   // chosen kernel variant per dtype (0: f32, 1: f64)
  float tune_ms[2][32] = {};    // autotuner timings (ms; 0 = not run / does not fit)
  DevTables dt{};
  DevBuf<int64_t> pstart, pp_ptr, nn_base, n1_base, n1q_base;
  DevBuf<int32_t> tile_f0[3], tile_nr[3], tile_map[3][2];
  DevBuf<int64_t> tpp_ptr[3];
  DevBuf<uint32_t> tpp_src[3];
  DevBuf<float> tpp_val32[3];
  DevBuf<double> tpp_val64[3];
  int ntiles[3] = {0, 0, 0};
  int nslots[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  DevBuf<int32_t> psec, pconj, nlen, nfirst, srcsec, pp_src, nn_emax, nn_src, pcls, p1_src, n1_emax;
  DevBuf<int32_t> woff, wrows, wstride, negk;
  DevBuf<int16_t> p1_meta;
  DevBuf<uint32_t> n1_pk, n1_pkv[3][2];
  DevBuf<uint8_t> n1_chunk;
  DevBuf<uint32_t> n1_pkr;
  DevBuf<float> pspe_sd, nspe_sd, pp_val32, nn_val32, W;
  DevBuf<double> pp_val64, nn_val64;
};

namespace {

constexpr int kTileRows[3] = {1, 2, 4};   // R of the kernel variants

int negk_of(const bs::Tables& t, int k) {
  return (int)(std::lower_bound(t.delta_of_class.begin(), t.delta_of_class.end(), -t.delta_of_class[k]) -
               t.delta_of_class.begin());
}
constexpr int kApplyBlock = 512;

// Slot order of one (sector, class) block of neutron entries (ELL, slot e of
// column n at pk[e*L + n]; pad = the block's zero entry).  The sum over a
// column's entries is order-free; they are placed in operator order.  A wave
// walks 64 consecutive columns slot by slot, reading per entry the staged
// source column and the coefficient row 2 * op + sign from LDS: in operator
// order the lanes at a slot mostly share the coefficient row (one broadcast)
// and, one operator mapping neighbouring determinants to neighbouring ones,
// read nearly consecutive source columns -- replaying the pf 6p6n slot
// patterns on the GPU, 19% fewer LDS bank-conflict cycles than source order.
void order_by_operator(uint32_t* pk, int L, int em, uint32_t pad) {
  if (em <= 1) return;
  std::vector<uint32_t> v;
  for (int n = 0; n < L; n++) {
    v.clear();
    for (int e = 0; e < em; e++) {
      const uint32_t x = pk[(int64_t)e * L + n];
      if (x != pad) v.push_back(x);
    }
    std::stable_sort(v.begin(), v.end(), [](uint32_t a, uint32_t b) {
      return (2u * ((a >> 16) & 0xffu) + ((a >> 24) & 1u)) < (2u * ((b >> 16) & 0xffu) + ((b >> 24) & 1u));
    });
    for (int e = 0; e < em; e++) pk[(int64_t)e * L + n] = e < (int)v.size() ? v[e] : pad;
  }
}

void upload_tables(bs_system* s) {
  const bs::Tables& t = s->t;
  const bs::Species& P = t.sp[0];
  const bs::Species& N = t.sp[1];
  const int nP = (int)P.mask.size(), nN = (int)N.mask.size();
  const int nd = t.ndelta;
  if ((int64_t)P.mask.size() >= INT32_MAX || (int64_t)N.mask.size() >= INT32_MAX)
    throw std::runtime_error("too many Slater determinants");
  std::vector<int32_t> psec(P.sdsec.begin(), P.sdsec.end());
  std::vector<int32_t> pconj(P.conj.begin(), P.conj.end());
  std::vector<int32_t> nlen(N.nsec), nfirst(N.nsec);
  int lmax = 1;
  for (int c = 0; c < N.nsec; c++) {
    nlen[c] = (int32_t)N.seccount[c];
    nfirst[c] = (int32_t)N.secfirst[c];
    lmax = std::max(lmax, nlen[c]);
  }
  // rows longer than the tile kernels' staged row, or Delta classes larger
  // than their coefficient table, leave only the row kernel (k_apply_rows)
  s->tiles_ok = lmax <= kMaxCols && 2 * (t.max_class + 1) <= kWbRows;
  if (lmax > 0xffff) throw std::runtime_error("neutron sector longer than 65535 SDs does not pack");
  std::vector<int32_t> srcsec((size_t)P.nsec * nd, -1);
  for (int sp = 0; sp < P.nsec; sp++)
    for (int k = 0; k < nd; k++)
      for (int q = 0; q < P.nsec; q++)
        if (P.secjz[q] == P.secjz[sp] - t.delta_of_class[k]) srcsec[(size_t)sp * nd + k] = q;
  // per-SD single-particle energies (used only when not subsumed)
  std::vector<float> pspe_sd(nP, 0.0f), nspe_sd(nN, 0.0f);
  for (int it = 0; it < 2; it++) {
    const bs::Species& S = t.sp[it];
    std::vector<float>& out = it == 0 ? pspe_sd : nspe_sd;
    const std::vector<float>& e = it == 0 ? t.pspe : t.nspe;
    for (size_t i = 0; i < S.mask.size(); i++) {
      float acc = 0.0f;
      for (uint64_t q = S.mask[i]; q; q &= q - 1) acc += e[t.sorb[__builtin_ctzll(q)]];
      out[i] = acc;
    }
  }
  // PP merged two-body (CSR by target)
  std::vector<float> ppv32(t.tb[0].val.size());
  for (size_t i = 0; i < ppv32.size(); i++) ppv32[i] = (float)t.tb[0].val[i];
  std::vector<int32_t> pp_src(t.tb[0].src.begin(), t.tb[0].src.end());
  for (int f = 0; f < nP; f++)
    for (int64_t e = t.tb[0].rowptr[f]; e < t.tb[0].rowptr[f + 1]; e++)
      if (P.sdsec[t.tb[0].src[e]] != P.sdsec[f]) throw std::runtime_error("PP jump leaves its sector");
  // NN merged two-body in ELL per neutron sector
  std::vector<int64_t> nn_base(N.nsec);
  std::vector<int32_t> nn_emax(N.nsec);
  int64_t tot = 0;
  for (int c = 0; c < N.nsec; c++) {
    int em = 0;
    for (int64_t n = N.secfirst[c]; n < N.secfirst[c] + N.seccount[c]; n++)
      em = std::max<int>(em, (int)(t.tb[1].rowptr[n + 1] - t.tb[1].rowptr[n]));
    nn_base[c] = tot;
    nn_emax[c] = em;
    tot += (int64_t)em * N.seccount[c];
  }
  std::vector<int32_t> nn_src(std::max<int64_t>(tot, 1), 0);
  s->nn_slot.assign(t.tb[1].src.size(), -1);
  std::vector<float> nnv32(std::max<int64_t>(tot, 1), 0.0f);
  std::vector<double> nnv64(std::max<int64_t>(tot, 1), 0.0);
  for (int c = 0; c < N.nsec; c++) {
    const int L = nlen[c];
    for (int64_t n = N.secfirst[c]; n < N.secfirst[c] + N.seccount[c]; n++) {
      const int64_t col = n - N.secfirst[c];
      int e = 0;
      for (int64_t q = t.tb[1].rowptr[n]; q < t.tb[1].rowptr[n + 1]; q++, e++) {
        int32_t src = t.tb[1].src[q];
        if (N.sdsec[src] != c) throw std::runtime_error("NN jump leaves its sector");
        const int64_t o = nn_base[c] + (int64_t)e * L + col;
        s->nn_slot[q] = o;
        nn_src[o] = (int32_t)(src - N.secfirst[c]);
        nnv64[o] = t.tb[1].val[q];
        nnv32[o] = (float)t.tb[1].val[q];
      }
    }
  }
  // proton one-body with class pointers
  const bs::OneBody& O0 = t.ob[0];
  if (O0.src.size() >= (size_t)INT32_MAX) throw std::runtime_error("too many proton one-body jumps");
  std::vector<int32_t> pcls((size_t)nP * (nd + 1));
  std::vector<int32_t> p1_src(O0.src.size());
  std::vector<int16_t> p1_meta(O0.src.size());
  for (int f = 0; f < nP; f++) {
    int64_t e = O0.rowptr[f];
    for (int k = 0; k <= nd; k++) {
      while (e < O0.rowptr[f + 1] && t.class_of_op[O0.op[e]] < k) e++;
      pcls[(size_t)f * (nd + 1) + k] = (int32_t)e;
    }
    pcls[(size_t)f * (nd + 1) + nd] = (int32_t)O0.rowptr[f + 1];
    for (int64_t q = O0.rowptr[f]; q < O0.rowptr[f + 1]; q++) {
      const int k = t.class_of_op[O0.op[q]];
      const int i = O0.src[q];
      if (srcsec[(size_t)P.sdsec[f] * nd + k] != P.sdsec[i]) throw std::runtime_error("proton jump source in the wrong sector");
      p1_src[q] = i;
      const int pl = t.local_of_op[O0.op[q]];
      if (pl > 255) throw std::runtime_error("operator class too large");
      p1_meta[q] = (int16_t)(pl | (O0.phase[q] < 0 ? 256 : 0));
    }
  }
  // neutron one-body: ELL per (target sector, class)
  const bs::OneBody& O1 = t.ob[1];
  std::vector<int64_t> n1_base((size_t)N.nsec * nd);
  std::vector<int32_t> n1_emax((size_t)N.nsec * nd);
  tot = 0;
  for (int c = 0; c < N.nsec; c++)
    for (int k = 0; k < nd; k++) {
      int em = 0;
      for (int64_t n = N.secfirst[c]; n < N.secfirst[c] + N.seccount[c]; n++) {
        int cnt = 0;
        for (int64_t q = O1.rowptr[n]; q < O1.rowptr[n + 1]; q++) cnt += t.class_of_op[O1.op[q]] == k;
        em = std::max(em, cnt);
      }
      n1_base[(size_t)c * nd + k] = tot;
      n1_emax[(size_t)c * nd + k] = em;
      tot += (int64_t)em * N.seccount[c];
    }
  std::vector<uint32_t> n1_pk(std::max<int64_t>(tot, 1), 0);
  std::vector<uint8_t> n1_chunk((size_t)N.nsec * nd * 64, 0);
  for (int c = 0; c < N.nsec; c++) {
    const int L = nlen[c];
    for (int k = 0; k < nd; k++) {
      const uint32_t pad = (uint32_t)t.class_size[k] << 16;  // zero column of the block
      const int64_t b = n1_base[(size_t)c * nd + k];
      const int em = n1_emax[(size_t)c * nd + k];
      for (int64_t n = N.secfirst[c]; n < N.secfirst[c] + N.seccount[c]; n++) {
        const int64_t col = n - N.secfirst[c];
        int e = 0;
        for (int64_t q = O1.rowptr[n]; q < O1.rowptr[n + 1]; q++) {
          if (t.class_of_op[O1.op[q]] != k) continue;
          const int32_t src = O1.src[q];
          int want = -1;
          for (int z = 0; z < N.nsec; z++)
            if (N.secjz[z] == N.secjz[c] - t.delta_of_class[k]) want = z;
          if (want < 0 || N.sdsec[src] != want) throw std::runtime_error("neutron jump source in the wrong sector");
          const int64_t sl = src - N.secfirst[want];
          const int nl = t.local_of_op[O1.op[q]];
          if (sl > 0xffff || nl >= 255) throw std::runtime_error("neutron jump does not pack");
          n1_pk[b + (int64_t)e * L + col] = (uint32_t)sl | ((uint32_t)nl << 16) | (O1.phase[q] < 0 ? (1u << 24) : 0u);
          e++;
        }
        if (s->tiles_ok) {
          uint8_t& cm = n1_chunk[((size_t)c * nd + k) * 64 + col / 64];
          if (e > 255) throw std::runtime_error("too many neutron jumps into one column");
          cm = (uint8_t)std::max<int>(cm, e);
        }
        for (; e < em; e++) n1_pk[b + (int64_t)e * L + col] = pad;
      }
    }
  }
  if (s->tiles_ok) {
    // (sector, class) blocks are disjoint: spread them over up to 16 host threads
    const int nblk = N.nsec * nd;
    std::atomic<int> next(0);
    auto work = [&]() {
      for (int b = next++; b < nblk; b = next++) {
        const int c = b / nd, k = b % nd;
        order_by_operator(n1_pk.data() + n1_base[(size_t)c * nd + k], nlen[c], n1_emax[(size_t)c * nd + k],
                          (uint32_t)t.class_size[k] << 16);
      }
    };
    const int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int i = 1; i < nt; i++) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
  }
  std::vector<int32_t> woff(nd), wrows(nd), wstride(nd), negk(nd);
  for (int k = 0; k < nd; k++) {
    int kn = (int)(std::lower_bound(t.delta_of_class.begin(), t.delta_of_class.end(), -t.delta_of_class[k]) -
                   t.delta_of_class.begin());
    negk[k] = kn;
    woff[k] = t.class_offset[k];
    wrows[k] = t.class_size[k];
    wstride[k] = t.class_size[kn];
  }
  // a proton jump of class k and the neutron jumps of class -k must meet on
  // the source row's columns: conj(source proton sector) == source neutron sector
  for (int sp = 0; sp < P.nsec; sp++)
    for (int k = 0; k < nd; k++) {
      int si = srcsec[(size_t)sp * nd + k];
      if (si < 0) continue;
      if (N.secjz[P.conj[si]] != N.secjz[P.conj[sp]] + t.delta_of_class[k])
        throw std::runtime_error("conjugate sectors inconsistent with Delta");
    }
  for (int f = 0; f < nP; f++)
    if (P.start[f] < 0 || P.start[f] + nlen[P.conj[P.sdsec[f]]] > t.dim) throw std::runtime_error("row outside the vector");
  // tiles for each R, and two workgroup->tile maps: identity (sector order)
  // and XCD-grouped (below)
  for (int v = 0; v < 3; v++) {
    std::vector<int32_t> f0, nr;
    for (int sp = 0; sp < P.nsec; sp++)
      for (int64_t f = P.secfirst[sp]; f < P.secfirst[sp] + P.seccount[sp]; f += kTileRows[v]) {
        f0.push_back((int32_t)f);
        int n = (int)std::min<int64_t>(kTileRows[v], P.secfirst[sp] + P.seccount[sp] - f);
        nr.push_back(n);
      }
    const int nt = (int)f0.size();
    s->tile_f0[v].upload(f0);
    s->tile_nr[v].upload(nr);
    {
      // PP source union per tile (sources of all R rows, each listed once)
      const int R = kTileRows[v];
      const bs::TwoBody& T = t.tb[0];
      std::vector<int64_t> ptr(nt + 1, 0);
      std::vector<uint32_t> src;
      std::vector<float> v32;
      std::vector<double> v64;
      std::vector<std::pair<int32_t, int64_t>> ent;  // (source SD, entry e of row r: e * 4 + r)
      if (R == 2) s->tpp_csr2.clear();
      for (int ti = 0; ti < nt; ti++) {
        ent.clear();
        for (int r = 0; r < nr[ti]; r++)
          for (int64_t e = T.rowptr[f0[ti] + r]; e < T.rowptr[f0[ti] + r + 1]; e++) ent.push_back({T.src[e], e * 4 + r});
        std::sort(ent.begin(), ent.end());
        for (size_t q = 0; q < ent.size();) {
          size_t u = q;
          const int32_t sd = ent[q].first;
          if ((uint32_t)sd >= (1u << 27)) throw std::runtime_error("too many proton SDs for the PP tile lists");
          uint32_t mask = 0;
          const size_t b = v32.size();
          v32.resize(b + R, 0.0f);
          v64.resize(b + R, 0.0);
          if (R == 2) s->tpp_csr2.insert(s->tpp_csr2.end(), {-1, -1});
          while (u < ent.size() && ent[u].first == sd) {
            // the row's merged value for this source (one entry per pair)
            const int r = (int)(ent[u].second & 3);
            const int64_t e = ent[u].second >> 2;
            v64[b + r] = T.val[e];
            v32[b + r] = (float)T.val[e];
            mask |= 1u << r;
            if (R == 2) {
              if (e > INT32_MAX) throw std::runtime_error("too many PP entries for the two-row union map");
              s->tpp_csr2[s->tpp_csr2.size() - 2 + r] = (int32_t)e;
            }
            u++;
          }
          src.push_back((uint32_t)sd | (mask << 27));
          q = u;
        }
        ptr[ti + 1] = (int64_t)src.size();
      }
      if (src.empty()) {
        src.push_back(0);
        v32.assign(R, 0.0f);
        v64.assign(R, 0.0);
      }
      s->tpp_ptr[v].upload(ptr);
      s->tpp_src[v].upload(src);
      s->tpp_val32[v].upload(v32);
      s->tpp_val64[v].upload(v64);
    }
    s->ntiles[v] = nt;
    std::vector<int32_t> ident(nt);
    for (int i = 0; i < nt; i++) ident[i] = i;
    s->tile_map[v][0].upload(ident);
    s->nslots[v][0] = nt;
    // XCD-grouped order: workgroup slot s runs on XCD s % 8 (round-robin
    // dispatch); slot 8j + x takes tile (j / kGroup) * 8 * kGroup + x * kGroup
    // + j % kGroup, so each XCD walks runs of kGroup consecutive tiles (which
    // share source rows in its L2) while all XCDs advance together
    constexpr int kXcd = 8, kGroup = 32;
    const int nsl = (nt + kXcd * kGroup - 1) / (kXcd * kGroup) * (kXcd * kGroup);
    std::vector<int32_t> xmap((size_t)nsl, -1);
    for (int sl = 0; sl < nsl; sl++) {
      const int x = sl % kXcd, j = sl / kXcd;
      const int tile = (j / kGroup) * kXcd * kGroup + x * kGroup + j % kGroup;
      xmap[sl] = tile < nt ? tile : -1;
    }
    s->tile_map[v][1].upload(xmap);
    s->nslots[v][1] = (int)xmap.size();
  }
  int cpt = 1;
  while (cpt * kApplyBlock < lmax) cpt *= 2;
  s->cpt = cpt;
  {
    std::vector<int32_t> rf, rc;
    for (int f = 0; f < nP; f++) {
      const int L = nlen[P.conj[P.sdsec[f]]];
      for (int c0 = 0; c0 < L; c0 += 256) {
        rf.push_back(f);
        rc.push_back(c0);
      }
    }
    if ((int64_t)rf.size() >= INT32_MAX) throw std::runtime_error("too many row-kernel workgroups");
    s->nrows_wg = (int)rf.size();
    if (rf.empty()) {
      rf.push_back(0);
      rc.push_back(0);
    }
    s->rows_f.upload(rf);
    s->rows_c.upload(rc);
  }

  s->pstart.upload(P.start);
  s->psec.upload(psec);
  s->pconj.upload(pconj);
  s->nlen.upload(nlen);
  s->nfirst.upload(nfirst);
  s->srcsec.upload(srcsec);
  s->pspe_sd.upload(pspe_sd);
  s->nspe_sd.upload(nspe_sd);
  s->pp_ptr.upload(t.tb[0].rowptr);
  s->pp_src.upload(pp_src);
  s->pp_val32.upload(ppv32);
  s->pp_val64.upload(t.tb[0].val);
  s->nn_base.upload(nn_base);
  s->nn_emax.upload(nn_emax);
  s->nn_src.upload(nn_src);
  s->nn_val32.upload(nnv32);
  s->nn_val64.upload(nnv64);
  s->pcls.upload(pcls);
  s->p1_src.upload(p1_src);
  s->p1_meta.upload(p1_meta);
  s->n1_base.upload(n1_base);
  s->n1_emax.upload(n1_emax);
  s->n1_pk.upload(n1_pk);
  s->n1_chunk.upload(n1_chunk);
  {
    std::vector<uint32_t> v(n1_pk.size());
    for (size_t i = 0; i < n1_pk.size(); i++) {
      const uint32_t pk = n1_pk[i];
      v[i] = (pk & 0xffffu) | ((2u * ((pk >> 16) & 0xffu) + ((pk >> 24) & 1u)) << 16);
    }
    s->n1_pkr.upload(v);
  }
  // the same entries as LDS byte offsets for each staged-row count G and
  // element type: x offset src*G*sizeof(T) (< 2^17) | Wb offset
  // (2*nl + sign)*G*4 << 17; padding (nl = class size) lands on a zero row.
  // Quad layout for 16-byte loads: per (sector, class) the slot count is
  // rounded up to a multiple of 4 and slot e of column n sits at
  // ((e/4)*L + n)*4 + e%4
  std::vector<int64_t> n1q_base((size_t)N.nsec * nd);
  int64_t qtot = 0;
  for (int c = 0; s->tiles_ok && c < N.nsec; c++)
    for (int k = 0; k < nd; k++) {
      n1q_base[(size_t)c * nd + k] = qtot;
      qtot += (int64_t)((n1_emax[(size_t)c * nd + k] + 3) / 4 * 4) * N.seccount[c];
    }
  std::vector<uint32_t> quad(std::max<int64_t>(qtot, 1), 0);
  for (int c = 0; s->tiles_ok && c < N.nsec; c++) {
    const int L = nlen[c];
    for (int k = 0; k < nd; k++) {
      const int em = n1_emax[(size_t)c * nd + k], em4 = (em + 3) / 4 * 4;
      const uint32_t pad = (uint32_t)t.class_size[k] << 16;
      for (int e = 0; e < em4; e++)
        for (int col = 0; col < L; col++)
          quad[n1q_base[(size_t)c * nd + k] + ((int64_t)(e / 4) * L + col) * 4 + e % 4] =
              e < em ? n1_pk[n1_base[(size_t)c * nd + k] + (int64_t)e * L + col] : pad;
    }
  }
  for (int gi = 0; gi < 3; gi++)
    for (int ti = 0; ti < 2; ti++) {
      const uint32_t G = 1u << gi, ts = ti == 0 ? 4u : 8u;
      std::vector<uint32_t> v(quad.size());
      for (size_t i = 0; i < quad.size(); i++) {
        const uint32_t pk = quad[i], sl = pk & 0xffffu, nl = (pk >> 16) & 0xffu, sg = (pk >> 24) & 1u;
        const uint32_t xo = sl * G * ts, wo = (2u * nl + sg) * G * 4u;
        if (s->tiles_ok && (xo >= (1u << 17) || wo >= (1u << 15)))
          throw std::runtime_error("neutron jump offsets do not pack");
        v[i] = xo | (wo << 17);
      }
      s->n1_pkv[gi][ti].upload(v);
    }
  s->n1q_base.upload(n1q_base);
  s->W.upload(t.wpn);
  s->woff.upload(woff);
  s->wrows.upload(wrows);
  s->wstride.upload(wstride);
  s->negk.upload(negk);

  DevTables& d = s->dt;
  d.dim = t.dim;
  d.nP = nP;
  d.nN = nN;
  d.ndelta = nd;
  d.call_spe = t.call_spe ? 1 : 0;
  d.lmax = lmax;
  d.wcols = t.max_class + 1;
  d.pstart = s->pstart.p;
  d.pspe_sd = s->pspe_sd.p;
  d.nspe_sd = s->nspe_sd.p;
  d.psec = s->psec.p;
  d.pconj = s->pconj.p;
  d.nlen = s->nlen.p;
  d.nfirst = s->nfirst.p;
  d.srcsec = s->srcsec.p;
  d.pp_ptr = s->pp_ptr.p;
  d.pp_src = s->pp_src.p;
  d.pp_val32 = s->pp_val32.p;
  d.pp_val64 = s->pp_val64.p;
  d.nn_base = s->nn_base.p;
  d.nn_emax = s->nn_emax.p;
  d.nn_src = s->nn_src.p;
  d.nn_val32 = s->nn_val32.p;
  d.nn_val64 = s->nn_val64.p;
  d.pcls = s->pcls.p;
  d.p1_src = s->p1_src.p;
  d.p1_meta = s->p1_meta.p;
  d.n1_base = s->n1_base.p;
  d.n1_emax = s->n1_emax.p;
  d.n1_pk = s->n1_pk.p;
  d.n1_chunk = s->n1_chunk.p;
  d.n1_pkr = s->n1_pkr.p;
  for (int gi = 0; gi < 3; gi++)
    for (int ti = 0; ti < 2; ti++) d.n1_pkv[gi][ti] = s->n1_pkv[gi][ti].p;
  d.n1q_base = s->n1q_base.p;
  d.W = s->W.p;
  d.woff = s->woff.p;
  d.wrows = s->wrows.p;
  d.wstride = s->wstride.p;
  d.negk = s->negk.p;
  s->uploaded = true;
}

// Kernel variants: (R rows per tile) x (G staged source rows).
constexpr int kNumShapes = 7;
// (R, G) shape x {identity, XCD} map x {staged after, loaded ahead (PF)}:
// variant v = pf * 2 * kNumShapes + map * kNumShapes + shape
constexpr int kNumVariants = 4 * kNumShapes;
constexpr int kVarR[kNumShapes] = {1, 1, 2, 2, 2, 4, 4};
constexpr int kVarG[kNumShapes] = {1, 2, 1, 2, 4, 2, 4};

template <typename T>
size_t lds_bytes_rg(const bs_system* s, int R, int G) {
  const int S = R > G ? R : G;
  return (size_t)S * kMaxCols * sizeof(T) + (size_t)kWbBytes * G;
}
template <typename T>
size_t lds_bytes(const bs_system* s, int v);

constexpr size_t kLdsLimit = 160 * 1024;

template <typename T, int CPT, int R, int G>
void launch_variant(bs_system* s, int map, bool pf, const T* x, T* y, int accumulate, hipStream_t st) {
  const int ri = R == 1 ? 0 : (R == 2 ? 1 : 2);
  DevTables d = s->dt;
  d.tile_f0 = s->tile_f0[ri].p;
  d.tile_nr = s->tile_nr[ri].p;
  d.tile_map = s->tile_map[ri][map].p;
  d.tpp_ptr = s->tpp_ptr[ri].p;
  d.tpp_src = s->tpp_src[ri].p;
  d.tpp_val32 = s->tpp_val32[ri].p;
  d.tpp_val64 = s->tpp_val64[ri].p;
  // static LDS (lds_bytes_rg).  Where it leaves room for one workgroup per
  // CU only (FP64 with four staged rows: 132 KB), the workgroup has 1024
  // threads with half the columns each, so a CU still holds 16 waves; every
  // column keeps its order of accumulation, so the result is the same.  The
  // PF forms walk a quad as two pairs (k_apply_tiles), which keeps the batch
  // loaded ahead in registers (pf 6p6n FP64 R4 G4 PF: 47.2 ms at 512 threads,
  // 58.1 at 1024 with quads, 42.7 with pairs)
  constexpr int S = R > G ? R : G;
  constexpr bool kWide = (size_t)kWbBytes * G + (size_t)S * kMaxCols * sizeof(T) > kLdsLimit / 2 && CPT >= 2;
  constexpr int B = kWide ? 2 * kApplyBlock : kApplyBlock;
  constexpr int C = kWide ? CPT / 2 : CPT;
  if (pf)
    hipLaunchKernelGGL((k_apply_tiles<T, B, C, R, G, true>), dim3(s->nslots[ri][map]), dim3(B), 0, st, d, x, y,
                       accumulate);
  else
    hipLaunchKernelGGL((k_apply_tiles<T, B, C, R, G, false>), dim3(s->nslots[ri][map]), dim3(B), 0, st, d, x, y,
                       accumulate);
  HIP_OK(hipGetLastError());
}

// the row kernel, after the tile variants
constexpr int kRowsVariant = kNumVariants;
constexpr int kAllVariants = kNumVariants + 1;

template <typename T>
size_t lds_bytes(const bs_system* s, int v) {
  if (v == kRowsVariant) return 0;
  if (!s->tiles_ok) return SIZE_MAX;
  return lds_bytes_rg<T>(s, kVarR[v % kNumShapes], kVarG[v % kNumShapes]);
}

template <typename T, int CPT>
void launch_cpt(bs_system* s, int v, const T* x, T* y, int accumulate, hipStream_t st) {
  const int m = (v / kNumShapes) % 2;
  const bool pf = v >= 2 * kNumShapes;
  switch (v % kNumShapes) {
    case 0: launch_variant<T, CPT, 1, 1>(s, m, pf, x, y, accumulate, st); break;
    case 1: launch_variant<T, CPT, 1, 2>(s, m, pf, x, y, accumulate, st); break;
    case 2: launch_variant<T, CPT, 2, 1>(s, m, pf, x, y, accumulate, st); break;
    case 3: launch_variant<T, CPT, 2, 2>(s, m, pf, x, y, accumulate, st); break;
    case 4: launch_variant<T, CPT, 2, 4>(s, m, pf, x, y, accumulate, st); break;
    case 5: launch_variant<T, CPT, 4, 2>(s, m, pf, x, y, accumulate, st); break;
    default: launch_variant<T, CPT, 4, 4>(s, m, pf, x, y, accumulate, st); break;
  }
}

template <typename T>
void launch_apply_v(bs_system* s, int v, const T* x, T* y, int accumulate, hipStream_t st) {
  if (v < 0 || v >= kAllVariants) throw std::runtime_error("bad kernel variant");
  if (lds_bytes<T>(s, v) > kLdsLimit) throw std::runtime_error("kernel variant does not fit in LDS");
  if (v == kRowsVariant) {
    if (s->nrows_wg > 0)
      hipLaunchKernelGGL(k_apply_rows<T>, dim3(s->nrows_wg), dim3(256), 0, st, s->dt, s->rows_f.p, s->rows_c.p, x, y,
                         accumulate);
    HIP_OK(hipGetLastError());
    return;
  }
  switch (s->cpt) {
    case 1: launch_cpt<T, 1>(s, v, x, y, accumulate, st); break;
    case 2: launch_cpt<T, 2>(s, v, x, y, accumulate, st); break;
    case 4: launch_cpt<T, 4>(s, v, x, y, accumulate, st); break;
    default: launch_cpt<T, 8>(s, v, x, y, accumulate, st); break;
  }
}

// Autotuner: time every variant that fits in LDS on this system (two launches
// each on a scratch vector) and keep the fastest.
template <typename T>
int autotune(bs_system* s, hipStream_t st) {
  const int di = sizeof(T) == 4 ? 0 : 1;
  const int64_t D = s->t.dim;
  DevBuf<T> x, y;
  x.alloc(D);
  y.alloc(D);
  k_fill<T><<<grid_for(D), kBlock, 0, st>>>(x.p, (T)0.5, D);
  HIP_OK(hipGetLastError());
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  float best = 1e30f;
  int bestv = -1;
  for (int v = 0; v < kAllVariants; v++) {
    if (lds_bytes<T>(s, v) > kLdsLimit) continue;
    launch_apply_v<T>(s, v, x.p, y.p, 0, st);
    HIP_OK(hipEventRecord(e0, st));
    launch_apply_v<T>(s, v, x.p, y.p, 0, st);
    HIP_OK(hipEventRecord(e1, st));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    s->tune_ms[di][v] = ms;
    if (ms < best) { best = ms; bestv = v; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (bestv < 0) throw std::runtime_error("no kernel variant fits in LDS");
  s->variant[di] = bestv;
  return bestv;
}

template <typename T>
void launch_apply(bs_system* s, const T* x, T* y, int accumulate, hipStream_t st) {
  const int di = sizeof(T) == 4 ? 0 : 1;
  if (s->variant[di] < 0) {
    // default: the largest tile that fits, until autotuned
    int v = 2 * kNumShapes - 1;
    while (v >= 0 && lds_bytes<T>(s, v) > kLdsLimit) v--;
    s->variant[di] = v >= 0 ? v : kRowsVariant;
  }
  launch_apply_v<T>(s, s->variant[di], x, y, accumulate, st);
}

// Lanczos workspace and loop
constexpr double kSmallBeta = 1e-4;  // smallbetatol, src/blanczosmain.f90:1460

// Strength-function extras: a pivot to start from (device, D elements of T;
// NULL = the uniform pivot 1/sqrt(D)), and for every Ritz value of the final
// projected matrix its value and |Y(0, k)|^2 ||pivot||^2.
template <typename T>
struct PivotRun {
  const T* pivot = nullptr;
  double* pivot_norm = nullptr;
  double* ritz = nullptr;
  double* strength = nullptr;
};

template <typename T>
int run_lanczos(bs_system* s, const bs_lanczos_params* p, double* alpha_out, double* beta_out,
                double* evals, void* evecs, double* resid, int* iters, const PivotRun<T>& pr = PivotRun<T>(),
                void* ext = nullptr, size_t ext_bytes = 0) {
  const int64_t D = s->t.dim;
  const hipStream_t st = nullptr;
  const int nkeep = p->nkeep;
  int niter = std::min<int64_t>(p->niter, D);
  if (nkeep < 1) return set_err("nkeep must be >= 1");
  if (niter < nkeep) return set_err("niter must be >= nkeep");
  // a Krylov space as large as the basis is exact after one cycle: no restart
  const int thick = p->niter >= D ? 0 : p->thick;
  if (thick < 0 || (thick > 0 && (thick < nkeep || thick > niter - 2)))
    return set_err("thick must be 0 or in [nkeep, niter - 2]");
  if (p->target < 0 || p->target > 2) return set_err("target must be 0 (lowest states), 1 or 2 (states nearest etarget)");
  if (p->target && p->thick <= 0) return set_err("targeting an energy needs a thick-restart run (thick > 0)");
  // the run is synchronous: it starts after all earlier device work, on
  // every stream (a pivot written on a caller's non-blocking stream included)
  HIP_OK(hipDeviceSynchronize());
  const int mtot = niter + 1;
  const int nb = grid_for(D) < 1024 ? grid_for(D) : 1024;
  // Lanczos vectors at a stride of whole 64-element blocks: every v_i starts
  // 256-byte aligned, so the streaming passes use 16-byte loads
  const int64_t ld = (D + 63) / 64 * 64;
  DevBuf<T> V, w;
  DevBuf<double> part, dots, ab;
  if (ext) {
    // caller's scratch (bs_lanczos_noalloc): the same five buffers carved
    // from it at 256-byte boundaries; the paths below that allocate (thick
    // restart, Ritz vectors, polish) are refused by the caller
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bV = up(sizeof(T) * (size_t)mtot * ld), bw = up(sizeof(T) * (size_t)ld);
    const size_t bp = up(sizeof(double) * (size_t)nb * mtot), bd = up(sizeof(double) * (size_t)mtot);
    const size_t ba = up(sizeof(double) * (2 * (size_t)mtot + 4));
    const size_t need = bV + bw + bp + bd + ba;
    if (need > ext_bytes)
      return set_err("this run needs " + std::to_string(need) + " bytes of scratch; the no-allocation entry point has " +
                     std::to_string(ext_bytes) + " (dim " + std::to_string(D) + ", " + std::to_string(niter) +
                     " iterations, " + std::to_string(sizeof(T)) + "-byte vectors)");
    if (thick || evecs || resid || p->polish) return set_err("the no-allocation run is a plain run without vectors");
    unsigned char* b = static_cast<unsigned char*>(ext);
    V.wrap(b, (size_t)mtot * ld);
    w.wrap(b + bV, ld);
    part.wrap(b + bV + bw, (size_t)nb * mtot);
    dots.wrap(b + bV + bw + bp, mtot);
    ab.wrap(b + bV + bw + bp + bd, 2 * (size_t)mtot + 4);
  } else {
    V.alloc((size_t)mtot * ld);
    w.alloc(D);
    part.alloc((size_t)nb * mtot);
    dots.alloc(mtot);
    ab.alloc(2 * (size_t)mtot + 4);  // alpha[0..], beta[mtot..], then the first flagged step (an int)
  }
  double* d_alpha = ab.p;
  double* d_beta = ab.p + mtot;
  const int g = grid_for(D);
  double pnorm = 1.0;
  if (pr.pivot) {
    // (src/blanczosmain.f90:1560-1563) the pivot normalized, its norm kept
    k_multidot<T><<<nb, kBlock, 0, st>>>(pr.pivot, D, 1, pr.pivot, D, part.p);
    k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, dots.p, 1);
    HIP_OK(hipMemcpy(&pnorm, dots.p, sizeof(double), hipMemcpyDeviceToHost));
    if (!(pnorm > 0.0) || !std::isfinite(pnorm)) return set_err("pivot vector is zero or not finite");
    k_normalize<T><<<g, kBlock, 0, st>>>(V.p, pr.pivot, dots.p, D);
  } else {
    k_fill<T><<<g, kBlock, 0, st>>>(V.p, (T)(1.0 / std::sqrt((double)D)), D);
  }
  HIP_OK(hipGetLastError());
  if (pr.pivot_norm) *pr.pivot_norm = pnorm;

  // v_{i+1} from w and beta_i, with the reference's breakdown handling
  // (src/blanczosmain.f90:1634-1658): beta_i below smallbetatol = 1e-4 of
  // the previous beta is a breakdown -- v_{i+1} becomes a seeded random
  // vector orthonormalized against the whole history
  // (random_restart_with_history, src/breorthog.f90:1471-1482) while beta_i
  // keeps its small value; beta_i below 0.1 of the previous one triggers the
  // orthogonality test (br_check_orthogonal, src/breorthog.f90:977-1045:
  // any |<v_h, v_{i+1}>| > 1e-5) and then a second Gram-Schmidt pass.  The
  // first step of a cycle has no previous beta; its scale is |H v_i|, which
  // is sqrt(alpha^2 + beta^2 + sum of the arrow couplings^2) there.  When the
  // Krylov space is the whole space (i + 1 == D) v_{i+1} is set to zero.
  int random_restarts = 0, second_passes = 0;
  // convergence test parameters (src/blanczosmain.f90:1469-1512)
  if (p->conv_test < 0 || p->conv_test > 3)
    return set_err("conv_test must be 0 (mean energy change), 1 (largest), 2 (mean Ritz-vector change) or 3 (largest)");
  if (p->conv_test >= 2 && p->target) return set_err("the Ritz-vector tests (conv_test 2, 3) take no target energy");
  if (p->ncheck_ex < 0) return set_err("ncheck_ex must be >= 0");
  const int itskip = std::max(10, (int)std::sqrt((double)nkeep));
  // a targeted run tests only the nkeep kept energies (src/blanczosmain.f90:1497-1499)
  const int ncheck = (int)std::min<int64_t>(D, (int64_t)nkeep + (p->target ? 0 : p->ncheck_ex));
  const int ntest = nkeep > 100 && thick == 0 ? (int)(std::sqrt((double)ncheck) + 1) : ncheck;
  // the test of convergence_test (src/blanczoslib.f90:2726-2770) between
  // two sets of Ritz values: test 0 sum_{j<nt} |dE_j| / sqrt(nt + 1) times
  // (1 + restarts), test 1 max_j |dE_j|
  auto energy_change = [&](const std::vector<double>& e0, const std::vector<double>& e1, int nt, int rst) {
    nt = std::min<int>(nt, (int)std::min(e0.size(), e1.size()));
    double sum = 0.0, mx = 0.0;
    for (int j = 0; j < nt; j++) {
      const double d = std::fabs(e0[j] - e1[j]);
      sum += d;
      mx = std::max(mx, d);
    }
    return p->conv_test == 1 ? mx : sum / std::sqrt((double)nt + 1.0) * (1.0 + rst);
  };
  // tests 2 and 3 (convergence_test's altconverge branch,
  // src/blanczoslib.f90:2741-2756): the first ncheck Ritz vectors of the
  // projected matrix (m x m, Y[k*m + j] component k of vector j) against the
  // previous solve's (mo x mo, zero beyond mo, as the reference's arrays are
  // zeroed before each solve), over components k < iter - 1:
  // vdot_j = sum_k Y[k][j] Yo[k][j]; test 2 sqrt(sum_j (|vdot_j| - 1)^2) /
  // ncheck, test 3 max_j ||vdot_j| - 1|.  nvec: columns that hold vectors (a
  // plain run of the reference solves for the lowest nkeep only,
  // src/blanczoslib.f90:230-234; the others count as zero)
  auto vector_change = [&](const std::vector<double>& Yn, int m, const std::vector<double>& Yo, int mo, int iter,
                           int nvec) {
    const int kmax = std::min(iter - 1, std::min(m, mo));
    double s2 = 0.0, mx = 0.0;
    for (int j = 0; j < ncheck; j++) {
      double vd = 0.0;
      if (j < nvec && j < m && j < mo)
        for (int k = 0; k < kmax; k++) vd += Yn[(size_t)k * m + j] * Yo[(size_t)k * mo + j];
      const double d = std::fabs(vd) - 1.0;
      s2 += d * d;
      mx = std::max(mx, std::fabs(d));
    }
    return p->conv_test == 3 ? mx : std::sqrt(s2) / ncheck;
  };
  // A step needs no host round trip: v_{i+1} = w / beta_i at once, and
  // k_bflag records on the device the first step whose beta falls below 0.1
  // of its scale (or is not finite).  Each time the host reads alpha/beta
  // anyway (a convergence check, a restart, the end) it first reads that
  // step; if there is one, the steps after it are discarded and it is redone
  // with the decision on the host (next_vector_host) -- the same decisions,
  // in the same order, as deciding at every step.
  int* d_first = reinterpret_cast<int*>(ab.p + 2 * (size_t)mtot + 2);
  auto reset_first = [&]() { k_fill<int><<<1, 64, 0, st>>>(d_first, INT_MAX, 1); };
  reset_first();
  auto orthonormalize = [&](T* v, int m) {
    k_multidot<T><<<nb, kBlock, 0, st>>>(V.p, ld, m, v, D, part.p);
    k_reduce<<<m, kBlock, 0, st>>>(part.p, nb, m, dots.p, 0);
    k_multiaxpy<T><<<g, kBlock, 0, st>>>(v, V.p, ld, m, dots.p, D);
    k_multidot<T><<<nb, kBlock, 0, st>>>(v, D, 1, v, D, part.p);
    k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, dots.p + m, 1);
    k_normalize<T><<<g, kBlock, 0, st>>>(v, v, dots.p + m, D);
  };
  auto next_vector = [&](int i, int kst) {
    T* vn = V.p + (int64_t)(i + 1) * ld;
    if (i + 1 >= D) {
      k_fill<T><<<g, kBlock, 0, st>>>(vn, T(0), D);
      return;
    }
    // the arrow couplings of a thick restart sit in beta[0:kst]
    k_bflag<<<1, 64, 0, st>>>(d_alpha, d_beta, i, kst, d_first);
    k_normalize<T><<<g, kBlock, 0, st>>>(vn, w.p, d_beta + i, D);
  };
  auto next_vector_host = [&](int i, int kst) {
    T* vn = V.p + (int64_t)(i + 1) * ld;
    double a = 0.0, bi = 0.0, bp = 0.0;
    HIP_OK(hipMemcpy(&a, d_alpha + i, sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&bi, d_beta + i, sizeof(double), hipMemcpyDeviceToHost));
    if (!std::isfinite(bi) || !std::isfinite(a))
      throw std::runtime_error("Lanczos step " + std::to_string(i) + ": alpha or beta is not finite");
    double scale;
    if (i == kst) {
      std::vector<double> arr(kst);
      if (kst) HIP_OK(hipMemcpy(arr.data(), d_beta, sizeof(double) * kst, hipMemcpyDeviceToHost));
      double s2 = a * a + bi * bi;
      for (int j = 0; j < kst; j++) s2 += arr[j] * arr[j];
      scale = std::sqrt(s2);
    } else {
      HIP_OK(hipMemcpy(&bp, d_beta + i - 1, sizeof(double), hipMemcpyDeviceToHost));
      scale = bp;
    }
    if (!(bi >= kSmallBeta * scale) || bi == 0.0) {
      k_pivot_random<T><<<g, kBlock, 0, st>>>(vn, D, 0x7a3f0e1d5c2bULL + 0x9E3779B9ULL * (uint64_t)(i + 1));
      k_multidot<T><<<nb, kBlock, 0, st>>>(vn, D, 1, vn, D, part.p);
      k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, dots.p, 1);
      k_normalize<T><<<g, kBlock, 0, st>>>(vn, vn, dots.p, D);
      orthonormalize(vn, i + 1);
      orthonormalize(vn, i + 1);
      random_restarts++;
      return;
    }
    // v_{i+1} = w_i / beta_i is already in place (next_vector ran first; w
    // itself has been overwritten by the later steps since)
    if (bi < 0.1 * scale) {
      const int m = i + 1;
      k_multidot<T><<<nb, kBlock, 0, st>>>(V.p, ld, m, vn, D, part.p);
      k_reduce<<<m, kBlock, 0, st>>>(part.p, nb, m, dots.p, 0);
      std::vector<double> d(m);
      HIP_OK(hipMemcpy(d.data(), dots.p, sizeof(double) * m, hipMemcpyDeviceToHost));
      double dmax = 0.0;
      for (double v : d) dmax = std::max(dmax, std::fabs(v));
      if (dmax > 1e-5) {
        orthonormalize(vn, m);
        second_passes++;
      }
    }
  };
  // before alpha/beta of steps < upto are used: the first step there that
  // needs the host's decision, redone now (the caller resumes after it), or -1
  int total_mv = 0;
  auto redo_flagged = [&](int upto, int kst) -> int {
    int f = INT_MAX;
    HIP_OK(hipMemcpy(&f, d_first, sizeof(int), hipMemcpyDeviceToHost));
    if (f >= upto) return -1;
    total_mv -= upto - 1 - f;  // the discarded steps f+1 .. upto-1
    reset_first();
    next_vector_host(f, kst);
    HIP_OK(hipGetLastError());
    return f;
  };
  auto check_finite = [&](const std::vector<double>& a, const std::vector<double>& b, int n) {
    for (int i = 0; i < n; i++)
      if (!std::isfinite(a[i]) || !std::isfinite(b[i]))
        throw std::runtime_error("Lanczos step " + std::to_string(i) + ": alpha or beta is not finite");
  };

  std::vector<double> alpha_h, beta_h;  // alpha/beta of the final projected matrix
  int kstart = 0;             // number of locked (thick-restart) vectors at the front
  std::vector<double> arrow;  // coupling of the locked vectors to v_kstart
  int restarts = 0;
  std::vector<double> e, Y;
  // one Lanczos step on v_i: w = H v_i, alpha_i, classical Gram-Schmidt
  // against v_0 .. v_{i-1}, beta_i, v_{i+1}
  auto step = [&](int i) {
    T* vi = V.p + (int64_t)i * ld;
    launch_apply<T>(s, vi, w.p, 0, st);
    total_mv++;
    k_multidot<T><<<nb, kBlock, 0, st>>>(vi, D, 1, w.p, D, part.p);
    k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, d_alpha + i, 0);
    k_sub_scaled<T><<<g, kBlock, 0, st>>>(w.p, vi, d_alpha + i, D);
    if (i > 0) {
      k_multidot<T><<<nb, kBlock, 0, st>>>(V.p, ld, i, w.p, D, part.p);
      k_reduce<<<i, kBlock, 0, st>>>(part.p, nb, i, dots.p, 0);
      k_multiaxpy<T><<<g, kBlock, 0, st>>>(w.p, V.p, ld, i, dots.p, D);
    }
    k_multidot<T><<<nb, kBlock, 0, st>>>(w.p, D, 1, w.p, D, part.p);
    k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, d_beta + i, 1);
    next_vector(i, kstart);
    HIP_OK(hipGetLastError());
  };
  // Ritz pairs of the projected matrix on v_0 .. v_{m-1}: the locked Ritz
  // values on the diagonal with the arrow couplings in column kstart
  // (src/blanczoslib.f90:143-222), then the tridiagonal part; ordered by
  // distance to the target when one is set (find_lanczos_eigenvalues)
  auto ritz = [&](int m) {
    alpha_h.assign(m, 0.0);
    beta_h.assign(m, 0.0);
    HIP_OK(hipMemcpy(alpha_h.data(), d_alpha, sizeof(double) * m, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(beta_h.data(), d_beta, sizeof(double) * m, hipMemcpyDeviceToHost));
    check_finite(alpha_h, beta_h, m);
    std::vector<double> A((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) A[(size_t)i * m + i] = alpha_h[i];
    for (int i = kstart; i + 1 < m; i++) {
      A[(size_t)i * m + i + 1] = beta_h[i];
      A[(size_t)(i + 1) * m + i] = beta_h[i];
    }
    for (int i = 0; i < kstart && kstart < m; i++) {
      A[(size_t)i * m + kstart] = arrow[i];
      A[(size_t)kstart * m + i] = arrow[i];
    }
    jacobi_eigen(m, A, e, Y);
    if (p->target) target_order(m, nkeep, p->etarget, e, Y);
  };
  int m = niter;  // size of the final projected matrix
  if (thick == 0) {
    // plain runs with a tolerance (options ld, lc): the reference's schedule
    // and test (lanczos_p, src/blanczosmain.f90:1489-1512, :1690-1724;
    // convergence_test, src/blanczoslib.f90:2660-2800): at every
    // itskipeig-th step (10, or sqrt(nkeep) if larger) beyond ncheck = nkeep
    // + ncheck_ex, the lowest ncheck Ritz values of the tridiagonal so far;
    // stop when, against the previous such check, the mean change of the
    // first ntest of them (sum |dE| / sqrt(ntest + 1), test 0) or their
    // largest change (test 1) is below tol
    std::vector<double> e_check, Y_check;  // the previous check's energies / Ritz vectors
    int m_check = 0;
    int i = 0;
    while (i < niter) {
      step(i);
      const int iter = i + 1;
      const bool check = p->tol > 0.0 && iter % itskip == 0 && iter > ncheck && iter < niter;
      if (check || iter == niter) {
        const int f = redo_flagged(iter, kstart);
        if (f >= 0 && f + 1 < iter) {
          i = f + 1;
          continue;
        }
      }
      if (check) {
        std::vector<double> a(iter), b(iter);
        HIP_OK(hipMemcpy(a.data(), d_alpha, sizeof(double) * iter, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(b.data(), d_beta, sizeof(double) * iter, hipMemcpyDeviceToHost));
        check_finite(a, b, iter);
        bool stop;
        if (p->conv_test >= 2) {
          std::vector<double> A((size_t)iter * iter, 0.0), ek, Yk;
          for (int r = 0; r < iter; r++) A[(size_t)r * iter + r] = a[r];
          for (int r = 0; r + 1 < iter; r++) A[(size_t)r * iter + r + 1] = A[(size_t)(r + 1) * iter + r] = b[r];
          jacobi_eigen(iter, A, ek, Yk);
          // (before the first check the earlier vectors are zero, as in the
          // reference's arrays, so its dots are 0)
          stop = vector_change(Yk, iter, Y_check, m_check, iter, nkeep) < p->tol;
          Y_check = std::move(Yk);
          m_check = iter;
        } else {
          std::vector<double> ek = tridiag_lowest(a, b, ncheck);
          stop = !e_check.empty() && energy_change(e_check, ek, ntest, 0) < p->tol;
          e_check = ek;
        }
        if (stop) {
          niter = iter;
          break;
        }
      }
      i++;
    }
    m = niter;
    ritz(m);
  } else {
    // thick restart (options td, tf, tc, tx), the reference's cycle
    // (lanczos_p, src/blanczosmain.f90:1690-1752; thick_restart_sub_p,
    // src/blanczoslib.f90:1656-1799): steps run to iter = niter - 1; the
    // restart keeps the `thick` leading Ritz vectors of the projected matrix
    // on the first niter - 2 vectors, couples them to v_{niter-1} (the next
    // start vector) by beta_{niter-2} Y[niter-2, k], and iter resumes at
    // thick.  The energies are taken at every itskipeig-th iter beyond
    // max(ncheck, nkeep) and at iter = niter - 1; the test (convergence_test)
    // compares them at every itskipeig-th iter beyond ncheck with the
    // previous iteration's, the mean change carrying the factor (1 + number
    // of restarts).  The run ends converged, at iter = niter (restart length
    // thick = niter - 1), or when a restart is due after max_restarts + 1
    // restarts.  A targeted run of the reference (tx) sets no count of tested
    // energies (ntest_converge keeps its initial 0, src/blanczosmain.f90:
    // 1497-1499), so its test passes at the first check; target = 2 tests
    // the nkeep kept energies instead.
    const int ntest_t = p->target == 1 ? 0 : ncheck;
    std::vector<double> ecur(ncheck, 0.0), eold(ncheck, 0.0);
    int m_solved = 0;  // size of the projected matrix of the last Ritz solve
    int iter = 0;
    bool finished = false;
    while (!finished) {
      iter++;
      step(iter - 1);
      const bool fly = iter % itskip == 0;
      if ((fly && iter > std::max(ncheck, nkeep)) || iter + 1 >= niter) {
        const int f = redo_flagged(iter, kstart);
        if (f >= 0 && f + 1 < iter) {
          iter = f + 1;
          continue;
        }
      }
      std::vector<double> Ylast;  // the Ritz vectors before this step's solve
      int mlast = m_solved;
      if ((fly && iter > std::max(ncheck, nkeep)) || iter + 1 == niter) {
        if (p->conv_test >= 2) Ylast = Y;
        ritz(iter);
        m_solved = iter;
        for (int j = 0; j < ncheck; j++) ecur[j] = j < iter ? e[j] : 0.0;
      }
      if (iter == niter) {
        finished = true;
      } else if (p->tol > 0.0) {
        if (iter > ncheck && fly) {
          if (p->conv_test >= 2) {
            // a check compares with the vectors of the solve before it (zero
            // before the first solve, as the reference's arrays)
            if (m_solved == iter && vector_change(Y, iter, Ylast, mlast, iter, ncheck) < p->tol) finished = true;
          } else if (energy_change(eold, ecur, ntest_t, restarts) < p->tol) {
            finished = true;
          }
        }
        eold = ecur;
      }
      if (finished || iter + 1 != niter) continue;
      if (restarts > p->max_restarts) break;
      const int mr = iter - 1;
      ritz(mr);
      m_solved = mr;
      for (int j = 0; j < ncheck; j++) ecur[j] = j < mr ? e[j] : 0.0;
      DevBuf<double> dY;
      std::vector<double> Yk((size_t)mr * thick);
      for (int j = 0; j < mr; j++)
        for (int k = 0; k < thick; k++) Yk[(size_t)j * thick + k] = Y[(size_t)j * mr + k];
      dY.upload(Yk);
      DevBuf<T> tmp;
      tmp.alloc((size_t)thick * ld);
      for (int k0 = 0; k0 < thick; k0 += 16) {
        int nk = std::min(16, thick - k0);
        k_combine_mfma<T, T><<<mfma_grid(D), 256, 0, st>>>(tmp.p + (int64_t)k0 * ld, ld, V.p, ld, mr, dY.p + k0,
                                                            thick, nk, D);
      }
      HIP_OK(hipMemcpyAsync(V.p, tmp.p, sizeof(T) * (size_t)thick * ld, hipMemcpyDeviceToDevice, st));
      if (mr != thick)
        HIP_OK(hipMemcpyAsync(V.p + (int64_t)thick * ld, V.p + (int64_t)mr * ld, sizeof(T) * D,
                              hipMemcpyDeviceToDevice, st));
      arrow.assign(thick, 0.0);
      std::vector<double> anew(thick);
      for (int k = 0; k < thick; k++) {
        arrow[k] = beta_h[mr - 1] * Y[(size_t)(mr - 1) * mr + k];
        anew[k] = e[k];
      }
      // the kept Ritz values and their couplings take the places of the
      // first thick alphas and betas, as in the reference's coefficients
      HIP_OK(hipMemcpy(d_alpha, anew.data(), sizeof(double) * thick, hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(d_beta, arrow.data(), sizeof(double) * thick, hipMemcpyHostToDevice));
      HIP_OK(hipStreamSynchronize(st));
      kstart = thick;
      restarts++;
      iter = thick;
    }
    m = iter;
    ritz(m);
  }
  for (int k = 0; k < nkeep; k++) evals[k] = e[k];
  // (src/blanczoslib.f90:1122-1136) E_k and eiglvec(1,k)^2 dnorm0^2
  for (int k = 0; k < m; k++) {
    if (pr.ritz) pr.ritz[k] = e[k];
    if (pr.strength) pr.strength[k] = Y[k] * Y[k] * pnorm * pnorm;
  }
  if (alpha_out) for (int i = 0; i < m; i++) alpha_out[i] = alpha_h[i];
  if (beta_out) for (int i = 0; i < m; i++) beta_out[i] = beta_h[i];
  if (evecs || resid || p->polish) {
    // Ritz vectors in FP64: x_k = sum_j Y[j,k] v_j
    DevBuf<double> dY, X64, HX, tmpd;
    std::vector<double> Yk((size_t)m * nkeep);
    for (int j = 0; j < m; j++)
      for (int k = 0; k < nkeep; k++) Yk[(size_t)j * nkeep + k] = Y[(size_t)j * m + k];
    dY.upload(Yk);
    X64.alloc((size_t)nkeep * D);
    HX.alloc(D);
    for (int k0 = 0; k0 < nkeep; k0 += 16)
      k_combine_mfma<T, double><<<mfma_grid(D), 256, 0, st>>>(X64.p + (int64_t)k0 * D, D, V.p, ld, m, dY.p + k0,
                                                               nkeep, std::min(16, nkeep - k0), D);
    HIP_OK(hipGetLastError());
    for (int k = 0; k < nkeep; k++) {
      double* xk = X64.p + (int64_t)k * D;
      // normalize, then FP64 Rayleigh quotient and residual
      k_multidot<double><<<nb, kBlock, 0, st>>>(xk, D, 1, xk, D, part.p);
      k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, dots.p, 1);
      k_normalize<double><<<g, kBlock, 0, st>>>(xk, xk, dots.p, D);
      if (p->polish || resid) {
        launch_apply<double>(s, xk, HX.p, 0, st);
        total_mv++;
        k_multidot<double><<<nb, kBlock, 0, st>>>(xk, D, 1, HX.p, D, part.p);
        k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, dots.p + 1, 0);
        double rq = 0.0;
        HIP_OK(hipMemcpy(&rq, dots.p + 1, sizeof(double), hipMemcpyDeviceToHost));
        if (p->polish) evals[k] = rq;
        if (resid) {
          k_sub_scaled<double><<<g, kBlock, 0, st>>>(HX.p, xk, dots.p + 1, D);
          k_multidot<double><<<nb, kBlock, 0, st>>>(HX.p, D, 1, HX.p, D, part.p);
          k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, dots.p + 2, 1);
          HIP_OK(hipMemcpy(&resid[k], dots.p + 2, sizeof(double), hipMemcpyDeviceToHost));
        }
      }
      if (evecs) {
        if (p->dtype == BS_F64)
          HIP_OK(hipMemcpyAsync((double*)evecs + (int64_t)k * D, xk, sizeof(double) * D, hipMemcpyDeviceToDevice, st));
        else
          k_convert<double, float><<<g, kBlock, 0, st>>>((float*)evecs + (int64_t)k * D, xk, D);
      }
    }
  }
  HIP_OK(hipStreamSynchronize(st));
  if (iters) *iters = total_mv;
  s->last_random_restarts = random_restarts;
  s->last_second_passes = second_passes;
  s->last_steps = m;
  s->last_thick_restarts = restarts;
  return 0;
}

bs_system* checked(bs_system* s) {
  if (!s) throw std::runtime_error("null system");
  return s;
}

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

void* scratch_ptr() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = ptr;
  }
  return g_scratch_dev[dev];
}

}  // namespace

namespace {
template <typename T>
double expectation(bs_system* o, const T* psi) {
  const int64_t D = o->t.dim;
  const int nb = std::min(grid_for(D), 1024);
  DevBuf<T> hy;
  DevBuf<double> part, out;
  hy.alloc(D);
  part.alloc(nb);
  out.alloc(2);
  launch_apply<T>(o, psi, hy.p, 0, nullptr);
  k_multidot<T><<<nb, kBlock>>>(psi, D, 1, hy.p, D, part.p);
  k_reduce<<<1, kBlock>>>(part.p, nb, 1, out.p, 0);
  k_multidot<T><<<nb, kBlock>>>(psi, D, 1, psi, D, part.p);
  k_reduce<<<1, kBlock>>>(part.p, nb, 1, out.p + 1, 0);
  HIP_OK(hipGetLastError());
  double h[2];
  HIP_OK(hipMemcpy(h, out.p, sizeof h, hipMemcpyDeviceToHost));
  return h[0] / h[1];
}
}  // namespace

namespace {

void build_density_tables(bs_system* s) {
  const bs::Tables& t = s->t;
  const bs::Species& P = t.sp[0];
  const bs::Species& N = t.sp[1];
  const int nd = t.ndelta;
  // PP: unmerged proton operators
  bs::TwoBodyOps op;
  bs::build_twobody_ops(t, 0, op);
  s->d_pp_ptr.upload(op.rowptr);
  s->d_pp_src.upload(op.src.empty() ? std::vector<int32_t>(1, 0) : op.src);
  s->d_pp_me.upload(op.me.empty() ? std::vector<int32_t>(1, 0) : op.me);
  s->d_pp_ph.upload(op.phase.empty() ? std::vector<int8_t>(1, 0) : op.phase);
  // NN: flat per-sector lists with compact element ids
  bs::TwoBodyOps on;
  bs::build_twobody_ops(t, 1, on);
  std::vector<int32_t> me_to_comp((size_t)t.npairxx * t.npairxx, -1);
  s->nn_comp_to_me.clear();
  std::vector<int64_t> off(N.nsec + 1, 0);
  std::vector<uint64_t> pk;
  std::vector<int32_t> mc;
  for (int c = 0; c < N.nsec; c++) {
    off[c] = (int64_t)pk.size();
    for (int64_t n = N.secfirst[c]; n < N.secfirst[c] + N.seccount[c]; n++)
      for (int64_t q = on.rowptr[n]; q < on.rowptr[n + 1]; q++) {
        const int64_t nf = n - N.secfirst[c], ns = on.src[q] - N.secfirst[c];
        if (ns < 0 || ns >= N.seccount[c]) throw std::runtime_error("NN operator leaves its sector");
        int32_t& cid = me_to_comp[on.me[q]];
        if (cid < 0) {
          cid = (int32_t)s->nn_comp_to_me.size();
          s->nn_comp_to_me.push_back(on.me[q]);
        }
        pk.push_back((uint64_t)nf | ((uint64_t)ns << 24) | (on.phase[q] < 0 ? (1ull << 48) : 0ull));
        mc.push_back(cid);
      }
  }
  off[N.nsec] = (int64_t)pk.size();
  if (pk.empty()) { pk.push_back(0); mc.push_back(0); }
  s->d_nn_off.upload(off);
  s->d_nn_pk.upload(pk);
  s->d_nn_mc.upload(mc);
  // PN: per (target neutron sector, class, neutron operator) lists
  const int mc1 = t.max_class + 1;
  std::vector<int64_t> poff((size_t)N.nsec * nd * mc1 + 1, 0);
  std::vector<std::vector<uint64_t>> bucket(mc1);
  std::vector<uint64_t> ppk;
  const bs::OneBody& O1 = t.ob[1];
  for (int c = 0; c < N.nsec; c++)
    for (int k = 0; k < nd; k++) {
      for (auto& b : bucket) b.clear();
      int want = -1;
      for (int z = 0; z < N.nsec; z++)
        if (N.secjz[z] == N.secjz[c] - t.delta_of_class[k]) want = z;
      for (int64_t n = N.secfirst[c]; n < N.secfirst[c] + N.seccount[c]; n++)
        for (int64_t q = O1.rowptr[n]; q < O1.rowptr[n + 1]; q++) {
          if (t.class_of_op[O1.op[q]] != k) continue;
          const int64_t nf = n - N.secfirst[c], ni = O1.src[q] - N.secfirst[want];
          bucket[t.local_of_op[O1.op[q]]].push_back((uint64_t)nf | ((uint64_t)ni << 24) |
                                                    (O1.phase[q] < 0 ? (1ull << 48) : 0ull));
        }
      for (int nl = 0; nl < mc1; nl++) {
        poff[((size_t)c * nd + k) * mc1 + nl] = (int64_t)ppk.size();
        ppk.insert(ppk.end(), bucket[nl].begin(), bucket[nl].end());
      }
    }
  poff[(size_t)N.nsec * nd * mc1] = (int64_t)ppk.size();
  // the kernel reads off[base + nl + 1]; the next base starts where this one ends
  if (ppk.empty()) ppk.push_back(0);
  s->d_pn_off.upload(poff);
  s->d_pn_pk.upload(ppk);
  DensTables& d = s->dn;
  d.pp_ptr = s->d_pp_ptr.p;
  d.pp_src = s->d_pp_src.p;
  d.pp_me = s->d_pp_me.p;
  d.pp_ph = s->d_pp_ph.p;
  d.nn_off = s->d_nn_off.p;
  d.nn_pk = s->d_nn_pk.p;
  d.nn_mc = s->d_nn_mc.p;
  d.nn_ncomp = (int)s->nn_comp_to_me.size();
  d.pn_off = s->d_pn_off.p;
  d.pn_pk = s->d_pn_pk.p;
  d.maxcls1 = mc1;
  (void)P;
  s->dens_ready = true;
}

template <typename T>
void run_density(bs_system* s, const T* psi, double* rho_pp, double* rho_nn, double* rho_pn, double* norm2) {
  if (!s->dens_ready) build_density_tables(s);
  const bs::Tables& t = s->t;
  const int64_t D = t.dim;
  const int npx = t.npairxx, nd = t.ndelta;
  DevBuf<unsigned long long> rpp, rnn, rw;
  DevBuf<double> part, nrm;
  rpp.alloc((size_t)std::max(npx, 1) * std::max(npx, 1));
  rnn.alloc(std::max<size_t>(s->nn_comp_to_me.size(), 1));
  rw.alloc(std::max<size_t>(t.wpn.size(), 1));
  HIP_OK(hipMemset(rpp.p, 0, rpp.n * sizeof(unsigned long long)));
  HIP_OK(hipMemset(rnn.p, 0, rnn.n * sizeof(unsigned long long)));
  HIP_OK(hipMemset(rw.p, 0, rw.n * sizeof(unsigned long long)));
  // |psi|^2 first: it bounds every element, and so fixes the fixed-point scale
  const int nb = std::min(grid_for(D), 1024);
  part.alloc(nb);
  nrm.alloc(1);
  k_multidot<T><<<nb, kBlock>>>(psi, D, 1, psi, D, part.p);
  k_reduce<<<1, kBlock>>>(part.p, nb, 1, nrm.p, 0);
  HIP_OK(hipGetLastError());
  double n2 = 0.0;
  HIP_OK(hipMemcpy(&n2, nrm.p, sizeof(double), hipMemcpyDeviceToHost));
  if (!std::isfinite(n2)) throw std::runtime_error("psi is not finite");
  // elements and partial sums below 64 |psi|^2, summed below 2^62
  int ex = 0;
  std::frexp(std::max(n2, 1e-300) * 64.0, &ex);
  const double scale = std::ldexp(1.0, std::min(62 - ex, 1000));
  DevTables d = s->dt;
  const int nP = d.nP;
  if (t.in.Z >= 2 && nP > 0) k_density_pp<T><<<nP, 256>>>(d, s->dn, psi, rpp.p, scale);
  if (t.in.N >= 2 && s->dn.nn_ncomp > 0) {
    d.tile_f0 = s->tile_f0[2].p;
    d.tile_nr = s->tile_nr[2].p;
    const size_t lds = (size_t)s->dn.nn_ncomp * sizeof(unsigned long long) + (size_t)4 * d.lmax * sizeof(T);
    if (lds <= kLdsLimit) {
      auto fn = k_density_nn<T, 4>;
      HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kLdsLimit));
      hipLaunchKernelGGL(fn, dim3(s->ntiles[2]), dim3(256), lds, nullptr, d, s->dn, psi, rnn.p, scale);
    } else if (nP > 0) {
      k_density_nn_g<T><<<nP, 256>>>(d, s->dn, psi, rnn.p, scale);
    }
  }
  if (t.in.Z >= 1 && t.in.N >= 1 && nP > 0) {
    const size_t lds = (size_t)5 * d.lmax * sizeof(T);
    if (lds <= kLdsLimit) {
      auto fn = k_density_pn<T, 4, false>;
      HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kLdsLimit));
      hipLaunchKernelGGL(fn, dim3(nP), dim3(256), lds, nullptr, d, s->dn, psi, rw.p, scale);
    } else {
      k_density_pn<T, 4, true><<<nP, 256>>>(d, s->dn, psi, rw.p, scale);
    }
  }
  HIP_OK(hipGetLastError());
  auto unfix = [&](const DevBuf<unsigned long long>& b, size_t n) {
    std::vector<unsigned long long> u(n);
    if (n) HIP_OK(hipMemcpy(u.data(), b.p, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
    std::vector<double> v(n);
    for (size_t i = 0; i < n; i++) v[i] = (double)(int64_t)u[i] / scale;
    return v;
  };
  if (norm2) *norm2 = n2;
  if (rho_pp && npx > 0) {
    const std::vector<double> v = unfix(rpp, (size_t)npx * npx);
    std::copy(v.begin(), v.end(), rho_pp);
  }
  if (rho_nn) {
    const std::vector<double> comp = unfix(rnn, s->nn_comp_to_me.size());
    std::fill(rho_nn, rho_nn + (size_t)npx * npx, 0.0);
    for (size_t i = 0; i < comp.size(); i++) rho_nn[s->nn_comp_to_me[i]] = comp[i];
  }
  if (rho_pn) {
    const std::vector<double> w = unfix(rw, t.wpn.size());
    const int ns = t.nsps, npn = ns * ns;
    std::fill(rho_pn, rho_pn + (size_t)npn * npn, 0.0);
    for (int a = 0; a < ns; a++)
      for (int c = 0; c < ns; c++) {
        const int k = t.class_of_op[a * ns + c], pl = t.local_of_op[a * ns + c];
        const int kn = negk_of(t, k), ws = t.class_size[kn];
        for (int b = 0; b < ns; b++)
          for (int dd = 0; dd < ns; dd++) {
            if (t.class_of_op[b * ns + dd] != kn) continue;
            const int nl = t.local_of_op[b * ns + dd];
            rho_pn[(size_t)(a * ns + b) * npn + (c * ns + dd)] = w[t.class_offset[k] + pl * ws + nl];
          }
      }
  }
  (void)nd;
}

}  // namespace

// ------------------------------------------------------------ batched mode
// K Hamiltonians on one basis (one interaction file each, same orbits,
// particle numbers and scaling).  The base system holds the structural jump
// tables (every structurally present two-body pair kept, whatever its value
// in any one interaction); per-Hamiltonian values are interleaved
// value[entry][k] so that each jump read serves all K vectors.
struct bs_batch {
  std::unique_ptr<bs_system> base;
  int K = 0;
  std::vector<std::string> intfiles;
  bool same_params = true;  // every interaction lists the same TBMEs as the first
  DevBuf<float> pp32, nn32, W, pspe, nspe, kt32;
  DevBuf<double> pp64, nn64, kt64;
  BatchVals bv{};
  // Lanczos workspace and its captured graphs: [0] the first Krylov cycle,
  // [1] a restarted cycle, [2] the thick restart, [3] Ritz vectors + polish
  struct Work {
    int dtype = 0, niter = 0, nkeep = 0, polish = 0, thick = 0;
    DevBuf<unsigned char> V, w, X64, HX, tmp;
    DevBuf<double> part, dots, alpha, beta, evals, Y, tdwork, rq, res, A, arrow, bnorm;
    DevBuf<int> flags;  // breakdown flags per Hamiltonian, then their maximum
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t stream = nullptr;
    ~Work() {
      for (auto& e : exec)
        if (e) (void)hipGraphExecDestroy(e);
      if (stream) (void)hipStreamDestroy(stream);
    }
  };
  std::unique_ptr<Work> work;
};

namespace {

constexpr int kMaxBatch = 64;

int batch_blocks(int64_t D) { return (int)std::min<int64_t>(std::max<int64_t>((D + 63) / 64, 1), 512); }

// Multi-vector kernels: variant 0 = one wave per basis state (lane = vector;
// small bases), 1 + 2*gi + map = k_apply_multi with G = 1 << gi staged rows
// and identity (map 0) or XCD-aware (map 1) workgroup order.
constexpr int kNumMulti = 7;

template <typename T, int CPT, int K, int G>
void launch_multi_kg(const bs_system* s, int map, const BatchVals& bv, const T* x, T* y, hipStream_t st) {
  DevTables d = s->dt;
  d.tile_f0 = s->tile_f0[0].p;
  d.tile_nr = s->tile_nr[0].p;
  d.tile_map = s->tile_map[0][map].p;
  // one Hamiltonian on all K vectors: coefficients staged once, not K times
  const bool sh = K > 1 && bv.vk == 0;
  const size_t lds = (size_t)G * K * d.lmax * sizeof(T) + (size_t)kWbRows * G * (sh ? 1 : K) * sizeof(float);
  if (lds > kLdsLimit) throw std::runtime_error("multi-vector variant does not fit in LDS");
  auto launch = [&](auto fn, bool& attr_set) {
    if (!attr_set) {
      HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kLdsLimit));
      attr_set = true;
    }
    hipLaunchKernelGGL(fn, dim3(s->nslots[0][map]), dim3(kApplyBlock), lds, st, d, bv, x, y);
  };
  static bool attr_sep = false, attr_sh = false;
  if constexpr (K > 1) {
    if (sh) launch(k_apply_multi<T, kApplyBlock, CPT, K, G, true>, attr_sh);
    else launch(k_apply_multi<T, kApplyBlock, CPT, K, G, false>, attr_sep);
  } else {
    launch(k_apply_multi<T, kApplyBlock, CPT, K, G, false>, attr_sep);
  }
  HIP_OK(hipGetLastError());
}

template <typename T, int CPT, int K>
void launch_multi_k(const bs_system* s, int v, const BatchVals& bv, const T* x, T* y, hipStream_t st) {
  const int gi = (v - 1) / 2, map = (v - 1) % 2;
  if (gi == 0) launch_multi_kg<T, CPT, K, 1>(s, map, bv, x, y, st);
  else launch_multi_kg<T, CPT, K, 2>(s, map, bv, x, y, st);
}

template <typename T, int CPT>
void launch_multi_cpt(const bs_system* s, int v, const BatchVals& bv, const T* x, T* y, hipStream_t st) {
  switch (bv.K) {
    case 1: launch_multi_k<T, CPT, 1>(s, v, bv, x, y, st); break;
    case 2: launch_multi_k<T, CPT, 2>(s, v, bv, x, y, st); break;
    case 4: launch_multi_k<T, CPT, 4>(s, v, bv, x, y, st); break;
    case 8: launch_multi_k<T, CPT, 8>(s, v, bv, x, y, st); break;
    default: throw std::runtime_error("multi-vector tile kernels take K = 1, 2, 4 or 8");
  }
}

bool multi_fits(const bs_system* s, int v, int K, size_t tsize, bool shared) {
  if (v == 0) return true;
  if (!s->tiles_ok) return false;
  // two-row tiles: FP32 with K = 2 or 4, FP64 with K = 2 (one Hamiltonian,
  // or one per vector with the batch's two-row union values)
  if (v >= 5) return (K == 2 || (K == 4 && tsize == 4)) && (shared || s->kt_ready);
  if (K != 1 && K != 2 && K != 4 && K != 8) return false;
  const int G = (v - 1) / 2 == 0 ? 1 : 2;
  return (size_t)G * K * s->dt.lmax * tsize + (size_t)kWbRows * G * K * sizeof(float) <= kLdsLimit;
}

// k_apply_ktiles on the two-row tiles (map 0: sector order, 1: XCD-grouped)
template <typename T, int K>
void launch_ktiles(const bs_system* s, int map, const BatchVals& bv, const T* x, T* y, hipStream_t st) {
  if (s->dt.lmax > kMaxCols) throw std::runtime_error("neutron rows too long for the two-row K-vector tiles");
  DevTables d = s->dt;
  d.tile_f0 = s->tile_f0[1].p;
  d.tile_nr = s->tile_nr[1].p;
  d.tile_map = s->tile_map[1][map].p;
  d.tpp_ptr = s->tpp_ptr[1].p;
  d.tpp_src = s->tpp_src[1].p;
  d.tpp_val32 = s->tpp_val32[1].p;
  d.tpp_val64 = s->tpp_val64[1].p;
  if (bv.vk == 0) {
    hipLaunchKernelGGL((k_apply_ktiles<T, K, true>), dim3(s->nslots[1][map]), dim3(1024), 0, st, d, bv, x, y);
  } else {
    if (!bv.kt32 || !bv.kt64) throw std::runtime_error("two-row union values missing for this batch");
    hipLaunchKernelGGL((k_apply_ktiles<T, K, false>), dim3(s->nslots[1][map]), dim3(1024), 0, st, d, bv, x, y);
  }
  HIP_OK(hipGetLastError());
}

template <typename T>
void launch_batched_v(const bs_system* s, int v, const BatchVals& bv, const T* x, T* y, hipStream_t st) {
  const DevTables& d = s->dt;
  if (bv.K < 1 || bv.K > kMaxBatch) throw std::runtime_error("batch size must be in 1..64");
  if (v == 0) {
    const int ny = (d.lmax + 3) / 4;
    hipLaunchKernelGGL(k_apply_batched<T>, dim3(d.nP, ny), dim3(256), 0, st, d, bv, x, y);
    HIP_OK(hipGetLastError());
    return;
  }
  if (v >= 5) {
    if (bv.K == 2) launch_ktiles<T, 2>(s, v - 5, bv, x, y, st);
    else if (bv.K == 4 && sizeof(T) == 4) launch_ktiles<T, sizeof(T) == 4 ? 4 : 2>(s, v - 5, bv, x, y, st);
    else throw std::runtime_error("the two-row K-vector tiles take K = 2 (FP32, FP64) or 4 (FP32)");
    return;
  }
  switch (s->cpt) {
    case 1: launch_multi_cpt<T, 1>(s, v, bv, x, y, st); break;
    case 2: launch_multi_cpt<T, 2>(s, v, bv, x, y, st); break;
    case 4: launch_multi_cpt<T, 4>(s, v, bv, x, y, st); break;
    default: launch_multi_cpt<T, 8>(s, v, bv, x, y, st); break;
  }
}

// The multi-vector variant for (dtype, K, shared values): timed on first use
// (two launches each on scratch vectors) and kept.  Never called under
// stream capture.
template <typename T>
int batch_variant(bs_system* s, const BatchVals& bv) {
  const int key = (sizeof(T) == 8 ? 1000 : 0) + (bv.vk ? 0 : 100) + bv.K;
  auto it = s->multi_choice.find(key);
  if (it != s->multi_choice.end()) return it->second;
  const int64_t n = s->t.dim * bv.K;
  DevBuf<T> x, y;
  x.alloc(n);
  y.alloc(n);
  k_fill<T><<<grid_for(n), kBlock>>>(x.p, (T)0.5, n);
  HIP_OK(hipGetLastError());
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  float best = 1e30f;
  int bestv = 0;
  for (int v = 0; v < kNumMulti; v++) {
    if (!multi_fits(s, v, bv.K, sizeof(T), bv.vk == 0)) continue;
    // one wave per basis state only pays on small bases: where the tiles fit
    // and the basis is large it is two orders of magnitude slower (8 s per
    // launch on pf 6p6n), so it is not timed there
    if (v == 0 && s->tiles_ok && s->t.dim > (int64_t)4000000) continue;
    launch_batched_v<T>(s, v, bv, x.p, y.p, nullptr);
    HIP_OK(hipEventRecord(e0, nullptr));
    launch_batched_v<T>(s, v, bv, x.p, y.p, nullptr);
    HIP_OK(hipEventRecord(e1, nullptr));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    s->multi_ms[key][v] = ms;
    if (ms < best) { best = ms; bestv = v; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s->multi_choice[key] = bestv;
  return bestv;
}

template <typename T>
void launch_batched(bs_system* s, const BatchVals& bv, const T* x, T* y, hipStream_t st) {
  launch_batched_v<T>(s, batch_variant<T>(s, bv), bv, x, y, st);
}

// the chosen variant, for launches recorded under stream capture
template <typename T>
void launch_batched_chosen(const bs_system* s, const BatchVals& bv, const T* x, T* y, hipStream_t st) {
  const int key = (sizeof(T) == 8 ? 1000 : 0) + (bv.vk ? 0 : 100) + bv.K;
  auto it = s->multi_choice.find(key);
  if (it == s->multi_choice.end()) throw std::runtime_error("multi-vector variant not chosen before capture");
  launch_batched_v<T>(s, it->second, bv, x, y, st);
}

void batch_upload_values(bs_batch* b) {
  const bs::Tables& t = b->base->t;
  const int K = b->K;
  const int64_t npp = (int64_t)t.tb[0].val.size();
  const int64_t nslot = (int64_t)std::max<size_t>(b->base->nn_val32.n, 1);
  const int64_t nw = (int64_t)t.wpn.size();
  const int64_t nP = (int64_t)t.sp[0].mask.size(), nN = (int64_t)t.sp[1].mask.size();
  std::vector<float> pp32((size_t)std::max<int64_t>(npp, 1) * K, 0.0f), nn32((size_t)nslot * K, 0.0f);
  std::vector<double> pp64(pp32.size(), 0.0), nn64(nn32.size(), 0.0);
  std::vector<float> W((size_t)std::max<int64_t>(nw, 1) * K, 0.0f);
  std::vector<float> ps((size_t)std::max<int64_t>(nP, 1) * K, 0.0f), ns((size_t)std::max<int64_t>(nN, 1) * K, 0.0f);
  // the two-row tiles' PP union values, [entry][row][K]
  const std::vector<int32_t>& csr2 = b->base->tpp_csr2;
  std::vector<float> kt32(std::max<size_t>(csr2.size(), 1) * K, 0.0f);
  std::vector<double> kt64(kt32.size(), 0.0);
  for (int k = 0; k < K; k++) {
    const bs::MeTables me = bs::build_me_tables(t, b->intfiles[k]);
    if (me.call_spe != t.call_spe) throw std::runtime_error("interactions disagree on the single-particle term");
    if (!me.same_lines) b->same_params = false;
    const std::vector<double> vp = bs::merged_values(t, 0, me.vpp);
    const std::vector<double> vn = bs::merged_values(t, 1, me.vnn);
    for (int64_t e = 0; e < npp; e++) {
      pp64[(size_t)e * K + k] = vp[e];
      pp32[(size_t)e * K + k] = (float)vp[e];
    }
    for (size_t i = 0; i < csr2.size(); i++)
      if (csr2[i] >= 0) {
        kt64[i * K + k] = vp[csr2[i]];
        kt32[i * K + k] = (float)vp[csr2[i]];
      }
    for (size_t q = 0; q < vn.size(); q++) {
      const int64_t o = b->base->nn_slot[q];
      nn64[(size_t)o * K + k] = vn[q];
      nn32[(size_t)o * K + k] = (float)vn[q];
    }
    if ((int64_t)me.wpn.size() != nw) throw std::runtime_error("pn element blocks differ between interactions");
    for (int64_t i = 0; i < nw; i++) W[(size_t)i * K + k] = me.wpn[i];
    for (int it = 0; it < 2; it++) {
      const bs::Species& S = t.sp[it];
      std::vector<float>& out = it == 0 ? ps : ns;
      const std::vector<float>& e = it == 0 ? me.pspe : me.nspe;
      for (size_t i = 0; i < S.mask.size(); i++) {
        float acc = 0.0f;
        for (uint64_t q = S.mask[i]; q; q &= q - 1) acc += e[t.sorb[__builtin_ctzll(q)]];
        out[i * K + k] = acc;
      }
    }
  }
  b->pp32.upload(pp32);
  b->pp64.upload(pp64);
  b->nn32.upload(nn32);
  b->nn64.upload(nn64);
  b->W.upload(W);
  b->pspe.upload(ps);
  b->nspe.upload(ns);
  b->kt32.upload(kt32);
  b->kt64.upload(kt64);
  b->base->kt_ready = !csr2.empty();
  BatchVals& v = b->bv;
  v.K = K;
  v.vs = K;
  v.vk = 1;
  v.pp32 = b->pp32.p;
  v.pp64 = b->pp64.p;
  v.nn32 = b->nn32.p;
  v.nn64 = b->nn64.p;
  v.W = b->W.p;
  v.pspe = b->pspe.p;
  v.nspe = b->nspe.p;
  v.kt32 = b->kt32.p;
  v.kt64 = b->kt64.p;
}

BatchVals shared_vals(const bs_system* s, int K) {
  BatchVals v{};
  v.K = K;
  v.vs = 1;
  v.vk = 0;
  v.pp32 = s->dt.pp_val32;
  v.pp64 = s->dt.pp_val64;
  v.nn32 = s->dt.nn_val32;
  v.nn64 = s->dt.nn_val64;
  v.W = s->dt.W;
  v.pspe = s->dt.pspe_sd;
  v.nspe = s->dt.nspe_sd;
  return v;
}

// A caller's pivot block for a block strength run (menu sb: the pivot block
// read from a .wfn, read_pivot_block_choose, src/blanczoslib.f90): K vectors
// [k][D] of dtype `dtype`; out: strength[i * nkeep + k] = |<Ritz k|pivot i>|^2
// and pnorm2[i] = |pivot i|^2.
struct BlockPivot {
  const void* piv = nullptr;
  int dtype = BS_F32;
  double* strength = nullptr;
  double* pnorm2 = nullptr;
};

// V[d * K + k] = piv[k * D + d]: K separate vectors into the interleaved block.
template <typename T, typename S, int K>
__global__ __launch_bounds__(kBlock) void k_load_block(T* __restrict__ V, const S* __restrict__ piv, int64_t D) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D * K; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = i / K;
    const int k = (int)(i - d * K);
    V[i] = (T)piv[(int64_t)k * D + d];
  }
}

// Block Lanczos (lanczos_block, src/blanczosmain.f90:1804-2153): blocks of K
// vectors, H applied to a block with the multi-vector kernel, full block
// classical Gram-Schmidt against every stored block (FP64 sums, one summed
// subtraction), each new block orthonormalized by the spectral decomposition
// of its Gram matrix (spectral_decomp in the reference), Ritz pairs from the
// block tridiagonal matrix, then the same FP64 normalization / Rayleigh
// quotient / residual as run_lanczos.  The pivot block is seeded
// pseudo-random (the reference draws it with random_number), or the caller's
// (bp): then the strength of each pivot in every Ritz state is returned too.
template <typename T, int K>
int run_block_lanczos(bs_system* s, const bs_lanczos_params* p, double* evals, void* evecs, double* resid,
                      int* iters, const BlockPivot* bp = nullptr) {
  const int64_t D = s->t.dim;
  const hipStream_t st = nullptr;
  const int nkeep = p->nkeep;
  if (nkeep < 1) return set_err("nkeep must be >= 1");
  if (D < K) return set_err("block larger than the basis");
  if (p->thick < 0) return set_err("thick must be >= 0");
  HIP_OK(hipDeviceSynchronize());  // synchronous: after all earlier device work, every stream
  // thick restart (option tb): kb whole blocks of Ritz vectors kept; a
  // Krylov space as large as the basis needs none
  const int kb = (p->thick > 0 && p->niter < D) ? p->thick / K : 0;
  // restart when (iter + 1) * K >= niter, iter = blocks in the projected
  // matrix (lanczos_block, src/blanczosmain.f90:2075-2086)
  const int iter_r = kb ? (p->niter + K - 1) / K - 1 : 0;
  int nblk = kb ? iter_r + 1 : (int)(std::min<int64_t>(p->niter, D) / K);
  if (kb) {
    if (kb * K < nkeep || kb + 1 >= iter_r)
      return set_err("thick restart needs nkeep <= (thick / block) * block and room for two more blocks before "
                     "the restart at niter");
    if (bp && bp->strength) return set_err("strengths are not defined after a thick restart");
    if (p->conv_test != 0 && p->conv_test != 1) return set_err("conv_test must be 0 or 1");
    if (p->ncheck_ex < 0) return set_err("ncheck_ex must be >= 0");
  } else if (nblk * K < nkeep) {
    return set_err("niter must give at least nkeep basis vectors (niter >= nkeep, multiple of block)");
  }
  const int64_t bsz = D * K;
  const int g = grid_for(D);
  const int nbx = std::min(g, 512);
  DevBuf<T> V, W;
  DevBuf<double> part, C, M;
  V.alloc((size_t)nblk * bsz);
  W.alloc(bsz);
  part.alloc((size_t)nbx * nblk * K * K);
  C.alloc((size_t)nblk * K * K);
  M.alloc(K * K);
  const BatchVals bv = shared_vals(s, K);
  auto gram = [&](const T* X, std::vector<double>& S) {
    k_block_dots<T, K><<<dim3(nbx, 1), kBlock, 0, st>>>(X, D, 1, X, part.p);
    k_reduce<<<K * K, kBlock, 0, st>>>(part.p, nbx, K * K, C.p, 0);
    HIP_OK(hipGetLastError());
    S.assign(K * K, 0.0);
    HIP_OK(hipMemcpy(S.data(), C.p, sizeof(double) * K * K, hipMemcpyDeviceToHost));
  };
  // X := X U diag(lambda)^(-1/2); returns B = diag(lambda)^(1/2) U^T (X_old = X_new B).
  // False when a direction of X is at the rounding level of H (relative to
  // floor): the stored span is then invariant to working precision.
  double hnorm = 0.0;
  const double eps = sizeof(T) == 4 ? 1.2e-7 : 2.3e-16;
  auto orthonormalize = [&](T* dst, const T* src, std::vector<double>& B, double floor) -> bool {
    std::vector<double> S, lam, U;
    gram(src, S);
    for (int a = 0; a < K; a++)
      for (int c = 0; c < a; c++) S[a * K + c] = S[c * K + a] = 0.5 * (S[a * K + c] + S[c * K + a]);
    jacobi_eigen(K, S, lam, U);
    if (!(lam[0] > 1e-28 * std::max(lam[K - 1], 1e-300))) return false;  // the block lost rank
    if (!(std::sqrt(lam[0]) > floor)) return false;
    std::vector<double> m(K * K);
    B.assign(K * K, 0.0);
    for (int a = 0; a < K; a++)
      for (int c = 0; c < K; c++) {
        m[a * K + c] = U[a * K + c] / std::sqrt(lam[c]);
        B[c * K + a] = std::sqrt(lam[c]) * U[a * K + c];
      }
    HIP_OK(hipMemcpy(M.p, m.data(), sizeof(double) * K * K, hipMemcpyHostToDevice));
    k_block_transform<T, K><<<g, kBlock, 0, st>>>(dst, src, D, M.p);
    HIP_OK(hipGetLastError());
    return true;
  };
  if (bp && bp->piv) {
    if (bp->dtype == BS_F32) k_load_block<T, float, K><<<g, kBlock, 0, st>>>(V.p, (const float*)bp->piv, D);
    else k_load_block<T, double, K><<<g, kBlock, 0, st>>>(V.p, (const double*)bp->piv, D);
    HIP_OK(hipGetLastError());
    std::vector<double> S0;
    gram(V.p, S0);
    if (bp->pnorm2)
      for (int k = 0; k < K; k++) bp->pnorm2[k] = S0[k * K + k];
  } else {
    k_pivot_random<T><<<g, kBlock, 0, st>>>(V.p, bsz, 0x5eed1234abcdULL);
    HIP_OK(hipGetLastError());
  }
  // pivot block = V_0 B0 (B0 from the spectral orthonormalization)
  std::vector<double> B0;
  if (!orthonormalize(V.p, V.p, B0, 0.0)) return set_err("pivot block is rank deficient");
  std::vector<std::vector<double>> A, B;  // diagonal and sub-diagonal blocks
  std::vector<double> e, Y, e_prev;
  int mv = 0, nb = 0;
  int ks = 0;               // kept blocks at the front after a thick restart
  std::vector<double> cpl;  // coupling of block ks to the kept vectors, [a][i], K x ks*K
  auto ritz = [&](int nbl) {
    const int n = nbl * K;
    std::vector<double> T_((size_t)n * n, 0.0);
    for (int j = 0; j < nbl; j++) {
      for (int a = 0; a < K; a++)
        for (int c = 0; c < K; c++) T_[(size_t)(j * K + a) * n + j * K + c] = A[j][a * K + c];
      if (j + 1 < nbl && j >= ks)
        for (int a = 0; a < K; a++)
          for (int c = 0; c < K; c++) {
            T_[(size_t)((j + 1) * K + a) * n + j * K + c] = B[j][a * K + c];
            T_[(size_t)(j * K + c) * n + (j + 1) * K + a] = B[j][a * K + c];
          }
    }
    if (ks > 0 && ks < nbl)
      for (int a = 0; a < K; a++)
        for (int i = 0; i < ks * K; i++) {
          T_[(size_t)(ks * K + a) * n + i] = cpl[(size_t)a * ks * K + i];
          T_[(size_t)i * n + ks * K + a] = cpl[(size_t)a * ks * K + i];
        }
    jacobi_eigen(n, T_, e, Y);
  };
  // thick restart: the reference's test (convergence_test with block_flag,
  // src/blanczoslib.f90:2726-2770) over ncheck = nkeep + ncheck_ex values at
  // every block iteration, the mean change carrying (1 + restarts)
  const int ncheck = kb ? (int)std::min<int64_t>(D, (int64_t)nkeep + p->ncheck_ex) : 0;
  std::vector<double> ecur(ncheck, 0.0), eold(ncheck, 0.0);
  int restarts = 0;
  auto energy_change = [&]() {
    double sum = 0.0, mx = 0.0;
    for (int i = 0; i < ncheck; i++) {
      const double d = std::fabs(eold[i] - ecur[i]);
      sum += d;
      mx = std::max(mx, d);
    }
    return p->conv_test == 1 ? mx : sum / std::sqrt((double)ncheck + 1.0) * (1.0 + restarts);
  };
  int j = 0;
  while (true) {
    T* Vj = V.p + (int64_t)j * bsz;
    launch_batched<T>(s, bv, Vj, W.p, st);
    mv += K;
    // C_i = V_i^T W for every stored block; A_j = C_j; W -= sum_i V_i C_i
    const int m = j + 1;
    k_block_dots<T, K><<<dim3(nbx, m), kBlock, 0, st>>>(V.p, D, m, W.p, part.p);
    k_reduce<<<m * K * K, kBlock, 0, st>>>(part.p, nbx, m * K * K, C.p, 0);
    HIP_OK(hipGetLastError());
    std::vector<double> Aj(K * K);
    HIP_OK(hipMemcpy(Aj.data(), C.p + (size_t)j * K * K, sizeof(double) * K * K, hipMemcpyDeviceToHost));
    for (int a = 0; a < K; a++)
      for (int c = 0; c < a; c++) Aj[a * K + c] = Aj[c * K + a] = 0.5 * (Aj[a * K + c] + Aj[c * K + a]);
    A.push_back(Aj);
    nb = j + 1;
    k_block_update<T, K><<<g, kBlock, 0, st>>>(W.p, V.p, D, m, C.p);
    // a second pass (classical Gram-Schmidt twice): once Ritz vectors converge,
    // H V_j lies almost in the stored span and one pass leaves a remainder
    // that is not orthogonal to it
    k_block_dots<T, K><<<dim3(nbx, m), kBlock, 0, st>>>(V.p, D, m, W.p, part.p);
    k_reduce<<<m * K * K, kBlock, 0, st>>>(part.p, nbx, m * K * K, C.p, 0);
    k_block_update<T, K><<<g, kBlock, 0, st>>>(W.p, V.p, D, m, C.p);
    HIP_OK(hipGetLastError());
    for (int a = 0; a < K; a++) hnorm = std::max(hnorm, std::fabs(Aj[a * K + a]));
    if (!kb) {
      if (p->tol > 0.0 && nb * K >= nkeep) {
        // (src/blanczoslib.f90:2660-2800) stop when no kept Ritz value moves by tol
        ritz(nb);
        if (!e_prev.empty()) {
          double dmax = 0.0;
          for (int k = 0; k < nkeep; k++) dmax = std::max(dmax, std::fabs(e[k] - e_prev[k]));
          if (dmax < p->tol) break;
        }
        e_prev.assign(e.begin(), e.begin() + nkeep);
      }
      if (j + 1 == nblk) break;
      std::vector<double> Bj;
      if (!orthonormalize(V.p + (int64_t)(j + 1) * bsz, W.p, Bj, 100.0 * eps * hnorm)) break;  // invariant span
      B.push_back(Bj);
      j++;
      continue;
    }
    // thick-restart block Lanczos: next block, energies once nb * K >= nkeep,
    // the test at every block iteration beyond ncheck (and the first), then
    // the restart when (nb + 1) * K >= niter
    std::vector<double> Bj;
    const bool next_ok = orthonormalize(V.p + (int64_t)nb * bsz, W.p, Bj, 100.0 * eps * hnorm);
    if (nb * K >= nkeep) {
      ritz(nb);
      for (int i = 0; i < ncheck; i++) ecur[i] = i < nb * K ? e[i] : 0.0;
    }
    bool finished = false;
    if (p->tol > 0.0) {
      if (nb * K > ncheck && nb > 1 && energy_change() < p->tol) finished = true;
      eold = ecur;
    }
    if (finished || !next_ok) break;
    B.push_back(Bj);
    if (nb < iter_r) {
      j = nb;
      continue;
    }
    if (restarts > p->max_restarts) break;
    // restart (block_thick_restart_sub_p, src/blanczoslib.f90:1929-2070):
    // kept Ritz vectors X = V Y[:, 0:kb*K] into blocks 0..kb-1, the new block
    // V_nb into block kb, coupled to them by B_{nb-1} Y[last block rows, 0:kb*K]
    ritz(nb);
    const int n = nb * K, nk = kb * K;
    {
      DevBuf<double> dY, X64;
      std::vector<double> Yk((size_t)n * nk);
      for (int c = 0; c < n; c++)
        for (int i = 0; i < nk; i++) Yk[(size_t)c * nk + i] = Y[(size_t)c * n + i];
      dY.upload(Yk);
      X64.alloc((size_t)nk * D);
      for (int k0 = 0; k0 < nk; k0 += 16)
        k_block_combine<T, K><<<g, kBlock, 0, st>>>(X64.p + (int64_t)k0 * D, V.p, D, n, dY.p + k0, nk,
                                                    std::min(16, nk - k0));
      for (int b = 0; b < kb; b++)
        k_load_block<T, double, K><<<g, kBlock, 0, st>>>(V.p + (int64_t)b * bsz, X64.p + (int64_t)b * K * D, D);
      HIP_OK(hipGetLastError());
      HIP_OK(hipMemcpyAsync(V.p + (int64_t)kb * bsz, V.p + (int64_t)nb * bsz, sizeof(T) * bsz,
                            hipMemcpyDeviceToDevice, st));
      HIP_OK(hipStreamSynchronize(st));
    }
    const std::vector<double>& Bl = B[nb - 1];
    cpl.assign((size_t)K * nk, 0.0);
    for (int a = 0; a < K; a++)
      for (int i = 0; i < nk; i++) {
        double v = 0.0;
        for (int c = 0; c < K; c++) v += Bl[a * K + c] * Y[(size_t)((nb - 1) * K + c) * n + i];
        cpl[(size_t)a * nk + i] = v;
      }
    A.resize(kb);
    B.assign(kb, std::vector<double>(K * K, 0.0));
    for (int b = 0; b < kb; b++) {
      A[b].assign(K * K, 0.0);
      for (int a = 0; a < K; a++) A[b][a * K + a] = e[b * K + a];
    }
    ks = kb;
    nb = kb;
    restarts++;
    j = kb;
  }
  if (nb * K < nkeep) return set_err("Krylov space smaller than nkeep");
  ritz(nb);
  const int n = nb * K;
  for (int k = 0; k < nkeep; k++) evals[k] = e[k];
  if (bp && bp->strength)
    for (int i = 0; i < K; i++)
      for (int k = 0; k < nkeep; k++) {
        double o = 0.0;  // <Ritz k | pivot i> = sum_c Y(c, k) B0(c, i)
        for (int c = 0; c < K; c++) o += Y[(size_t)c * n + k] * B0[c * K + i];
        bp->strength[(size_t)i * nkeep + k] = o * o;
      }
  if (evecs || resid || p->polish) {
    const int nb1 = grid_for(D) < 1024 ? grid_for(D) : 1024;
    DevBuf<double> dY, X64, HX, part1, dots;
    std::vector<double> Yk((size_t)n * nkeep);
    for (int c = 0; c < n; c++)
      for (int k = 0; k < nkeep; k++) Yk[(size_t)c * nkeep + k] = Y[(size_t)c * n + k];
    dY.upload(Yk);
    X64.alloc((size_t)nkeep * D);
    HX.alloc(D);
    part1.alloc(nb1);
    dots.alloc(4);
    for (int k0 = 0; k0 < nkeep; k0 += 16)
      k_block_combine<T, K><<<g, kBlock, 0, st>>>(X64.p + (int64_t)k0 * D, V.p, D, n, dY.p + k0, nkeep,
                                                  std::min(16, nkeep - k0));
    HIP_OK(hipGetLastError());
    for (int k = 0; k < nkeep; k++) {
      double* xk = X64.p + (int64_t)k * D;
      k_multidot<double><<<nb1, kBlock, 0, st>>>(xk, D, 1, xk, D, part1.p);
      k_reduce<<<1, kBlock, 0, st>>>(part1.p, nb1, 1, dots.p, 1);
      k_normalize<double><<<g, kBlock, 0, st>>>(xk, xk, dots.p, D);
      if (p->polish || resid) {
        launch_apply<double>(s, xk, HX.p, 0, st);
        mv++;
        k_multidot<double><<<nb1, kBlock, 0, st>>>(xk, D, 1, HX.p, D, part1.p);
        k_reduce<<<1, kBlock, 0, st>>>(part1.p, nb1, 1, dots.p + 1, 0);
        double rq = 0.0;
        HIP_OK(hipMemcpy(&rq, dots.p + 1, sizeof(double), hipMemcpyDeviceToHost));
        if (p->polish) evals[k] = rq;
        if (resid) {
          k_sub_scaled<double><<<g, kBlock, 0, st>>>(HX.p, xk, dots.p + 1, D);
          k_multidot<double><<<nb1, kBlock, 0, st>>>(HX.p, D, 1, HX.p, D, part1.p);
          k_reduce<<<1, kBlock, 0, st>>>(part1.p, nb1, 1, dots.p + 2, 1);
          HIP_OK(hipMemcpy(&resid[k], dots.p + 2, sizeof(double), hipMemcpyDeviceToHost));
        }
      }
      if (evecs) {
        if (p->dtype == BS_F64)
          HIP_OK(hipMemcpyAsync((double*)evecs + (int64_t)k * D, xk, sizeof(double) * D, hipMemcpyDeviceToDevice, st));
        else
          k_convert<double, float><<<g, kBlock, 0, st>>>((float*)evecs + (int64_t)k * D, xk, D);
      }
    }
  }
  HIP_OK(hipStreamSynchronize(st));
  if (iters) *iters = mv;
  s->last_random_restarts = 0;
  s->last_second_passes = 0;
  s->last_steps = nb * K;
  s->last_thick_restarts = restarts;
  return 0;
}

template <typename T>
int run_block_lanczos_k(bs_system* s, const bs_lanczos_params* p, int K, double* evals, void* evecs,
                        double* resid, int* iters, const BlockPivot* bp = nullptr) {
  switch (K) {
    case 1: return run_block_lanczos<T, 1>(s, p, evals, evecs, resid, iters, bp);
    case 2: return run_block_lanczos<T, 2>(s, p, evals, evecs, resid, iters, bp);
    case 3: return run_block_lanczos<T, 3>(s, p, evals, evecs, resid, iters, bp);
    case 4: return run_block_lanczos<T, 4>(s, p, evals, evecs, resid, iters, bp);
    case 5: return run_block_lanczos<T, 5>(s, p, evals, evecs, resid, iters, bp);
    case 6: return run_block_lanczos<T, 6>(s, p, evals, evecs, resid, iters, bp);
    case 7: return run_block_lanczos<T, 7>(s, p, evals, evecs, resid, iters, bp);
    case 8: return run_block_lanczos<T, 8>(s, p, evals, evecs, resid, iters, bp);
    default: return set_err("block size must be in 1..8");
  }
}


// Lanczos steps i0..i1-1 for all K Hamiltonians in lockstep (FP64
// alpha/beta and Gram-Schmidt dots per Hamiltonian), enqueued on st with no
// host synchronisation and no memset, so that it can be captured.
template <typename T>
void enqueue_batch_steps(bs_batch* b, bs_batch::Work& w, int i0, int i1, hipStream_t st) {
  const bs_system* s = b->base.get();
  const int K = b->K;
  const int64_t D = s->t.dim, DK = D * K;
  const int nb = batch_blocks(D);
  const int g = grid_for(DK);
  T* V = (T*)w.V.p;
  T* wv = (T*)w.w.p;
  if (i0 == 0) k_fill<T><<<g, kBlock, 0, st>>>(V, (T)(1.0 / std::sqrt((double)D)), DK);
  for (int i = i0; i < i1; i++) {
    T* vi = V + (int64_t)i * DK;
    launch_batched_chosen<T>(s, b->bv, vi, wv, st);
    k_bmultidot<T><<<nb, 256, 0, st>>>(vi, DK, 1, wv, D, K, w.part.p);
    k_breduce<<<(K + 255) / 256, 256, 0, st>>>(w.part.p, nb, 1, K, w.alpha.p + (int64_t)i * K, 0);
    k_bsub_scaled<T><<<g, kBlock, 0, st>>>(wv, vi, w.alpha.p + (int64_t)i * K, D, K);
    if (i > 0) {
      k_bmultidot<T><<<nb, 256, 0, st>>>(V, DK, i, wv, D, K, w.part.p);
      k_breduce<<<(i * K + 255) / 256, 256, 0, st>>>(w.part.p, nb, i, K, w.dots.p, 0);
      k_bmultiaxpy<T><<<g, kBlock, 0, st>>>(wv, V, DK, i, w.dots.p, D, K);
    }
    k_bmultidot<T><<<nb, 256, 0, st>>>(wv, DK, 1, wv, D, K, w.part.p);
    k_breduce<<<(K + 255) / 256, 256, 0, st>>>(w.part.p, nb, 1, K, w.beta.p + (int64_t)i * K, 1);
    // v_{i+1} with the breakdown handling of k_bbreak: random restarts
    // (flag 2) are orthonormalized against v_0..v_i twice, members whose
    // beta fell below 0.1 of the previous one (flag 1) once if they lost
    // orthogonality; the gated kernels return at once when no member broke
    T* vn = V + (int64_t)(i + 1) * DK;
    const int m = i + 1;
    k_bbreak<<<1, 64, 0, st>>>(w.alpha.p, w.beta.p, w.arrow.p, i, i0, K, w.flags.p);
    k_bnext<T><<<g, kBlock, 0, st>>>(vn, wv, w.beta.p + (int64_t)i * K, D, K, w.flags.p,
                                     0x7a3f0e1d5c2bULL + 0x9E3779B9ULL * (uint64_t)(i + 1), i + 1 >= D ? 1 : 0);
    if (i + 1 < D) {
      for (int pass = 0; pass < 3; pass++) {
        const int minflag = pass == 2 ? 2 : 1;
        if (pass == 0) {
          // normalize the random vectors first
          k_bmultidot_g<T><<<nb, 256, 0, st>>>(vn, DK, 1, vn, D, K, w.part.p, w.flags.p, 2);
          k_breduce_g<<<(K + 255) / 256, 256, 0, st>>>(w.part.p, nb, 1, K, w.bnorm.p, 1, w.flags.p, 2);
          k_bnormalize_m<T><<<g, kBlock, 0, st>>>(vn, w.bnorm.p, D, K, w.flags.p, 2);
          continue;
        }
        k_bmultidot_g<T><<<nb, 256, 0, st>>>(V, DK, m, vn, D, K, w.part.p, w.flags.p, minflag);
        k_breduce_g<<<(m * K + 255) / 256, 256, 0, st>>>(w.part.p, nb, m, K, w.dots.p, 0, w.flags.p, minflag);
        if (pass == 1) k_bloss<<<1, 64, 0, st>>>(w.dots.p, m, K, w.flags.p);
        k_bmultiaxpy_m<T><<<g, kBlock, 0, st>>>(vn, V, DK, m, w.dots.p, D, K, w.flags.p, minflag);
        k_bmultidot_g<T><<<nb, 256, 0, st>>>(vn, DK, 1, vn, D, K, w.part.p, w.flags.p, minflag);
        k_breduce_g<<<(K + 255) / 256, 256, 0, st>>>(w.part.p, nb, 1, K, w.bnorm.p, 1, w.flags.p, minflag);
        k_bnormalize_m<T><<<g, kBlock, 0, st>>>(vn, w.bnorm.p, D, K, w.flags.p, minflag);
      }
    }
  }
  HIP_OK(hipGetLastError());
}

// Projected matrix of a (restarted) Krylov cycle for each Hamiltonian: the
// locked Ritz values on the diagonal with their couplings in column kstart
// (src/blanczoslib.f90:143-222), then the tridiagonal part.  A[k][m][m].
__global__ void k_projected(const double* __restrict__ alpha, const double* __restrict__ beta,
                            const double* __restrict__ arrow, int kstart, int m, int K, double* __restrict__ A) {
  const int k = blockIdx.y;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < m * m; idx += gridDim.x * blockDim.x) {
    const int i = idx / m, j = idx - i * m;
    double v = 0.0;
    if (i == j) v = alpha[(int64_t)i * K + k];
    else if (i >= kstart && j >= kstart && (j == i + 1 || i == j + 1)) v = beta[(int64_t)min(i, j) * K + k];
    else if (j == kstart && i < kstart) v = arrow[(int64_t)i * K + k];
    else if (i == kstart && j < kstart) v = arrow[(int64_t)j * K + k];
    A[((int64_t)k * m + i) * m + j] = v;
  }
}

// Lowest nout eigenpairs of each A[k] (cyclic Jacobi, FP64, one thread per
// matrix; work holds m*m per matrix).  evals[k*nout + n], Y[(k*m + i)*nout + n].
__global__ void k_jacobi_ritz(double* __restrict__ A, int m, int K, int nout, double* __restrict__ evals,
                              double* __restrict__ Y, double* __restrict__ work) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double* a = A + (int64_t)k * m * m;
  double* v = work + (int64_t)k * m * m;
  for (int i = 0; i < m * m; i++) v[i] = 0.0;
  for (int i = 0; i < m; i++) v[i * m + i] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) {
        const double t = a[i * m + j] * a[i * m + j];
        tot += t;
        if (i != j) off += t;
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < m - 1; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = a[p * m + q];
        if (fabs(apq) < 1e-300) continue;
        const double theta = (a[q * m + q] - a[p * m + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int r = 0; r < m; r++) {
          const double arp = a[r * m + p], arq = a[r * m + q];
          a[r * m + p] = c * arp - sn * arq;
          a[r * m + q] = sn * arp + c * arq;
        }
        for (int r = 0; r < m; r++) {
          const double apr = a[p * m + r], aqr = a[q * m + r];
          a[p * m + r] = c * apr - sn * aqr;
          a[q * m + r] = sn * apr + c * aqr;
        }
        for (int r = 0; r < m; r++) {
          const double vrp = v[r * m + p], vrq = v[r * m + q];
          v[r * m + p] = c * vrp - sn * vrq;
          v[r * m + q] = sn * vrp + c * vrq;
        }
      }
  }
  // the nout smallest diagonal entries, ascending (selection)
  unsigned long long used[4] = {0, 0, 0, 0};  // m <= 256
  for (int n = 0; n < nout; n++) {
    int best = -1;
    for (int i = 0; i < m; i++) {
      if ((used[i >> 6] >> (i & 63)) & 1ull) continue;
      if (best < 0 || a[i * m + i] < a[best * m + best]) best = i;
    }
    used[best >> 6] |= 1ull << (best & 63);
    evals[(int64_t)k * nout + n] = a[best * m + best];
    for (int i = 0; i < m; i++) Y[((int64_t)k * m + i) * nout + n] = v[i * m + best];
  }
}

// After a thick restart: locked Ritz values become alpha[0..thick), their
// couplings to the carried-over Krylov vector arrow[j] = beta_{m-1} Y[m-1, j].
__global__ void k_restart_coeffs(double* __restrict__ alpha, const double* __restrict__ beta,
                                 double* __restrict__ arrow, const double* __restrict__ evals,
                                 const double* __restrict__ Y, int m, int K, int nout, int thick) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= thick * K) return;
  const int j = idx / K, k = idx - j * K;
  alpha[(int64_t)j * K + k] = evals[(int64_t)k * nout + j];
  arrow[(int64_t)j * K + k] = beta[(int64_t)(m - 1) * K + k] * Y[((int64_t)k * m + m - 1) * nout + j];
}

template <typename T>
__global__ void k_copy(T* __restrict__ dst, const T* __restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Ritz pairs of the current projected matrices (Sturm bisection / inverse
// iteration for a plain tridiagonal cycle, Jacobi for an arrowhead one).
void enqueue_batch_ritz(bs_batch::Work& w, int K, int kstart, int nout, hipStream_t st) {
  const int m = w.niter;
  if (kstart == 0 && w.thick == 0) {
    k_tridiag_ritz<<<(K + 63) / 64, 64, 0, st>>>(w.alpha.p, w.beta.p, m, K, nout, w.evals.p, w.Y.p, w.tdwork.p);
  } else {
    k_projected<<<dim3((m * m + 255) / 256, K), 256, 0, st>>>(w.alpha.p, w.beta.p, w.arrow.p, kstart, m, K, w.A.p);
    k_jacobi_ritz<<<(K + 63) / 64, 64, 0, st>>>(w.A.p, m, K, nout, w.evals.p, w.Y.p, w.tdwork.p);
  }
  HIP_OK(hipGetLastError());
}

// thick restart: V[0:thick] = V[0:m] Y[:, 0:thick], V[thick] = V[m]
template <typename T>
void enqueue_batch_restart(bs_batch* b, bs_batch::Work& w, int nout, hipStream_t st) {
  const int K = b->K, m = w.niter, thick = w.thick;
  const int64_t DK = b->base->t.dim * K;
  const int g = grid_for(DK);
  T* V = (T*)w.V.p;
  T* tmp = (T*)w.tmp.p;
  k_bcombine<T, T><<<g, kBlock, 0, st>>>(tmp, V, DK, m, w.Y.p, nout, thick, b->base->t.dim, K);
  k_copy<T><<<grid_for((int64_t)thick * DK), kBlock, 0, st>>>(V, tmp, (int64_t)thick * DK);
  k_copy<T><<<g, kBlock, 0, st>>>(V + (int64_t)thick * DK, V + (int64_t)m * DK, DK);
  k_restart_coeffs<<<(thick * K + 255) / 256, 256, 0, st>>>(w.alpha.p, w.beta.p, w.arrow.p, w.evals.p, w.Y.p, m, K,
                                                            nout, thick);
  HIP_OK(hipGetLastError());
}

// Ritz vectors in FP64 from the last Ritz solve, normalised, then the FP64
// Rayleigh quotient and residual (polish)
template <typename T>
void enqueue_batch_finish(bs_batch* b, bs_batch::Work& w, int nout, hipStream_t st) {
  const bs_system* s = b->base.get();
  const int K = b->K, nkeep = w.nkeep;
  const int64_t D = s->t.dim, DK = D * K;
  const int nb = batch_blocks(D);
  const int g = grid_for(DK);
  double* X = (double*)w.X64.p;
  double* HX = (double*)w.HX.p;
  k_bcombine<T, double><<<g, kBlock, 0, st>>>(X, (const T*)w.V.p, DK, w.niter, w.Y.p, nout, nkeep, D, K);
  HIP_OK(hipGetLastError());
  for (int q = 0; q < nkeep; q++) {
    double* xq = X + (int64_t)q * DK;
    k_bmultidot<double><<<nb, 256, 0, st>>>(xq, DK, 1, xq, D, K, w.part.p);
    k_breduce<<<(K + 255) / 256, 256, 0, st>>>(w.part.p, nb, 1, K, w.dots.p, 1);
    k_bnormalize<double><<<g, kBlock, 0, st>>>(xq, xq, w.dots.p, D, K);
    if (w.polish) {
      launch_batched_chosen<double>(s, b->bv, xq, HX, st);
      k_bmultidot<double><<<nb, 256, 0, st>>>(xq, DK, 1, HX, D, K, w.part.p);
      k_breduce<<<(K + 255) / 256, 256, 0, st>>>(w.part.p, nb, 1, K, w.rq.p + (int64_t)q * K, 0);
      k_bsub_scaled<double><<<g, kBlock, 0, st>>>(HX, xq, w.rq.p + (int64_t)q * K, D, K);
      k_bmultidot<double><<<nb, 256, 0, st>>>(HX, DK, 1, HX, D, K, w.part.p);
      k_breduce<<<(K + 255) / 256, 256, 0, st>>>(w.part.p, nb, 1, K, w.res.p + (int64_t)q * K, 1);
    }
  }
  HIP_OK(hipGetLastError());
}

template <typename F>
hipGraphExec_t capture(hipStream_t st, F&& body) {
  hipGraph_t graph = nullptr;
  HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(st, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HIP_OK(hipStreamEndCapture(st, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_OK(ie);
  return exec;
}

template <typename T>
void run_batch_lanczos(bs_batch* b, const bs_lanczos_params* p, double* alpha, double* beta, double* evals,
                       void* evecs, double* resid, int* iters) {
  const int K = b->K;
  const int64_t D = b->base->t.dim, DK = D * K;
  const int niter = (int)std::min<int64_t>(p->niter, D);
  const int nkeep = p->nkeep;
  // a Krylov space as large as the basis is exact after one cycle: no restart
  const int thick = p->niter >= D ? 0 : p->thick;
  if (nkeep < 1 || nkeep > 16) throw std::runtime_error("nkeep must be in 1..16");
  if (niter < nkeep) throw std::runtime_error("niter must be >= nkeep");
  if (p->target) throw std::runtime_error("the batched run keeps the lowest states (target must be 0)");
  if (thick < 0 || (thick > 0 && (thick < nkeep || thick >= niter || thick > 16)))
    throw std::runtime_error("thick must be 0 or in [nkeep, min(niter, 17))");
  if (thick > 0 && niter > 256) throw std::runtime_error("thick restart takes niter <= 256");
  const int nout = std::max(nkeep, thick);
  const int dt = sizeof(T) == 4 ? BS_F32 : BS_F64;
  std::unique_ptr<bs_batch::Work>& wp = b->work;
  if (!wp || wp->dtype != dt || wp->niter != niter || wp->nkeep != nkeep || wp->polish != (p->polish ? 1 : 0) ||
      wp->thick != thick) {
    wp.reset(new bs_batch::Work());
    bs_batch::Work& w = *wp;
    w.dtype = dt;
    w.niter = niter;
    w.nkeep = nkeep;
    w.thick = thick;
    w.polish = p->polish ? 1 : 0;
    const int nb = batch_blocks(D);
    w.V.alloc((size_t)(niter + 1) * DK * sizeof(T));
    w.w.alloc((size_t)DK * sizeof(T));
    w.X64.alloc((size_t)nkeep * DK * sizeof(double));
    w.HX.alloc((size_t)DK * sizeof(double));
    if (thick) w.tmp.alloc((size_t)thick * DK * sizeof(T));
    w.part.alloc((size_t)nb * niter * K);
    w.dots.alloc((size_t)niter * K);
    w.alpha.alloc((size_t)niter * K);
    w.beta.alloc((size_t)niter * K);
    w.evals.alloc((size_t)K * nout);
    w.Y.alloc((size_t)K * niter * nout);
    w.tdwork.alloc((size_t)K * std::max(6 * niter, thick ? niter * niter : 0));
    w.rq.alloc((size_t)nkeep * K);
    w.res.alloc((size_t)nkeep * K);
    if (thick) w.A.alloc((size_t)K * niter * niter);
    w.arrow.alloc((size_t)std::max(thick, 1) * K);
    w.bnorm.alloc((size_t)K);
    w.flags.alloc((size_t)K + 1);
    HIP_OK(hipStreamCreate(&w.stream));
    batch_variant<T>(b->base.get(), b->bv);
    if (w.polish) batch_variant<double>(b->base.get(), b->bv);
    const hipStream_t st = w.stream;
    w.exec[0] = capture(st, [&] {
      enqueue_batch_steps<T>(b, w, 0, niter, st);
      enqueue_batch_ritz(w, K, 0, nout, st);
    });
    if (thick) {
      w.exec[1] = capture(st, [&] {
        enqueue_batch_steps<T>(b, w, thick, niter, st);
        enqueue_batch_ritz(w, K, thick, nout, st);
      });
      w.exec[2] = capture(st, [&] { enqueue_batch_restart<T>(b, w, nout, st); });
    }
    w.exec[3] = capture(st, [&] { enqueue_batch_finish<T>(b, w, nout, st); });
  }
  bs_batch::Work& w = *wp;
  const hipStream_t st = w.stream;
  int matvecs = niter;
  HIP_OK(hipGraphLaunch(w.exec[0], st));
  std::vector<double> ev((size_t)K * nout), prev;
  HIP_OK(hipStreamSynchronize(st));
  HIP_OK(hipMemcpy(ev.data(), w.evals.p, sizeof(double) * ev.size(), hipMemcpyDeviceToHost));
  for (int restarts = 0; thick > 0 && restarts < p->max_restarts; restarts++) {
    if (p->tol > 0.0 && !prev.empty()) {
      double dmax = 0.0;
      for (int k = 0; k < K; k++)
        for (int q = 0; q < nkeep; q++)
          dmax = std::max(dmax, std::fabs(ev[(size_t)k * nout + q] - prev[(size_t)k * nout + q]));
      if (dmax < p->tol) break;
    }
    prev = ev;
    HIP_OK(hipGraphLaunch(w.exec[2], st));
    HIP_OK(hipGraphLaunch(w.exec[1], st));
    matvecs += niter - thick;
    HIP_OK(hipStreamSynchronize(st));
    HIP_OK(hipMemcpy(ev.data(), w.evals.p, sizeof(double) * ev.size(), hipMemcpyDeviceToHost));
  }
  HIP_OK(hipGraphLaunch(w.exec[3], st));
  HIP_OK(hipStreamSynchronize(st));
  std::vector<double> rq((size_t)nkeep * K), rs((size_t)nkeep * K);
  if (w.polish) {
    HIP_OK(hipMemcpy(rq.data(), w.rq.p, sizeof(double) * rq.size(), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(rs.data(), w.res.p, sizeof(double) * rs.size(), hipMemcpyDeviceToHost));
  }
  if (alpha || beta) {
    std::vector<double> a((size_t)niter * K), bb((size_t)niter * K);
    HIP_OK(hipMemcpy(a.data(), w.alpha.p, sizeof(double) * a.size(), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(bb.data(), w.beta.p, sizeof(double) * bb.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < K; k++)
      for (int i = 0; i < niter; i++) {
        if (alpha) alpha[(size_t)k * niter + i] = a[(size_t)i * K + k];
        if (beta) beta[(size_t)k * niter + i] = bb[(size_t)i * K + k];
      }
  }
  for (int k = 0; k < K; k++)
    for (int q = 0; q < nkeep; q++) {
      evals[(size_t)k * nkeep + q] = w.polish ? rq[(size_t)q * K + k] : ev[(size_t)k * nout + q];
      if (resid) resid[(size_t)k * nkeep + q] = w.polish ? rs[(size_t)q * K + k] : -1.0;
    }
  if (evecs) {
    if (p->dtype == BS_F64)
      HIP_OK(hipMemcpy(evecs, w.X64.p, sizeof(double) * (size_t)nkeep * DK, hipMemcpyDeviceToDevice));
    else {
      const int64_t n = (int64_t)nkeep * DK;
      k_convert<double, float><<<grid_for(n), kBlock, 0, st>>>((float*)evecs, (const double*)w.X64.p, n);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(st));
    }
  }
  if (iters) *iters = matvecs + (w.polish ? nkeep : 0);
}

// lane k of an interleaved set -> contiguous FP64 vector
template <typename T>
__global__ void k_gather_lane(double* __restrict__ out, const T* __restrict__ in, int64_t D, int K, int k) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (double)in[i * K + k];
}

}  // namespace

extern "C" {

const char* bs_last_error(void) { return g_err.c_str(); }

int bs_init_host(const char* sps, const char* intf, int Z, int N, int jz2, float spscale, float a0, float a,
                 float x, bs_system** out) {
  try {
    if (!out) return set_err("null output pointer");
    std::unique_ptr<bs_system> s(new bs_system());
    if (!sps || !intf) return set_err("null file name");
    bs::ModelInput in;
    bs::read_sps(sps, in);
    s->t = bs::build_tables(in.orbits, intf, Z, N, jz2, spscale, a0, a, x, true);
    *out = s.release();
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_upload(bs_system* s) {
  try {
    checked(s);
    if (!s->uploaded) upload_tables(s);
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_init(const char* sps, const char* intf, int Z, int N, int jz2, float spscale, float a0, float a,
            float x, bs_system** out) {
  int r = bs_init_host(sps, intf, Z, N, jz2, spscale, a0, a, x, out);
  if (r) return r;
  r = bs_upload(*out);
  if (r) {
    delete *out;
    *out = nullptr;
  }
  return r;
}

int bs_finalize(bs_system* s) {
  delete s;
  return 0;
}

int bs_dim(const bs_system* s, int64_t* dim) {
  if (!s || !dim) return set_err("null argument");
  *dim = s->t.dim;
  return 0;
}

int bs_num_sd(const bs_system* s, int it, int64_t* nsd) {
  if (!s || !nsd || it < 0 || it > 1) return set_err("bad argument");
  *nsd = (int64_t)s->t.sp[it].mask.size();
  return 0;
}

int bs_basis(const bs_system* s, int it, uint64_t* mask, int64_t* start) {
  if (!s || it < 0 || it > 1) return set_err("bad argument");
  const bs::Species& S = s->t.sp[it];
  if (mask) std::memcpy(mask, S.mask.data(), sizeof(uint64_t) * S.mask.size());
  if (start) std::memcpy(start, S.start.data(), sizeof(int64_t) * S.start.size());
  return 0;
}

int bs_parity(const bs_system* s, int* parity) {
  if (!s || !parity) return set_err("null argument");
  // one-parity spaces only: every SD has parity (-1)^(l * (Z + N))
  const int l = s->t.in.orbits.empty() ? 0 : s->t.in.orbits[0].l;
  *parity = ((l * (s->t.in.Z + s->t.in.N)) % 2) ? -1 : 1;
  return 0;
}

int bs_num_sps(const bs_system* s, int* nsps) {
  if (!s || !nsps) return set_err("null argument");
  *nsps = s->t.nsps;
  return 0;
}

int bs_sps(const bs_system* s, int* orb, int* j2, int* m2) {
  if (!s) return set_err("null argument");
  for (int i = 0; i < s->t.nsps; i++) {
    if (orb) orb[i] = s->t.sorb[i];
    if (j2) j2[i] = s->t.sj2[i];
    if (m2) m2[i] = s->t.sm2[i];
  }
  return 0;
}

int bs_me_table(const bs_system* s, int kind, int* npair, float* table) {
  if (!s || !npair || kind < 0 || kind > 2) return set_err("bad argument");
  const std::vector<float>& v = kind == 0 ? s->t.vpp : kind == 1 ? s->t.vnn : s->t.vpn;
  *npair = kind == 2 ? s->t.npairpn : s->t.npairxx;
  if (table) std::memcpy(table, v.data(), sizeof(float) * v.size());
  return 0;
}

int bs_op_counts(const bs_system* s, int64_t out[8]) {
  if (!s || !out) return set_err("null argument");
  out[0] = s->t.nops_spe;
  out[1] = s->t.nops_pp;
  out[2] = s->t.nops_nn;
  out[3] = s->t.nops_pn;
  out[4] = s->t.call_spe ? s->t.dim : 0;
  out[5] = s->t.ref_pp;
  out[6] = s->t.ref_nn;
  out[7] = s->t.nops_pn;
  return 0;
}

int bs_jumps(const bs_system* s, int it, int order, int64_t* count, int32_t* isd, int32_t* fsd, int32_t* op,
             int8_t* phase) {
  try {
    if (!s || !count) return set_err("null argument");
    if (it != 0 && it != 1) return set_err("species must be 0 (protons) or 1 (neutrons)");
    if (order != 1 && order != 2) return set_err("order must be 1 or 2");
    const bs::Tables& t = s->t;
    const bool want = isd || fsd || op || phase;
    if (order == 1) {
      const bs::OneBody& O = t.ob[it];
      const int64_t nsd = O.rowptr.empty() ? 0 : (int64_t)O.rowptr.size() - 1;
      const int64_t n = nsd ? O.rowptr[nsd] : 0;
      if (want && *count < n) return set_err("count is smaller than the number of jumps");
      *count = n;
      if (!want) return 0;
      for (int64_t f = 0; f < nsd; f++)
        for (int64_t e = O.rowptr[f]; e < O.rowptr[f + 1]; e++) {
          if (isd) isd[e] = O.src[e];
          if (fsd) fsd[e] = (int32_t)f;
          if (op) op[e] = O.op[e];
          if (phase) phase[e] = O.phase[e];
        }
      return 0;
    }
    bs::TwoBodyOps T;
    bs::build_twobody_ops(t, it, T);
    const int64_t nsd = (int64_t)T.rowptr.size() - 1;
    const int64_t n = nsd > 0 ? T.rowptr[nsd] : 0;
    if (want && *count < n) return set_err("count is smaller than the number of jumps");
    *count = n;
    if (!want) return 0;
    for (int64_t f = 0; f < nsd; f++)
      for (int64_t e = T.rowptr[f]; e < T.rowptr[f + 1]; e++) {
        if (isd) isd[e] = T.src[e];
        if (fsd) fsd[e] = (int32_t)f;
        if (op) op[e] = T.me[e];
        if (phase) phase[e] = T.phase[e];
      }
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_apply_h(bs_system* s, const void* x, void* y, int dtype, int accumulate, void* stream) {
  try {
    checked(s);
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (!x || !y) return set_err("null vector");
    if (x == y) return set_err("x and y must not overlap");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == BS_F32) launch_apply<float>(s, (const float*)x, (float*)y, accumulate ? 1 : 0, st);
    else if (dtype == BS_F64) launch_apply<double>(s, (const double*)x, (double*)y, accumulate ? 1 : 0, st);
    else return set_err("dtype must be BS_F32 or BS_F64");
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_num_variants(int* n) {
  if (!n) return set_err("null argument");
  *n = kAllVariants;
  return 0;
}

int bs_variant_info(int v, int* rows, int* staged, int* xcd_map, int* prefetch) {
  if (v < 0 || v >= kAllVariants) return set_err("bad variant");
  if (v == kRowsVariant) {
    // the row kernel: one proton row, nothing staged
    if (rows) *rows = 1;
    if (staged) *staged = 0;
    if (xcd_map) *xcd_map = 0;
    if (prefetch) *prefetch = 0;
    return 0;
  }
  if (rows) *rows = kVarR[v % kNumShapes];
  if (staged) *staged = kVarG[v % kNumShapes];
  if (xcd_map) *xcd_map = (v / kNumShapes) % 2;
  if (prefetch) *prefetch = v >= 2 * kNumShapes;
  return 0;
}

int bs_set_variant(bs_system* s, int dtype, int v) {
  try {
    checked(s);
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    if (v < -1 || v >= kAllVariants) return set_err("bad variant");
    if (v >= 0) {
      size_t need = dtype == BS_F32 ? lds_bytes<float>(s, v) : lds_bytes<double>(s, v);
      if (need > kLdsLimit) return set_err("variant does not fit in LDS for this system");
    }
    s->variant[dtype == BS_F32 ? 0 : 1] = v;
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_get_variant(const bs_system* s, int dtype, int* v) {
  if (!s || !v || (dtype != BS_F32 && dtype != BS_F64)) return set_err("bad argument");
  *v = s->variant[dtype == BS_F32 ? 0 : 1];
  return 0;
}

int bs_autotune(bs_system* s, int dtype, int* v) {
  try {
    checked(s);
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    int r;
    if (dtype == BS_F32) r = autotune<float>(s, nullptr);
    else if (dtype == BS_F64) r = autotune<double>(s, nullptr);
    else return set_err("dtype must be BS_F32 or BS_F64");
    if (v) *v = r;
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_apply_h_host(bs_system* s, const void* x, void* y, int dtype, int accumulate) {
  try {
    checked(s);
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (!x || !y) return set_err("null vector");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    const size_t bytes = (size_t)s->t.dim * (dtype == BS_F32 ? 4 : 8);
    if (s->hostx.n < bytes) {
      s->hostx.alloc(bytes);
      s->hosty.alloc(bytes);
    }
    HIP_OK(hipMemcpy(s->hostx.p, x, bytes, hipMemcpyHostToDevice));
    if (accumulate) HIP_OK(hipMemcpy(s->hosty.p, y, bytes, hipMemcpyHostToDevice));
    if (dtype == BS_F32) launch_apply<float>(s, (const float*)s->hostx.p, (float*)s->hosty.p, accumulate ? 1 : 0, nullptr);
    else launch_apply<double>(s, (const double*)s->hostx.p, (double*)s->hosty.p, accumulate ? 1 : 0, nullptr);
    HIP_OK(hipMemcpy(y, s->hosty.p, bytes, hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}


int bs_observable(bs_system* s, int which, const void* psi, int dtype, double* value) {
  try {
    checked(s);
    if (which != 0 && which != 1) return set_err("which must be 0 (J^2) or 1 (T^2)");
    if (!psi || !value) return set_err("null argument");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    HIP_OK(hipDeviceSynchronize());  // psi may come from a caller's non-blocking stream
    if (!s->obs[which]) {
      std::unique_ptr<bs_system> o(new bs_system());
      o->t = bs::build_tables_text(s->t.in.orbits, bs::observable_int_text(s->t.in.orbits, which), s->t.in.Z,
                                   s->t.in.N, s->t.in.jz2, true);
      if (o->t.dim != s->t.dim) return set_err("observable basis differs");
      upload_tables(o.get());
      s->obs[which] = std::move(o);
    }
    if (dtype == BS_F32) *value = expectation<float>(s->obs[which].get(), (const float*)psi);
    else *value = expectation<double>(s->obs[which].get(), (const double*)psi);
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}


int bs_operator_matrix(bs_system* s, const char* intf, float spscale, float a0, float a, float x, const void* vecs,
                       int nvec, int dtype, double* out) {
  try {
    checked(s);
    if (!vecs || !out || nvec < 1) return set_err("null argument or nvec < 1");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    // the operator as an interaction on the same orbits, nucleus and 2Jz, or
    // the system's own Hamiltonian when no file is given
    bs_system o;
    bs_system* op = s;
    if (intf) {
      o.t = bs::build_tables(s->t.in.orbits, intf, s->t.in.Z, s->t.in.N, s->t.in.jz2, spscale, a0, a, x, true);
      if (o.t.dim != s->t.dim) return set_err("operator basis differs");
      upload_tables(&o);
      op = &o;
    } else if (!s->uploaded) {
      return set_err("tables not uploaded (call bs_upload)");
    }
    const int64_t D = s->t.dim;
    const size_t es = dtype == BS_F32 ? 4 : 8;
    const int nb = std::min(grid_for(D), 1024);
    DevBuf<unsigned char> hv;
    DevBuf<double> part, res;
    hv.alloc(D * es);
    part.alloc((size_t)nb * nvec);
    res.alloc(nvec);
    for (int i = 0; i < nvec; i++) {
      const void* vi = static_cast<const char*>(vecs) + (size_t)i * D * es;
      // (src/bexpect.f90:136-160) O v_i, then <v_f | O v_i> for every f
      if (dtype == BS_F32) {
        launch_apply<float>(op, (const float*)vi, (float*)hv.p, 0, nullptr);
        k_multidot<float><<<nb, kBlock>>>((const float*)vecs, D, nvec, (const float*)hv.p, D, part.p);
      } else {
        launch_apply<double>(op, (const double*)vi, (double*)hv.p, 0, nullptr);
        k_multidot<double><<<nb, kBlock>>>((const double*)vecs, D, nvec, (const double*)hv.p, D, part.p);
      }
      k_reduce<<<nvec, kBlock>>>(part.p, nb, nvec, res.p, 0);
      HIP_OK(hipGetLastError());
      HIP_OK(hipMemcpy(out + (size_t)i * nvec, res.p, sizeof(double) * nvec, hipMemcpyDeviceToHost));
    }
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_apply_onebody(bs_system* s, const char* opme, const void* x, void* y, int nvec, int dtype, int* jop,
                     int* top) {
  try {
    checked(s);
    if (!opme || !x || !y || nvec < 1) return set_err("null argument or nvec < 1");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    const bs::OneBodyOp op = bs::read_opme(opme, s->t);
    if (jop) *jop = op.jop;
    if (top) *top = op.top;
    bs_system o;
    o.t = bs::onebody_tables(s->t, op);
    upload_tables(&o);
    const int64_t D = s->t.dim;
    const size_t es = dtype == BS_F32 ? 4 : 8;
    // (src/bdenslib1.f90:1666-1689) each state: y = O x
    for (int k = 0; k < nvec; k++) {
      const char* xk = static_cast<const char*>(x) + (size_t)k * D * es;
      char* yk = static_cast<char*>(y) + (size_t)k * D * es;
      if (dtype == BS_F32) launch_apply<float>(&o, (const float*)xk, (float*)yk, 0, nullptr);
      else launch_apply<double>(&o, (const double*)xk, (double*)yk, 0, nullptr);
    }
    HIP_OK(hipDeviceSynchronize());
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_num_params(const bs_system* s, int* nspe, int* ntbme) {
  if (!s) return set_err("null system");
  if (nspe) *nspe = (int)s->t.in.orbits.size();
  if (ntbme) *ntbme = (int)s->t.lines.size();
  return 0;
}

int bs_density(bs_system* s, const void* psi, int dtype, double* rho_pp, double* rho_nn, double* rho_pn,
               double* norm2) {
  try {
    checked(s);
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (!psi) return set_err("null vector");
    HIP_OK(hipDeviceSynchronize());  // psi may come from a caller's non-blocking stream
    if (dtype == BS_F32) run_density<float>(s, (const float*)psi, rho_pp, rho_nn, rho_pn, norm2);
    else if (dtype == BS_F64) run_density<double>(s, (const double*)psi, rho_pp, rho_nn, rho_pn, norm2);
    else return set_err("dtype must be BS_F32 or BS_F64");
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_gradient(bs_system* s, const void* psi, int dtype, double* grad) {
  try {
    checked(s);
    if (!grad) return set_err("null gradient");
    const bs::Tables& t = s->t;
    const size_t npx2 = (size_t)t.npairxx * t.npairxx, npn2 = (size_t)t.npairpn * t.npairpn;
    std::vector<double> rho[3] = {std::vector<double>(npx2), std::vector<double>(npx2), std::vector<double>(npn2)};
    double n2 = 0.0;
    int rc = bs_density(s, psi, dtype, rho[0].data(), rho[1].data(), rho[2].data(), &n2);
    if (rc) return rc;
    if (!(n2 > 0.0)) return set_err("zero vector");
    if (!s->jac_ready) {
      s->jac = bs::param_jacobian(t);
      s->jac_ready = true;
    }
    const int norb = (int)t.in.orbits.size();
    for (size_t p = 0; p < s->jac.size(); p++) {
      double g = 0.0;
      for (int k = 0; k < 3; k++)
        for (size_t i = 0; i < s->jac[p].idx[k].size(); i++) g += (double)s->jac[p].coef[k][i] * rho[k][s->jac[p].idx[k][i]];
      grad[p] = g / n2;
    }
    if (t.call_spe) {
      // single particle: the SPEs act through the diagonal, dE/deps_a = spescale * <n_a>
      const int64_t D = t.dim;
      std::vector<double> h(D);
      if (dtype == BS_F32) {
        std::vector<float> f(D);
        HIP_OK(hipMemcpy(f.data(), psi, sizeof(float) * D, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < D; i++) h[i] = f[i];
      } else {
        HIP_OK(hipMemcpy(h.data(), psi, sizeof(double) * D, hipMemcpyDeviceToHost));
      }
      std::vector<double> occ(norb, 0.0);
      for (int it = 0; it < 2; it++) {
        const bs::Species& S = t.sp[it];
        const bs::Species& C = t.sp[1 - it];
        for (size_t sd = 0; sd < S.mask.size(); sd++) {
          double w = 0.0;
          const int cs = S.conj[S.sdsec[sd]];
          for (int64_t o = C.secfirst[cs]; o < C.secfirst[cs] + C.seccount[cs]; o++) {
            const int64_t st = it == 0 ? S.start[sd] + C.start[o] : C.start[o] + S.start[sd];
            w += h[st] * h[st];
          }
          for (uint64_t q = S.mask[sd]; q; q &= q - 1) occ[t.sorb[__builtin_ctzll(q)]] += w;
        }
      }
      for (int a = 0; a < norb; a++) grad[a] += (double)t.in.spscale * occ[a] / n2;
    }
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_apply_block(bs_system* s, const void* x, void* y, int K, int dtype) {
  try {
    checked(s);
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (!x || !y) return set_err("null vector");
    if (K < 1 || K > kMaxBatch) return set_err("K must be in 1..64");
    const BatchVals v = shared_vals(s, K);
    if (dtype == BS_F32) launch_batched<float>(s, v, (const float*)x, (float*)y, nullptr);
    else if (dtype == BS_F64) launch_batched<double>(s, v, (const double*)x, (double*)y, nullptr);
    else return set_err("dtype must be BS_F32 or BS_F64");
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_multi_variant(bs_system* s, int dtype, int K, int shared, int v) {
  try {
    checked(s);
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    if (K < 1 || K > kMaxBatch) return set_err("K must be in 1..64");
    if (v < 0 || v >= kNumMulti) return set_err("no such multi-vector variant");
    if (!multi_fits(s, v, K, dtype == BS_F32 ? 4 : 8, shared != 0))
      return set_err("variant does not fit for this system and K");
    const int key = (dtype == BS_F64 ? 1000 : 0) + (shared ? 100 : 0) + K;
    s->multi_choice[key] = v;
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_multi_info(const bs_system* s, int dtype, int K, int shared, int* chosen, float* ms) {
  if (!s) return set_err("null system");
  const int key = (dtype == BS_F64 ? 1000 : 0) + (shared ? 100 : 0) + K;
  auto it = s->multi_choice.find(key);
  if (chosen) *chosen = it == s->multi_choice.end() ? -1 : it->second;
  if (ms) {
    auto jt = s->multi_ms.find(key);
    for (int v = 0; v < kNumMulti; v++) ms[v] = jt == s->multi_ms.end() ? 0.0f : jt->second[v];
  }
  return 0;
}

int bs_batch_set_variant(bs_batch* b, int dtype, int v) {
  if (!b) return set_err("null batch");
  b->work.reset();  // a captured graph holds the old choice
  return bs_multi_variant(b->base.get(), dtype, b->K, 0, v);
}

int bs_batch_init(const char* sps, const char* const* intfiles, int K, int Z, int N, int jz2, float spscale,
                  float a0, float a, float x, bs_batch** out) {
  try {
    if (!out || !sps || !intfiles) return set_err("null argument");
    *out = nullptr;
    if (K < 1 || K > kMaxBatch) return set_err("K must be in 1..64");
    for (int k = 0; k < K; k++)
      if (!intfiles[k]) return set_err("null interaction file name");
    std::unique_ptr<bs_batch> b(new bs_batch());
    b->K = K;
    for (int k = 0; k < K; k++) b->intfiles.push_back(intfiles[k]);
    b->base.reset(new bs_system());
    bs::ModelInput in;
    bs::read_sps(sps, in);
    b->base->t = bs::build_tables_structural(in.orbits, intfiles[0], Z, N, jz2, spscale, a0, a, x);
    upload_tables(b->base.get());
    batch_upload_values(b.get());
    HIP_OK(hipDeviceSynchronize());
    *out = b.release();
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_batch_set_interactions(bs_batch* b, const char* const* intfiles) {
  try {
    if (!b || !intfiles) return set_err("null argument");
    for (int k = 0; k < b->K; k++)
      if (!intfiles[k]) return set_err("null interaction file name");
    const std::vector<std::string> old = b->intfiles;
    const bool old_same = b->same_params;
    b->intfiles.assign(intfiles, intfiles + b->K);
    b->same_params = true;
    try {
      // every file is read and checked before any device buffer changes
      batch_upload_values(b);
    } catch (...) {
      b->intfiles = old;
      b->same_params = old_same;
      throw;
    }
    b->work.reset();  // captured graphs hold the old value buffers
    HIP_OK(hipDeviceSynchronize());
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_batch_finalize(bs_batch* b) {
  delete b;
  return 0;
}

int bs_batch_size(const bs_batch* b, int* K, int64_t* dim) {
  if (!b) return set_err("null batch");
  if (K) *K = b->K;
  if (dim) *dim = b->base->t.dim;
  return 0;
}

bs_system* bs_batch_system(bs_batch* b) { return b ? b->base.get() : nullptr; }

int bs_batch_apply(bs_batch* b, const void* x, void* y, int dtype) {
  try {
    if (!b) return set_err("null batch");
    if (!x || !y) return set_err("null vector");
    if (dtype == BS_F32) launch_batched<float>(b->base.get(), b->bv, (const float*)x, (float*)y, nullptr);
    else if (dtype == BS_F64) launch_batched<double>(b->base.get(), b->bv, (const double*)x, (double*)y, nullptr);
    else return set_err("dtype must be BS_F32 or BS_F64");
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_batch_lanczos(bs_batch* b, const bs_lanczos_params* p, double* alpha, double* beta, double* evals,
                     void* evecs, double* resid, int* iters) {
  try {
    if (!b || !p || !evals) return set_err("null argument");
    if (p->dtype == BS_F32) run_batch_lanczos<float>(b, p, alpha, beta, evals, evecs, resid, iters);
    else if (p->dtype == BS_F64) run_batch_lanczos<double>(b, p, alpha, beta, evals, evecs, resid, iters);
    else return set_err("dtype must be BS_F32 or BS_F64");
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_batch_gradient(bs_batch* b, const void* psi, int dtype, double* grad) {
  try {
    if (!b || !psi || !grad) return set_err("null argument");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    if (!b->same_params) return set_err("interactions list different two-body elements; gradients need one list");
    bs_system* s = b->base.get();
    const int64_t D = s->t.dim;
    const int np = (int)(s->t.in.orbits.size() + s->t.lines.size());
    DevBuf<double> lane;
    lane.alloc(D);
    for (int k = 0; k < b->K; k++) {
      if (dtype == BS_F32)
        k_gather_lane<float><<<grid_for(D), kBlock>>>(lane.p, (const float*)psi, D, b->K, k);
      else
        k_gather_lane<double><<<grid_for(D), kBlock>>>(lane.p, (const double*)psi, D, b->K, k);
      HIP_OK(hipGetLastError());
      const int rc = bs_gradient(s, lane.p, BS_F64, grad + (size_t)k * np);
      if (rc) return rc;
    }
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_write_wfn(const bs_system* s, const char* path, int nkeep, const float* energy, const float* xj,
                 const float* xt2, const void* vecs, int dtype, int on_device) {
  try {
    if (!s || !path || !vecs || (nkeep > 0 && (!energy || !xj || !xt2))) return set_err("null argument");
    if (nkeep < 0) return set_err("nkeep must be >= 0");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    const int64_t D = s->t.dim;
    const std::string h = bs::wfn_header(s->t);
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_err(std::string("cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    const int32_t nk = nkeep;
    ok = ok && std::fwrite(&nk, 4, 1, f) == 1;
    std::vector<float> v32(D);
    std::vector<double> v64(dtype == BS_F64 ? D : 0);
    for (int k = 0; k < nkeep && ok; k++) {
      // (src/bwfnlib.f90:124-125) magic, 1-based index, E, J, T(T+1), then
      // the vector in single precision
      const int32_t head[2] = {27182818, k + 1};
      const float vals[3] = {energy[k], xj[k], xt2[k]};
      ok = std::fwrite(head, 4, 2, f) == 2 && std::fwrite(vals, 4, 3, f) == 3;
      const size_t es = dtype == BS_F32 ? 4 : 8;
      const char* src = static_cast<const char*>(vecs) + (size_t)k * D * es;
      if (dtype == BS_F32) {
        if (on_device) HIP_OK(hipMemcpy(v32.data(), src, 4 * (size_t)D, hipMemcpyDeviceToHost));
        else std::memcpy(v32.data(), src, 4 * (size_t)D);
      } else {
        if (on_device) HIP_OK(hipMemcpy(v64.data(), src, 8 * (size_t)D, hipMemcpyDeviceToHost));
        else std::memcpy(v64.data(), src, 8 * (size_t)D);
        for (int64_t i = 0; i < D; i++) v32[i] = (float)v64[i];
      }
      ok = ok && std::fwrite(v32.data(), 4, (size_t)D, f) == (size_t)D;
    }
    if (!ok) return set_err(std::string("write failed: ") + path);
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_init_wfn_host(const char* wfn, const char* intf, float spscale, float a0, float a, float x,
                     bs_system** out) {
  try {
    if (!wfn || !out) return set_err("null argument");
    *out = nullptr;
    const bs::WfnInfo w = bs::read_wfn_info(wfn);
    std::unique_ptr<bs_system> s(new bs_system());
    if (intf) {
      s->t = bs::build_tables(w.orbits, intf, w.Z, w.N, w.jz2, spscale, a0, a, x, true);
    } else {
      // no interaction: the basis and its jumps only (H = 0)
      std::string zero = "0";
      for (size_t i = 0; i < w.orbits.size(); i++) zero += " 0";
      s->t = bs::build_tables_text(w.orbits, zero + "\n", w.Z, w.N, w.jz2, true);
    }
    // the rebuilt basis must be the one the vectors are stored in
    if (bs::wfn_header(s->t) != w.header)
      return set_err(std::string("basis built from ") + wfn +
                     " differs from the one it records (parity, W truncation or orbit order)");
    *out = s.release();
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_init_wfn(const char* wfn, const char* intf, float spscale, float a0, float a, float x, bs_system** out) {
  int r = bs_init_wfn_host(wfn, intf, spscale, a0, a, x, out);
  if (r) return r;
  r = bs_upload(*out);
  if (r) {
    delete *out;
    *out = nullptr;
  }
  return r;
}

int bs_read_wfn(const bs_system* s, const char* path, int first, int count, float* energy, float* xj, float* xt2,
                void* vecs, int dtype, int on_device, int* nkeep) {
  try {
    if (!s || !path) return set_err("null argument");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    const bs::WfnInfo w = bs::read_wfn_info(path);
    if (w.header != bs::wfn_header(s->t))
      return set_err(std::string("the basis of ") + path + " is not this system's basis");
    if (nkeep) *nkeep = w.nkeep;
    if (count == 0) return 0;
    if (first < 0 || count < 0 || (int64_t)first + count > w.nkeep)
      return set_err("states " + std::to_string(first + 1) + ".." + std::to_string(first + count) +
                     " requested, file holds " + std::to_string(w.nkeep));
    if (!vecs || !energy || !xj || !xt2) return set_err("null output");
    const int64_t D = w.dim;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return set_err(std::string("cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    std::vector<unsigned char> raw((size_t)D * w.elem);
    std::vector<float> v32(dtype == BS_F32 ? D : 0);
    std::vector<double> v64(dtype == BS_F64 ? D : 0);
    const int64_t reclen = 20 + (int64_t)w.elem * D;
    for (int q = 0; q < count; q++) {
      const int k = first + q;
      // (src/bwfnlib.f90:926-993 position, :159-273 record and its checks)
      if (fseeko(f, (off_t)(w.offset + (int64_t)k * reclen), SEEK_SET) != 0) return set_err("seek failed");
      int32_t head[2];
      float ejt[3];
      if (std::fread(head, 4, 2, f) != 2 || std::fread(ejt, 4, 3, f) != 3 ||
          std::fread(raw.data(), 1, raw.size(), f) != raw.size())
        return set_err(std::string("read failed: ") + path);
      if (head[0] != 27182818) return set_err("vector " + std::to_string(k + 1) + " is missing magic number");
      if (head[1] != k + 1) return set_err("read wrong vector: wanted " + std::to_string(k + 1) + ", found " +
                                            std::to_string(head[1]));
      energy[q] = ejt[0];
      xj[q] = ejt[1];
      xt2[q] = ejt[2];
      const size_t es = dtype == BS_F32 ? 4 : 8;
      void* host;
      if (dtype == BS_F32) {
        if (w.elem == 4) std::memcpy(v32.data(), raw.data(), raw.size());
        else
          for (int64_t i = 0; i < D; i++) {
            double d;
            std::memcpy(&d, &raw[8 * i], 8);
            v32[i] = (float)d;
          }
        host = v32.data();
      } else {
        if (w.elem == 8) std::memcpy(v64.data(), raw.data(), raw.size());
        else
          for (int64_t i = 0; i < D; i++) {
            float g;
            std::memcpy(&g, &raw[4 * i], 4);
            v64[i] = g;
          }
        host = v64.data();
      }
      char* dst = static_cast<char*>(vecs) + (size_t)q * D * es;
      if (on_device) HIP_OK(hipMemcpy(dst, host, es * (size_t)D, hipMemcpyHostToDevice));
      else std::memcpy(dst, host, es * (size_t)D);
    }
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_autotune_times(const bs_system* s, int dtype, float* ms) {
  if (!s || !ms || (dtype != BS_F32 && dtype != BS_F64)) return set_err("bad argument");
  for (int v = 0; v < kAllVariants; v++) ms[v] = s->tune_ms[dtype == BS_F32 ? 0 : 1][v];
  return 0;
}

int bs_strength(bs_system* s, const bs_lanczos_params* p, const void* pivot, double* pivot_norm, double* ritz,
                double* strength, double* alpha, double* beta, double* evals, void* evecs, double* resid,
                int* iters) {
  try {
    checked(s);
    if (!p || !evals || !pivot) return set_err("null argument");
    if (p->thick != 0) return set_err("strength functions need a fixed run (thick = 0)");
    if (p->dtype != BS_F32 && p->dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (p->dtype == BS_F32) {
      PivotRun<float> pr{(const float*)pivot, pivot_norm, ritz, strength};
      return run_lanczos<float>(s, p, alpha, beta, evals, evecs, resid, iters, pr);
    }
    if (p->dtype == BS_F64) {
      PivotRun<double> pr{(const double*)pivot, pivot_norm, ritz, strength};
      return run_lanczos<double>(s, p, alpha, beta, evals, evecs, resid, iters, pr);
    }
    return set_err("dtype must be BS_F32 or BS_F64");
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_exact(bs_system* s, int nkeep, int dtype, double* evals, void* evecs, double* resid, int* iters) {
  try {
    checked(s);
    if (!evals) return set_err("null argument");
    if (dtype != BS_F32 && dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    const int64_t D = s->t.dim;
    if (nkeep < 1 || nkeep > D) return set_err("nkeep must be in 1..dim");
    if (D > 65536) return set_err("exact diagonalization is limited to dim <= 65536 (dense H in HBM)");
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    const hipStream_t st = nullptr;
    // (src/blanczosmain.f90:2171-2471) H column by column: here 64 unit
    // vectors at a time through the multi-vector kernel, FP64
    DevBuf<double> Hd, X, Y, W, E;
    DevBuf<int> info;
    Hd.alloc((size_t)D * D);
    X.alloc((size_t)D * 64);
    Y.alloc((size_t)D * 64);
    W.alloc(D);
    E.alloc(D);
    info.alloc(1);
    const BatchVals bv = shared_vals(s, 64);
    const int g = grid_for(D * 64);
    for (int64_t j0 = 0; j0 < D; j0 += 64) {
      const int kb = (int)std::min<int64_t>(64, D - j0);
      k_unit_block<<<g, kBlock, 0, st>>>(X.p, D, j0, kb);
      launch_batched<double>(s, bv, X.p, Y.p, st);
      k_scatter_cols<<<g, kBlock, 0, st>>>(Hd.p, D, Y.p, j0, kb);
      HIP_OK(hipGetLastError());
    }
    // the dense symmetric eigenproblem (the reference calls DSYEV): rocSOLVER
    // divide and conquer, eigenvectors in the columns of H
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return set_err("rocblas_create_handle failed");
    std::unique_ptr<std::remove_pointer<rocblas_handle>::type, rocblas_status (*)(rocblas_handle)> guard(
        h, rocblas_destroy_handle);
    if (rocblas_set_stream(h, st) != rocblas_status_success) return set_err("rocblas_set_stream failed");
    if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)D, Hd.p, (rocblas_int)D, W.p,
                         E.p, info.p) != rocblas_status_success)
      return set_err("rocsolver_dsyevd failed");
    int inf = 0;
    HIP_OK(hipMemcpy(&inf, info.p, sizeof(int), hipMemcpyDeviceToHost));
    if (inf != 0) return set_err("rocsolver_dsyevd did not converge (info " + std::to_string(inf) + ")");
    std::vector<double> ev(D);
    HIP_OK(hipMemcpy(ev.data(), W.p, sizeof(double) * D, hipMemcpyDeviceToHost));
    for (int k = 0; k < nkeep; k++) evals[k] = ev[k];
    const int gd = grid_for(D);
    const int nb1 = gd < 1024 ? gd : 1024;
    DevBuf<double> HX, part1, dots;
    HX.alloc(D);
    part1.alloc(nb1);
    dots.alloc(4);
    for (int k = 0; k < nkeep; k++) {
      double* xk = Hd.p + (int64_t)k * D;
      if (resid) {
        // ||H x - E x|| of the eigenvector, one more apply
        launch_apply<double>(s, xk, HX.p, 0, st);
        HIP_OK(hipMemcpy(dots.p + 1, &evals[k], sizeof(double), hipMemcpyHostToDevice));
        k_sub_scaled<double><<<gd, kBlock, 0, st>>>(HX.p, xk, dots.p + 1, D);
        k_multidot<double><<<nb1, kBlock, 0, st>>>(HX.p, D, 1, HX.p, D, part1.p);
        k_reduce<<<1, kBlock, 0, st>>>(part1.p, nb1, 1, dots.p + 2, 1);
        HIP_OK(hipMemcpy(&resid[k], dots.p + 2, sizeof(double), hipMemcpyDeviceToHost));
      }
      if (evecs) {
        if (dtype == BS_F64)
          HIP_OK(hipMemcpyAsync((double*)evecs + (int64_t)k * D, xk, sizeof(double) * D, hipMemcpyDeviceToDevice, st));
        else
          k_convert<double, float><<<gd, kBlock, 0, st>>>((float*)evecs + (int64_t)k * D, xk, D);
      }
    }
    HIP_OK(hipStreamSynchronize(st));
    if (iters) *iters = (int)D;
    return 0;
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_block_lanczos(bs_system* s, const bs_lanczos_params* p, int block, double* evals, void* evecs,
                     double* resid, int* iters) {
  try {
    checked(s);
    if (!p || !evals) return set_err("null argument");
    if (block < 1 || block > 8) return set_err("block size must be in 1..8");
    if (p->dtype != BS_F32 && p->dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (p->dtype == BS_F32) return run_block_lanczos_k<float>(s, p, block, evals, evecs, resid, iters);
    return run_block_lanczos_k<double>(s, p, block, evals, evecs, resid, iters);
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_block_strength(bs_system* s, const bs_lanczos_params* p, int block, const void* pivots, int pivot_dtype,
                      double* pnorm2, double* strength, double* evals, void* evecs, double* resid, int* iters) {
  try {
    checked(s);
    if (!p || !evals || !pivots) return set_err("null argument");
    if (block < 1 || block > 8) return set_err("block size must be in 1..8");
    if (p->dtype != BS_F32 && p->dtype != BS_F64) return set_err("dtype must be BS_F32 or BS_F64");
    if (pivot_dtype != BS_F32 && pivot_dtype != BS_F64) return set_err("pivot dtype must be BS_F32 or BS_F64");
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    BlockPivot bp;
    bp.piv = pivots;
    bp.dtype = pivot_dtype;
    bp.strength = strength;
    bp.pnorm2 = pnorm2;
    if (p->dtype == BS_F32) return run_block_lanczos_k<float>(s, p, block, evals, evecs, resid, iters, &bp);
    return run_block_lanczos_k<double>(s, p, block, evals, evecs, resid, iters, &bp);
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_lanczos(bs_system* s, const bs_lanczos_params* p, double* alpha, double* beta, double* evals,
               void* evecs, double* resid, int* iters) {
  try {
    checked(s);
    if (!p || !evals) return set_err("null argument");
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (p->dtype == BS_F32) return run_lanczos<float>(s, p, alpha, beta, evals, evecs, resid, iters);
    if (p->dtype == BS_F64) return run_lanczos<double>(s, p, alpha, beta, evals, evecs, resid, iters);
    return set_err("dtype must be BS_F32 or BS_F64");
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_lanczos_noalloc(bs_system* s, const bs_lanczos_params* p, double* alpha, double* beta, double* evals,
                       int* iters) {
  try {
    checked(s);
    if (!p || !evals) return set_err("null argument");
    if (!s->uploaded) return set_err("tables not uploaded (call bs_upload)");
    if (p->thick != 0) return set_err("bs_lanczos_noalloc runs plain Lanczos only (thick = 0)");
    if (p->polish != 0) return set_err("bs_lanczos_noalloc has no polish (polish = 0)");
    void* ptr = scratch_ptr();
    if (p->dtype == BS_F32)
      return run_lanczos<float>(s, p, alpha, beta, evals, nullptr, nullptr, iters, PivotRun<float>(), ptr,
                                kScratchBytes);
    if (p->dtype == BS_F64)
      return run_lanczos<double>(s, p, alpha, beta, evals, nullptr, nullptr, iters, PivotRun<double>(), ptr,
                                 kScratchBytes);
    return set_err("dtype must be BS_F32 or BS_F64");
  } catch (const std::exception& e) {
    return set_err(e.what());
  }
}

int bs_lanczos_stats(const bs_system* s, int* random_restarts, int* second_passes, int* steps,
                     int* thick_restarts) {
  if (!s) return set_err("null system");
  if (thick_restarts) *thick_restarts = s->last_thick_restarts;
  if (random_restarts) *random_restarts = s->last_random_restarts;
  if (second_passes) *second_passes = s->last_second_passes;
  if (steps) *steps = s->last_steps;
  return 0;
}

}  // extern "C"
