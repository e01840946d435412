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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bigstick_amd.h"
#include "bs_tables.hpp"

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
  const int64_t* pstart;     // [nP] row offset of proton SD
  const int32_t* rowlen;     // [nP] neutron row length
  const int32_t* nfirst;     // [nP] first neutron SD (global) of the row's sector
  const float* pspe_sd;      // [nP]
  const float* nspe_sd;      // [nN]
  // PP merged CSR by target proton SD
  const int64_t* pp_ptr;
  const int32_t* pp_src;
  const float* pp_val32;
  const double* pp_val64;
  // NN merged CSR by target neutron SD; src is local (within sector)
  const int64_t* nn_ptr;
  const int32_t* nn_srcl;
  const float* nn_val32;
  const double* nn_val64;
  // proton one-body: class pointers per row [nP][ndelta+1] (absolute), entries
  const int32_t* pcls;
  const int32_t* p1_src;     // proton source SD (global)
  const int16_t* p1_meta;    // pp_local | (phase<0) << 8
  // neutron one-body: class pointers per neutron SD [nN][ndelta+1], packed entries
  const int32_t* ncls;
  const uint32_t* n1_pk;     // src_local (16b) | nn_local (8b) << 16 | (phase<0) << 24
  // pn element blocks
  const float* W;
  const int32_t* woff;       // [ndelta]
  const int32_t* wstride;    // [ndelta] = size of class -k
  const int32_t* negk;       // [ndelta]
};

template <typename T> __device__ __forceinline__ T vload(const float* a, const double* b, int64_t i);
template <> __device__ __forceinline__ float vload<float>(const float* a, const double*, int64_t i) { return a[i]; }
template <> __device__ __forceinline__ double vload<double>(const float*, const double* b, int64_t i) { return b[i]; }

// One workgroup per (proton SD row f, column chunk).  CPT columns per thread.
template <typename T, int CPT>
__global__ __launch_bounds__(kBlock) void k_apply_rows(DevTables tb, const T* __restrict__ x,
                                                       T* __restrict__ y, int accumulate) {
  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t base_f = tb.pstart[f];
  const int L = tb.rowlen[f];
  const int nf0 = tb.nfirst[f];
  const int c0 = blockIdx.y * (kBlock * CPT);
  T acc[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) acc[q] = T(0);

  if (tb.call_spe) {
    const T ep = (T)tb.pspe_sd[f];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int c = c0 + tid + q * kBlock;
      if (c < L) acc[q] += (ep + (T)tb.nspe_sd[nf0 + c]) * x[base_f + c];
    }
  }
  // PP: row AXPYs
  for (int64_t e = tb.pp_ptr[f]; e < tb.pp_ptr[f + 1]; e++) {
    const int64_t bi = tb.pstart[tb.pp_src[e]];
    const T v = vload<T>(tb.pp_val32, tb.pp_val64, e);
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int c = c0 + tid + q * kBlock;
      if (c < L) acc[q] += v * x[bi + c];
    }
  }
  // NN: gathers within row f
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int c = c0 + tid + q * kBlock;
    if (c < L) {
      const int n = nf0 + c;
      T s = T(0);
      for (int64_t e = tb.nn_ptr[n]; e < tb.nn_ptr[n + 1]; e++)
        s += vload<T>(tb.nn_val32, tb.nn_val64, e) * x[base_f + tb.nn_srcl[e]];
      acc[q] += s;
    }
  }
  // PN
  const int nd1 = tb.ndelta + 1;
  for (int k = 0; k < tb.ndelta; k++) {
    const int p0 = tb.pcls[(int64_t)f * nd1 + k], p1 = tb.pcls[(int64_t)f * nd1 + k + 1];
    if (p0 == p1) continue;
    const int kn = tb.negk[k];
    const float* Wk = tb.W + tb.woff[k];
    const int ws = tb.wstride[k];
    for (int pj = p0; pj < p1; pj++) {
      const int64_t bi = tb.pstart[tb.p1_src[pj]];
      const int meta = tb.p1_meta[pj];
      const float* Wr = Wk + (meta & 0xff) * ws;
      const T php = (meta >> 8) ? T(-1) : T(1);
#pragma unroll
      for (int q = 0; q < CPT; q++) {
        const int c = c0 + tid + q * kBlock;
        if (c < L) {
          const int64_t n = nf0 + c;
          const int e0 = tb.ncls[n * nd1 + kn], e1 = tb.ncls[n * nd1 + kn + 1];
          T s = T(0);
          for (int e = e0; e < e1; e++) {
            const uint32_t pk = tb.n1_pk[e];
            const T w = (T)Wr[(pk >> 16) & 0xff];
            const T xv = x[bi + (pk & 0xffff)];
            s += ((pk >> 24) ? -w : w) * xv;
          }
          acc[q] += php * s;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int c = c0 + tid + q * kBlock;
    if (c < L) {
      if (accumulate) y[base_f + c] += acc[q];
      else y[base_f + c] = acc[q];
    }
  }
}

// ----------------------------------------------------------- vector kernels
// Partial sums of <V_j, w> for j in [0, m): block b writes part[b*m + j].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_multidot(const T* __restrict__ V, int64_t ld, int m,
                                                     const T* __restrict__ w, int64_t D,
                                                     double* __restrict__ part) {
  __shared__ double red[8][kBlock / 64];
  const int64_t chunk = (D + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < D ? lo + chunk : D;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int j0 = 0; j0 < m; j0 += 8) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
      const double wi = (double)w[i];
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (j0 + u < m) a[u] += (double)V[(int64_t)(j0 + u) * ld + i] * wi;
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

// w -= T(sum_j d_j V_j)   (src/breorthog.f90:880-889)
template <typename T>
__global__ void k_multiaxpy(T* __restrict__ w, const T* __restrict__ V, int64_t ld, int m,
                            const double* __restrict__ d, int64_t D) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x) {
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

template <typename T, typename U>
__global__ void k_convert(U* __restrict__ out, const T* __restrict__ in, int64_t D) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < D; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (U)in[i];
}

// ---------------------------------------------------------------- helpers
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIP_OK(hipMalloc(&p, count * sizeof(T)));
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
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

int grid_for(int64_t D) {
  int64_t g = (D + kBlock - 1) / kBlock;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 4096);
}

}  // namespace

struct bs_system {
  bs::Tables t;
  bool uploaded = false;
  int cpt = 1, nchunk = 1;
  DevTables dt{};
  DevBuf<int64_t> pstart, pp_ptr, nn_ptr;
  DevBuf<int32_t> rowlen, nfirst, pp_src, nn_srcl, pcls, p1_src, ncls, woff, wstride, negk;
  DevBuf<int16_t> p1_meta;
  DevBuf<uint32_t> n1_pk;
  DevBuf<float> pspe_sd, nspe_sd, pp_val32, nn_val32, W;
  DevBuf<double> pp_val64, nn_val64;
};

namespace {

void upload_tables(bs_system* s) {
  const bs::Tables& t = s->t;
  const bs::Species& P = t.sp[0];
  const bs::Species& N = t.sp[1];
  const int nP = (int)P.mask.size(), nN = (int)N.mask.size();
  const int nd = t.ndelta, ns = t.nsps;
  std::vector<int64_t> pstart(P.start);
  std::vector<int32_t> rowlen(nP), nfirst(nP);
  for (int f = 0; f < nP; f++) {
    int c = P.conj[P.sdsec[f]];
    rowlen[f] = (int32_t)N.seccount[c];
    nfirst[f] = (int32_t)N.secfirst[c];
  }
  int maxL = 1;
  for (int f = 0; f < nP; f++) maxL = std::max(maxL, rowlen[f]);
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
  // PP / NN merged two-body
  std::vector<float> ppv32(t.tb[0].val.size()), nnv32(t.tb[1].val.size());
  for (size_t i = 0; i < ppv32.size(); i++) ppv32[i] = (float)t.tb[0].val[i];
  for (size_t i = 0; i < nnv32.size(); i++) nnv32[i] = (float)t.tb[1].val[i];
  std::vector<int32_t> nn_srcl(t.tb[1].src.size());
  for (int n = 0; n < nN; n++)
    for (int64_t e = t.tb[1].rowptr[n]; e < t.tb[1].rowptr[n + 1]; e++) {
      int32_t src = t.tb[1].src[e];
      if (N.sdsec[src] != N.sdsec[n]) throw std::runtime_error("NN jump leaves its sector");
      nn_srcl[e] = (int32_t)(src - N.secfirst[N.sdsec[n]]);
    }
  std::vector<int32_t> pp_src(t.tb[0].src.begin(), t.tb[0].src.end());
  for (int f = 0; f < nP; f++)
    for (int64_t e = t.tb[0].rowptr[f]; e < t.tb[0].rowptr[f + 1]; e++)
      if (P.sdsec[t.tb[0].src[e]] != P.sdsec[f]) throw std::runtime_error("PP jump leaves its sector");
  // proton one-body with class pointers
  const bs::OneBody& O0 = t.ob[0];
  std::vector<int32_t> pcls((size_t)nP * (nd + 1));
  std::vector<int32_t> p1_src(O0.src.size());
  std::vector<int16_t> p1_meta(O0.src.size());
  if (O0.src.size() >= (size_t)INT32_MAX) throw std::runtime_error("too many proton one-body jumps");
  for (int f = 0; f < nP; f++) {
    int64_t e = O0.rowptr[f];
    for (int k = 0; k <= nd; k++) {
      while (e < O0.rowptr[f + 1] && t.class_of_op[O0.op[e]] < k) e++;
      pcls[(size_t)f * (nd + 1) + k] = (int32_t)e;
    }
    pcls[(size_t)f * (nd + 1) + nd] = (int32_t)O0.rowptr[f + 1];
    for (int64_t q = O0.rowptr[f]; q < O0.rowptr[f + 1]; q++) {
      p1_src[q] = O0.src[q];
      int pl = t.local_of_op[O0.op[q]];
      if (pl > 255) throw std::runtime_error("operator class too large");
      p1_meta[q] = (int16_t)(pl | (O0.phase[q] < 0 ? 256 : 0));
    }
  }
  // neutron one-body packed entries, source index local to the source sector
  const bs::OneBody& O1 = t.ob[1];
  std::vector<int32_t> ncls((size_t)nN * (nd + 1));
  std::vector<uint32_t> n1_pk(O1.src.size());
  if (O1.src.size() >= (size_t)INT32_MAX) throw std::runtime_error("too many neutron one-body jumps");
  for (int n = 0; n < nN; n++) {
    int64_t e = O1.rowptr[n];
    for (int k = 0; k <= nd; k++) {
      while (e < O1.rowptr[n + 1] && t.class_of_op[O1.op[e]] < k) e++;
      ncls[(size_t)n * (nd + 1) + k] = (int32_t)e;
    }
    ncls[(size_t)n * (nd + 1) + nd] = (int32_t)O1.rowptr[n + 1];
    for (int64_t q = O1.rowptr[n]; q < O1.rowptr[n + 1]; q++) {
      int32_t src = O1.src[q];
      int64_t sl = src - N.secfirst[N.sdsec[src]];
      int nl = t.local_of_op[O1.op[q]];
      if (sl > 0xffff || nl > 255) throw std::runtime_error("neutron row too long for packed jumps");
      n1_pk[q] = (uint32_t)sl | ((uint32_t)nl << 16) | (O1.phase[q] < 0 ? (1u << 24) : 0u);
    }
  }
  std::vector<int32_t> woff(nd), wstride(nd), negk(nd);
  for (int k = 0; k < nd; k++) {
    int kn = (int)(std::lower_bound(t.delta_of_class.begin(), t.delta_of_class.end(), -t.delta_of_class[k]) -
                   t.delta_of_class.begin());
    negk[k] = kn;
    woff[k] = t.class_offset[k];
    wstride[k] = t.class_size[kn];
  }
  // Host-side validation of everything the kernel dereferences, O(entries):
  // rows inside the vector; two-body sources in the target's sector; every
  // one-body source in the sector that Delta(2m) of its class implies, so
  // that a proton jump of class k and the neutron jumps of class -k always
  // meet on the source row's columns.
  auto sec_of_jz = [](const bs::Species& S, int jz) {
    for (int q = 0; q < S.nsec; q++)
      if (S.secjz[q] == jz) return q;
    return -1;
  };
  for (int f = 0; f < nP; f++) {
    if (pstart[f] < 0 || pstart[f] + rowlen[f] > t.dim) throw std::runtime_error("row outside the vector");
    for (int k = 0; k < nd; k++)
      for (int q = pcls[(size_t)f * (nd + 1) + k]; q < pcls[(size_t)f * (nd + 1) + k + 1]; q++) {
        int i = p1_src[q];
        if (i < 0 || i >= nP) throw std::runtime_error("proton jump source out of range");
        if (P.secjz[P.sdsec[i]] != P.secjz[P.sdsec[f]] - t.delta_of_class[k])
          throw std::runtime_error("proton jump source in the wrong sector");
        if ((p1_meta[q] & 0xff) >= t.class_size[k]) throw std::runtime_error("bad pp index");
        if (N.secjz[P.conj[P.sdsec[i]]] != N.secjz[P.conj[P.sdsec[f]]] + t.delta_of_class[k])
          throw std::runtime_error("conjugate sectors inconsistent with Delta");
      }
  }
  for (int n = 0; n < nN; n++)
    for (int k = 0; k < nd; k++)
      for (int q = ncls[(size_t)n * (nd + 1) + k]; q < ncls[(size_t)n * (nd + 1) + k + 1]; q++) {
        int src = O1.src[q];
        int want = sec_of_jz(N, N.secjz[N.sdsec[n]] - t.delta_of_class[k]);
        if (want < 0 || N.sdsec[src] != want) throw std::runtime_error("neutron jump source in the wrong sector");
        if ((int)((n1_pk[q] >> 16) & 0xff) >= t.class_size[k]) throw std::runtime_error("bad nn index");
      }
  (void)ns;
  int cpt = 1;
  while (cpt * kBlock < maxL && cpt < 16) cpt *= 2;
  s->cpt = cpt;
  s->nchunk = (maxL + cpt * kBlock - 1) / (cpt * kBlock);

  s->pstart.upload(pstart);
  s->rowlen.upload(rowlen);
  s->nfirst.upload(nfirst);
  s->pspe_sd.upload(pspe_sd);
  s->nspe_sd.upload(nspe_sd);
  s->pp_ptr.upload(t.tb[0].rowptr);
  s->pp_src.upload(pp_src);
  s->pp_val32.upload(ppv32);
  s->pp_val64.upload(t.tb[0].val);
  s->nn_ptr.upload(t.tb[1].rowptr);
  s->nn_srcl.upload(nn_srcl);
  s->nn_val32.upload(nnv32);
  s->nn_val64.upload(t.tb[1].val);
  s->pcls.upload(pcls);
  s->p1_src.upload(p1_src);
  s->p1_meta.upload(p1_meta);
  s->ncls.upload(ncls);
  s->n1_pk.upload(n1_pk);
  s->W.upload(t.wpn);
  s->woff.upload(woff);
  s->wstride.upload(wstride);
  s->negk.upload(negk);

  DevTables& d = s->dt;
  d.dim = t.dim;
  d.nP = nP;
  d.nN = nN;
  d.ndelta = nd;
  d.call_spe = t.call_spe ? 1 : 0;
  d.pstart = s->pstart.p;
  d.rowlen = s->rowlen.p;
  d.nfirst = s->nfirst.p;
  d.pspe_sd = s->pspe_sd.p;
  d.nspe_sd = s->nspe_sd.p;
  d.pp_ptr = s->pp_ptr.p;
  d.pp_src = s->pp_src.p;
  d.pp_val32 = s->pp_val32.p;
  d.pp_val64 = s->pp_val64.p;
  d.nn_ptr = s->nn_ptr.p;
  d.nn_srcl = s->nn_srcl.p;
  d.nn_val32 = s->nn_val32.p;
  d.nn_val64 = s->nn_val64.p;
  d.pcls = s->pcls.p;
  d.p1_src = s->p1_src.p;
  d.p1_meta = s->p1_meta.p;
  d.ncls = s->ncls.p;
  d.n1_pk = s->n1_pk.p;
  d.W = s->W.p;
  d.woff = s->woff.p;
  d.wstride = s->wstride.p;
  d.negk = s->negk.p;
  s->uploaded = true;
}

template <typename T>
void launch_apply(bs_system* s, const T* x, T* y, int accumulate, hipStream_t st) {
  dim3 grid(s->dt.nP, s->nchunk);
  switch (s->cpt) {
    case 1: k_apply_rows<T, 1><<<grid, kBlock, 0, st>>>(s->dt, x, y, accumulate); break;
    case 2: k_apply_rows<T, 2><<<grid, kBlock, 0, st>>>(s->dt, x, y, accumulate); break;
    case 4: k_apply_rows<T, 4><<<grid, kBlock, 0, st>>>(s->dt, x, y, accumulate); break;
    case 8: k_apply_rows<T, 8><<<grid, kBlock, 0, st>>>(s->dt, x, y, accumulate); break;
    default: k_apply_rows<T, 16><<<grid, kBlock, 0, st>>>(s->dt, x, y, accumulate); break;
  }
  HIP_OK(hipGetLastError());
}

// Lanczos workspace and loop
template <typename T>
int run_lanczos(bs_system* s, const bs_lanczos_params* p, double* alpha_out, double* beta_out,
                double* evals, void* evecs, double* resid, int* iters) {
  const int64_t D = s->t.dim;
  const hipStream_t st = nullptr;
  const int nkeep = p->nkeep;
  int niter = std::min<int64_t>(p->niter, D);
  if (nkeep < 1 || nkeep > 16) return set_err("nkeep must be in 1..16");
  if (niter < nkeep) return set_err("niter must be >= nkeep");
  const int thick = p->thick;
  if (thick < 0 || (thick > 0 && (thick < nkeep || thick >= niter))) return set_err("thick must be 0 or in [nkeep, niter)");
  const int mtot = niter + 1;
  const int nb = grid_for(D) < 1024 ? grid_for(D) : 1024;
  DevBuf<T> V, w;
  DevBuf<double> part, dots, ab;
  V.alloc((size_t)mtot * D);
  w.alloc(D);
  part.alloc((size_t)nb * mtot);
  dots.alloc(mtot);
  ab.alloc(2 * (size_t)mtot + 2);  // alpha[0..], beta[mtot..]
  double* d_alpha = ab.p;
  double* d_beta = ab.p + mtot;
  const int g = grid_for(D);
  k_fill<T><<<g, kBlock, 0, st>>>(V.p, (T)(1.0 / std::sqrt((double)D)), D);
  HIP_OK(hipGetLastError());

  std::vector<double> alpha_h, beta_h;  // current arrowhead/tridiagonal data
  std::vector<double> e_prev;
  int total_mv = 0;
  int kstart = 0;           // number of locked (thick-restart) vectors at the front
  std::vector<double> arrow;  // coupling of the locked vectors to v_kstart
  int restarts = 0;
  std::vector<double> e, Y;
  while (true) {
    for (int i = kstart; i < niter; i++) {
      T* vi = V.p + (int64_t)i * D;
      launch_apply<T>(s, vi, w.p, 0, st);
      total_mv++;
      // alpha = <v_i, w>; w -= alpha v_i
      k_multidot<T><<<nb, kBlock, 0, st>>>(vi, D, 1, w.p, D, part.p);
      k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, d_alpha + i, 0);
      k_sub_scaled<T><<<g, kBlock, 0, st>>>(w.p, vi, d_alpha + i, D);
      // classical Gram-Schmidt against v_0 .. v_{i-1}
      if (i > 0) {
        k_multidot<T><<<nb, kBlock, 0, st>>>(V.p, D, i, w.p, D, part.p);
        k_reduce<<<i, kBlock, 0, st>>>(part.p, nb, i, dots.p, 0);
        k_multiaxpy<T><<<g, kBlock, 0, st>>>(w.p, V.p, D, i, dots.p, D);
      }
      k_multidot<T><<<nb, kBlock, 0, st>>>(w.p, D, 1, w.p, D, part.p);
      k_reduce<<<1, kBlock, 0, st>>>(part.p, nb, 1, d_beta + i, 1);
      k_normalize<T><<<g, kBlock, 0, st>>>(V.p + (int64_t)(i + 1) * D, w.p, d_beta + i, D);
      HIP_OK(hipGetLastError());
    }
    alpha_h.assign(mtot, 0.0);
    beta_h.assign(mtot, 0.0);
    HIP_OK(hipMemcpy(alpha_h.data(), d_alpha, sizeof(double) * niter, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(beta_h.data(), d_beta, sizeof(double) * niter, hipMemcpyDeviceToHost));
    // projected matrix: locked Ritz values on the diagonal with the arrow
    // couplings in column kstart (src/blanczoslib.f90:143-222), then the
    // tridiagonal part
    const int m = niter;
    std::vector<double> A((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) A[(size_t)i * m + i] = alpha_h[i];
    for (int i = kstart; i + 1 < m; i++) {
      A[(size_t)i * m + i + 1] = beta_h[i];
      A[(size_t)(i + 1) * m + i] = beta_h[i];
    }
    for (int i = 0; i < kstart; i++) {
      A[(size_t)i * m + i] = alpha_h[i];
      A[(size_t)i * m + kstart] = arrow[i];
      A[(size_t)kstart * m + i] = arrow[i];
    }
    jacobi_eigen(m, A, e, Y);
    bool done = (thick == 0) || restarts >= p->max_restarts;
    if (!done && p->tol > 0.0 && !e_prev.empty()) {
      double dmax = 0.0;
      for (int k = 0; k < nkeep; k++) dmax = std::max(dmax, std::fabs(e[k] - e_prev[k]));
      if (dmax < p->tol) done = true;
    }
    e_prev.assign(e.begin(), e.begin() + nkeep);
    if (done) break;
    // thick restart (src/blanczoslib.f90:1656-1799): keep `thick` Ritz
    // vectors, V[0:thick] = V[0:m] Y[:, 0:thick]; couplings beta_m * Y[m-1, k]
    {
      DevBuf<double> dY;
      std::vector<double> Yk((size_t)m * thick);
      for (int j = 0; j < m; j++)
        for (int k = 0; k < thick; k++) Yk[(size_t)j * thick + k] = Y[(size_t)j * m + k];
      dY.upload(Yk);
      DevBuf<T> tmp;
      tmp.alloc((size_t)thick * D);
      for (int k0 = 0; k0 < thick; k0 += 16) {
        int nk = std::min(16, thick - k0);
        k_combine<T, T><<<g, kBlock, 0, st>>>(tmp.p + (int64_t)k0 * D, D, V.p, D, m, dY.p + k0, thick, nk, D);
      }
      HIP_OK(hipMemcpyAsync(V.p, tmp.p, sizeof(T) * (size_t)thick * D, hipMemcpyDeviceToDevice, st));
      // the last Krylov vector becomes v_thick
      HIP_OK(hipMemcpyAsync(V.p + (int64_t)thick * D, V.p + (int64_t)m * D, sizeof(T) * D,
                            hipMemcpyDeviceToDevice, st));
      arrow.assign(thick, 0.0);
      std::vector<double> anew(thick);
      for (int k = 0; k < thick; k++) {
        arrow[k] = beta_h[m - 1] * Y[(size_t)(m - 1) * m + k];
        anew[k] = e[k];
      }
      HIP_OK(hipMemcpy(d_alpha, anew.data(), sizeof(double) * thick, hipMemcpyHostToDevice));
      HIP_OK(hipStreamSynchronize(st));
    }
    kstart = thick;
    restarts++;
  }
  for (int k = 0; k < nkeep; k++) evals[k] = e[k];
  if (alpha_out) for (int i = 0; i < niter; i++) alpha_out[i] = alpha_h[i];
  if (beta_out) for (int i = 0; i < niter; i++) beta_out[i] = beta_h[i];
  if (evecs || resid || p->polish) {
    // Ritz vectors in FP64: x_k = sum_j Y[j,k] v_j
    const int m = niter;
    DevBuf<double> dY, X64, HX, tmpd;
    std::vector<double> Yk((size_t)m * nkeep);
    for (int j = 0; j < m; j++)
      for (int k = 0; k < nkeep; k++) Yk[(size_t)j * nkeep + k] = Y[(size_t)j * m + k];
    dY.upload(Yk);
    X64.alloc((size_t)nkeep * D);
    HX.alloc(D);
    k_combine<T, double><<<g, kBlock, 0, st>>>(X64.p, D, V.p, D, m, dY.p, nkeep, nkeep, D);
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
  return 0;
}

bs_system* checked(bs_system* s) {
  if (!s) throw std::runtime_error("null system");
  return s;
}

}  // namespace

extern "C" {

const char* bs_last_error(void) { return g_err.c_str(); }

int bs_init_host(const char* sps, const char* intf, int Z, int N, int jz2, float spscale, float a0, float a,
                 float x, bs_system** out) {
  try {
    if (!out) return set_err("null output pointer");
    std::unique_ptr<bs_system> s(new bs_system());
    s->t = bs::build_tables(sps, intf, Z, N, jz2, spscale, a0, a, x, true);
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

}  // extern "C"
