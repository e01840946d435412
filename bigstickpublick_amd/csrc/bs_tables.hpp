// Host-side model tables for the M-scheme Hamiltonian (built once, uploaded to
// HBM).  Native C++; no GPU calls.
//
// Reference correspondence (file:line in the reference tree):
//   .sps / .int reading      src/binput.f90:43, src/binput_tbme.f90:898-1990
//   s.p. state order         src/bspstates.f90:108-330
//   basis order              src/bbasis_haikus.f90:327-411, src/bbasis_blocks.f90,
//                            src/bbasis_limits.f90:64-291, src/bbasislib.f90:189-677
//   SPE subsumption          src/bsppot.f90:20-240
//   FP32 m-scheme MEs        src/btbme.f90:1209-1558, src/btbmepn.f90:177-594,
//                            src/libra.f:320-521 (CLEB)
// The jump tables are this library's own design (output-stationary CSR by
// target Slater determinant) and replace src/bjumplib_*.f90 / bsortjumpslib.f90.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace bs {

constexpr int kMaxOrb = 16;
constexpr int kMaxSps = 64;

struct Orbit { int n, l, j2, w; };

struct ModelInput {
  std::vector<Orbit> orbits;
  int Z = 0, N = 0, jz2 = 0;
  float spscale = 1.0f, vscale = 1.0f;
};

// One species' Slater determinants, in the reference's order.
struct Species {
  int np = 0;
  int nsec = 0;
  std::vector<int> secjz;            // 2*Jz of each sector
  std::vector<int64_t> secfirst;     // first SD index of each sector
  std::vector<int64_t> seccount;     // SDs in each sector
  std::vector<int> conj;             // conjugate sector (other species)
  std::vector<uint64_t> mask;        // occupation bits over combined labels
  std::vector<int> sdsec;            // sector of each SD
  std::vector<int64_t> start;        // pstart / nstart (0-based)
};

// Output-stationary CSR of one-body operators a+_a a_c on one species:
// for each target SD f, the entries whose source SD i satisfies
// a+_a a_c |i> = phase |f>, ordered by source sector, then source index.
struct OneBody {
  std::vector<int64_t> rowptr;        // nSD+1
  std::vector<int32_t> src;           // source SD index (global within species)
  std::vector<int16_t> op;            // a*nsps + c
  std::vector<int8_t> phase;
};

// Output-stationary CSR of two-body operators a+_i a+_j a_l a_k on one species,
// merged per (target, source) SD pair; `val` is the summed matrix element
// times phase (FP64 sum of the FP32 elements).
struct TwoBody {
  std::vector<int64_t> rowptr;
  std::vector<int32_t> src;
  std::vector<double> val;
};

// Unmerged two-body operator list (for densities / gradients): every
// (target, source, pair-pair element index, phase).
struct TwoBodyOps {
  std::vector<int64_t> rowptr;
  std::vector<int32_t> src;
  std::vector<int32_t> me;            // cpair*npair + dpair (dense pair table index)
  std::vector<int8_t> phase;
};

// One TBME line of the .int file, in file order (a b c d J T V).
struct TbmeLine { int a, b, c, d, J, T; float v; };

// d(uncoupled element)/d(parameter) for one parameter: sparse lists into the
// dense pp / nn / pn element tables.
struct ParamColumn {
  std::vector<int32_t> idx[3];
  std::vector<float> coef[3];
};

struct Tables {
  ModelInput in;
  std::vector<float> spe_raw;          // SPEs as read (before spescale)
  std::vector<TbmeLine> lines;         // TBME lines as read
  // single-particle states, combined label order (left m<0 first, then m>=0)
  int nsps = 0, nL = 0, nR = 0;
  std::vector<int> sorb, sj2, sm2;
  std::vector<int> ssorb, ssm2;        // the same states before the left/right split (sorted on |m|)
  Species sp[2];
  int64_t dim = 0;
  bool call_spe = false;
  std::vector<float> pspe, nspe;       // per orbit (after scaling)
  // FP32 uncoupled MEs: pp/nn over XX pair index c1*(c1-1)/2+c2 (c1>c2),
  // pn over (cp*nsps + cn); dense npair x npair.
  int npairxx = 0, npairpn = 0;
  std::vector<float> vpp, vnn, vpn;
  // coupled FP32 values after reading and subsumption (for inspection)
  std::vector<float> cpp, cnn, cpn;
  // jump tables
  OneBody ob[2];
  TwoBody tb[2];
  // operator-op regrouping for PN: class of Delta(2m) for op a*nsps+c
  int ndelta = 0;                      // number of Delta classes
  std::vector<int> delta_of_class;     // 2*(m_a - m_c)
  std::vector<int> class_of_op;        // nsps*nsps -> class
  std::vector<int> local_of_op;        // nsps*nsps -> index within class
  std::vector<int> class_size;
  std::vector<int> class_offset;       // offset of class W block
  std::vector<float> wpn;              // per class: [pp_local][nn_local] with
                                       // nn class = class of -Delta
  int max_class = 0;
  // counts
  int64_t nops_pp = 0, nops_nn = 0, nops_pn = 0, nops_spe = 0;
  // the reference's own operation count (src/bigstick_main.f90:1150-1192):
  // two-body operator applications with diagonal pair terms counted twice
  int64_t ref_pp = 0, ref_nn = 0;
  // keep every structurally present two-body (target, source) pair, whatever
  // its summed value (needed when several interactions share the tables)
  bool keep_structural = false;
  std::vector<int64_t> tb_unmerged[2];  // per SD: unmerged two-body ops into it (+ diagonal again)
};

// Errors are reported by throwing std::runtime_error with a message.
void read_sps(const std::string& path, ModelInput& in);
// The builders take the orbit list of an .sps file (read_sps) or of a .wfn
// header (read_wfn_header).
Tables build_tables(const std::vector<Orbit>& orbits, const std::string& intfile, int Z, int N, int jz2,
                    float spscale, float a0, float a, float x, bool with_jumps = true);
// Element tables of another interaction on the same orbits, particle numbers
// and scaling (no basis, no jumps), with the pn elements in the Delta-class
// block layout of `base`.
struct MeTables {
  std::vector<float> vpp, vnn, vpn, wpn, pspe, nspe;
  bool call_spe = false;
  bool same_lines = false;  // same TBME list (a b c d J T, file order) as base
};
MeTables build_me_tables(const Tables& base, const std::string& intfile);
// Merged two-body values of element table v (dense pair-pair) in the entry
// order of base.tb[it]; base must have been built with keep_structural.
std::vector<double> merged_values(const Tables& base, int it, const std::vector<float>& v);
Tables build_tables_structural(const std::vector<Orbit>& orbits, const std::string& intfile, int Z, int N,
                               int jz2, float spscale, float a0, float a, float x);

// Jacobian of the FP32 uncoupled elements with respect to the .int parameters
// (SPEs first, then TBME lines in file order): built from unit interactions
// through the same reading, subsumption and uncoupling steps.
std::vector<ParamColumn> param_jacobian(const Tables& t);

// Same as build_tables with the interaction given as the text of an iso .int
// file (scaling 1 1 1 1).
Tables build_tables_text(const std::vector<Orbit>& orbits, const std::string& inttext, int Z, int N, int jz2,
                         bool with_jumps = true);
// Iso interaction text of J^2 (which = 0) or T^2 (which = 1) on these orbits.
std::string observable_int_text(const std::vector<Orbit>& orbits, int which);
void build_twobody_ops(const Tables& t, int it, TwoBodyOps& out);
// Header of the reference's binary .wfn file for this basis, up to and
// including the basis dimension (src/bwfnlib.f90:373-541); the vectors start
// 4 bytes after it (the count of vectors sits in between).
std::string wfn_header(const Tables& t);
// Basis recorded in a .wfn header (the fields read by read_wfn_header,
// src/bwfnlib.f90:678-904): orbits of species 1, Z, N, 2Jz, dimension, the
// header bytes up to the dimension, the vector offset and the vector count.
// Vectors are stored in the writer's Lanczos precision (4 or 8 bytes per
// amplitude, src/bwfnlib.f90:963-968), found from the file size.
struct WfnInfo {
  int Z = 0, N = 0, jz2 = 0, nkeep = 0, elem = 4;
  std::vector<Orbit> orbits;
  int64_t dim = 0, offset = 0;
  std::string header;
};
WfnInfo read_wfn_info(const std::string& path);
float cleb(int j1, int m1, int j2, int m2, int j3, int m3);

// A one-body operator read from an .opme file (readin1bodyop,
// src/bdenslib1.f90:1823-2071) and decoupled to its M = 0 m-scheme
// coefficients (decouple1bodyop, :2444-2597): p / n hold, over combined
// labels, [a * nsps + b] = the FP32 coefficient of a+_a a_b.
struct OneBodyOp {
  int jop = 0, top = 0;
  bool pn = false;
  std::vector<float> p, n;
};
OneBodyOp read_opme(const std::string& path, const Tables& t);
// Tables that apply the operator with the engine's PP and NN sections: the
// one-body jumps of `base` with Delta m = 0, merged per (target, source) SD
// pair; no PN part, no SPEs.
Tables onebody_tables(const Tables& base, const OneBodyOp& op);

}  // namespace bs
