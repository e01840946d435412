/*
 * bigstick_amd.h -- C ABI of the MI355X M-scheme Hamiltonian engine.
 *
 * Plain C types only; every function returns 0 on success or a negative error
 * code (message in bs_last_error()).  Device pointers are HIP device memory on
 * the current device; `stream` is a hipStream_t (NULL = default stream).
 * bs_apply_h and bs_apply_block are asynchronous on the caller's stream; the
 * solvers and the calls that return host results (bs_lanczos, bs_strength,
 * bs_block_lanczos, bs_density, bs_gradient, bs_observable, ...) are
 * synchronous: they start after all earlier device work on every stream, so
 * inputs written on a caller's non-blocking stream are complete, and their
 * device outputs are complete when they return.
 * Calls on one system are not reentrant (the reference's apply routine is not
 * either, src/bapplyhlibOMP.f90:934).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   bs_init      setup chain of src/bigstick_main.f90:305-676: get_orbit_info
 *                (src/binput.f90:43), define_system (:649), basismaster
 *                (src/bbasislib.f90:92), master_readin_tbmes
 *                (src/binput_tbme.f90:58), subsume_sp_pot_into_2body
 *                (src/bsppot.f90:39), jumpmaster (src/bjumplib_master.f90:342),
 *                uncoupleXXtbme / count_uncouplepn (src/btbme.f90:1209,
 *                src/btbmepn.f90:177)
 *   bs_apply_h   applyHbundled_omp('n') (src/bapplyhlibOMP.f90:882) and
 *                applyHbundled_g (src/bapplyhlib1.f90:56): vec2 (+)= H vec1
 *   bs_lanczos   lanczos_p (src/blanczosmain.f90:1413-1793) with the br_*
 *                reorthogonalization (src/breorthog.f90:683-891),
 *                find_lanczos_eigenvalues (src/blanczoslib.f90:143-376) and
 *                thick_restart_sub_p (src/blanczoslib.f90:1656-1799)
 *   bs_finalize  end of run deallocation
 */
#ifndef BIGSTICK_AMD_H
#define BIGSTICK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bs_system bs_system;

#define BS_F32 4
#define BS_F64 8

/* Build the basis, jump tables and FP32 matrix elements from an iso .sps and
 * an OXBASH iso .int file for Z protons and N neutrons at 2*Jz = jz2, with the
 * reference's scaling line (spescale, A0, A, x): TBMEs scaled by (A0/A)^x, or
 * by A0 when A or x is zero (src/binput_tbme.f90:1024-1031).  Host only; no
 * device memory is touched. */
int bs_init_host(const char *sps_path, const char *int_path, int Z, int N, int jz2,
                 float spscale, float a0, float a, float x, bs_system **out);

/* Copy the tables to the current HIP device (once).  Required before any
 * device call. */
int bs_upload(bs_system *sys);

/* bs_init_host + bs_upload. */
int bs_init(const char *sps_path, const char *int_path, int Z, int N, int jz2,
            float spscale, float a0, float a, float x, bs_system **out);

int bs_finalize(bs_system *sys);
const char *bs_last_error(void);

/* Parity (+1 / -1) of every basis state (the spaces this engine builds have
 * one parity).  Replaces find_parity (src/boutputlib.f90:1673-1873), which
 * reads it from the vector's components. */
int bs_parity(const bs_system *sys, int *parity);

/* ---- host-side inspection (no device needed) ---- */
int bs_dim(const bs_system *sys, int64_t *dim);
/* it = 0 protons, 1 neutrons */
int bs_num_sd(const bs_system *sys, int it, int64_t *nsd);
/* occupation masks over the reference's single-particle labels (left m<0
 * states first, then m>=0; bit k = label k+1) and the 0-based offsets
 * pstart/nstart with state = pstart[p] + nstart[n] (src/bbasislib.f90:539-677). */
int bs_basis(const bs_system *sys, int it, uint64_t *mask, int64_t *start);
int bs_num_sps(const bs_system *sys, int *nsps);
/* single-particle labels: orbit index (0-based, .sps order), 2j, 2m */
int bs_sps(const bs_system *sys, int *orb, int *j2, int *m2);
/* FP32 uncoupled matrix elements, dense npair x npair row-major:
 * kind 0 pp, 1 nn (pair index c1*(c1-1)/2+c2, c1>c2 labels, 0-based),
 * kind 2 pn (pair index cp*nsps+cn).  Pass table=NULL to query npair. */
int bs_me_table(const bs_system *sys, int kind, int *npair, float *table);
/* Work counts per H x.  out[0..3] = SPE, PP, NN, PN updates done by this
 * engine (two-body terms merged per target/source pair); out[4..7] = the
 * same terms counted the reference's way, whose sum is the "total #
 * operations" printed at src/bigstick_main.f90:1192 (two-body operator
 * applications, diagonal pair terms twice). */
int bs_op_counts(const bs_system *sys, int64_t out[8]);

/* One- (order 1) or two-body (order 2) jumps of species it (0 protons, 1
 * neutrons) as this engine builds them, one entry per operator application
 * on a Slater determinant: isd -> fsd (0-based SD indices of the species, in
 * the reference's SD order), op (order 1: a*nsps + c for a+_a a_c; order 2:
 * the dense pair-pair index cpair*npair + dpair of bs_me_table for
 * a+_i a+_j a_l a_k) and the fermion phase.  Both directions are listed; the
 * reference stores one triangle and applies the rest as transposes: proton
 * one-body jumps with Jz(isd) >= Jz(fsd), neutron one-body jumps with
 * Jz(isd) <= Jz(fsd), two-body jumps with isd >= fsd.  Pass NULL arrays to get
 * the count in *count;
 * otherwise *count is the arrays' capacity on entry and the count on exit.
 * Replaces the jump lists of jumpmaster / weld (src/bjumplib_master.f90:
 * 342-431, src/bjumplib_weld.f90:32-818) as dumped by write_out_jumps
 * (src/boutputlib.f90:1258-1351). */
int bs_jumps(const bs_system *sys, int it, int order, int64_t *count, int32_t *isd, int32_t *fsd, int32_t *op,
             int8_t *phase);

/* ---- device ---- */
/* y = H x (accumulate = 0) or y += H x (accumulate = 1).  x, y: device
 * arrays of dim elements, dtype BS_F32 (FP32 vectors, FP32 accumulation) or
 * BS_F64 (FP64 vectors and accumulation, FP32 matrix elements promoted).
 * x and y must not overlap. */
int bs_apply_h(bs_system *sys, const void *x, void *y, int dtype, int accumulate, void *stream);

/* Kernel variants of the H x kernel: a tile of `rows` proton rows per
 * workgroup, `staged` source rows held in LDS at a time, the
 * workgroup-to-tile order (xcd_map 0: sector order; 1: runs of consecutive
 * tiles on the workgroups that the dispatcher places on one XCD, so each
 * XCD's L2 holds their shared source rows), and prefetch (1: the next batch
 * of source rows is loaded into registers while the current one is applied).
 * The last variant (staged = 0) is the row kernel: one thread per output
 * state, every gather from global memory, no LDS; it is the only one on
 * spaces whose neutron sectors exceed 4096 SDs or whose Delta-m operator
 * classes exceed 63 operators (up to 65535 SDs and 254 operators).
 * bs_autotune times every variant that fits in LDS on this system and keeps
 * the fastest for that dtype; bs_set_variant forces one (-1 = default
 * choice). */
int bs_num_variants(int *n);
int bs_variant_info(int variant, int *rows, int *staged, int *xcd_map, int *prefetch);
int bs_set_variant(bs_system *sys, int dtype, int variant);
int bs_get_variant(const bs_system *sys, int dtype, int *variant);
int bs_autotune(bs_system *sys, int dtype, int *variant);
/* Per-variant kernel times (ms) of the last bs_autotune for dtype; 0 for
 * variants that were not timed.  ms has bs_num_variants entries. */
int bs_autotune_times(const bs_system *sys, int dtype, float *ms);

/* Host-pointer form of bs_apply_h for hosts that keep vectors in host memory
 * (e.g. the reference's Fortran vec1/vec2): copies x (and y when
 * accumulating) to the device, applies H, copies y back; synchronous.  The
 * PCIe copies dominate its cost; device-resident callers use bs_apply_h. */
int bs_apply_h_host(bs_system *sys, const void *x, void *y, int dtype, int accumulate);

/* <psi|O|psi>/<psi|psi> for O = J^2 (which = 0) or T^2 (which = 1) on the
 * same basis, psi a device vector of dim elements (the reference's
 * applyobsbundled, src/bapplyobslib1.f90:110-197).  The operator tables are
 * built on first use as the iso interaction J^2 = sum_i j_i^2 +
 * sum_{i<j} 2 j_i.j_j (and likewise T^2) and applied with the same kernel. */
int bs_observable(bs_system *sys, int which, const void *psi, int dtype, double *value);

/* Parameters of the interaction: nspe single-particle energies (one per
 * orbit) followed by ntbme two-body matrix elements, in .int file order. */
int bs_num_params(const bs_system *sys, int *nspe, int *ntbme);

/* Adjoint pass over the same jump tables as bs_apply_h (the analogue of
 * applyh{PP,NN,PN}bundled_den, src/bdenslib4.f90:129-725):
 * rho[me] = <psi| O_me |psi> for every uncoupled operator, with
 * H = sum_me v_me O_me and v_me the dense tables of bs_me_table.  psi is a
 * device vector (dtype, dim); rho_* are host FP64 arrays of npair*npair
 * (any may be NULL); norm2 (may be NULL) receives <psi|psi>.  Not normalized. */
int bs_density(bs_system *sys, const void *psi, int dtype, double *rho_pp, double *rho_nn,
               double *rho_pn, double *norm2);

/* Hellmann-Feynman gradient dE/dtheta = <psi|dH/dtheta|psi>/<psi|psi> for all
 * interaction parameters (grad has nspe + ntbme entries, order of
 * bs_num_params), from bs_density and the Jacobian of the FP32 elements with
 * respect to the .int values (built through the same reading, subsumption and
 * uncoupling steps; the scaling line is included). */
int bs_gradient(bs_system *sys, const void *psi, int dtype, double *grad);

typedef struct {
  int nkeep;        /* eigenpairs wanted (<= niter; <= 16 for the batched solve) */
  int niter;        /* Lanczos iterations (plain) or restart length (thick) */
  int thick;        /* 0: plain Lanczos; >0: vectors kept at each thick restart, <= niter - 2.
                       The cycle is the reference's (lanczos_p, src/blanczosmain.f90:1690-1752;
                       thick_restart_sub_p, src/blanczoslib.f90:1656-1799): steps to iter =
                       niter - 1, restart from the projection on the first niter - 2 vectors
                       with v_{niter-1} as the next start vector, iter resuming at thick */
  int max_restarts; /* thick restart: a restart due after max_restarts + 1 of them ends the run */
  int dtype;        /* BS_F32: FP32 vectors/matvec, FP64 coefficients; BS_F64: all FP64 */
  double tol;       /* convergence tolerance (0: run all iterations); the test is conv_test's */
  int polish;       /* 1: FP64 Rayleigh quotient / residual polish of the Ritz vectors */
  int conv_test;    /* the reference's energy tests (convergence_test, src/blanczoslib.f90:2660-2800)
                       over ncheck = nkeep + ncheck_ex Ritz values: 0 = their mean change
                       sum |dE| / sqrt(ncheck + 1) (times 1 + restarts in a thick run), 1 = their
                       largest change.  A plain run checks every 10th step (sqrt(nkeep) if larger)
                       beyond ncheck, as lanczos_p does (src/blanczosmain.f90:1690-1724); a thick
                       run takes the energies at every such step of a cycle beyond max(ncheck,
                       nkeep) and at iter = niter - 1, and compares at every such step beyond
                       ncheck with the previous iteration's energies (a restart's own Ritz
                       values after a restart).  2, 3 = the reference's Ritz-vector tests
                       (altconverge, src/blanczoslib.f90:2741-2756): vdot_j of the first ncheck
                       Ritz vectors of the projected matrix with the previous solve's (components
                       below iter - 1, zero beyond the earlier size), 2 = sqrt(sum_j (|vdot_j| -
                       1)^2) / ncheck, 3 = max_j ||vdot_j| - 1|; a plain run has vectors for the
                       lowest nkeep only, as the reference (the ncheck_ex extra ones count as
                       zero, so with ncheck_ex > 0 the test never passes); not with a target. */
  int ncheck_ex;    /* additional Ritz values in the test (the reference's ld uses 5) */
  int target;       /* 0: the lowest states.  1, 2: the Ritz pairs nearest etarget (option tx,
                       txchoice 1, src/blanczoslib.f90:277-327; thick > 0): every Ritz solve orders
                       them by |E - etarget|, the thick restart keeps the thick nearest, evals
                       receives the nkeep nearest in ascending energy.  1 tests convergence as the
                       reference's tx does, over no energies (its count is never set on that path,
                       src/blanczosmain.f90:1497-1499), so the run stops at the first check;
                       2 tests the nkeep kept energies */
  double etarget;   /* target energy (MeV) when target != 0 */
} bs_lanczos_params;

/* Lanczos from the reference's uniform pivot 1/sqrt(D)
 * (src/blanczoslib.f90:825-831) with full FP64-accumulated classical
 * Gram-Schmidt.  alpha/beta receive the coefficients of the last Krylov
 * sequence (up to niter entries each, may be NULL; after a thick restart the
 * kept Ritz values and their couplings come first, then the last cycle's
 * steps, as in the reference's .lcoef); evals receives nkeep Ritz
 * values; evecs (device, may be NULL) receives nkeep Ritz vectors of dim
 * elements in dtype, row-major; resid (host, may be NULL) receives the FP64
 * residual norms ||H v - E v||.  iters (may be NULL) receives the total number
 * of matvecs. */
int bs_lanczos(bs_system *sys, const bs_lanczos_params *p, double *alpha, double *beta,
               double *evals, void *evecs, double *resid, int *iters);

/* bs_lanczos with no device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Plain Lanczos only
 * (p->thick = 0, p->polish = 0; no vectors or residuals); alpha, beta, evals
 * and iters as in bs_lanczos.  The Lanczos vectors and partial sums live in a
 * fixed device array of 128 MiB per GPU (a static device symbol, held from
 * the library's load until the process ends): the run needs about
 * (niter + 2) * dim * (4 or 8) bytes, and a larger request is refused with an
 * error that states both sizes.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU race; each call initializes what it reads. */
int bs_lanczos_noalloc(bs_system *sys, const bs_lanczos_params *p, double *alpha, double *beta,
                       double *evals, int *iters);

/* Breakdown handling of the last bs_lanczos / bs_strength run on sys, as the
 * reference does it inside lanczos_p (src/blanczosmain.f90:1634-1658):
 * random_restarts counts steps whose beta fell below 1e-4 of the previous
 * beta (|H v| on a cycle's first step) and whose next vector was replaced by
 * a seeded random vector orthonormalized against the history
 * (random_restart_with_history, src/breorthog.f90:1471-1482), beta itself
 * kept; second_passes counts steps whose beta fell below 0.1 of the previous
 * one and whose new vector then overlapped a stored one by more than 1e-5
 * (br_check_orthogonal, src/breorthog.f90:977-1045) and was
 * re-orthogonalized; steps receives the number of Lanczos steps of the last
 * Krylov sequence (where a converged plain run stopped: the length of the
 * reference's .lcoef; for a thick-restart run the size of the final projected
 * matrix, the reference's "iterations"), thick_restarts the number of thick
 * restarts taken (the reference's total = steps + niter * thick_restarts).
 * Any pointer may be NULL. */
int bs_lanczos_stats(const bs_system *sys, int *random_restarts, int *second_passes, int *steps,
                     int *thick_restarts);

/* Exact diagonalization: the dense H (dim <= 65536, dim^2 doubles in HBM)
 * built 64 columns at a time with the multi-vector kernel in FP64, all
 * eigenpairs by rocSOLVER's symmetric divide-and-conquer solver; evals (nkeep
 * lowest), evecs (device, may be NULL: nkeep vectors in dtype), resid (may be
 * NULL: ||H x - E x|| of each), iters (may be NULL: dim, the matvec count).
 * Replaces exactdiag_p, option ex (src/blanczosmain.f90:2171-2471, :754-763),
 * which builds H with dim calls of applyHbundled_g and calls DSYEV. */
int bs_exact(bs_system *sys, int nkeep, int dtype, double *evals, void *evecs, double *resid, int *iters);

/* Block Lanczos: blocks of `block` (1..8) vectors, H applied to a block with
 * the multi-vector kernel, full block Gram-Schmidt against every stored
 * block (FP64 sums), each new block orthonormalized by the spectral
 * decomposition of its Gram matrix, Ritz pairs of the block tridiagonal
 * matrix.  p->niter is the basis size (rounded down to a multiple of block),
 * p->tol > 0 stops when no kept Ritz value moves by more than tol between
 * block iterations; evals, evecs, resid, iters and the FP64 polish as in
 * bs_lanczos.  The pivot block is seeded pseudo-random.  p->thick > 0 is
 * option tb (src/blanczosmain.f90:1179-1212, block_thick_restart_sub_p,
 * src/blanczoslib.f90:1929-2070): when the next block would bring the basis
 * to niter vectors, (thick / block) whole blocks of Ritz vectors are kept
 * (nkeep <= that count), the new block is coupled to them by its beta block
 * times the last block rows of their eigenvectors, and the run goes on; at
 * most max_restarts + 1 restarts; the test is then the reference's energy
 * test p->conv_test over nkeep + p->ncheck_ex values at every block iteration
 * (mean change times 1 + restarts; tb uses 0, 2 and 1e-3).
 * bs_lanczos_stats reports the final basis size and the restarts.
 * Replaces lanczos_block with option bf / bd / tb
 * (src/blanczosmain.f90:1804-2153, :822-872). */
int bs_block_lanczos(bs_system *sys, const bs_lanczos_params *p, int block, double *evals, void *evecs,
                     double *resid, int *iters);

/* Block strength function: bs_block_lanczos started from the caller's pivot
 * block instead of a random one: `pivots` holds `block` device vectors one
 * after another ([k][dim], pivot_dtype BS_F32 or BS_F64; not modified).
 * pnorm2 (may be NULL) receives |pivot i|^2; strength (may be NULL)
 * receives strength[i * nkeep + k] = |<Ritz k|pivot i>|^2 for every pivot i
 * and kept Ritz state k; evals, evecs, resid, iters as in bs_block_lanczos
 * (p->polish = 0 gives the Ritz values themselves, as the reference prints).
 * Replaces menu sb / sbs (src/bmenu.f90:612-627): read_pivot_block_choose
 * (src/blanczoslib.f90:522-707), lanczos_block from that block
 * (src/blanczosmain.f90:1945-1953) and the per-pivot Energy/Strength tables
 * (src/blanczoslib.f90:1146-1181). */
int bs_block_strength(bs_system *sys, const bs_lanczos_params *p, int block, const void *pivots, int pivot_dtype,
                      double *pnorm2, double *strength, double *evals, void *evecs, double *resid, int *iters);

/* Strength function: bs_lanczos started from `pivot` (device, dim elements
 * in p->dtype; not modified) instead of the uniform vector, with a fixed
 * number of iterations (p->thick must be 0).  pivot_norm receives ||pivot||;
 * ritz and strength (niter entries each, may be NULL) receive every Ritz
 * value E_k of the final tridiagonal matrix and its strength
 * |Y(1,k)|^2 ||pivot||^2, which sum to ||pivot||^2.  The other outputs are
 * those of bs_lanczos.  Replaces menu s: readpivot as the starting vector
 * (src/bwfnlib.f90:298-366, src/blanczosmain.f90:1560-1568) and the
 * Energy/Strength table of lanczos_output (src/blanczoslib.f90:1122-1142). */
int bs_strength(bs_system *sys, const bs_lanczos_params *p, const void *pivot, double *pivot_norm,
                double *ritz, double *strength, double *alpha, double *beta, double *evals, void *evecs,
                double *resid, int *iters);

/* y_k = O x_k (overwritten) for nvec device vectors of dim elements, O the
 * one-body operator of an .opme file (first line iso, pns or xpn; orbits
 * checked against the system's; charge-changing xpn lines skipped) in its
 * M = 0 m-scheme form.  jop, top (may be NULL) receive the operator's J and
 * T (2 for pns/xpn).  Replaces menu o: readin1bodyop + decouple1bodyop
 * (src/bdenslib1.f90:1823-2071, :2444-2597) and applicator1b's
 * master_apply1bop per state (:1645-1693, :2612-2707). */
int bs_apply_onebody(bs_system *sys, const char *opme, const void *x, void *y, int nvec, int dtype, int *jop,
                     int *top);

/* Matrix <v_f | O | v_i> of a scalar two-body operator O, given as an iso .int
 * file with its scaling line (intf = NULL: the system's own Hamiltonian),
 * between nvec vectors (device, dim elements each, consecutive) on this
 * system's basis: out[i*nvec + f].  Replaces the hme table of
 * hmatrixelements_boss (menu h, src/bexpect.f90:84-180), which applies the
 * operator with applyHbundled_g to each stored state and overlaps it with
 * the others; its diagonal on normalized states is the <H> column of
 * expectator_p (menu x, src/bexpect.f90:16-74). */
int bs_operator_matrix(bs_system *sys, const char *intf, float spscale, float a0, float a, float x,
                       const void *vecs, int nvec, int dtype, double *out);

/* Write the reference's binary .wfn file (src/bwfnlib.f90:77-134 vectors,
 * :373-541 header; stream access, little-endian): the basis header, nkeep,
 * then per state magic, index, energy, J, T(T+1) and the vector in single
 * precision.  vecs holds nkeep vectors of dim elements in dtype, on the
 * device (on_device = 1) or the host (0).  Replaces wfn_writeeigenvec +
 * write_wfn_header for the single-file, one-fragment case. */
int bs_write_wfn(const bs_system *sys, const char *path, int nkeep, const float *energy, const float *xj,
                 const float *xt2, const void *vecs, int dtype, int on_device);

/* Build and upload a system on the basis recorded in a .wfn file (orbits, Z,
 * N, 2Jz from its header) with the interaction intf and its scaling line
 * (intf = NULL: no interaction, H = 0, for operators applied to the states).
 * Fails when the basis rebuilt from those fields does not reproduce the
 * file's header byte for byte (a parity, W or ordering this engine does not
 * build).  Replaces wfn_ropen_file + read_wfn_header (src/bwfnlib.f90:1619-
 * 1669, :678-904), the basis set-up of menu options x and h
 * (src/bmenu.f90:407-421, :520-526). */
int bs_init_wfn(const char *wfn, const char *intf, float spscale, float a0, float a, float x,
                bs_system **sys);
/* bs_init_wfn without the upload (host tables only; bs_upload completes it). */
int bs_init_wfn_host(const char *wfn, const char *intf, float spscale, float a0, float a, float x,
                     bs_system **sys);

/* Read states first .. first+count-1 (0-based) of a .wfn file on this
 * system's basis: energy, J, T(T+1) (count entries each) and the vectors
 * (count*dim elements in dtype, on the device when on_device = 1, else on
 * the host).  nkeep (may be NULL) receives the number of states in the file;
 * count = 0 only reads it.  Vectors stored in single or double precision are
 * both accepted.  Errors: a header that is not this basis, a state range
 * beyond nkeep, a record without its magic number or with the wrong index.
 * Replaces wfn_read_nkeep + wfn_readeigenvec (src/bwfnlib.f90:1052-1065,
 * :159-273). */
int bs_read_wfn(const bs_system *sys, const char *path, int first, int count, float *energy, float *xj,
                float *xt2, void *vecs, int dtype, int on_device, int *nkeep);

/* ---------------------------------------------------------------- block and batched
 * Interleaved multi-vector layout: element k of basis state i at [i*K + k],
 * K <= 64. */

/* y = H x for K vectors of the same Hamiltonian (device pointers, y is
 * overwritten).  Replaces the block apply of the reference
 * (src/bapplyhlibOMP_block.f90:43-887, vectors block(ivec,istate)). */
int bs_apply_block(bs_system *sys, const void *x, void *y, int K, int dtype);

/* Multi-vector kernel variant for (dtype, K, shared: 1 = one Hamiltonian as
 * in bs_apply_block, 0 = one per vector as in bs_batch_apply): 0 = one wave
 * per basis state, 1..4 = one proton row per workgroup with G = 1 (1, 2) or 2
 * (3, 4) staged source rows, identity (odd) or XCD-aware (even) order; 5, 6 =
 * two proton rows per workgroup (FP32 K = 2 or 4, FP64 K = 2; one Hamiltonian
 * per vector only on a batch's system), identity (5) or XCD-aware (6) order.  Without a call the fastest is timed
 * and kept on first use.  Nonzero if it does not fit. */
int bs_multi_variant(bs_system *sys, int dtype, int K, int shared, int v);

/* The variant in use for (dtype, K, shared) (-1 if none yet) and the
 * autotuner's timings in ms for variants 0..6 (0 = not timed). */
int bs_multi_info(const bs_system *sys, int dtype, int K, int shared, int *chosen, float *ms);

/* K Hamiltonians on one basis: one iso .int file each, same orbits, particle
 * numbers and scaling line (src/binput_tbme.f90:898-1250 per file).  Jump
 * tables are built once; values are stored per Hamiltonian.  No counterpart
 * in the reference, which diagonalizes one Hamiltonian per run. */
typedef struct bs_batch bs_batch;
int bs_batch_init(const char *sps, const char *const *intfiles, int K, int Z, int N, int jz2,
                  float spscale, float a0, float a, float x, bs_batch **out);
int bs_batch_finalize(bs_batch *b);
int bs_batch_size(const bs_batch *b, int *K, int64_t *dim);
/* The system holding the shared basis and jump tables (values of the first
 * interaction); owned by the batch. */
bs_system *bs_batch_system(bs_batch *b);
/* Replace the K interactions' values by those of K other .int files on the
 * same orbits, particle numbers and scaling, read as bs_batch_init reads
 * them (src/binput_tbme.f90:898-1250 per file); the basis and jump tables
 * are kept, so an optimisation loop over interaction parameters builds them
 * once.  The tables are structural (every pair the orbits allow, zero
 * values included), so any values fit them.  On an error (a file that cannot
 * be read, a different single-particle term) the batch keeps its previous
 * values.  Drops the captured Lanczos graphs; must not overlap other calls
 * on the batch.  No counterpart in the reference, which reads one
 * interaction per run. */
int bs_batch_set_interactions(bs_batch *b, const char *const *intfiles);
/* bs_multi_variant for the batch's own K (drops a captured Lanczos graph). */
int bs_batch_set_variant(bs_batch *b, int dtype, int v);
/* y_k = H_k x_k, interleaved device vectors. */
int bs_batch_apply(bs_batch *b, const void *x, void *y, int dtype);
/* Lanczos on all K Hamiltonians in lockstep (the recurrence of bs_lanczos,
 * plain or thick restart with thick <= 16 and niter <= 256; p->tol stops the
 * thick run when no kept Ritz value of any Hamiltonian moves by more than tol
 * over a restart), captured once into HIP graphs (first cycle, restarted
 * cycle, restart, Ritz vectors + polish) and replayed on later calls with the
 * same parameters; breakdown handled on the device per Hamiltonian as in
 * bs_lanczos; Ritz values and vectors by FP64 Sturm bisection and inverse
 * iteration (plain) or Jacobi on the arrowhead matrix (thick) on the
 * device.  alpha/beta (may be NULL):
 * [k*min(niter,dim) + i]; evals[k*nkeep + n];
 * evecs (device, may be NULL): nkeep interleaved sets [n][i*K + k] in dtype;
 * resid[k*nkeep + n] (may be NULL; -1 without polish). */
int bs_batch_lanczos(bs_batch *b, const bs_lanczos_params *p, double *alpha, double *beta, double *evals,
                     void *evecs, double *resid, int *iters);
/* Hellmann-Feynman gradients of <psi_k|H_k|psi_k>/<psi_k|psi_k> for every
 * Hamiltonian: grad[k*nparams + p], parameters as in bs_gradient. */
int bs_batch_gradient(bs_batch *b, const void *psi, int dtype, double *grad);

#ifdef __cplusplus
}
#endif
#endif
