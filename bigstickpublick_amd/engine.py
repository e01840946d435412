"""ctypes binding of libbigstick_amd.so (include/bigstick_amd.h).

The library must have been built (bigstickpublick_amd.build); nothing here
falls back to another implementation: a missing library is an error."""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbigstick_amd.so")
BS_F32, BS_F64 = 4, 8

_lib = None


class _Params(ctypes.Structure):
    _fields_ = [("nkeep", ctypes.c_int), ("niter", ctypes.c_int), ("thick", ctypes.c_int),
                ("max_restarts", ctypes.c_int), ("dtype", ctypes.c_int), ("tol", ctypes.c_double),
                ("polish", ctypes.c_int), ("conv_test", ctypes.c_int), ("ncheck_ex", ctypes.c_int),
                ("target", ctypes.c_int), ("etarget", ctypes.c_double)]


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libbigstick_amd.so is not built; run __graft_entry__.build()")
    # torch's HIP runtime sets the device up before this library makes its
    # first HIP call: in a process whose first HIP call comes from the library,
    # torch finds no GPU afterwards
    import torch
    torch.cuda.is_available()
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64p = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)
    f32 = ctypes.c_float
    L.bs_last_error.restype = ctypes.c_char_p
    for name in ("bs_init_host", "bs_init"):
        fn = getattr(L, name)
        fn.argtypes = [ctypes.c_char_p, ctypes.c_char_p, i32, i32, i32, f32, f32, f32, f32,
                       ctypes.POINTER(vp)]
    L.bs_upload.argtypes = [vp]
    L.bs_finalize.argtypes = [vp]
    L.bs_dim.argtypes = [vp, i64p]
    L.bs_num_sd.argtypes = [vp, i32, i64p]
    L.bs_basis.argtypes = [vp, i32, vp, vp]
    L.bs_num_sps.argtypes = [vp, ctypes.POINTER(i32)]
    L.bs_sps.argtypes = [vp, vp, vp, vp]
    L.bs_me_table.argtypes = [vp, i32, ctypes.POINTER(i32), vp]
    L.bs_op_counts.argtypes = [vp, vp]
    L.bs_apply_h.argtypes = [vp, vp, vp, i32, i32, vp]
    L.bs_jumps.argtypes = [vp, i32, i32, i64p, vp, vp, vp, vp]
    L.bs_num_variants.argtypes = [ctypes.POINTER(i32)]
    L.bs_variant_info.argtypes = [i32] + [ctypes.POINTER(i32)] * 4
    L.bs_set_variant.argtypes = [vp, i32, i32]
    L.bs_get_variant.argtypes = [vp, i32, ctypes.POINTER(i32)]
    L.bs_autotune.argtypes = [vp, i32, ctypes.POINTER(i32)]
    L.bs_autotune_times.argtypes = [vp, i32, vp]
    L.bs_apply_h_host.argtypes = [vp, vp, vp, i32, i32]
    L.bs_num_params.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.bs_density.argtypes = [vp, vp, i32, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
    L.bs_gradient.argtypes = [vp, vp, i32, vp]
    L.bs_observable.argtypes = [vp, i32, vp, i32, ctypes.POINTER(ctypes.c_double)]
    L.bs_lanczos.argtypes = [vp, ctypes.POINTER(_Params), vp, vp, vp, vp, vp,
                             ctypes.POINTER(i32)]
    L.bs_lanczos_noalloc.argtypes = [vp, ctypes.POINTER(_Params), vp, vp, vp, ctypes.POINTER(i32)]
    L.bs_lanczos_stats.argtypes = [vp] + [ctypes.POINTER(i32)] * 4
    L.bs_block_lanczos.argtypes = [vp, ctypes.POINTER(_Params), i32, vp, vp, vp, ctypes.POINTER(i32)]
    L.bs_block_strength.argtypes = [vp, ctypes.POINTER(_Params), i32, vp, i32, vp, vp, vp, vp, vp,
                                    ctypes.POINTER(i32)]
    L.bs_exact.argtypes = [vp, i32, i32, vp, vp, vp, ctypes.POINTER(i32)]
    L.bs_strength.argtypes = [vp, ctypes.POINTER(_Params), vp, ctypes.POINTER(ctypes.c_double), vp, vp, vp, vp,
                              vp, vp, vp, ctypes.POINTER(i32)]
    L.bs_apply_block.argtypes = [vp, vp, vp, i32, i32]
    L.bs_write_wfn.argtypes = [vp, ctypes.c_char_p, i32, vp, vp, vp, vp, i32, i32]
    L.bs_operator_matrix.argtypes = [vp, ctypes.c_char_p, f32, f32, f32, f32, vp, i32, i32, vp]
    for name in ("bs_init_wfn_host", "bs_init_wfn"):
        getattr(L, name).argtypes = [ctypes.c_char_p, ctypes.c_char_p, f32, f32, f32, f32, ctypes.POINTER(vp)]
    L.bs_read_wfn.argtypes = [vp, ctypes.c_char_p, i32, i32, vp, vp, vp, vp, i32, i32, ctypes.POINTER(i32)]
    L.bs_apply_onebody.argtypes = [vp, ctypes.c_char_p, vp, vp, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.bs_multi_variant.argtypes = [vp, i32, i32, i32, i32]
    L.bs_batch_set_variant.argtypes = [vp, i32, i32]
    L.bs_multi_info.argtypes = [vp, i32, i32, i32, ctypes.POINTER(i32), vp]
    L.bs_batch_init.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), i32, i32, i32, i32,
                                f32, f32, f32, f32, ctypes.POINTER(vp)]
    L.bs_batch_finalize.argtypes = [vp]
    L.bs_batch_size.argtypes = [vp, ctypes.POINTER(i32), i64p]
    L.bs_batch_system.argtypes = [vp]
    L.bs_batch_system.restype = vp
    L.bs_batch_apply.argtypes = [vp, vp, vp, i32]
    L.bs_batch_lanczos.argtypes = [vp, ctypes.POINTER(_Params), vp, vp, vp, vp, vp, ctypes.POINTER(i32)]
    L.bs_batch_gradient.argtypes = [vp, vp, i32, vp]
    L.bs_batch_set_interactions.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError("bigstick_amd: " + load_library().bs_last_error().decode())


def _dev(t, name, dim=None, shape=None, dtype=None, numel=None, contiguous=True):
    """Checks a tensor handed to a device entry point before any library call:
    a CUDA float32/float64 tensor (of `dtype` when given) with `numel`
    elements or the given shape or trailing dimension `dim`, contiguous when
    it is written by the library.  Raises ValueError otherwise, so that a host
    tensor or a short buffer never reaches a kernel."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError("%s must be a torch tensor" % name)
    if not t.is_cuda:
        raise ValueError("%s must be a CUDA tensor" % name)
    if t.dtype not in (torch.float32, torch.float64):
        raise ValueError("%s must be float32 or float64" % name)
    if dtype is not None and t.dtype != dtype:
        raise ValueError("%s must have dtype %s (got %s)" % (name, dtype, t.dtype))
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError("%s must have shape %s (got %s)" % (name, tuple(shape), tuple(t.shape)))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s must have %d elements (got %d)" % (name, numel, t.numel()))
    if dim is not None and (t.dim() == 0 or t.shape[-1] != dim or t.numel() % dim):
        raise ValueError("%s must have dim = %d columns" % (name, dim))
    if contiguous and not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    return t


def _out_like(y, x, name="y"):
    """An output buffer for x: a new tensor, or the caller's y checked to have
    x's dtype, device and size."""
    import torch
    if y is None:
        return torch.empty_like(x)
    _dev(y, name, dtype=x.dtype, numel=x.numel())
    if y.device != x.device:
        raise ValueError("%s must be on %s" % (name, x.device))
    return y


@dataclass
class LanczosResult:
    evals: np.ndarray
    alpha: np.ndarray
    beta: np.ndarray
    resid: np.ndarray
    matvecs: int
    evecs: object = None


@dataclass
class StrengthResult:
    ritz: np.ndarray       # every Ritz value of the fixed run
    strength: np.ndarray   # |Y(1,k)|^2 ||pivot||^2 for each
    pivot_norm: float
    evals: np.ndarray      # the nkeep lowest, polished as in lanczos()
    alpha: np.ndarray
    beta: np.ndarray
    resid: np.ndarray
    matvecs: int
    evecs: object = None


class System:
    """One nucleus in one model space: tables built once on the host, kept in
    HBM after upload()."""

    def __init__(self, sps, intfile, Z, N, jz2=0, scaling=(1.0, 1.0, 1.0, 1.0), upload=False):
        L = load_library()
        h = ctypes.c_void_p()
        _check(L.bs_init_host(str(sps).encode(), str(intfile).encode(), Z, N, jz2,
                              *[float(v) for v in scaling], ctypes.byref(h)))
        self._adopt(h, upload)

    @classmethod
    def from_wfn(cls, wfn, intfile=None, scaling=(1.0, 1.0, 1.0, 1.0), upload=False):
        """The basis recorded in a .wfn file (orbits, Z, N, 2Jz) with the
        interaction `intfile` (None: no interaction); fails unless the rebuilt
        basis reproduces the file's header."""
        L = load_library()
        h = ctypes.c_void_p()
        name = None if intfile is None else str(intfile).encode()
        _check(L.bs_init_wfn_host(str(wfn).encode(), name, *[float(v) for v in scaling], ctypes.byref(h)))
        self = cls.__new__(cls)
        self._adopt(h, upload)
        return self

    def _adopt(self, h, upload):
        L = load_library()
        self._lib = L  # held so that close() still works at interpreter exit
        self.h = h
        d = ctypes.c_int64()
        _check(L.bs_dim(h, ctypes.byref(d)))
        self.dim = d.value
        n = ctypes.c_int()
        _check(L.bs_num_sps(h, ctypes.byref(n)))
        self.nsps = n.value
        self.uploaded = False
        if upload:
            self.upload()

    def close(self):
        if getattr(self, "h", None):
            self._lib.bs_finalize(self.h)
            self.h = None

    __del__ = close

    def upload(self):
        if not self.uploaded:
            _check(load_library().bs_upload(self.h))
            self.uploaded = True

    def basis(self, it):
        L = load_library()
        n = ctypes.c_int64()
        _check(L.bs_num_sd(self.h, it, ctypes.byref(n)))
        mask = np.zeros(n.value, np.uint64)
        start = np.zeros(n.value, np.int64)
        _check(L.bs_basis(self.h, it, mask.ctypes.data, start.ctypes.data))
        return mask, start

    def me_table(self, kind):
        L = load_library()
        n = ctypes.c_int()
        _check(L.bs_me_table(self.h, kind, ctypes.byref(n), None))
        t = np.zeros((n.value, n.value), np.float32)
        _check(L.bs_me_table(self.h, kind, ctypes.byref(n), t.ctypes.data))
        return t

    def op_counts(self):
        out = np.zeros(8, np.int64)
        _check(load_library().bs_op_counts(self.h, out.ctypes.data))
        return dict(spe=int(out[0]), pp=int(out[1]), nn=int(out[2]), pn=int(out[3]),
                    ref_spe=int(out[4]), ref_pp=int(out[5]), ref_nn=int(out[6]),
                    ref_pn=int(out[7]), ref_total=int(out[4:].sum()))

    def jumps(self, it, order):
        """(isd, fsd, op, phase) arrays of the engine's one- (order 1) or
        two-body (order 2) jumps of species it, every direction."""
        L = load_library()
        n = ctypes.c_int64(0)
        _check(L.bs_jumps(self.h, it, order, ctypes.byref(n), None, None, None, None))
        isd, fsd, op = (np.zeros(n.value, np.int32) for _ in range(3))
        ph = np.zeros(n.value, np.int8)
        _check(L.bs_jumps(self.h, it, order, ctypes.byref(n), isd.ctypes.data, fsd.ctypes.data, op.ctypes.data,
                          ph.ctypes.data))
        return isd, fsd, op, ph

    # ---- device side: torch tensors are used only as device buffers ----
    def apply(self, x, y=None, accumulate=False, stream=None):
        import torch
        x = _dev(x, "x", numel=self.dim, contiguous=False).contiguous()
        y = _out_like(y, x)
        if y.data_ptr() == x.data_ptr():
            raise ValueError("x and y must not overlap")
        self.upload()
        dt = BS_F32 if x.dtype == torch.float32 else BS_F64
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(load_library().bs_apply_h(self.h, ctypes.c_void_p(x.data_ptr()),
                                         ctypes.c_void_p(y.data_ptr()), dt, int(accumulate),
                                         ctypes.c_void_p(st)))
        return y

    @staticmethod
    def variants():
        L = load_library()
        n = ctypes.c_int()
        _check(L.bs_num_variants(ctypes.byref(n)))
        out = []
        for v in range(n.value):
            r, g, m, pf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _check(L.bs_variant_info(v, ctypes.byref(r), ctypes.byref(g), ctypes.byref(m), ctypes.byref(pf)))
            out.append((r.value, g.value, m.value, pf.value))
        return out

    def set_variant(self, dtype, v):
        """Force a kernel variant for dtype ('f32'/'f64'); returns False if it
        does not fit in LDS for this system."""
        self.upload()
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        rc = load_library().bs_set_variant(self.h, dt, v)
        return rc == 0

    def get_variant(self, dtype):
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        v = ctypes.c_int()
        _check(load_library().bs_get_variant(self.h, dt, ctypes.byref(v)))
        return v.value

    def autotune(self, dtype="f32"):
        self.upload()
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        v = ctypes.c_int()
        _check(load_library().bs_autotune(self.h, dt, ctypes.byref(v)))
        return v.value

    def apply_host(self, x, accumulate=False, y=None):
        """numpy in, numpy out (host-pointer entry point, PCIe included)."""
        self.upload()
        x = np.ascontiguousarray(x)
        if x.dtype not in (np.float32, np.float64) or x.size != self.dim:
            raise ValueError("x must be float32/float64 of length dim")
        if y is None:
            y = np.zeros_like(x)
        dt = BS_F32 if x.dtype == np.float32 else BS_F64
        _check(load_library().bs_apply_h_host(self.h, x.ctypes.data, y.ctypes.data, dt, int(accumulate)))
        return y

    def observable(self, which, psi):
        """<psi|J^2|psi> (which='J2') or <psi|T^2|psi> (which='T2') for a CUDA vector."""
        psi = _dev(psi, "psi", numel=self.dim, contiguous=False).contiguous()
        self.upload()
        dt = BS_F32 if psi.dtype.itemsize == 4 else BS_F64
        v = ctypes.c_double()
        _check(load_library().bs_observable(self.h, 0 if which == "J2" else 1,
                                            ctypes.c_void_p(psi.data_ptr()), dt, ctypes.byref(v)))
        return v.value

    def num_params(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(load_library().bs_num_params(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def density(self, psi):
        """(rho_pp, rho_nn, rho_pn, <psi|psi>) for a CUDA vector psi."""
        psi = _dev(psi, "psi", numel=self.dim, contiguous=False).contiguous()
        self.upload()
        npx = self.me_table(0).shape[0]
        npn = self.me_table(2).shape[0]
        rpp = np.zeros((npx, npx))
        rnn = np.zeros((npx, npx))
        rpn = np.zeros((npn, npn))
        n2 = ctypes.c_double()
        dt = BS_F32 if psi.dtype.itemsize == 4 else BS_F64
        _check(load_library().bs_density(self.h, ctypes.c_void_p(psi.data_ptr()), dt, rpp.ctypes.data,
                                         rnn.ctypes.data, rpn.ctypes.data, ctypes.byref(n2)))
        return rpp, rnn, rpn, n2.value

    def gradient(self, psi):
        """dE/dtheta for theta = (SPEs, TBME lines in .int order), psi a CUDA vector."""
        psi = _dev(psi, "psi", numel=self.dim, contiguous=False).contiguous()
        self.upload()
        nspe, ntb = self.num_params()
        g = np.zeros(nspe + ntb)
        dt = BS_F32 if psi.dtype.itemsize == 4 else BS_F64
        _check(load_library().bs_gradient(self.h, ctypes.c_void_p(psi.data_ptr()), dt, g.ctypes.data))
        return g

    def autotune_times(self, dtype="f32"):
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        out = np.zeros(len(self.variants()), np.float32)
        _check(load_library().bs_autotune_times(self.h, dt, out.ctypes.data))
        return out

    def lanczos(self, nkeep, niter, thick=0, max_restarts=0, dtype="f32", tol=0.0, polish=True,
                want_vectors=False, conv_test=1, ncheck_ex=0, target=None, target_test="reference"):
        """Lanczos from the uniform pivot.  With tol > 0 the run stops by the
        reference's test conv_test (0: mean energy change, 1: largest energy
        change, 2 / 3: mean / largest change of the Ritz vectors)
        over nkeep + ncheck_ex Ritz values (the reference's ld is conv_test=0,
        ncheck_ex=5, tol=1e-3).  target (MeV, thick > 0): keep the Ritz
        pairs nearest that energy instead of the lowest (option tx);
        target_test "reference" tests convergence as the reference's tx does
        (over no energies: the first check stops the run), "kept" over the
        nkeep kept energies."""
        import torch
        self.upload()
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        if target_test not in ("reference", "kept"):
            raise ValueError("target_test must be 'reference' or 'kept'")
        mode = 0 if target is None else (1 if target_test == "reference" else 2)
        p = _Params(nkeep, niter, thick, max_restarts, dt, tol, int(polish), conv_test, ncheck_ex,
                    mode, 0.0 if target is None else float(target))
        m = min(niter, self.dim)
        alpha = np.zeros(m)
        beta = np.zeros(m)
        ev = np.zeros(nkeep)
        res = np.zeros(nkeep)
        it = ctypes.c_int()
        vec = None
        if want_vectors:
            vec = torch.empty((nkeep, self.dim), device="cuda",
                              dtype=torch.float32 if dt == BS_F32 else torch.float64)
        torch.cuda.synchronize()
        _check(load_library().bs_lanczos(self.h, ctypes.byref(p), alpha.ctypes.data,
                                         beta.ctypes.data, ev.ctypes.data,
                                         ctypes.c_void_p(vec.data_ptr()) if vec is not None else None,
                                         res.ctypes.data, ctypes.byref(it)))
        return LanczosResult(ev, alpha, beta, res, it.value, vec)

    def lanczos_stats(self):
        """(random restarts, second Gram-Schmidt passes, Lanczos steps of the
        last Krylov sequence, thick restarts) of the last lanczos() or
        strength() run."""
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(load_library().bs_lanczos_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                               ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def exact(self, nkeep, dtype="f64", want_vectors=False):
        """Exact diagonalization of the dense H (dim <= 65536): the nkeep
        lowest eigenpairs and their residuals."""
        import torch
        self.upload()
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        ev = np.zeros(nkeep)
        res = np.zeros(nkeep)
        it = ctypes.c_int()
        vec = None
        if want_vectors:
            vec = torch.empty((nkeep, self.dim), device="cuda",
                              dtype=torch.float32 if dt == BS_F32 else torch.float64)
        torch.cuda.synchronize()
        _check(load_library().bs_exact(self.h, nkeep, dt, ev.ctypes.data,
                                       ctypes.c_void_p(vec.data_ptr()) if vec is not None else None,
                                       res.ctypes.data, ctypes.byref(it)))
        return LanczosResult(ev, None, None, res, it.value, vec)

    def block_lanczos(self, nkeep, niter, block, dtype="f32", tol=0.0, polish=True, want_vectors=False,
                      thick=0, max_restarts=0, conv_test=0, ncheck_ex=0):
        """Block Lanczos with `block` vectors per block and a basis of niter
        vectors (fixed, or stopped at tol); alpha/beta are not returned.
        thick > 0 (option tb): restart when the basis would reach niter,
        keeping (thick // block) whole blocks of Ritz vectors, at most
        max_restarts + 1 restarts, the reference's energy test conv_test over
        nkeep + ncheck_ex values at every block iteration (tb: 0, 2, 1e-3)."""
        import torch
        self.upload()
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        p = _Params(nkeep, niter, thick, max_restarts, dt, tol, int(polish), conv_test, ncheck_ex)
        ev = np.zeros(nkeep)
        res = np.zeros(nkeep)
        it = ctypes.c_int()
        vec = None
        if want_vectors:
            vec = torch.empty((nkeep, self.dim), device="cuda",
                              dtype=torch.float32 if dt == BS_F32 else torch.float64)
        torch.cuda.synchronize()
        _check(load_library().bs_block_lanczos(self.h, ctypes.byref(p), block, ev.ctypes.data,
                                               ctypes.c_void_p(vec.data_ptr()) if vec is not None else None,
                                               res.ctypes.data, ctypes.byref(it)))
        return LanczosResult(ev, None, None, res, it.value, vec)

    def block_strength(self, pivots, nblocks, dtype="f64", polish=False):
        """Block strength function: block Lanczos from the pivot block
        `pivots` (a (K, dim) CUDA tensor, float32 or float64) with nblocks
        blocks; returns (Ritz values, strengths (K, K*nblocks) with
        strength[i, k] = |<Ritz k|pivot i>|^2, |pivot i|^2, matvecs)."""
        import torch
        if not isinstance(pivots, torch.Tensor) or pivots.dim() != 2:
            raise ValueError("pivots must be a (K, dim) CUDA tensor")
        K = int(pivots.shape[0])
        pv = _dev(pivots, "pivots", shape=(K, self.dim), contiguous=False).contiguous()
        pdt = BS_F32 if pv.dtype == torch.float32 else BS_F64
        self.upload()
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        n = K * nblocks
        p = _Params(n, n, 0, 0, dt, 0.0, int(polish))
        ev, st, pn2 = np.zeros(n), np.zeros(K * n), np.zeros(K)
        it = ctypes.c_int()
        torch.cuda.synchronize()
        _check(load_library().bs_block_strength(self.h, ctypes.byref(p), K, ctypes.c_void_p(pv.data_ptr()), pdt,
                                                pn2.ctypes.data, st.ctypes.data, ev.ctypes.data, None, None,
                                                ctypes.byref(it)))
        return ev, st.reshape(K, n), pn2, it.value

    def strength(self, pivot, nkeep, niter, polish=True, want_vectors=False):
        """Strength function: a fixed Lanczos run of niter steps started from
        `pivot` (a CUDA float32/float64 tensor of dim elements; its dtype is
        the run's vector precision)."""
        import torch
        pv = _dev(pivot, "pivot", numel=self.dim, contiguous=False).contiguous().reshape(-1)
        self.upload()
        dt = BS_F32 if pv.dtype == torch.float32 else BS_F64
        p = _Params(nkeep, niter, 0, 0, dt, 0.0, int(polish))
        m = min(niter, self.dim)
        ritz, stren, alpha, beta = (np.zeros(m) for _ in range(4))
        ev = np.zeros(nkeep)
        res = np.zeros(nkeep)
        pn = ctypes.c_double()
        it = ctypes.c_int()
        vec = torch.empty((nkeep, self.dim), device="cuda", dtype=pv.dtype) if want_vectors else None
        torch.cuda.synchronize()
        _check(load_library().bs_strength(self.h, ctypes.byref(p), ctypes.c_void_p(pv.data_ptr()), ctypes.byref(pn),
                                          ritz.ctypes.data, stren.ctypes.data, alpha.ctypes.data, beta.ctypes.data,
                                          ev.ctypes.data,
                                          ctypes.c_void_p(vec.data_ptr()) if vec is not None else None,
                                          res.ctypes.data, ctypes.byref(it)))
        return StrengthResult(ritz, stren, pn.value, ev, alpha, beta, res, it.value, vec)

    def set_block_variant(self, dtype, K, v):
        """Force multi-vector kernel variant v (0..6) for apply_block with K
        vectors; False if it does not fit."""
        self.upload()
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        return load_library().bs_multi_variant(self.h, dt, K, 1, v) == 0

    def block_variant_info(self, dtype, K):
        """(multi-vector variant in use for apply_block with K vectors or -1,
        the autotuner's ms per variant)."""
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        c = ctypes.c_int()
        ms = np.zeros(7, np.float32)
        _check(load_library().bs_multi_info(self.h, dt, K, 1, ctypes.byref(c), ms.ctypes.data))
        return c.value, ms

    def operator_matrix(self, intfile, vecs, scaling=(1.0, 1.0, 1.0, 1.0)):
        """<v_f|O|v_i> for the two-body operator in `intfile` (iso .int; None:
        this system's Hamiltonian) and an (nvec, dim) CUDA tensor of vectors;
        returns an (nvec, nvec) array [i, f]."""
        v = _dev(vecs, "vecs", dim=self.dim, contiguous=False).contiguous()
        n = v.numel() // self.dim
        self.upload()
        out = np.zeros((n, n))
        dt = BS_F32 if v.dtype.itemsize == 4 else BS_F64
        name = None if intfile is None else str(intfile).encode()
        _check(load_library().bs_operator_matrix(self.h, name, *[float(s) for s in scaling],
                                                 ctypes.c_void_p(v.data_ptr()), n, dt, out.ctypes.data))
        return out

    def apply_onebody(self, opme, x):
        """O x for the one-body operator of an .opme file and a (dim,) or
        (nvec, dim) CUDA float32/float64 tensor; returns (y, J, T)."""
        import torch
        xv = _dev(x, "x", dim=self.dim, contiguous=False).contiguous()
        self.upload()
        y = torch.empty_like(xv)
        nvec = xv.numel() // self.dim
        j, t = ctypes.c_int(), ctypes.c_int()
        _check(load_library().bs_apply_onebody(self.h, str(opme).encode(), ctypes.c_void_p(xv.data_ptr()),
                                               ctypes.c_void_p(y.data_ptr()), nvec,
                                               BS_F32 if xv.dtype == torch.float32 else BS_F64,
                                               ctypes.byref(j), ctypes.byref(t)))
        return y, j.value, t.value

    def wfn_count(self, path):
        """Number of states in a .wfn file on this basis."""
        n = ctypes.c_int()
        _check(load_library().bs_read_wfn(self.h, str(path).encode(), 0, 0, None, None, None, None, BS_F32, 0,
                                          ctypes.byref(n)))
        return n.value

    def read_wfn(self, path, first=0, count=None, dtype="f32", device=False):
        """States first..first+count-1 of a .wfn file on this basis: energy,
        J, T(T+1) (float32 arrays) and the vectors, an (count, dim) numpy
        array, or a CUDA tensor when device=True."""
        n = self.wfn_count(path)
        if count is None:
            count = n - first
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        e, j, t = (np.zeros(count, np.float32) for _ in range(3))
        if device:
            import torch
            self.upload()
            v = torch.empty((count, self.dim), dtype=torch.float32 if dt == BS_F32 else torch.float64,
                            device="cuda")
            ptr = v.data_ptr()
        else:
            v = np.zeros((count, self.dim), np.float32 if dt == BS_F32 else np.float64)
            ptr = v.ctypes.data
        nk = ctypes.c_int()
        _check(load_library().bs_read_wfn(self.h, str(path).encode(), first, count, e.ctypes.data, j.ctypes.data,
                                          t.ctypes.data, ctypes.c_void_p(ptr), dt, 1 if device else 0,
                                          ctypes.byref(nk)))
        return e, j, t, v

    def write_wfn(self, path, energy, xj, xt2, vecs):
        """Write the reference's binary .wfn: energies, J, T(T+1) (length
        nkeep) and vecs, an (nkeep, dim) numpy array or CUDA tensor."""
        e = np.ascontiguousarray(energy, dtype=np.float32)
        j = np.ascontiguousarray(xj, dtype=np.float32)
        t = np.ascontiguousarray(xt2, dtype=np.float32)
        nk = len(e)
        if isinstance(vecs, np.ndarray):
            v = np.ascontiguousarray(vecs)
            if v.dtype not in (np.float32, np.float64) or v.size != nk * self.dim:
                raise ValueError("vecs must be float32/float64 of shape (nkeep, dim)")
            ptr, dt, dev = v.ctypes.data, (BS_F32 if v.dtype == np.float32 else BS_F64), 0
        else:
            v = _dev(vecs, "vecs", numel=nk * self.dim, contiguous=False).contiguous()
            ptr, dt, dev = v.data_ptr(), (BS_F32 if v.dtype.itemsize == 4 else BS_F64), 1
        _check(load_library().bs_write_wfn(self.h, str(path).encode(), nk, e.ctypes.data, j.ctypes.data,
                                           t.ctypes.data, ctypes.c_void_p(ptr), dt, dev))

    def apply_block(self, x, y=None):
        """y[i, k] = (H x)[i, k] for a (dim, K) CUDA tensor of K vectors, K <= 64."""
        import torch
        if not isinstance(x, torch.Tensor) or x.dim() != 2:
            raise ValueError("x must be a (dim, K) CUDA float32/float64 tensor")
        x = _dev(x, "x", shape=(self.dim, x.shape[1]), contiguous=False).contiguous()
        if not 1 <= x.shape[1] <= 64:
            raise ValueError("K must be in 1..64")
        y = _out_like(y, x)
        if y.data_ptr() == x.data_ptr():
            raise ValueError("x and y must not overlap")
        self.upload()
        dt = BS_F32 if x.dtype == torch.float32 else BS_F64
        _check(load_library().bs_apply_block(self.h, ctypes.c_void_p(x.data_ptr()),
                                             ctypes.c_void_p(y.data_ptr()), x.shape[1], dt))
        return y


@dataclass
class BatchResult:
    evals: np.ndarray      # (K, nkeep)
    alpha: np.ndarray      # (K, niter)
    beta: np.ndarray       # (K, niter)
    resid: np.ndarray      # (K, nkeep)
    matvecs: int
    evecs: object = None   # (nkeep, dim, K) CUDA tensor


class Batch:
    """K Hamiltonians (one .int file each) on one basis, K <= 64: shared jump
    tables, interleaved (dim, K) vectors."""

    def __init__(self, sps, intfiles, Z, N, jz2=0, scaling=(1.0, 1.0, 1.0, 1.0)):
        L = load_library()
        names = (ctypes.c_char_p * len(intfiles))(*[str(f).encode() for f in intfiles])
        h = ctypes.c_void_p()
        _check(L.bs_batch_init(str(sps).encode(), names, len(intfiles), Z, N, jz2,
                               *[float(v) for v in scaling], ctypes.byref(h)))
        self._lib = L  # held so that close() still works at interpreter exit
        self.h = h
        k, d = ctypes.c_int(), ctypes.c_int64()
        _check(L.bs_batch_size(h, ctypes.byref(k), ctypes.byref(d)))
        self.K, self.dim = k.value, d.value
        self.intfiles = list(intfiles)

    def close(self):
        if getattr(self, "h", None):
            self._lib.bs_batch_finalize(self.h)
            self.h = None

    def set_interactions(self, intfiles):
        """Take the values of K other .int files (same orbits, particle numbers
        and scaling) on the same basis and jump tables, which are not rebuilt.
        On an error the batch keeps its previous values."""
        if len(intfiles) != self.K:
            raise ValueError("need %d interaction files (got %d)" % (self.K, len(intfiles)))
        names = (ctypes.c_char_p * self.K)(*[str(f).encode() for f in intfiles])
        _check(load_library().bs_batch_set_interactions(self.h, names))
        self.intfiles = list(intfiles)

    __del__ = close

    def num_params(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        sysh = load_library().bs_batch_system(self.h)
        _check(load_library().bs_num_params(sysh, ctypes.byref(a), ctypes.byref(b)))
        return a.value + b.value

    def set_variant(self, dtype, v):
        """Force multi-vector kernel variant v (0..6); False if it does not fit."""
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        return load_library().bs_batch_set_variant(self.h, dt, v) == 0

    def variant_info(self, dtype="f32"):
        """(chosen multi-vector variant or -1, autotuner ms per variant)."""
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        L = load_library()
        c = ctypes.c_int()
        ms = np.zeros(7, np.float32)
        _check(L.bs_multi_info(L.bs_batch_system(self.h), dt, self.K, 0, ctypes.byref(c), ms.ctypes.data))
        return c.value, ms

    def apply(self, x, y=None):
        """y[:, k] = H_k x[:, k] for a (dim, K) CUDA tensor."""
        import torch
        x = _dev(x, "x", shape=(self.dim, self.K), contiguous=False).contiguous()
        y = _out_like(y, x)
        if y.data_ptr() == x.data_ptr():
            raise ValueError("x and y must not overlap")
        dt = BS_F32 if x.dtype == torch.float32 else BS_F64
        torch.cuda.synchronize()
        _check(load_library().bs_batch_apply(self.h, ctypes.c_void_p(x.data_ptr()),
                                             ctypes.c_void_p(y.data_ptr()), dt))
        torch.cuda.synchronize()
        return y

    def lanczos(self, nkeep, niter, thick=0, max_restarts=0, dtype="f32", tol=0.0, polish=True,
                want_vectors=False):
        import torch
        dt = BS_F32 if dtype in ("f32", "float32") else BS_F64
        p = _Params(nkeep, niter, thick, max_restarts, dt, tol, int(polish))
        ev = np.zeros((self.K, nkeep))
        res = np.zeros((self.K, nkeep))
        m = min(niter, self.dim)
        alpha = np.zeros((self.K, m))
        beta = np.zeros((self.K, m))
        it = ctypes.c_int()
        vec = None
        if want_vectors:
            vec = torch.empty((nkeep, self.dim, self.K), device="cuda",
                              dtype=torch.float32 if dt == BS_F32 else torch.float64)
        torch.cuda.synchronize()
        _check(load_library().bs_batch_lanczos(self.h, ctypes.byref(p), alpha.ctypes.data, beta.ctypes.data,
                                               ev.ctypes.data,
                                               ctypes.c_void_p(vec.data_ptr()) if vec is not None else None,
                                               res.ctypes.data, ctypes.byref(it)))
        return BatchResult(ev, alpha, beta, res, it.value, vec)

    def gradient(self, psi):
        """dE_k/dtheta for each Hamiltonian: psi is a (dim, K) CUDA tensor, one
        eigenvector per Hamiltonian; returns (K, nparams)."""
        import torch
        psi = _dev(psi, "psi", shape=(self.dim, self.K), contiguous=False).contiguous()
        g = np.zeros((self.K, self.num_params()))
        dt = BS_F32 if psi.dtype == torch.float32 else BS_F64
        torch.cuda.synchronize()
        _check(load_library().bs_batch_gradient(self.h, ctypes.c_void_p(psi.data_ptr()), dt, g.ctypes.data))
        return g
