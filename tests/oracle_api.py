"""ctypes binding of the C oracle (oracle/bsoracle.c).  Test infrastructure:
only tests/, the smoke check and bench.py's CPU-baseline leg import this."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "libbsoracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB)
        L.bso_create.restype = ctypes.c_void_p
        L.bso_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                 ctypes.POINTER(ctypes.c_int)]
        L.bso_free.argtypes = [ctypes.c_void_p]
        L.bso_dim.restype = ctypes.c_int64
        L.bso_dim.argtypes = [ctypes.c_void_p]
        L.bso_nsd.restype = ctypes.c_int64
        L.bso_nsd.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.bso_nsps.argtypes = [ctypes.c_void_p]
        L.bso_nleft.argtypes = [ctypes.c_void_p]
        L.bso_call_spe.argtypes = [ctypes.c_void_p]
        L.bso_sds.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p]
        L.bso_sps.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 3
        L.bso_me_table.restype = ctypes.POINTER(ctypes.c_float)
        L.bso_me_table.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.bso_coupled.restype = ctypes.c_float
        L.bso_coupled.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6
        L.bso_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.bso_apply_rows.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64]
        L.bso_port_create.restype = ctypes.c_void_p
        L.bso_port_create.argtypes = [ctypes.c_void_p]
        L.bso_port_free.argtypes = [ctypes.c_void_p]
        L.bso_port_apply.restype = ctypes.c_int64
        L.bso_port_apply.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64] * 2
        L.bso_port_count.restype = ctypes.c_int64
        L.bso_port_count.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]
        L.bso_cleb.restype = ctypes.c_float
        L.bso_cleb.argtypes = [ctypes.c_int] * 6
        L.bso_opme.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.bso_apply_onebody.argtypes = [ctypes.c_void_p] * 5
        _lib = L
    return _lib


class Oracle:
    """FP64 reference Hamiltonian for one (sps, int, Z, N, 2Jz) system."""

    def __init__(self, sps, intf, Z, N, jz2=0, spscale=1.0, vscale=1.0):
        L = lib()
        err = ctypes.c_int(0)
        self.h = L.bso_create(sps.encode(), intf.encode(), Z, N, jz2, spscale, vscale,
                              ctypes.byref(err))
        if not self.h:
            raise RuntimeError("oracle failed to build system (code %d)" % err.value)
        self.dim = L.bso_dim(self.h)
        self.nsps = L.bso_nsps(self.h)
        self.nleft = L.bso_nleft(self.h)
        self.call_spe = L.bso_call_spe(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().bso_free(self.h)
            self.h = None

    def sds(self, it):
        n = lib().bso_nsd(self.h, it)
        mask = np.zeros(n, np.uint64)
        start = np.zeros(n, np.int64)
        sec = np.zeros(n, np.int32)
        lib().bso_sds(self.h, it, mask.ctypes.data, start.ctypes.data, sec.ctypes.data)
        return mask, start, sec

    def sps(self):
        n = self.nsps
        a = [np.zeros(n, np.int32) for _ in range(3)]
        lib().bso_sps(self.h, *[x.ctypes.data for x in a])
        return a

    def me_table(self, kind):
        n = ctypes.c_int(0)
        p = lib().bso_me_table(self.h, kind, ctypes.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value, n.value)).copy()

    def coupled(self, kind, a, b, c, d, J):
        return lib().bso_coupled(self.h, kind, a, b, c, d, J)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().bso_apply(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def apply_rows(self, x, rows):
        """(H x) on the states of the listed proton SDs (FP64 brute force, the
        same arithmetic as apply); returns the full-length y with only those
        states set."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        r = np.ascontiguousarray(rows, dtype=np.int64)
        if lib().bso_apply_rows(self.h, x.ctypes.data, y.ctypes.data, r.ctypes.data, len(r)) != 0:
            raise ValueError("proton row out of range")
        return y

    def opme(self, path):
        """m-scheme coefficients (op_p, op_n: nsps x nsps, [a, b] multiplies
        a+_a a_b), J and T of the one-body operator in an .opme file."""
        n = self.nsps
        op_p, op_n = np.zeros((n, n)), np.zeros((n, n))
        j, t = ctypes.c_int(), ctypes.c_int()
        rc = lib().bso_opme(self.h, str(path).encode(), op_p.ctypes.data, op_n.ctypes.data, ctypes.byref(j),
                            ctypes.byref(t))
        if rc != 0:
            raise RuntimeError("oracle .opme read failed (code %d)" % rc)
        return op_p, op_n, j.value, t.value

    def apply_onebody(self, op_p, op_n, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().bso_apply_onebody(self.h, np.ascontiguousarray(op_p).ctypes.data,
                                np.ascontiguousarray(op_n).ctypes.data, x.ctypes.data, y.ctypes.data)
        return y

    def port(self):
        """Factorized FP32 CPU port (OpenMP) sharing this oracle's tables."""
        return Port(self)

    def dense(self):
        D = self.dim
        H = np.zeros((D, D))
        for i in range(D):
            e = np.zeros(D)
            e[i] = 1.0
            H[:, i] = self.apply(e)
        return H


def read_hamdat(path):
    """ham.dat written by the reference's 'wh' option (src/blanczosmain.f90:2253-2294):
    first line D, then `i j H_ij` (1-based, i<=j)."""
    with open(path) as f:
        D = int(f.readline())
        data = np.loadtxt(f)
    return D, data


def read_basis_dump(path, nsps):
    """The reference's .basis dump (src/boutputlib.f90:473-700): for every basis
    state its proton and neutron SD indices and occupied labels (protons
    1..nsps, neutrons nsps+1..2nsps).  Returns per-species dicts SD index ->
    mask and the (state, pSD, nSD) triples, all 0-based."""
    pm, nm, triples = {}, {}, []

    def records(f):
        # A record longer than the reference's format wraps by Fortran format
        # reversion: the continuation line repeats the leading text
        # ("State:", ", pSD:", ", nSD:") around the remaining labels.
        cur = None
        for line in f:
            if line.startswith("State:") and "Occupied:" in line:
                if cur is not None:
                    yield cur
                cur = line.strip()
            elif line.startswith("State:") and cur is not None:
                extra = line.replace("State:", " ").replace(", pSD:", " ").replace(", nSD:", " ")
                cur = cur + " " + extra.strip()
            elif cur is not None and line.strip() and "!" not in line:
                cur = cur + " " + line.strip()
        if cur is not None:
            yield cur

    with open(path) as f:
        for line in records(f):
            head, occ = line.split("Occupied:")
            parts = head.replace(",", " ").split()
            s, p, n = int(parts[1]) - 1, int(parts[3]) - 1, int(parts[5]) - 1
            labels = [int(t) for t in occ.split()]
            pmask = 0
            nmask = 0
            for lab in labels:
                if lab <= nsps:
                    pmask |= 1 << (lab - 1)
                else:
                    nmask |= 1 << (lab - 1 - nsps)
            pm[p] = pmask
            nm[n] = nmask
            triples.append((s, p, n))
    return pm, nm, np.array(triples, dtype=np.int64)


class Port:
    """The factorized CPU restatement of the reference's apply (bso_port_*)."""

    def __init__(self, oracle):
        self.o = oracle
        self.h = lib().bso_port_create(oracle.h)
        self.nP = lib().bso_nsd(oracle.h, 0)

    def __del__(self):
        if getattr(self, "h", None):
            lib().bso_port_free(self.h)
            self.h = None

    def count(self, lo=0, hi=None):
        return lib().bso_port_count(self.h, lo, self.nP if hi is None else hi)

    def apply(self, x, lo=0, hi=None):
        """FP32 y = H x for proton rows lo..hi (others left zero)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros_like(x)
        n = lib().bso_port_apply(self.h, x.ctypes.data, y.ctypes.data, lo,
                                 self.nP if hi is None else hi)
        return y, n
