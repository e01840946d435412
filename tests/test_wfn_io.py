"""Reading the reference's .wfn files back (bs_init_wfn_host, bs_read_wfn) on
the CPU: the reference's own files, partial reads, both record precisions,
write -> read round trips and the reader's refusals."""
import os
import shutil
import struct

import numpy as np
import pytest

from bigstickpublick_amd.engine import System
from bigstickpublick_amd.inputs import small_case, write_case
from wfn_io import read_wfn

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WFN = [("sd2p2n", "sd", 2, 2, 0, 77), ("pf2p1n_m3", "pf", 2, 1, 3, 83)]


def _case(name, space, Z, N, jz2, seed, d):
    case = small_case(space, Z, N, jz2, seed=seed)
    case.name = name
    write_case(case, d)
    return os.path.join(d, space + ".sps"), os.path.join(d, name + ".int")


@pytest.mark.parametrize("name,space,Z,N,jz2,seed", WFN)
def test_read_reference_wfn(name, space, Z, N, jz2, seed, tmp_path):
    """Every state of the reference's file comes back bit for bit, through a
    system built from the .sps and through one built from the file itself."""
    sps, intf = _case(name, space, Z, N, jz2, seed, str(tmp_path))
    path = os.path.join(GOLD, name + ".wfn")
    ref = read_wfn(path)
    s = System(sps, intf, Z, N, jz2)
    w = System.from_wfn(path, intf)
    assert w.dim == s.dim == ref["dim"]
    for it in (0, 1):
        for a, b in zip(s.basis(it), w.basis(it)):
            assert np.array_equal(a, b)
    for sysm in (s, w):
        assert sysm.wfn_count(path) == ref["nkeep"]
        e, j, t, v = sysm.read_wfn(path)
        for q, st in enumerate(ref["states"]):
            assert e[q] == st["e"] and j[q] == st["xj"] and t[q] == st["xt2"]
            assert np.array_equal(v[q], st["vec"])
    # a middle slice, and FP64 output (exact widening of the stored FP32)
    e, j, t, v = w.read_wfn(path, first=2, count=2, dtype="f64")
    assert v.dtype == np.float64 and v.shape == (2, ref["dim"])
    for q in range(2):
        assert e[q] == ref["states"][2 + q]["e"]
        assert np.array_equal(v[q], ref["states"][2 + q]["vec"].astype(np.float64))


def _wfn_f64_copy(src, dst):
    """The same file with each vector stored in double precision (the layout
    a reference built with 8-byte Lanczos vectors writes)."""
    ref = read_wfn(src)
    with open(dst, "wb") as f:
        f.write(ref["header"])
        f.write(struct.pack("<i", ref["nkeep"]))
        for st in ref["states"]:
            f.write(struct.pack("<2i3f", st["magic"], st["index"], st["e"], st["xj"], st["xt2"]))
            f.write((st["vec"].astype(np.float64) * (1 + 2.0 ** -30)).astype("<f8").tobytes())
    return ref


def test_read_double_precision_records(tmp_path):
    sps, intf = _case(*WFN[0], str(tmp_path))
    p = str(tmp_path / "d.wfn")
    ref = _wfn_f64_copy(os.path.join(GOLD, "sd2p2n.wfn"), p)
    w = System.from_wfn(p, intf)
    _, _, _, v64 = w.read_wfn(p, dtype="f64")
    _, _, _, v32 = w.read_wfn(p, dtype="f32")
    for q, st in enumerate(ref["states"]):
        exact = st["vec"].astype(np.float64) * (1 + 2.0 ** -30)
        assert np.array_equal(v64[q], exact)
        assert np.array_equal(v32[q], exact.astype(np.float32))


@pytest.mark.parametrize("vec_dtype", [np.float32, np.float64])
def test_write_then_read_round_trip(vec_dtype, tmp_path):
    sps, intf = _case(*WFN[1], str(tmp_path))
    s = System(sps, intf, 2, 1, 3)
    rng = np.random.default_rng(3)
    nk = 3
    vecs = rng.standard_normal((nk, s.dim)).astype(vec_dtype)
    e = np.array([-1.5, -1.25, 0.5], np.float32)
    j = np.array([1.5, 2.5, 3.5], np.float32)
    t = np.array([0.75, 0.75, 3.75], np.float32)
    p = str(tmp_path / "rt.wfn")
    s.write_wfn(p, e, j, t, vecs)
    ge, gj, gt, gv = System.from_wfn(p, intf).read_wfn(p, dtype="f64" if vec_dtype == np.float64 else "f32")
    assert np.array_equal(ge, e) and np.array_equal(gj, j) and np.array_equal(gt, t)
    # the file holds single precision: FP64 vectors come back rounded once
    assert np.array_equal(gv, vecs.astype(np.float32).astype(vec_dtype))


def test_reader_refusals(tmp_path):
    sps, intf = _case(*WFN[0], str(tmp_path))
    src = os.path.join(GOLD, "sd2p2n.wfn")
    s = System(sps, intf, 2, 2, 0)
    ref = read_wfn(src)
    rec = 20 + 4 * ref["dim"]
    off = len(ref["header"]) + 4

    def bad(mutate, msg, reader=None):
        p = str(tmp_path / "bad.wfn")
        b = bytearray(open(src, "rb").read())
        b = mutate(b)
        open(p, "wb").write(bytes(b))
        with pytest.raises(RuntimeError, match=msg):
            (reader or (lambda q: s.read_wfn(q)))(p)

    def put(b, at, v):
        b[at:at + 4] = struct.pack("<i", v)
        return b

    bad(lambda b: put(b, 0, 12345), "bad magic number")
    bad(lambda b: put(b, off + 2 * rec, 7), "vector 3 is missing magic number")
    bad(lambda b: put(b, off + rec + 4, 5), "read wrong vector")
    bad(lambda b: b[:-8], "file size does not match")
    bad(lambda b: b, "states 5..6 requested, file holds 5", lambda q: s.read_wfn(q, first=4, count=2))
    # a file on another basis: other 2Jz, other nucleus
    other = System(sps, intf, 2, 2, 2)
    with pytest.raises(RuntimeError, match="is not this system's basis"):
        other.read_wfn(src)
    with pytest.raises(RuntimeError, match="is not this system's basis"):
        System(sps, intf, 3, 1, 0).read_wfn(src)
    with pytest.raises(RuntimeError, match="cannot open"):
        s.read_wfn(str(tmp_path / "missing.wfn"))
    # an .int whose orbits do not match the file's basis
    d2 = str(tmp_path / "pf")
    _, pf_int = _case(*WFN[1], d2)
    with pytest.raises(RuntimeError):
        System.from_wfn(src, pf_int)
    shutil.rmtree(d2)


def test_empty_wfn_round_trip(tmp_path):
    """A .wfn with no states: written, counted as 0, and a read of one state
    refused."""
    sps, intf = _case(*WFN[0], str(tmp_path))
    s = System(sps, intf, 2, 2, 0)
    p = str(tmp_path / "empty.wfn")
    s.write_wfn(p, [], [], [], np.zeros((0, s.dim), np.float32))
    assert s.wfn_count(p) == 0
    assert System.from_wfn(p).wfn_count(p) == 0
    with pytest.raises(RuntimeError, match="file holds 0"):
        s.read_wfn(p, first=0, count=1)
