"""Product host tables (libbigstick_amd.so, no GPU needed) against the oracle:
basis order and FP32 matrix elements bit-identical; the exported C ABI covers
every function declared in include/bigstick_amd.h."""
import ctypes
import os
import re

import numpy as np
import pytest

from bigstickpublick_amd.engine import LIB_PATH, System
from golden_cases import SMALL, config_fixture, paths, small_fixture
from oracle_api import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "bigstick_amd.h")).read()
    names = set(re.findall(r"^\s*(?:int|const char \*)\s*\**\s*(bs_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 12
    lib = ctypes.CDLL(LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n


@pytest.mark.parametrize("name", SMALL)
def test_tables_match_oracle(name, casedir):
    case, d, fx = small_fixture(name, casedir)
    sps, intf = paths(case, d)
    s = System(sps, intf, case.Z, case.N, case.jz2)
    o = Oracle(sps, intf, case.Z, case.N, case.jz2)
    assert s.dim == o.dim == int(fx["dim"])
    for it in (0, 1):
        m, st = s.basis(it)
        om, ost, _ = o.sds(it)
        np.testing.assert_array_equal(m, om)
        np.testing.assert_array_equal(st, ost)
    for kind in (0, 1, 2):
        a = s.me_table(kind)
        b = o.me_table(kind)
        assert a.shape == b.shape
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), "FP32 bits differ"


@pytest.mark.parametrize("cfg", [0, 1, 3])
def test_config_tables(cfg, casedir):
    case, d, fx = config_fixture(cfg, casedir)
    s = System(*paths(case, d), case.Z, case.N, case.jz2)
    assert s.dim == int(fx["dim"])
    m, st = s.basis(0)
    np.testing.assert_array_equal(m, fx["pmask"])
    np.testing.assert_array_equal(st, fx["pstart"])
    m, st = s.basis(1)
    np.testing.assert_array_equal(m, fx["nmask"])
    np.testing.assert_array_equal(st, fx["nstart"])
    c = s.op_counts()
    assert c["pp"] > 0 and c["nn"] > 0 and c["pn"] > 0
    # the reference's printed operation counts (src/bigstick_main.f90:1150-1192)
    assert c["ref_pp"] == int(fx["nops_pp"])
    assert c["ref_nn"] == int(fx["nops_nn"])
    assert c["ref_pn"] == int(fx["nops_pn"])
    assert c["ref_total"] == int(fx["nops_total"])


def test_bad_inputs_raise(casedir):
    case, d, fx = small_fixture("sd2p2n", casedir)
    sps, intf = paths(case, d)
    with pytest.raises(RuntimeError):
        System(sps + ".missing", intf, 2, 2, 0)
    with pytest.raises(RuntimeError):
        System(sps, intf, 2, 2, 1)  # 2Jz parity mismatch
    bad = intf.replace(".int", "_badT.int")
    lines = open(intf).read().splitlines()
    p = lines[2].split()
    p[5] = "2"
    lines[2] = " ".join(p)
    open(bad, "w").write("\n".join(lines) + "\n")
    with pytest.raises(RuntimeError):
        System(sps, bad, 2, 2, 0)


@pytest.mark.parametrize("name,Z,N,jz2,seed", [("sd2p2n", 2, 2, 0, 77), ("pf2p1n_m3", 2, 1, 3, 83)])
def test_wfn_writer_reproduces_reference_file(name, Z, N, jz2, seed, casedir, tmp_path):
    """bs_write_wfn, given the reference's own energies, J, T(T+1) and vectors,
    writes the reference's .wfn byte for byte (header from this engine's
    tables: orbits, |m|-sorted s.p. states, 2Jz, dimension)."""
    from bigstickpublick_amd import System
    from bigstickpublick_amd.inputs import small_case, write_case
    from wfn_io import read_wfn
    ref = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".wfn")
    r = read_wfn(ref)
    assert r["end"] == r["size"] and all(s["magic"] == 27182818 for s in r["states"])
    case = small_case("sd" if name.startswith("sd") else "pf", Z, N, jz2, seed=seed)
    case.name = name
    d = os.path.join(casedir, "wfn_" + name)
    write_case(case, d)
    s = System(os.path.join(d, case.space + ".sps"), os.path.join(d, name + ".int"), Z, N, jz2)
    assert s.dim == r["dim"]
    out = str(tmp_path / (name + ".wfn"))
    st = r["states"]
    s.write_wfn(out, [x["e"] for x in st], [x["xj"] for x in st], [x["xt2"] for x in st],
                np.stack([x["vec"] for x in st]))
    assert open(out, "rb").read() == open(ref, "rb").read()


def test_entry_points_refuse_bad_arguments(casedir, tmp_path):
    """Argument checks that return before any device work: a nonzero code and
    a message, never a crash."""
    import ctypes
    from bigstickpublick_amd.engine import BS_F32, load_library
    L = load_library()
    case, d, fx = small_fixture("sd2p2n", casedir)
    sps, intf = paths(case, d)
    s = System(sps, intf, 2, 2, 0)  # host tables only
    # device entry points on a system that was never uploaded
    assert L.bs_apply_block(s.h, None, None, 4, BS_F32) != 0
    assert L.bs_multi_variant(s.h, BS_F32, 4, 1, 1) != 0
    # batches: K out of range, missing names
    h = ctypes.c_void_p()
    names = (ctypes.c_char_p * 1)(intf.encode())
    assert L.bs_batch_init(sps.encode(), names, 0, 2, 2, 0, 1, 1, 1, 1, ctypes.byref(h)) != 0
    assert L.bs_batch_init(sps.encode(), names, 65, 2, 2, 0, 1, 1, 1, 1, ctypes.byref(h)) != 0
    nulls = (ctypes.c_char_p * 2)(intf.encode(), None)
    assert L.bs_batch_init(sps.encode(), nulls, 2, 2, 2, 0, 1, 1, 1, 1, ctypes.byref(h)) != 0
    assert b"null" in L.bs_last_error()
    assert L.bs_batch_apply(None, None, None, BS_F32) != 0
    # .wfn: bad dtype, unwritable path
    v = np.zeros((1, s.dim), np.float32)
    e = np.zeros(1, np.float32)
    assert L.bs_write_wfn(s.h, str(tmp_path / "a.wfn").encode(), 1, e.ctypes.data, e.ctypes.data,
                          e.ctypes.data, v.ctypes.data, 3, 0) != 0
    assert L.bs_write_wfn(s.h, str(tmp_path / "no" / "such" / "dir.wfn").encode(), 1, e.ctypes.data,
                          e.ctypes.data, e.ctypes.data, v.ctypes.data, BS_F32, 0) != 0
    assert b"cannot open" in L.bs_last_error()
    # strength function: thick restart, bad dtype, null pivot, no upload
    from bigstickpublick_amd.engine import _Params
    ev = np.zeros(4)
    pn = ctypes.c_double()
    it = ctypes.c_int()
    for prm, piv, msg in ((_Params(2, 10, 4, 1, BS_F32, 0.0, 1), v.ctypes.data, b"thick = 0"),
                          (_Params(2, 10, 0, 0, 3, 0.0, 1), v.ctypes.data, b"dtype"),
                          (_Params(2, 10, 0, 0, BS_F32, 0.0, 1), None, b"null"),
                          (_Params(2, 10, 0, 0, BS_F32, 0.0, 1), v.ctypes.data, b"not uploaded")):
        assert L.bs_strength(s.h, ctypes.byref(prm), piv, ctypes.byref(pn), None, None, None, None,
                             ev.ctypes.data, None, None, ctypes.byref(it)) != 0
        assert msg in L.bs_last_error(), L.bs_last_error()
    # block Lanczos: block size out of range, bad dtype, no upload
    for blk, prm, msg in ((0, _Params(2, 10, 0, 0, BS_F32, 0.0, 1), b"block size"),
                          (9, _Params(2, 10, 0, 0, BS_F32, 0.0, 1), b"block size"),
                          (2, _Params(2, 10, 0, 0, 5, 0.0, 1), b"dtype"),
                          (2, _Params(2, 10, 0, 0, BS_F32, 0.0, 1), b"not uploaded")):
        assert L.bs_block_lanczos(s.h, ctypes.byref(prm), blk, ev.ctypes.data, None, None, ctypes.byref(it)) != 0
        assert msg in L.bs_last_error(), L.bs_last_error()
    # block strength: no pivots, block size, pivot dtype, no upload
    piv = np.zeros(4, np.float32)
    for blk, pv, pdt, msg in ((2, None, BS_F32, b"null argument"), (9, piv.ctypes.data, BS_F32, b"block size"),
                              (2, piv.ctypes.data, 7, b"pivot dtype"), (2, piv.ctypes.data, BS_F32, b"not uploaded")):
        prm = _Params(2, 10, 0, 0, BS_F32, 0.0, 0)
        assert L.bs_block_strength(s.h, ctypes.byref(prm), blk, pv, pdt, None, None, ev.ctypes.data, None, None,
                                   ctypes.byref(it)) != 0
        assert msg in L.bs_last_error(), L.bs_last_error()


def test_config4_operation_count(tmp_path):
    """pf 6p6n M=0: the host tables (no GPU) reproduce the reference's own
    printed operation counts per type and in total (tests/golden/config4.npz,
    written from its option-li run; src/bigstick_main.f90:1150-1192)."""
    from bigstickpublick_amd.inputs import config_case, write_case
    case = config_case(4)
    write_case(case, str(tmp_path))
    fx = dict(np.load(os.path.join(ROOT, "tests", "golden", "config4.npz")))
    s = System(*paths(case, str(tmp_path)), case.Z, case.N, case.jz2)
    assert s.dim == 109954620
    c = s.op_counts()
    assert c["ref_total"] == int(fx["nops_total"]) == 152771809712
    for k in ("pp", "nn", "pn"):
        assert c["ref_" + k] == int(fx["nops_" + k]), k
    s.close()
