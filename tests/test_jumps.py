"""The engine's jump lists against the reference's own.

bs_jumps exports every one- and two-body operator application on the Slater
determinants of one species, as the engine builds them (isd -> fsd, op,
phase).  The reference's print_jumps dump (write_out_jumps,
src/boutputlib.f90:1258-1351, switched on in its build by oracle/ref_hooks.c;
tests/golden/jumps_*.npz, written by tests/golden/make_golden.py) holds the
jumps it stores:

* proton one-body jumps with Jz(initial) >= Jz(final), neutron one-body jumps
  with Jz(initial) <= Jz(final) -- the reference keeps one triangle of sector
  jumps and applies the other as its transpose (PN forward / backward);
* two-body jumps (always sector-diagonal) with initial SD >= final SD.

On those triangles the two lists are equal as multisets of (initial SD, final
SD), bit for bit, on every system the dump can be written for (N >= 2: the
dump routine prints the neutron two-body op array beside each jump and faults
when it was never allocated).  The dump prints no phases and its op column is
that neutron array whatever the block, so ops and phases are pinned instead by
second quantization: every exported jump, applied to its initial SD's
occupation, gives its final SD with its phase.  No GPU needed."""
import ctypes
import os

import numpy as np
import pytest

from bigstickpublick_amd.engine import System, load_library
from bigstickpublick_amd.inputs import config_case, write_case
from golden_cases import GOLD, paths, small_fixture

CASES = ["sd2p2n", "sd1p2n_m1", "sd0p3n_m3", "pf2p2n", "config0"]


def _system(name, casedir):
    if name == "config0":
        case = config_case(0)
        d = os.path.join(casedir, "jcfg0")
        write_case(case, d)
    else:
        case, d, _ = small_fixture(name, casedir)
    return System(*paths(case, d), case.Z, case.N, case.jz2), case


def _m2(s):
    L = load_library()
    n = ctypes.c_int()
    L.bs_num_sps(s.h, ctypes.byref(n))
    o, j, m = (np.zeros(n.value, np.int32) for _ in range(3))
    L.bs_sps(s.h, o.ctypes.data, j.ctypes.data, m.ctypes.data)
    return m


def _jz(masks, m2):
    bits = (masks[:, None] >> np.arange(len(m2), dtype=np.uint64)[None, :]) & np.uint64(1)
    return (bits.astype(np.int64) * m2[None, :]).sum(axis=1)


@pytest.mark.parametrize("name", CASES)
def test_jumps_equal_reference_dump(name, casedir):
    s, case = _system(name, casedir)
    fx = dict(np.load(os.path.join(GOLD, "jumps_%s.npz" % name)))
    m2 = _m2(s)
    nonempty = 0
    for it, sp in ((0, "p"), (1, "n")):
        masks, _ = s.basis(it)
        jz = _jz(masks, m2)
        for order in (1, 2):
            key = "%s%d" % (sp, order)
            isd, fsd, op, ph = s.jumps(it, order)
            if order == 1:
                keep = jz[isd] >= jz[fsd] if it == 0 else jz[isd] <= jz[fsd]
            else:
                keep = isd >= fsd
            ours = np.stack([isd[keep] + 1, fsd[keep] + 1], axis=1)
            if key not in fx:
                # the reference builds one-body jumps only with both species
                # present, two-body jumps only with two or more particles
                assert len(ours) == 0 or (order == 1 and (case.Z == 0 or case.N == 0)), key
                continue
            ref = fx[key]
            ours = ours[np.lexsort((ours[:, 1], ours[:, 0]))]
            ref = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
            assert ours.shape == ref.shape, (key, ours.shape, ref.shape)
            assert np.array_equal(ours, ref), key
            nonempty += 1
    assert nonempty >= 1
    s.close()


def _fsign(mask, c):
    return -1 if bin(mask & ((1 << c) - 1)).count("1") & 1 else 1


@pytest.mark.parametrize("name", ["sd2p2n", "pf2p2n", "sd1p2n_m1"])
def test_jump_ops_and_phases_by_second_quantization(name, casedir):
    s, case = _system(name, casedir)
    ns = len(_m2(s))
    npair = ns * (ns - 1) // 2
    for it in (0, 1):
        masks, _ = s.basis(it)
        index = {int(m): i for i, m in enumerate(masks)}
        for order in (1, 2):
            isd, fsd, op, ph = s.jumps(it, order)
            for q in range(len(isd)):
                m = int(masks[isd[q]])
                o = int(op[q])
                if order == 1:
                    a, c = divmod(o, ns)
                    assert m >> c & 1
                    sg = _fsign(m, c)
                    m &= ~(1 << c)
                    assert not m >> a & 1
                    sg *= _fsign(m, a)
                    m |= 1 << a
                else:
                    cp, dp = divmod(o, npair)
                    i = int((1 + np.sqrt(1 + 8 * cp)) / 2)
                    j = cp - i * (i - 1) // 2
                    k = int((1 + np.sqrt(1 + 8 * dp)) / 2)
                    l_ = dp - k * (k - 1) // 2
                    assert i > j and k > l_
                    sg = 1
                    for c in (k, l_):  # a_l a_k: k first
                        assert m >> c & 1
                        sg *= _fsign(m, c)
                        m &= ~(1 << c)
                    for a in (j, i):  # then a+_j, a+_i
                        assert not m >> a & 1
                        sg *= _fsign(m, a)
                        m |= 1 << a
                assert index[m] == fsd[q], (it, order, q)
                assert sg == ph[q], (it, order, q)
    s.close()
