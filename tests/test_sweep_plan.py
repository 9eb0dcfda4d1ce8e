"""The launch plan of the two-step sweep (slbd_sweep_plan), checked on the
host without a GPU: a simulation of which step of which plane every buffer
slot holds shows that each launch reads exactly the values the one-step
kernel would (step t-1 around a plane for phase 1, step t for phase 2),
writes only slots inside the second buffer's plane window, and that after
the plan the live buffer holds step t+1 on every plane.  The arithmetic per
node is the one-step kernel's, so a correct order makes the sweep
bit-identical to two single steps (checked on the GPU in
test_gpu_sweep.py)."""
import ctypes as C
import os

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "sunlightlb_amd", "libsunlightlb.so")


def plan(esi, P):
    L = C.CDLL(LIB)
    L.slbd_sweep_plan.restype = C.c_int
    L.slbd_sweep_plan.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]
    n = L.slbd_sweep_plan(esi, P, None, 0)
    rows = (C.c_int * (5 * max(n, 1)))()
    assert L.slbd_sweep_plan(esi, P, rows, n) == n
    return [tuple(rows[5 * q:5 * q + 5]) for q in range(n)]


def slot(i, ring):
    """The kernel's plane slot (plane_slot in slb_dev.hip), restated from its
    definition: planes 0 and 1 fixed, the rest cycling through ring slots."""
    return i if (ring == 0 or i < 2) else 2 + (i - 2) % ring


def simulate(esi, P, rows):
    A = {i: ("A", i, 0) for i in range(esi)}      # live buffer: plane i holds step 0
    B = {}                                        # second buffer: slot -> (plane, step)
    done = {1: set(), 2: set()}
    for phase, i0, np_, sring, dring in rows:
        planes = [(i0 + q) % esi for q in range(np_)]
        assert np_ >= 1
        writes = {}
        for i in planes:
            for d in (-1, 0, 1):
                ii = (i + d) % esi
                if phase == 1:
                    assert sring == 0 and A[ii] == ("A", ii, 0), (phase, i, ii, A[ii])
                else:
                    s = slot(ii, sring)
                    assert 0 <= s < esi
                    assert B.get(s) == (ii, 1), (phase, i, ii, s, B.get(s))
            if phase == 1:
                s = slot(i, dring)
                assert dring > 0 and s < dring + 2 <= esi
                writes[("B", s)] = (i, 1)
            else:
                assert dring == 0
                writes[("A", i)] = ("A", i, 2)
            assert i not in done[phase], f"plane {i} computed twice in phase {phase}"
            done[phase].add(i)
        # a launch reads one buffer and writes the other: apply after the reads
        for (buf, key), v in writes.items():
            if buf == "A":
                A[key] = v
            else:
                B[key] = v
    assert all(A[i] == ("A", i, 2) for i in range(esi))
    assert done[1] == done[2] == set(range(esi))


@pytest.mark.parametrize("P", [4, 5, 7, 8, 16, 24, 32])
def test_sweep_plan_dependencies(P):
    for esi in range(2 * P + 2, 2 * P + 2 + 3 * P + 5):
        simulate(esi, P, plan(esi, P))
    for esi in (256, 257, 300, 512):
        if esi >= 2 * P + 2:
            simulate(esi, P, plan(esi, P))


def test_sweep_plan_declines_small_lattices():
    assert plan(9, 4) == []        # needs esi >= 2P + 2
    assert plan(64, 3) == []       # chunks of fewer than 4 planes


def test_simulation_catches_a_misordered_plan():
    """The checker itself: a plan whose closing launch runs before the last
    chunk, or whose window is one plane too small, must fail."""
    rows = plan(40, 8)
    bad = [rows[-1]] + rows[:-1]
    with pytest.raises(AssertionError):
        simulate(40, 8, bad)
    small = [(ph, i0, n, s - 1 if s else 0, d - 1 if d else 0) for ph, i0, n, s, d in rows]
    with pytest.raises(AssertionError):
        simulate(40, 8, small)
