"""GPU parity of the batched path: every (box, right-hand side) problem of a
batch equals the same problem run alone through the single-lattice API, FP64
to 1e-12 relative (populations, object forces and torques); FP32 to 1e-5 (the
batched kernel forms the wall term c.U in FP32 from per-object rows, the
single-lattice kernel in FP64 from the per-node vspread field)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

BOXES = [
    dict(dims=(10, 12, 9), tau=0.8, spheres=[(2.5, 4.2, 5.1, 4.0), (2.0, 7.5, 2.5, 6.0)], wall=None),
    dict(dims=(16, 11, 13), tau=0.95, spheres=[(3.0, 8.0, 5.0, 6.5)], wall=("z", 0)),
    dict(dims=(8, 8, 8), tau=0.7, spheres=[(2.2, 3.5, 3.5, 3.5)], wall=None),
]
NRHS = 3


def build(box, prec=64):
    import sunlightlb_amd as s
    lat = s.Lattice(*box["dims"], precision=prec)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.objects_moving = box["tau"], 1, 1
    if box["wall"]:
        L.insertwall(h, box["wall"][0].encode(), box["wall"][1])
    for sp in box["spheres"]:
        L.insertsph(h, *sp)
    L.syncobjs(h)
    return lat


def velocities(nobj, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.01, (NRHS, nobj, 6))


def run_alone(box, vel, steps, prec=64):
    lat = build(box, prec)
    objs = lat.objects()
    for o, row in zip(objs, vel):
        o.velx, o.vely, o.velz, o.angx, o.angy, o.angz = row
    lat.L.vspread(lat.h)
    lat.step(steps)
    lat.L.measure(lat.h)
    f = np.array([[o.frcx, o.frcy, o.frcz] for o in lat.objects()])
    t = np.array([[o.trqx, o.trqy, o.trqz] for o in lat.objects()])
    a = lat.pops()
    lat.free()
    return a, f, t


def make_batch(L, prec=64):
    b = L.slb_batch_new(prec)
    vels = []
    for q, box in enumerate(BOXES):
        lat = build(box, prec)
        nobj = len(lat.objects())
        vel = velocities(nobj, 100 + q)
        vels.append(vel)
        flat = np.ascontiguousarray(vel.reshape(-1))
        assert L.slb_batch_add(b, lat.h, NRHS, flat.ctypes.data_as(C.POINTER(C.c_double))) == q
        lat.free()
    assert L.slb_batch_commit(b) == 0
    return b, vels


def check_batch(L, b, vels, steps, tol=1e-12, prec=64):
    L.slb_batch_run(b, steps)
    L.slb_batch_measure(b)
    for q, box in enumerate(BOXES):
        N = int(np.prod(box["dims"]))
        for r in range(NRHS):
            a_ref, f_ref, t_ref = run_alone(box, vels[q][r], steps, prec)
            nobj = len(f_ref)
            frc = np.zeros(3 * nobj)
            trq = np.zeros(3 * nobj)
            got = L.slb_batch_forces(b, q, r, frc.ctypes.data_as(C.POINTER(C.c_double)),
                                     trq.ctypes.data_as(C.POINTER(C.c_double)))
            assert got == nobj
            a = np.zeros(15 * N)
            L.slb_batch_get_pops(b, q, r, a.ctypes.data_as(C.POINTER(C.c_double)))
            assert rel_l2(a, a_ref) < tol, (q, r)
            assert rel_l2(frc, f_ref) < tol, (q, r)
            assert rel_l2(trq, t_ref) < tol, (q, r)


@pytest.mark.parametrize("steps", [40, 70])
def test_batch_matches_single_lattices(slb, steps):
    b, vels = make_batch(slb)
    try:
        check_batch(slb, b, vels, steps)
    finally:
        slb.slb_batch_free(b)


@pytest.mark.parametrize("cfg", [(256, 1, 0), (128, 2, 8), (512, 4, 0)])
def test_batch_fp32_matches_single_lattices(slb, cfg):
    b, vels = make_batch(slb, prec=32)
    try:
        slb.slb_batch_set_config(b, *cfg)
        check_batch(slb, b, vels, 45, tol=1e-5, prec=32)
    finally:
        slb.slb_batch_free(b)


@pytest.mark.parametrize("cfg", [(64, 1, 0), (128, 2, 8), (512, 4, 0), (256, 1, 3)])
def test_batch_launch_shapes_agree(slb, cfg):
    b, vels = make_batch(slb)
    try:
        slb.slb_batch_set_config(b, *cfg)
        got = (C.c_int * 4)()
        slb.slb_batch_get_config(b, got)
        assert tuple(got[:3]) == cfg
        check_batch(slb, b, vels, 35)
    finally:
        slb.slb_batch_free(b)


def test_batch_autotune_cache(slb, tmp_path):
    path = str(tmp_path / "tune.txt").encode()
    b, vels = make_batch(slb)
    try:
        assert slb.slb_batch_autotune(b, path) == 0
        cfg = (C.c_int * 4)()
        slb.slb_batch_get_config(b, cfg)
        assert cfg[0] in (64, 128, 256, 512) and cfg[1] in (1, 2, 4)
        check_batch(slb, b, vels, 33)      # tuning left the state at rest
    finally:
        slb.slb_batch_free(b)
    b2, _ = make_batch(slb)
    try:
        assert slb.slb_batch_autotune(b2, path) == 1   # same shape class: cache hit
        cfg2 = (C.c_int * 4)()
        slb.slb_batch_get_config(b2, cfg2)
        assert tuple(cfg2[:3]) == tuple(cfg[:3])
    finally:
        slb.slb_batch_free(b2)
    assert len(open(path.decode()).read().strip().splitlines()) == 1


def test_batch_graph_replay_after_odd_run(slb):
    """A captured step loop replays the buffers it was captured with: after
    an odd number of plain steps the other buffer is live, and a replay
    must use the loop captured from that buffer.  33 + 32 steps equal 65
    steps in one call and the single-lattice run."""
    b1, vels = make_batch(slb)
    b2, _ = make_batch(slb)
    try:
        slb.slb_batch_run(b1, 33)      # one graph replay + one plain step
        slb.slb_batch_run(b1, 32)      # replay from the other buffer
        slb.slb_batch_measure(b1)
        slb.slb_batch_run(b2, 65)
        slb.slb_batch_measure(b2)
        for q, box in enumerate(BOXES):
            N = int(np.prod(box["dims"]))
            for r in range(NRHS):
                a1, a2 = np.zeros(15 * N), np.zeros(15 * N)
                slb.slb_batch_get_pops(b1, q, r, a1.ctypes.data_as(C.POINTER(C.c_double)))
                slb.slb_batch_get_pops(b2, q, r, a2.ctypes.data_as(C.POINTER(C.c_double)))
                assert np.array_equal(a1, a2), (q, r)
                a_ref, f_ref, t_ref = run_alone(box, vels[q][r], 65)
                assert rel_l2(a1, a_ref) < 1e-12, (q, r)
                nobj = len(f_ref)
                frc, trq = np.zeros(3 * nobj), np.zeros(3 * nobj)
                slb.slb_batch_forces(b1, q, r, frc.ctypes.data_as(C.POINTER(C.c_double)),
                                     trq.ctypes.data_as(C.POINTER(C.c_double)))
                assert rel_l2(frc, f_ref) < 1e-12 and rel_l2(trq, t_ref) < 1e-12, (q, r)
    finally:
        slb.slb_batch_free(b1)
        slb.slb_batch_free(b2)


@pytest.mark.parametrize("cfg", [(256, 1, 0), (128, 4, 8)])
def test_batch_against_reference_fixture(slb, cfg):
    """The batched kernel against the reference library directly (no
    single-lattice run in between): the moving sphere + moving wall case of
    the golden fixture (src/lbkernel.c:330-454 moving-wall bounce-back,
    :464-618 its force), populations after 40 steps and object force and
    torque after measure(), FP64 to 1e-12.  A second right-hand side with
    every velocity doubled must give exactly doubled results (Stokes flow
    from rest is linear in the wall velocities)."""
    from conftest import load_golden
    g = load_golden("moving.npz")
    dims = tuple(g["meta"]["dims"])
    import sunlightlb_amd as s
    lat = s.Lattice(*dims, precision=64)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.skip_solid, lat.s.objects_moving = 0.9, 1, 0, 1
    wall = L.insertwall(h, b"z", 0)
    sph = L.insertsph(h, 3.1, 6.3, 5.4, 6.2)
    L.syncobjs(h)
    objs = lat.objects()                     # list order: sphere (head), wall
    assert len(objs) == 2 and objs[0].posx == 6.3
    v1 = np.zeros((2, 6))
    v1[0] = [0.012, -0.004, 0.0, 0.0, 0.003, 0.002]
    v1[1] = [-0.005, 0.0, 0.0, 0.0, 0.0, 0.0]
    vel = np.stack([v1, 2.0 * v1])
    b = L.slb_batch_new(64)
    try:
        flat = np.ascontiguousarray(vel.reshape(-1))
        assert L.slb_batch_add(b, h, 2, flat.ctypes.data_as(C.POINTER(C.c_double))) == 0
        lat.free()
        assert L.slb_batch_commit(b) == 0
        L.slb_batch_set_config(b, *cfg)
        L.slb_batch_run(b, 40)
        N = int(np.prod(dims))
        for r, scale in ((0, 1.0), (1, 2.0)):
            a = np.zeros(15 * N)
            L.slb_batch_get_pops(b, 0, r, a.ctypes.data_as(C.POINTER(C.c_double)))
            assert rel_l2(a, scale * g["a_40"]) < 1e-12, r
        L.slb_batch_measure(b)
        for r, scale in ((0, 1.0), (1, 2.0)):
            frc, trq = np.zeros(6), np.zeros(6)
            assert L.slb_batch_forces(b, 0, r, frc.ctypes.data_as(C.POINTER(C.c_double)),
                                      trq.ctypes.data_as(C.POINTER(C.c_double))) == 2
            ref = g["meas_objs"]
            assert rel_l2(frc, scale * ref[:, 9:12]) < 1e-12, r
            assert rel_l2(trq, scale * ref[:, 12:15]) < 1e-12, r
    finally:
        L.slb_batch_free(b)
