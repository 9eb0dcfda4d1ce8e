"""The two-step sweep (two time steps per pass over the planes, the
intermediate step held in a sliding plane window) against the one-step
kernel: populations bit-identical after odd and even step counts, with
moving walls, skip_solid, Navier-Stokes and per-node body forces, in FP64
and FP32; and FP64 against the CPU oracle.  The sweep is chosen per call
from SLB_SWEEP (planes per chunk; 0 = one launch per step)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def build(dims, prec, moving=True, nonuniform=True):
    import sunlightlb_amd as s
    lat = s.Lattice(*dims, precision=prec)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.skip_solid = 0.7, 0, 1
    lat.s.objects_moving = 1 if moving else 0
    w = L.insertwall(h, b"y", 2)
    s1 = L.insertsph(h, 5.3, 17.2, 9.5, 12.7)
    s2 = L.insertsph(h, 4.1, 3.0, 20.1, 30.2)        # cut by the i = 0 face
    if moving:
        s1.contents.velz, s1.contents.angx = 0.02, 0.003
        s2.contents.velx = -0.01
        w.contents.velx = 0.005
        L.vspread(h)
    L.zerobf(h)
    if nonuniform:
        rng = np.random.default_rng(11)
        lat.sync()
        solid = lat.solid()
        for nm, sc in (("bx", 2e-4), ("by", 5e-5), ("bz", 1e-4)):
            v = s.view(getattr(lat.s, nm), lat.N)
            v[:] = np.where(solid, 0.0, sc * (1 + rng.random(lat.N)))
        L.slb_sync_device(h)
    else:
        L.addbf(h, 1e-4, 0.0, 2e-5)
    return lat


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("planes", [4, 5, 8])
def test_sweep_bit_identical_to_single_steps(prec, planes, monkeypatch):
    dims = (37, 32, 32)          # 1024-node planes; 37 = 4 chunks of 8 + 5
    print(f"sweep: prec {prec} planes {planes} dims {dims}", flush=True)
    monkeypatch.setenv("SLB_SWEEP", "0")
    ref = build(dims, prec)
    ref.step(1)
    ref.L.slb_run(ref.h, 41)
    a_ref = ref.pops()
    monkeypatch.setenv("SLB_SWEEP", str(planes))
    sw = build(dims, prec)
    sw.step(1)
    sw.L.slb_run(sw.h, 41)
    a_sw = sw.pops()
    assert sw.s.step == ref.s.step == 42
    assert np.array_equal(a_sw, a_ref), (
        f"max diff {np.abs(a_sw - a_ref).max():.3e}")
    # measure() after the sweep: the force-collecting steps run one at a
    # time; per-object sums use FP64 atomics (summation order varies at the
    # last bit), node forces are bit-identical
    ref.L.measure(ref.h)
    sw.L.measure(sw.h)
    for nm in ("fx", "fy", "fz"):
        assert np.array_equal(ref.field(nm), sw.field(nm))
    for o1, o2 in zip(ref.objects(), sw.objects()):
        f1 = np.array([o1.frcx, o1.frcy, o1.frcz, o1.trqx, o1.trqy, o1.trqz])
        f2 = np.array([o2.frcx, o2.frcy, o2.frcz, o2.trqx, o2.trqy, o2.trqz])
        assert np.abs(f1 - f2).max() <= 1e-13 * np.abs(f1).max()
    ref.free()
    sw.free()


@pytest.mark.parametrize("prec", [64, 32])
def test_sweep_uniform_force_static(prec, monkeypatch):
    dims = (24, 32, 64)
    monkeypatch.setenv("SLB_SWEEP", "0")
    ref = build(dims, prec, moving=False, nonuniform=False)
    ref.L.slb_run(ref.h, 30)
    monkeypatch.setenv("SLB_SWEEP", "6")
    sw = build(dims, prec, moving=False, nonuniform=False)
    sw.L.slb_run(sw.h, 30)
    assert np.array_equal(sw.pops(), ref.pops())
    ref.free()
    sw.free()


def test_sweep_fp64_matches_oracle(oracle, monkeypatch):
    """The sweep in FP64 against the CPU oracle (Navier-Stokes, skip_solid,
    moving walls, per-node body force)."""
    monkeypatch.setenv("SLB_SWEEP", "4")
    dims = (20, 32, 32)
    lat = build(dims, 64)
    lat.L.slb_run(lat.h, 24)
    from test_gpu_parity import dense_objid, objrows
    objid = dense_objid(lat)
    mask = oracle.makemask(objid, dims)
    rows = objrows(lat)
    u = oracle.vspread(objid, dims, rows[:, 0:3], rows[:, 3:6], rows[:, 6:9])
    bb = [lat.field("bx"), lat.field("by"), lat.field("bz")]
    ao = np.zeros(lat.N * 15)
    oracle.lbstep(ao, mask, objid, dims, bb, 0.7, 0, 1, u=u, nsteps=24)
    assert rel_l2(lat.pops(), ao) < 1e-12
    lat.free()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("dims", [(12, 16, 64), (9, 8, 128), (6, 6, 256), (1, 1, 512), (3, 2, 256),
                                  (2, 1024, 512)])
def test_lds_staged_step_bit_identical(prec, dims, monkeypatch):
    """k_step_lds (the k pulls staged in LDS; rows of 64-512 nodes; planes of
    2^19 nodes and more in the XCD-contiguous block order) against k_step: populations after odd and even step counts and the node forces
    of measure(), moving walls, Navier-Stokes, per-node body force."""
    import sunlightlb_amd as s

    def run(lds):
        monkeypatch.setenv("SLB_STEP_LDS", lds)
        lat = s.Lattice(*dims, precision=prec)
        L, h = lat.L, lat.h
        lat.s.tau, lat.s.stokes_flow, lat.s.objects_moving = 0.7, 0, 1
        sph = L.insertsph(h, 3.2, dims[0] / 2, dims[1] / 2, 5.5)
        sph2 = L.insertsph(h, 2.5, 1.0, 2.0, dims[2] - 2.0)     # wraps in k
        sph.contents.velx, sph.contents.angz = 0.01, 0.002
        sph2.contents.velz = -0.02
        L.vspread(h)
        L.zerobf(h)
        lat.sync()
        rng = np.random.default_rng(4)
        solid = lat.solid()
        for nm, sc in (("bx", 2e-4), ("by", 5e-5), ("bz", 1e-4)):
            s.view(getattr(lat.s, nm), lat.N)[:] = np.where(solid, 0.0, sc * (1 + rng.random(lat.N)))
        L.slb_sync_device(h)
        L.slb_run(h, 17)
        a = lat.pops()
        L.measure(h)
        f = np.stack([lat.field(n) for n in ("fx", "fy", "fz")])
        out = (a, lat.pops(), f)
        lat.free()
        return out

    ref, got = run("0"), run("1")
    for r, g in zip(ref, got):
        assert np.array_equal(r, g)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("dims", [(5, 96, 96), (3, 10, 100), (2, 6, 160), (4, 8, 200), (7, 5, 48),
                                  (3, 7, 97), (2, 3, 33), (2, 4, 64), (2, 5, 300), (6, 32, 32),
                                  (2, 3, 1), (1, 1, 7), (2, 2, 2), (2, 1, 3)])
def test_row_staged_step_bit_identical(prec, dims, monkeypatch):
    """k_step_ldsr (the k pulls staged in LDS for rows of any length up to
    512 nodes, as many whole rows per block as divide the plane; forced on
    every shape with SLB_STEP_LDS=2, including rows of 1-3 nodes whose k
    pulls wrap onto the node itself) against k_step: populations after odd
    and even step counts and the node forces of measure(), moving walls,
    Navier-Stokes, per-node body force."""
    import sunlightlb_amd as s

    def run(mode):
        monkeypatch.setenv("SLB_STEP_LDS", mode)
        lat = s.Lattice(*dims, precision=prec)
        L, h = lat.L, lat.h
        lat.s.tau, lat.s.stokes_flow, lat.s.objects_moving = 0.7, 0, 1
        rng = np.random.default_rng(sum(dims))
        for q in range(6):
            r = float(rng.uniform(0.6, 3.5))
            c = [float(rng.uniform(-1.0, n)) for n in dims]
            sph = L.insertsph(h, r, *c)
            if sph:
                sph.contents.velx, sph.contents.vely = 0.01 * (q + 1), -0.004
                sph.contents.velz, sph.contents.angz = 0.006, 0.002
        L.vspread(h)
        L.zerobf(h)
        lat.sync()
        solid = lat.solid()
        for nm, sc in (("bx", 2e-4), ("by", 5e-5), ("bz", 1e-4)):
            s.view(getattr(lat.s, nm), lat.N)[:] = np.where(solid, 0.0, sc * (1 + rng.random(lat.N)))
        L.slb_sync_device(h)
        lat.set_pops(1e-3 * rng.standard_normal(lat.N * 15))
        L.slb_run(h, 17)
        a = lat.pops()
        L.measure(h)
        f = np.stack([lat.field(n) for n in ("fx", "fy", "fz")])
        out = (a, lat.pops(), f)
        lat.free()
        return out

    ref, got = run("0"), run("2")
    for r, g in zip(ref, got):
        assert np.array_equal(r, g)


@pytest.mark.parametrize("dims,want", [
    ((16, 256, 256), "k_step_lds"),      # rows of 256
    ((32, 96, 96), "k_step_ldsr"),       # 4 rows of 96 per block, blocks on 128-byte lines
    ((8, 200, 200), "k_step_ldsr"),      # 2 rows of 200, off the lines but cache-resident
    ((64, 200, 200), "k_step"),          # off the lines and streamed from HBM
    ((8, 96, 96), "k_step"),             # fewer than 2^18 nodes
    ((4, 4, 1024), "k_step"),            # rows longer than 512
    ((400, 7, 97), "k_step"),            # one row of 97 in a 128-thread block: 76% used
])
def test_step_kernel_policy(dims, want, monkeypatch):
    """The step kernel a lattice launches (slb_step_kernel), by row length,
    lattice size, alignment of the blocks and cache residency; SLB_STEP_LDS=0
    always gives k_step."""
    import sunlightlb_amd as s
    lat = s.Lattice(*dims, precision=32)
    assert lat.L.slb_step_kernel(lat.h).decode() == want
    lat.free()
    monkeypatch.setenv("SLB_STEP_LDS", "0")
    lat = s.Lattice(*dims, precision=32)
    assert lat.L.slb_step_kernel(lat.h).decode() == "k_step"
    lat.free()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("dims,stokes", [((12, 16, 64), 0), ((9, 8, 128), 1), ((6, 6, 256), 0),
                                         ((1, 1, 512), 0), ((3, 2, 256), 1)])
def test_moment_run_bit_identical(prec, dims, stokes, monkeypatch):
    """Runs on moment storage (SLB_STEP_MOM=1: the first step stores the 10
    post-collision moments, the middle steps pull populations rebuilt from
    the source nodes' moments, the last stores populations) against the
    population step: populations after runs of 17, 2 and 16 steps and the
    node forces of measure(), moving walls, Navier-Stokes or Stokes,
    per-node or uniform body force."""
    import sunlightlb_amd as s

    def run(mom):
        monkeypatch.setenv("SLB_STEP_MOM", mom)
        lat = s.Lattice(*dims, precision=prec)
        L, h = lat.L, lat.h
        lat.s.tau, lat.s.stokes_flow, lat.s.objects_moving = 0.7, stokes, 1
        sph = L.insertsph(h, 3.2, dims[0] / 2, dims[1] / 2, 5.5)
        sph2 = L.insertsph(h, 2.5, 1.0, 2.0, dims[2] - 2.0)     # wraps in k
        sph.contents.velx, sph.contents.angz = 0.01, 0.002
        sph2.contents.velz = -0.02
        L.vspread(h)
        L.zerobf(h)
        rng = np.random.default_rng(7)
        if stokes:
            L.addbf(h, 2e-5, -1e-5, 3e-5)
        else:
            lat.sync()
            solid = lat.solid()
            for nm, sc in (("bx", 2e-4), ("by", 5e-5), ("bz", 1e-4)):
                s.view(getattr(lat.s, nm), lat.N)[:] = np.where(solid, 0.0,
                                                                sc * (1 + rng.random(lat.N)))
            L.slb_sync_device(h)
        lat.set_pops(1e-3 * rng.standard_normal(lat.N * 15))
        out = []
        for n in (17, 2, 16):
            L.slb_run(h, n)
            out.append(lat.pops())
        L.measure(h)
        out.append(np.stack([lat.field(n) for n in ("fx", "fy", "fz")]))
        out.append(lat.pops())
        lat.free()
        return out

    ref, got = run("0"), run("1")
    for r, g in zip(ref, got):
        assert np.array_equal(r, g)
