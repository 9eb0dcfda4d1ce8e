"""GPU tests at the sizes of the benchmark workloads (the bench's own
generators and seeds), beyond the small fixture cases:

* the 256^3 permeability pack (random sequential addition of radius-8
  spheres to lattice solid fraction 0.35): sphere centres equal to an
  independent restatement of the generator, solid map and boundary-link
  mask equal to the reference library's (SHA-256 digests written by the
  reference from the same centres), and FP32 vs FP64 permeability at
  residual 1e-6 within 1e-5;
* the 128^3 passive-scalar pack (radius 6, solid fraction 0.25) at Peclet
  10 with adsorption 0.01: 20 moment-propagation steps of the GPU against
  the CPU oracle at full size (FP64, 1e-12), and 5000 steps conserving
  fluid + adsorbed tracer (FP64), with FP32 within its stated tolerance;
* the 512-box batch of the resistance workload (6 right-hand sides per
  box, one launch per step): every 6x6 resistance matrix symmetric, 8
  boxes x 6 right-hand sides against the single-lattice path, 2 boxes
  against the live CPU oracle (FP64, 1e-12).
"""
import ctypes as C
import gzip
import hashlib
import json
import os
import time

import numpy as np
import pytest

import bench
from conftest import GOLDEN, rel_l2

pytestmark = pytest.mark.gpu

SEED = 20261202          # bench.py's default seed
ETA = 72.0 * (0.75 - 0.5) / 3.0


def lib():
    import sunlightlb_amd as s
    return s.lib()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def log(msg):
    print(msg, flush=True)


# ---------------------------------------------------------------- 256^3 pack

@pytest.mark.parametrize("flat", [False, True])
def test_pack256_tables_match_reference(flat):
    """The 256^3 benchmark pack (slb_rsa_pack + flattenobj, or
    slb_rsa_pack_flat with the sphere images counted and rasterised on the
    GPU) against the reference library's tables of the same centres."""
    with gzip.open(os.path.join(GOLDEN, "pack256.json.gz"), "rt") as f:
        rec = json.load(f)
    import sunlightlb_amd as s
    L = lib()
    lat = s.Lattice(256, 256, 256, precision=32)
    h = lat.h
    cap = 4096
    cen = np.zeros(3 * cap)
    phi = C.c_double(0.0)
    pack = L.slb_rsa_pack_flat if flat else L.slb_rsa_pack
    n = pack(h, 8.0, 0.35, SEED, 50_000_000, cen.ctypes.data_as(C.POINTER(C.c_double)), cap,
             C.byref(phi))
    log(f"pack256: {n} spheres, phi {phi.value:.5f}")
    assert n == rec["spheres"]
    assert np.array_equal(cen[:3 * n].reshape(n, 3), np.array(rec["centres"]))
    assert phi.value == rec["solid"] / 256 ** 3
    if not flat:
        L.flattenobj(h)
    L.syncobjs(h)
    solid = lat.solid()
    assert int(solid.sum()) == rec["solid"]
    assert sha(np.where(solid, 0, -1)) == rec["flat_objid_sha256"]
    assert sha(lat.mask()) == rec["flat_mask_sha256"]
    assert L.countfluid(h) == rec["countfluid"]
    assert L.countmask(h, b"links") == rec["countmask_links"]
    lat.free()


@pytest.mark.parametrize("dims,R,phi", [((128, 96, 112), 6.0, 0.3), ((64, 64, 64), 9.0, 0.33),
                                        ((32, 40, 36), 18.0, 0.6)])
def test_rsa_pack_flat_equals_pack_and_flatten(dims, R, phi):
    """slb_rsa_pack_flat against slb_rsa_pack + flattenobj on a lattice that
    already holds an object (a wall, merged into the flattened object by
    both): same sphere count, centres, solid fraction, node map, object
    position and next id.  In the last case the radius exceeds half the box,
    so the periodic images of one sphere overlap each other."""
    import sunlightlb_amd as s
    L = lib()
    out = []
    for flat in (False, True):
        lat = s.Lattice(*dims, precision=32)
        h = lat.h
        L.insertwall(h, b"y", 5)
        cen = np.zeros(3 * 4096)
        ph = C.c_double(0.0)
        pack = L.slb_rsa_pack_flat if flat else L.slb_rsa_pack
        n = pack(h, R, phi, SEED + 3, 5_000_000, cen.ctypes.data_as(C.POINTER(C.c_double)),
                 4096, C.byref(ph))
        if not flat:
            L.flattenobj(h)
        objs = lat.objects()
        out.append((n, cen[:3 * n].copy(), ph.value, lat.solid(), len(objs),
                    (objs[0].posx, objs[0].posy, objs[0].posz), lat.s.next_id))
        lat.free()
    a, b = out
    assert a[0] == b[0] and a[0] >= (1 if R > min(dims) / 2 else 10)
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]
    assert np.array_equal(a[3], b[3])
    assert a[4] == b[4] == 1
    assert a[5] == b[5]
    assert a[6] == b[6]


def solve_pack256(L, prec):
    os.environ["SLB_PRECISION"] = str(prec)
    try:
        h, nsph, phi = bench.build_pack(L, 256, 8.0, 0.35, SEED, 1e-6)
    finally:
        os.environ.pop("SLB_PRECISION", None)
    r = C.c_double(0.0)
    t0 = time.perf_counter()
    steps = L.slb_run_to_residual(h, 1e-6, 100, 20000, C.byref(r))
    L.measure(h)
    k = h.contents.vmeanx * ETA / 1e-6
    log(f"pack256 fp{prec}: {steps} steps to residual {r.value:.3e} "
        f"({time.perf_counter() - t0:.2f} s), permeability {k:.10g}")
    L.slb_destroy(h)
    return steps, r.value, k


def test_pack256_fp32_permeability_within_1e5_of_fp64():
    """The north star's accuracy bar at full size: FP32 shifted populations
    against the FP64 path, both run to the velocity residual 1e-6, Darcy
    permeability k = vmean eta / F (examples/rcp_perm.pl:161-167)."""
    L = lib()
    s64, r64, k64 = solve_pack256(L, 64)
    s32, r32, k32 = solve_pack256(L, 32)
    assert 0.0 <= r64 < 1e-6 and 0.0 <= r32 < 1e-6
    assert abs(k32 - k64) / abs(k64) < 1e-5, (k32, k64)
    assert 1.0 < k64 < 10.0            # lattice units; a 0.35 pack of R 8


@pytest.mark.parametrize("prec", [32, 64])
def test_pack96_row_staged_step_bit_identical(prec, monkeypatch):
    """A 96^3 pack (rows of 96 nodes: the staged step with 4 whole rows per
    block, k_step_ldsr, chosen by default) against the plain k_step on the
    same pack: populations after 60 steps and the forces and mean velocity of
    measure() bit-identical."""
    L = lib()

    def run(lds):
        monkeypatch.setenv("SLB_STEP_LDS", lds)
        monkeypatch.setenv("SLB_PRECISION", str(prec))
        h, nsph, phi = bench.build_pack(L, 96, 3.0, 0.35, SEED, 1e-6)
        kern = L.slb_step_kernel(h).decode()
        L.slb_run(h, 60)
        from sunlightlb_amd import Lattice
        lat = Lattice.wrap(h)
        a = lat.pops()
        L.measure(h)
        out = (kern, a, np.stack([lat.field(n) for n in ("fx", "fy", "fz")]),
               (h.contents.vmeanx, h.contents.vmeany, h.contents.vmeanz))
        L.slb_destroy(h)
        return out

    ref, got = run("0"), run("1")
    assert ref[0] == "k_step" and got[0] == "k_step_ldsr"
    assert np.array_equal(ref[1], got[1])
    assert np.array_equal(ref[2], got[2])
    assert ref[3] == got[3]


# ---------------------------------------------------------- 128^3 MP workload

def mp_setup(L, prec, pads=0.01, peclet=10.0):
    """bench.py's passive-scalar workload: flow of the 128^3 pack solved to
    residual 1e-6, scaled to the Peclet number, uniform unit tracer on fluid
    nodes, partial adsorption, displacement moments along x."""
    os.environ["SLB_PRECISION"] = str(prec)
    try:
        h, nsph, phi = bench.build_pack(L, 128, 6.0, 0.25, SEED + 1, 1e-6)
    finally:
        os.environ.pop("SLB_PRECISION", None)
    r = C.c_double(0.0)
    L.slb_run_to_residual(h, 1e-6, 100, 20000, C.byref(r))
    L.measure(h)
    porosity = L.countfluid(h) / 128.0 ** 3
    U = h.contents.vmeanx / porosity       # interstitial mean speed
    D = 1.0 / 6.0                          # (1 - alpha)/6, alpha = 0
    L.slb_scale_pops(h, peclet * D / (2.0 * 6.0) / U)
    mp = L.newmp(h, b"null")
    mp.contents.alpha = 0.0
    L.mpinit(mp)
    L.mpall(mp, 1.0)
    L.slb_mp_set_adsorption(mp, pads)
    L.slb_mp_track_moments(mp, 0)
    return h, mp


@pytest.mark.parametrize("factors", ["0", "1"])
def test_mp128_matches_oracle_at_full_size(oracle, factors, monkeypatch):
    """20 moment-propagation steps at 128^3 (src/mpkernel.c:47-368 with the
    partial-adsorption rule and displacement moments) against the CPU
    oracle on the same populations, density and tables: FP64, 1e-12; the
    gather of pulled populations and the link factors on compact records."""
    monkeypatch.setenv("SLB_MP_WEIGHTS", factors)
    import sunlightlb_amd as s
    from test_gpu_parity import _dense_objid_fast
    L = lib()
    h, mp = mp_setup(L, 64)
    lat = s.Lattice.wrap(h)
    a = lat.pops().reshape(-1)
    objid = _dense_objid_fast(lat)
    mask = lat.mask()
    L.slb_mp_run(mp, 20)
    out = (C.c_double * 6)()
    L.slb_mp_sums(mp, out)
    L.slb_mp_sync_host(mp)
    p_gpu = s.view(mp.contents.p, lat.N).copy()
    dims = (128, 128, 128)
    t0 = time.perf_counter()
    rho = oracle.mprhoinit(a, objid)
    p = (objid < 0).astype(np.float64)
    m1, m2 = np.zeros(lat.N), np.zeros(lat.N)
    for _ in range(20):
        oracle.mpstep(a, mask, objid, rho, dims, p, alpha=0.0, pads=0.01, m1=m1, m2=m2, axis=0)
    log(f"mp128: oracle 20 steps in {time.perf_counter() - t0:.1f} s")
    fl = objid < 0
    exp = [p[fl].sum(), p[~fl].sum(), m1[fl].sum(), m2[fl].sum()]
    np.testing.assert_allclose([out[0], out[1], out[3], out[4]], exp, rtol=1e-12)
    assert rel_l2(p_gpu, p) < 1e-12
    L.slb_destroy(h)


def mp_run5000(L, prec):
    h, mp = mp_setup(L, prec)
    s0 = (C.c_double * 6)()
    L.slb_mp_sums(mp, s0)
    L.slb_mp_run(mp, 5000)
    s1 = (C.c_double * 6)()
    L.slb_mp_sums(mp, s1)
    L.slb_destroy(h)
    ptot = s1[0]
    mean = s1[3] / ptot
    return dict(total0=s0[0] + s0[1], fluid=s1[0], solid=s1[1], surviving=s1[0] / s0[0],
                mean=mean, var=s1[4] / ptot - mean * mean)


def test_mp128_5000_steps_conserve_and_fp32_within_tolerance():
    """5000 steps of the workload: fluid + adsorbed tracer is conserved (the
    partial-adsorption rule moves what it removes from the fluid onto the
    solid node), FP64 to 1e-10 relative; the FP32 run's surviving fraction
    and displacement moments within their stated tolerances of FP64."""
    L = lib()
    r64 = mp_run5000(L, 64)
    r32 = mp_run5000(L, 32)
    log(f"mp128 5000 steps fp64 {r64}\nfp32 {r32}")
    tot64 = r64["fluid"] + r64["solid"]
    assert abs(tot64 - r64["total0"]) / r64["total0"] < 1e-10
    assert 0.05 < r64["surviving"] < 0.9
    assert abs(r32["surviving"] - r64["surviving"]) / r64["surviving"] < 1e-4
    assert abs(r32["mean"] - r64["mean"]) / abs(r64["mean"]) < 1e-5
    assert abs(r32["var"] - r64["var"]) / r64["var"] < 1e-4


# ---------------------------------------------------------- 512-box batch

U0, TAU = 1e-3, 1.0


def boxes512():
    return bench.make_boxes(512, SEED + 7)


def test_batch512_resistance_matrices_symmetric():
    """The resistance workload as the bench runs it (512 boxes x 6
    right-hand sides, FP32 populations, FP64 force sums, 232 steps +
    measure): every 6x6 resistance matrix is symmetric (Stokes-flow
    reciprocity) and its translational diagonal is a drag."""
    L = lib()
    boxes = boxes512()
    t0 = time.perf_counter()
    b, rows = bench.build_batch(L, 32, boxes, U0, TAU)
    assert L.slb_batch_commit(b) == 0
    log(f"batch512: set up in {time.perf_counter() - t0:.1f} s, "
        f"{L.slb_batch_nodes(b)} nodes per step")
    L.slb_batch_run(b, 232)
    L.slb_batch_measure(b)
    mats = bench.resistance_matrices(L, b, rows, U0)
    L.slb_batch_free(b)
    asym = np.array([np.abs(m - m.T).max() / np.abs(m).max() for m in mats])
    log(f"batch512: asymmetry median {np.median(asym):.3e} max {asym.max():.3e}")
    assert len(mats) == 512
    assert asym.max() < 1e-5          # measured 3.5e-7
    assert all((np.diag(m)[:3] < 0).all() for m in mats)   # force opposes the motion


def box_lattice(L, Lb, sph, vel, prec):
    """One box and right-hand side as a single lattice: objects as
    bench.build_batch inserts them, velocities from vel (list order),
    vspread."""
    import sunlightlb_amd as s
    lat = s.Lattice(Lb, Lb, Lb, precision=prec)
    h = lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.objects_moving = TAU, 1, 1
    for q in range(len(sph) - 1, -1, -1):
        bench.insert_periodic(L, h, Lb, sph[q][0], sph[q][1:])
    L.syncobjs(h)
    objs = lat.objects()
    for o, row in zip(objs, vel):
        o.velx, o.vely, o.velz, o.angx, o.angy, o.angz = row
    L.vspread(h)
    return lat


def single_lattice_forces(L, Lb, sph, vel, steps, prec):
    """The box run alone: steps, measure; returns the lattice, object forces
    and torques."""
    lat = box_lattice(L, Lb, sph, vel, prec)
    h = lat.h
    lat.step(steps)
    L.measure(h)
    f = np.array([[o.frcx, o.frcy, o.frcz] for o in lat.objects()])
    t = np.array([[o.trqx, o.trqy, o.trqz] for o in lat.objects()])
    return lat, f, t


def batch_subset(L, boxes, prec, steps):
    b, rows = bench.build_batch(L, prec, boxes, U0, TAU)
    assert L.slb_batch_commit(b) == 0
    L.slb_batch_run(b, steps)
    L.slb_batch_measure(b)
    return b, rows


def test_batch512_boxes_match_single_lattices():
    """8 of the workload's boxes, all 6 right-hand sides, batched (FP64)
    against the single-lattice path run alone: populations, object forces
    and torques to 1e-12."""
    L = lib()
    boxes = [boxes512()[q] for q in (0, 37, 101, 202, 303, 404, 470, 511)]
    steps = 40
    b, rows = batch_subset(L, boxes, 64, steps)
    try:
        for q, (Lb, sph) in enumerate(boxes):
            mrows, nobj, R0 = rows[q]
            vel = bench.rhs_velocities(nobj, mrows, R0, U0)
            for k in range(6):
                lat, f_ref, t_ref = single_lattice_forces(L, Lb, sph, vel[k], steps, 64)
                a_ref = lat.pops().reshape(-1)
                lat.free()
                frc, trq = np.zeros(3 * nobj), np.zeros(3 * nobj)
                L.slb_batch_forces(b, q, k, frc.ctypes.data_as(C.POINTER(C.c_double)),
                                   trq.ctypes.data_as(C.POINTER(C.c_double)))
                a = np.zeros(15 * Lb ** 3)
                L.slb_batch_get_pops(b, q, k, a.ctypes.data_as(C.POINTER(C.c_double)))
                assert rel_l2(a, a_ref) < 1e-12, (q, k)
                assert rel_l2(frc, f_ref) < 1e-12, (q, k)
                assert rel_l2(trq, t_ref) < 1e-12, (q, k)
    finally:
        L.slb_batch_free(b)


def test_batch512_boxes_match_live_oracle(oracle):
    """2 of the workload's boxes, batched (FP64), against the CPU oracle:
    moving-wall bounce-back with the vspread wall velocity of every image
    object (src/lbkernel.c:330-454, src/lbobjs.c:565-589), then measure()'s
    per-object force and torque (src/lbkernel.c:464-618,
    src/lbmeasure.c:236-263), 1e-12."""
    from test_gpu_parity import dense_objid, objrows
    L = lib()
    pick = (5, 260)
    boxes = [boxes512()[q] for q in pick]
    steps = 30
    b, rows = batch_subset(L, boxes, 64, steps)
    try:
        for q, (Lb, sph) in enumerate(boxes):
            mrows, nobj, R0 = rows[q]
            vel = bench.rhs_velocities(nobj, mrows, R0, U0)
            for k in (0, 4):                   # a translation and a rotation
                lat = box_lattice(L, Lb, sph, vel[k], 64)
                dims = (Lb, Lb, Lb)
                objid = dense_objid(lat)
                mask = oracle.makemask(objid, dims)
                assert (mask == lat.mask()).all()
                orow = objrows(lat)
                lat.free()
                u = oracle.vspread(objid, dims, orow[:, 0:3], vel[k][:, 0:3], vel[k][:, 3:6])
                zero = np.zeros(Lb ** 3)
                ao = np.zeros(15 * Lb ** 3)
                oracle.lbstep(ao, mask, objid, dims, [zero, zero, zero], TAU, 1, 0, u=u,
                              nsteps=steps)
                a = np.zeros(15 * Lb ** 3)
                res = oracle.measure(ao, mask, objid, dims, [zero, zero, zero], TAU, 1, 0,
                                     orow[:, 0:3], u=u)
                frc, trq = np.zeros(3 * nobj), np.zeros(3 * nobj)
                L.slb_batch_forces(b, q, k, frc.ctypes.data_as(C.POINTER(C.c_double)),
                                   trq.ctypes.data_as(C.POINTER(C.c_double)))
                L.slb_batch_get_pops(b, q, k, a.ctypes.data_as(C.POINTER(C.c_double)))
                assert rel_l2(a, ao) < 1e-12, (pick[q], k)      # after measure's 2 steps
                assert rel_l2(frc, res["frc"]) < 1e-12, (pick[q], k)
                assert rel_l2(trq, res["trq"]) < 1e-12, (pick[q], k)
                log(f"batch512 box {pick[q]} ({Lb}^3, {nobj} objects) rhs {k}: "
                    f"force {rel_l2(frc, res['frc']):.1e}")
    finally:
        L.slb_batch_free(b)


def test_mp128_peclet_scaling_closed_form():
    """The Peclet scaling of the passive-scalar workload (Pe = U 2R / D, U the
    interstitial mean speed, D = (1 - alpha)/6): without adsorption a
    uniform tracer's mean displacement grows as U t, so after scaling the
    flow to Pe 10 it is 10 D / (2R) per step (measured 2.7e-4 off after 1000
    steps: the tracer drifts with the populations' own momentum, U is the
    force-corrected mean; bound 1e-3)."""
    L = lib()
    h, mp = mp_setup(L, 64, pads=0.0)
    L.slb_mp_track_moments(mp, 0)
    L.slb_mp_run(mp, 1000)
    s1 = (C.c_double * 6)()
    L.slb_mp_sums(mp, s1)
    L.slb_destroy(h)
    mean = s1[3] / s1[0]
    expect = 10.0 * (1.0 / 6.0) / 12.0 * 1000
    log(f"mp128 Pe 10, no adsorption: mean displacement {mean:.6f}, U t = {expect:.6f}")
    assert abs(mean - expect) < 1e-3 * expect


# ---------------------------------------------------------- learned warm start

WEIGHTS = os.path.join(os.path.dirname(GOLDEN), "..", "sunlightlb_amd", "weights",
                       "warm_unet128.bin")


def test_committed_weights_solve_unseen_packs():
    """The committed network (bench.py --warm-only --warm-save, trained on
    2000 GPU-solved 128^3 packs) loaded with slb_warm_load: two unseen 128^3
    packs solved with slb_warm_solve (start policy and fallback) reach
    residual 1e-6 with the cold solve's permeability within 1e-5; without
    the safeguards the warm start reaches 1e-4 in fewer steps than the cold
    solve (the gain the network gives at moderate tolerances)."""
    import numpy as np
    L = lib()
    net = L.slb_warm_load(os.path.abspath(WEIGHTS).encode())
    assert net, "weights file missing or refused"
    params = bench.warm_params(np.random.default_rng(SEED + 1234), 2)
    ratios = []
    for (solid, phi), p in zip(bench.warm_maps(L, 128, params, 4), params):
        res = {}
        for mode in ("cold", "warm", "warm_trace"):
            h = bench.warm_lattice(L, 128, solid, 1e-6)
            if mode == "cold":
                steps, hist, r = bench.solve_loop(L, h, 1e-6, 100, 40000)
            elif mode == "warm":
                rr, uw = C.c_double(0.0), C.c_int(0)
                steps = L.slb_warm_solve(net, h, 1e-6, 100, 40000, C.byref(rr), C.byref(uw))
                r = rr.value
            else:
                L.slb_warm_set_mode(net, 1)
                L.slb_warm_start(net, h)
                steps, hist, r = bench.solve_loop(L, h, 1e-4, 100, 40000)
            L.measure(h)
            res[mode] = (steps, r, h.contents.vmeanx)
            if mode == "cold":
                cold_hist = hist
            L.slb_destroy(h)
        cold_to_1e4 = next((i + 1) * 100 for i, x in enumerate(cold_hist) if 0 <= x < 1e-4)
        log(f"pack R {p[0]:.2f} phi {phi:.3f}: cold {res['cold'][0]} steps, warm solve "
            f"{res['warm'][0]}, to 1e-4 cold {cold_to_1e4} warm {res['warm_trace'][0]}")
        assert 0 <= res["warm"][1] < 1e-6 and 0 <= res["cold"][1] < 1e-6
        kc, kw = res["cold"][2], res["warm"][2]
        assert abs(kw - kc) < 1e-5 * abs(kc)
        ratios.append(cold_to_1e4 / res["warm_trace"][0])
    assert min(ratios) > 1.0, ratios
    L.slb_warm_free(net)
