"""GPU parity: the HIP path, called through the C API (libsunlightlb.so),
against the reference library's golden fixtures and the CPU oracle.

Bars: tables (solid map, boundary-link mask) bit-exact; FP64 populations,
forces, torques and mean velocities within 1e-12 relative (L2 over the
field); FP32 within the per-quantity tolerance written in each test.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, rel_l2

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def lattice(dims, prec=64):
    import sunlightlb_amd as s
    return s.Lattice(*dims, precision=prec)


def objrows(lat):
    return np.array([[o.posx, o.posy, o.posz, o.velx, o.vely, o.velz, o.angx, o.angy, o.angz,
                      o.frcx, o.frcy, o.frcz, o.trqx, o.trqy, o.trqz] for o in lat.objects()])


def dense_objid(lat):
    """Dense object index in list order (head = 0), -1 fluid, from obj[]."""
    s = lat.s
    addrs = []
    p = s.obj_head
    while p:
        addrs.append(C.addressof(p.contents))
        p = p.contents.next_obj
    o = np.ctypeslib.as_array(C.cast(s.obj, C.POINTER(C.c_uint64)), shape=(lat.N,))
    out = np.full(lat.N, -1, dtype=np.int32)
    for k, a in enumerate(addrs):
        out[o == a] = k
    return out


def build_static(g, prec=64):
    m = g["meta"]
    lat = lattice(tuple(m["dims"]), prec)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.skip_solid = m["tau"], m["stokes"], m["skip"]
    for w in m["walls"]:
        L.insertwall(h, w[0].encode(), int(w[1]))
    for sp in m["spheres"]:
        L.insertsph(h, *sp)
    L.zerobf(h)
    L.addbf(h, *m["bf"])
    return lat


@pytest.mark.parametrize("name", ["static_stokes", "static_ns_skip"])
def test_static_fp64_matches_reference(name):
    g = load_golden(f"{name}.npz")
    lat = build_static(g)
    L, h = lat.L, lat.h
    done = 0
    for k in g["meta"]["steps"]:
        lat.step(k - done)
        done = k
        assert rel_l2(lat.pops(), g[f"a_{k}"]) < TOL64, k
    assert (lat.mask() == g["mask"]).all()
    assert (dense_objid(lat) == g["objid"]).all()
    assert L.countfluid(h) == int(g["countfluid"])
    assert L.countmask(h, b"links") == int(g["countmask_links"])
    assert L.countmask(h, b"faces") == int(g["countmask_faces"])
    L.measure(h)
    vm = np.array([lat.s.vmeanx, lat.s.vmeany, lat.s.vmeanz])
    assert rel_l2(vm, g["meas_vmean"]) < TOL64
    for c in ("fx", "fy", "fz"):
        assert rel_l2(lat.field(c), g["meas_" + c]) < TOL64
    rows = objrows(lat)
    assert rel_l2(rows[:, 9:12], g["meas_objs"][:, 9:12]) < TOL64
    assert rel_l2(rows[:, 12:15], g["meas_objs"][:, 12:15]) < TOL64
    assert rel_l2(lat.pops(), g["a_after_measure"]) < TOL64
    assert abs(L.discrf(h) - float(g["discrf"])) < 1e-12
    assert lat.s.step == done + 2 and lat.s.status == 49


def test_tables_bit_exact():
    g = load_golden("tables.npz")
    lat = lattice(tuple(g["meta"]["dims"]))
    L, h = lat.L, lat.h
    placed = [bool(L.insertsph(h, *sp)) for sp in g["spheres"]]
    assert placed == list(g["placed"].astype(bool))
    assert (dense_objid(lat) == g["sph_objid"]).all()
    L.insertcyl(h, b"y", 2.6, 15.3, 4.7)
    L.insertcyl(h, b"zh", 9.5, 10.2, 8.1)
    L.insertslab(h, 1.5, 4.2, -1.0, 6.5, 17.3, 30.0)
    L.insertwall(h, b"x", 20)
    L.syncobjs(h)
    assert (dense_objid(lat) == g["objid"]).all()
    assert (lat.mask() == g["mask"]).all()
    assert lat.s.next_id == int(g["next_id"])
    np.testing.assert_array_equal(objrows(lat)[:, :9], g["objs"][:, :9])
    L.flattenobj(h)
    L.syncobjs(h)
    np.testing.assert_allclose(objrows(lat)[:, :3], g["flat_objs"][:, :3], rtol=1e-15)
    L.invertobj(h)
    L.syncobjs(h)
    assert (dense_objid(lat) == g["inv_objid"]).all()
    assert (lat.mask() == g["inv_mask"]).all()


def test_moving_walls_match_reference():
    g = load_golden("moving.npz")
    lat = lattice(tuple(g["meta"]["dims"]))
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.skip_solid, lat.s.objects_moving = 0.9, 1, 0, 1
    wall = L.insertwall(h, b"z", 0)
    sph = L.insertsph(h, 3.1, 6.3, 5.4, 6.2)
    sph.contents.velx, sph.contents.vely = 0.012, -0.004
    sph.contents.angy, sph.contents.angz = 0.003, 0.002
    wall.contents.velx = -0.005
    L.vspread(h)
    u = np.stack([lat.field("ux"), lat.field("uy"), lat.field("uz")])
    assert rel_l2(u, g["u"]) < 1e-15
    lat.step(40)
    assert rel_l2(lat.pops(), g["a_40"]) < TOL64
    L.measure(h)
    rows = objrows(lat)
    assert rel_l2(rows[:, 9:12], g["meas_objs"][:, 9:12]) < TOL64
    assert rel_l2(rows[:, 12:15], g["meas_objs"][:, 12:15]) < TOL64
    vm = np.array([lat.s.vmeanx, lat.s.vmeany, lat.s.vmeanz])
    assert rel_l2(vm, g["meas_vmean"]) < TOL64


@pytest.mark.parametrize("stepping", ["mpstep", "run", "run_factors"])
@pytest.mark.parametrize("variant", ["reflect", "adsorb", "moving"])
def test_mp_matches_reference(variant, stepping, monkeypatch):
    """25 MP steps against the reference library's fields: one mpstep call
    per step, one slb_mp_run call gathering pulled populations, or one whose
    steps share link factors computed once from the frozen flow and run on
    compact records (SLB_MP_WEIGHTS=1 forces that form for any run)."""
    monkeypatch.setenv("SLB_MP_WEIGHTS", "1" if stepping == "run_factors" else "0")
    g = load_golden("mp.npz")
    dims = tuple(g["meta"]["dims"])
    lat = lattice(dims)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.skip_solid = 0.8, 1, 0
    if variant == "moving":
        lat.s.objects_moving = 1
    sph = L.insertsph(h, 2.8, 5.1, 4.3, 5.4)
    if variant == "moving":
        sph.contents.velx = 0.01
        sph.contents.angz = 0.004
        L.vspread(h)
    L.zerobf(h)
    L.addbf(h, 2e-4, 5e-5, 0.0)
    lat.step(30)
    mp = L.newmp(h, b"null")
    mp.contents.alpha = 0.1
    L.mpinit(mp)
    L.mpall(mp, 1.0)
    if variant == "adsorb":
        L.adall(mp)
    assert rel_l2(lat.pops(), g[f"{variant}_a"]) < TOL64
    assert rel_l2(lat.field("rho"), g[f"{variant}_rho"]) < 1e-15
    if stepping == "mpstep":
        for _ in range(25):
            L.mpstep(mp)
    else:
        L.slb_mp_run(mp, 25)        # "run" and "run_factors"
    L.slb_mp_sync_host(mp)
    import sunlightlb_amd as s
    p = s.view(mp.contents.p, lat.N).copy()
    assert rel_l2(p, g[f"{variant}_p"]) < TOL64
    ad = L.adtotal(mp, b"null", 1)
    assert abs(ad - float(g[f"{variant}_adtotal"])) <= 1e-12 * max(1.0, abs(ad))
    assert mp.contents.step == 25


def test_lbloadvel_matches_reference(tmp_path):
    g = load_golden("loadvel.npz")
    dims = (7, 9, 8)
    path = str(tmp_path / "vel.bin")
    np.column_stack([g["rho"], g["u"]]).astype(np.float64).tofile(path)
    for stokes in (0, 1):
        lat = lattice(dims)
        lat.s.stokes_flow = stokes
        lat.L.insertsph(lat.h, 2.2, 3.0, 4.0, 3.5)
        lat.L.lbloadvel(lat.h, path.encode(), b"binary")
        assert rel_l2(lat.pops(), g[f"a_stokes{stokes}"]) < TOL64


def test_config1_fp64_and_fp32():
    """32^3 sphere, 2000 steps: FP64 drag and vmean within 1e-12 of the
    reference; FP32 (shifted populations, FP64 reductions) within 1e-5."""
    with open(os.path.join(GOLDEN, "config1.json")) as f:
        ref = json.load(f)
    for prec, tol in ((64, TOL64), (32, 1e-5)):
        lat = lattice((32, 32, 32), prec)
        L, h = lat.L, lat.h
        lat.s.stokes_flow = 1
        sph = L.insertsph(h, 4.5, 15.5, 15.5, 15.5)
        L.zerobf(h)
        L.addbf(h, 1e-5, 0.0, 0.0)
        lat.step(2000)
        L.measure(h)
        r = ref["fp64"]
        e_drag = abs(sph.contents.frcx - r["frcx"]) / abs(r["frcx"])
        e_v = abs(lat.s.vmeanx - r["vmean"][0]) / abs(r["vmean"][0])
        print(f"config1 fp{prec}: drag {sph.contents.frcx:.17g} rel.err {e_drag:.3e}; "
              f"vmean {lat.s.vmeanx:.17g} rel.err {e_v:.3e}")
        assert e_drag < tol and e_v < tol


def test_separate_phases_equal_fused_step():
    g = load_golden("static_stokes.npz")
    a = build_static(g)
    b = build_static(g)
    for _ in range(7):
        a.step(1)
        b.L.bbl(b.h)
        b.L.propagate(b.h)
        b.L.collide(b.h)
    assert rel_l2(a.pops(), b.pops()) < 1e-15
    assert a.s.step == b.s.step and b.s.status == 49


def test_against_live_oracle_ns_moving_skip(oracle):
    """A case no fixture covers: Navier-Stokes, skip_solid, moving sphere
    and wall, non-uniform per-node body force, on a lattice with a 2-wide
    axis.  GPU FP64 vs the CPU oracle."""
    dims = (16, 2, 19)
    lat = lattice(dims)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.skip_solid, lat.s.objects_moving = 0.66, 0, 1, 1
    w = L.insertwall(h, b"x", 3)
    s1 = L.insertsph(h, 3.4, 9.2, 0.5, 8.7)
    s1.contents.velz, s1.contents.angx, s1.contents.angy = 0.02, 0.003, -0.002
    w.contents.vely = 0.01
    L.vspread(h)
    rng = np.random.default_rng(5)
    N = lat.N
    L.zerobf(h)
    lat.sync()
    import sunlightlb_amd as s
    solid = lat.solid()
    for nm, sc in (("bx", 2e-4), ("by", 5e-5), ("bz", 1e-4)):
        v = s.view(getattr(lat.s, nm), N)
        v[:] = np.where(solid, 0.0, sc * (1 + rng.random(N)))
    L.slb_sync_device(h)
    lat.step(60)
    L.measure(h)
    # oracle on the same inputs
    objid = dense_objid(lat)
    mask = oracle.makemask(objid, dims)
    assert (lat.mask() == mask).all()
    rows = objrows(lat)
    u = oracle.vspread(objid, dims, rows[:, 0:3], rows[:, 3:6], rows[:, 6:9])
    bb = [lat.field("bx"), lat.field("by"), lat.field("bz")]
    ao = np.zeros(N * 15)
    oracle.lbstep(ao, mask, objid, dims, bb, 0.66, 0, 1, u=u, nsteps=60)
    res = oracle.measure(ao, mask, objid, dims, bb, 0.66, 0, 1, rows[:, 0:3], u=u)
    assert rel_l2(lat.pops(), ao) < TOL64
    assert rel_l2(rows[:, 9:12], res["frc"]) < TOL64
    assert rel_l2(rows[:, 12:15], res["trq"]) < TOL64
    vm = np.array([lat.s.vmeanx, lat.s.vmeany, lat.s.vmeanz])
    assert rel_l2(vm, res["vmean"]) < TOL64


@pytest.mark.parametrize("factors", ["0", "1"])
def test_mp_partial_adsorption_and_moments_against_oracle(oracle, factors, monkeypatch):
    """Partial adsorption 0.2 and displacement moments, 20 steps in one run:
    the gather of pulled populations (SLB_MP_WEIGHTS=0) and the link factors
    on compact records (SLB_MP_WEIGHTS=1) against the CPU oracle."""
    monkeypatch.setenv("SLB_MP_WEIGHTS", factors)
    g = load_golden("mp.npz")
    dims = tuple(g["meta"]["dims"])
    lat = lattice(dims)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow = 0.8, 1
    L.insertsph(h, 2.8, 5.1, 4.3, 5.4)
    L.zerobf(h)
    L.addbf(h, 2e-4, 5e-5, 0.0)
    lat.step(30)
    mp = L.newmp(h, b"null")
    mp.contents.alpha = 0.1
    L.mpinit(mp)
    L.mpall(mp, 1.0)
    L.slb_mp_set_adsorption(mp, 0.2)
    L.slb_mp_track_moments(mp, 0)
    L.slb_mp_run(mp, 20)
    out = (C.c_double * 6)()
    L.slb_mp_sums(mp, out)
    a = lat.pops().reshape(-1)
    objid = dense_objid(lat)
    mask = lat.mask()
    rho = oracle.mprhoinit(a, objid)
    p = (objid < 0).astype(np.float64)
    m1, m2 = np.zeros(lat.N), np.zeros(lat.N)
    for _ in range(20):
        oracle.mpstep(a, mask, objid, rho, dims, p, alpha=0.1, pads=0.2, m1=m1, m2=m2, axis=0)
    fl = objid < 0
    exp = [p[fl].sum(), p[~fl].sum(), m1[fl].sum(), m2[fl].sum()]
    got = [out[0], out[1], out[3], out[4]]
    np.testing.assert_allclose(got, exp, rtol=1e-12)
    L.slb_mp_sync_host(mp)
    import sunlightlb_amd as s
    assert rel_l2(s.view(mp.contents.p, lat.N), p) < TOL64


def test_fp32_static_error_is_small():
    g = load_golden("static_stokes.npz")
    lat = build_static(g, prec=32)
    lat.step(50)
    err = rel_l2(lat.pops(), g["a_50"])
    print(f"static_stokes fp32 population rel. L2 error after 50 steps: {err:.3e}")
    assert err < 1e-5


def test_checkpoint_round_trip(tmp_path):
    """lbsave then lbload reproduces the state up to the float storage of
    the moment file, and lbsavevel writes the reference text format."""
    g = load_golden("static_stokes.npz")
    lat = build_static(g)
    lat.step(20)
    f = str(tmp_path / "x.dms").encode()
    lat.L.lbsave(lat.h, f)
    a_saved = lat.pops()
    lat2 = build_static(g)
    lat2.L.lbload(lat2.h, f)
    assert rel_l2(lat2.pops(), a_saved) < 1e-6
    v = str(tmp_path / "x.vel").encode()
    lat.L.lbsavevel(lat.h, v, b"ascii")
    rows = open(v.decode()).read().split("\n")
    assert len(rows[0].split()) == 7 and rows[0].split()[4:] == ["0", "0", "0"]


def _packs():
    import gzip
    with gzip.open(os.path.join(GOLDEN, "packs.json.gz"), "rt") as f:
        return json.load(f)


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def _dense_objid_fast(lat):
    """dense_objid by a sorted lookup (packs have hundreds of objects)."""
    s = lat.s
    addrs = []
    p = s.obj_head
    while p:
        addrs.append(C.addressof(p.contents))
        p = p.contents.next_obj
    o = np.ctypeslib.as_array(C.cast(s.obj, C.POINTER(C.c_uint64)), shape=(lat.N,))
    out = np.full(lat.N, -1, dtype=np.int32)
    if not addrs:
        return out
    a = np.array(addrs, dtype=np.uint64)
    order = np.argsort(a)
    pos = np.minimum(np.searchsorted(a[order], o), len(a) - 1)
    hit = (o != 0) & (a[order][pos] == o)
    out[hit] = order[pos[hit]].astype(np.int32)
    assert (hit == (o != 0)).all()
    return out


@pytest.mark.parametrize("name", ["pack128", "pack_noncubic", "box0", "box1", "box2", "box3",
                                  "box4", "box5"])
def test_tables_at_workload_sizes_bit_exact(name):
    """Workload-size geometries through the C API: the 128^3 radius-6 pack of
    the passive-scalar workload, a non-cubic mixed-radius pack and boxes like
    the batched ones.  Solid map (host) and boundary-link mask (computed on
    the GPU) equal the reference library's bit for bit (SHA-256 of the int32
    tables), before and after flattenobj, with countfluid and countmask."""
    case = _packs()[name]
    rec = case["record"]
    lat = lattice(tuple(rec["dims"]))
    L, h = lat.L, lat.h
    placed = [1 if L.insertsph(h, *row) else 0 for row in case["inserts"]]
    assert placed == case["placed_list"]
    L.syncobjs(h)
    dense = _dense_objid_fast(lat)
    mask = lat.mask()
    assert int((dense >= 0).sum()) == rec["solid"]
    assert int((mask != 0).sum()) == rec["mask_nonzero"]
    assert _sha(dense) == rec["objid_sha256"]
    assert _sha(mask) == rec["mask_sha256"]
    assert L.countfluid(h) == rec["countfluid"]
    assert L.countmask(h, b"links") == rec["countmask_links"]
    assert L.countmask(h, b"faces") == rec["countmask_faces"]
    if "flat_mask_sha256" in rec:
        L.flattenobj(h)
        L.syncobjs(h)
        assert _sha(_dense_objid_fast(lat)) == rec["flat_objid_sha256"]
        assert _sha(lat.mask()) == rec["flat_mask_sha256"]
    lat.free()


@pytest.mark.parametrize("prec", [64, 32])
def test_64bit_index_path_is_bit_identical(prec, monkeypatch):
    """Lattices of more than 2^31 populations index with 64-bit integers;
    SLB_INDEX64=1 forces that path on a small lattice.  A Navier-Stokes,
    skip-solid run with a moving wall (the fused step with the wall term),
    measure() (the force-accumulating step) and an MP run with displacement
    moments give bit-identical results on both paths (object force and torque
    sums, FP64 atomics in a varying order, to 1e-13)."""
    import sunlightlb_amd as s
    g = load_golden("static_ns_skip.npz")
    runs = []
    for env in ("0", "1"):
        monkeypatch.setenv("SLB_INDEX64", env)
        lat = build_static(g, prec)
        L, h = lat.L, lat.h
        lat.s.objects_moving = 1
        objs = lat.objects()
        objs[-1].velx = 0.003          # the wall (inserted first, listed last)
        objs[0].angz = 0.001
        L.vspread(h)
        lat.step(20)
        L.measure(h)
        pops = lat.pops()
        rows = objrows(lat)
        vmean = (lat.s.vmeanx, lat.s.vmeany, lat.s.vmeanz)
        lat.s.objects_moving = 0
        mp = L.newmp(h, b"null")
        mp.contents.alpha = 0.05
        L.mpinit(mp)
        L.mpall(mp, 1.0)
        L.slb_mp_track_moments(mp, 0)
        L.slb_mp_run(mp, 12)
        sums = (C.c_double * 6)()
        L.slb_mp_sums(mp, sums)
        L.slb_mp_sync_host(mp)
        p = s.view(mp.contents.p, lat.N).copy()
        runs.append((pops, rows, vmean, p, list(sums)))
        lat.free()
    a, b = runs
    assert np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64))
    # object sums are FP64 atomics (their order varies between runs): 1e-13
    scale = np.abs(a[1]).max()
    assert np.abs(a[1] - b[1]).max() <= 1e-13 * scale
    assert a[2] == b[2]
    assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))
    assert a[4] == b[4]
    assert np.abs(a[1][:, 9:15]).max() > 0      # forces and torques were measured
