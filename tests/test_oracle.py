"""Pin the CPU oracle (oracle/slb_oracle.c) before trusting it.

Checks against (1) the golden fixtures written by the reference library
itself (tests/golden/make_golden.py), (2) closed forms and invariants that a
dropped term, wrong sign or transposed index would break.  CPU only.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, rel_l2
from oracle import C3, OPP, W

TOL = 1e-12


def body(objid, bf):
    fl = (objid < 0).astype(np.float64)
    return [fl * bf[0], fl * bf[1], fl * bf[2]]


def objpos_from(rows):
    return np.ascontiguousarray(rows[:, 0:3])


# --------------------------------------------------------------------------
# tables: bit-exact

def test_insertsph_and_makemask_bit_exact(oracle):
    g = load_golden("tables.npz")
    dims = tuple(g["meta"]["dims"])
    N = int(np.prod(dims))
    objid = np.full(N, -1, dtype=np.int32)
    ids = []
    for k, sp in enumerate(g["spheres"]):
        ns, _ = oracle.insertsph(objid, dims, k, *sp)
        ids.append(k if ns else None)
    placed = np.array([i is not None for i in ids])
    assert (placed == g["placed"].astype(bool)).all()
    # reference objid is the list order (newest first) of placed spheres
    order = [k for k in ids if k is not None][::-1]
    remap = np.full(len(g["spheres"]), -1, dtype=np.int32)
    for dense, k in enumerate(order):
        remap[k] = dense
    mine = np.where(objid >= 0, remap[np.maximum(objid, 0)], -1)
    assert (mine == g["sph_objid"]).all()
    # makemask over the full geometry (spheres + cylinder + hole + slab + wall)
    assert (oracle.makemask(g["objid"], dims) == g["mask"]).all()
    assert (oracle.makemask(g["inv_objid"], dims) == g["inv_mask"]).all()


def test_makemask_brute_force_single_node(oracle):
    """One solid node in open fluid: all 14 link bits; on 1-wide axes the node
    is its own neighbour along that axis, so those bits must drop."""
    dims = (5, 4, 6)
    objid = np.full(np.prod(dims), -1, dtype=np.int32)
    objid[(2 * 4 + 1) * 6 + 3] = 0
    m = oracle.makemask(objid, dims)
    assert m[(2 * 4 + 1) * 6 + 3] == 0x3FFF and (m != 0).sum() == 1
    dims = (1, 4, 6)
    objid = np.full(np.prod(dims), -1, dtype=np.int32)
    objid[1 * 6 + 3] = 0
    m = oracle.makemask(objid, dims)
    # only the pure-x links (c_1, c_2) land on the solid node itself; the
    # diagonals still move in y and z
    assert m[1 * 6 + 3] == 0x3FFF & ~0x3


# --------------------------------------------------------------------------
# stepping against the reference

@pytest.mark.parametrize("name", ["static_stokes", "static_ns_skip"])
def test_static_steps_match_reference(oracle, name):
    g = load_golden(f"{name}.npz")
    m = g["meta"]
    dims = tuple(m["dims"])
    N = int(np.prod(dims))
    objid, mask = g["objid"], g["mask"]
    assert (oracle.makemask(objid, dims) == mask).all()
    b = body(objid, m["bf"])
    a = np.zeros(N * 15)
    done = 0
    for k in m["steps"]:
        oracle.lbstep(a, mask, objid, dims, b, m["tau"], m["stokes"], m["skip"], nsteps=k - done)
        done = k
        assert rel_l2(a, g[f"a_{k}"]) < TOL, k
    res = oracle.measure(a, mask, objid, dims, b, m["tau"], m["stokes"], m["skip"],
                         objpos_from(g["meas_objs"]))
    assert rel_l2(res["vmean"], g["meas_vmean"]) < TOL
    for c in ("fx", "fy", "fz"):
        assert rel_l2(res[c], g["meas_" + c]) < TOL
    assert rel_l2(res["frc"], g["meas_objs"][:, 9:12]) < TOL
    assert rel_l2(res["trq"], g["meas_objs"][:, 12:15]) < TOL
    assert rel_l2(a, g["a_after_measure"]) < TOL


def test_moving_walls_match_reference(oracle):
    g = load_golden("moving.npz")
    dims = tuple(g["meta"]["dims"])
    N = int(np.prod(dims))
    objid, mask = g["objid"], g["mask"]
    rows = g["meas_objs"]
    u = oracle.vspread(objid, dims, rows[:, 0:3], rows[:, 3:6], rows[:, 6:9])
    assert rel_l2(np.stack(u), g["u"]) < 1e-15
    a = np.zeros(N * 15)
    b = body(objid, (0, 0, 0))
    oracle.lbstep(a, mask, objid, dims, b, 0.9, 1, 0, u=u, nsteps=40)
    assert rel_l2(a, g["a_40"]) < TOL
    res = oracle.measure(a, mask, objid, dims, b, 0.9, 1, 0, rows[:, 0:3], u=u)
    assert rel_l2(res["frc"], rows[:, 9:12]) < TOL
    assert rel_l2(res["trq"], rows[:, 12:15]) < TOL
    assert rel_l2(res["vmean"], g["meas_vmean"]) < TOL


@pytest.mark.parametrize("variant", ["reflect", "adsorb", "moving"])
def test_mp_matches_reference(oracle, variant):
    g = load_golden("mp.npz")
    dims = tuple(g["meta"]["dims"])
    a, mask, objid = g[f"{variant}_a"], g[f"{variant}_mask"], g[f"{variant}_objid"]
    a = np.ascontiguousarray(a).reshape(-1)
    rho = oracle.mprhoinit(a, objid)
    assert rel_l2(rho, g[f"{variant}_rho"]) < 1e-15
    p = (objid < 0).astype(np.float64)
    kw = {}
    if variant == "moving":
        kw = dict(moving=True, objtab=g["moving_objtab"])
    for _ in range(25):
        oracle.mpstep(a, mask, objid, rho, dims, p, alpha=0.1, **kw)
    assert rel_l2(p, g[f"{variant}_p"]) < TOL
    ads = p[(mask & 0x4000) != 0].sum()
    assert abs(ads - g[f"{variant}_adtotal"]) <= 1e-12 * max(1.0, abs(ads))


def test_equilibrium_rebuild_matches_lbloadvel(oracle):
    g = load_golden("loadvel.npz")
    dims = (7, 9, 8)
    N = int(np.prod(dims))
    for stokes in (0, 1):
        a = np.zeros(N * 15)
        u = g["u"]
        oracle.L.orc_equilibrium(N, a, np.ascontiguousarray(g["rho"]),
                                 np.ascontiguousarray(u[:, 0]), np.ascontiguousarray(u[:, 1]),
                                 np.ascontiguousarray(u[:, 2]), 72.0, stokes)
        objid = g[f"objid_stokes{stokes}"]
        z = np.zeros(N)
        oracle.L.orc_collide(N, a, objid, z, z, z, 0.75, 72.0, stokes, 0)
        assert rel_l2(a, g[f"a_stokes{stokes}"]) < TOL


def test_config1_drag_and_vmean(oracle):
    """Config-1 case (32^3 sphere, 2000 steps): drag and mean velocity."""
    with open(os.path.join(GOLDEN, "config1.json")) as f:
        ref = json.load(f)["fp64"]
    dims = (32, 32, 32)
    N = 32 ** 3
    objid = np.full(N, -1, dtype=np.int32)
    oracle.insertsph(objid, dims, 0, 4.5, 15.5, 15.5, 15.5)
    mask = oracle.makemask(objid, dims)
    b = body(objid, (1e-5, 0, 0))
    a = np.zeros(N * 15)
    oracle.lbstep(a, mask, objid, dims, b, 0.75, 1, 0, nsteps=2000)
    res = oracle.measure(a, mask, objid, dims, b, 0.75, 1, 0, np.array([[15.5, 15.5, 15.5]]))
    assert abs(res["frc"][0, 0] - ref["frcx"]) < TOL * abs(ref["frcx"])
    assert abs(res["vmean"][0] - ref["vmean"][0]) < TOL * abs(ref["vmean"][0])


# --------------------------------------------------------------------------
# closed forms and invariants (independent of the reference run)

def test_collide_conserves_mass_and_momentum_and_is_idempotent_on_equilibrium(oracle):
    rng = np.random.default_rng(3)
    N = 50
    a = rng.normal(0, 0.1, 15 * N)
    objid = np.full(N, -1, dtype=np.int32)
    z = np.zeros(N)
    before = a.reshape(N, 15).copy()
    oracle.L.orc_collide(N, a, objid, z, z, z, 0.8, 72.0, 1, 0)
    after = a.reshape(N, 15)
    np.testing.assert_allclose(after.sum(1), before.sum(1), atol=1e-14)
    np.testing.assert_allclose(after @ C3, before @ C3, atol=1e-14)
    # tau = 1 (omega = 1) Stokes: stress relaxes to zero -> equilibrium at rest
    a2 = after.reshape(-1).copy()
    oracle.L.orc_collide(N, a2, objid, z, z, z, 1.0, 72.0, 1, 0)
    eq = a2.reshape(N, 15)
    d = eq.sum(1)
    j = eq @ C3
    expect = (W / 72.0)[None, :] * (d[:, None] + 3.0 * (j @ C3.T))
    np.testing.assert_allclose(eq, expect, atol=1e-14)
    # an equilibrium is a fixed point of a second collision
    a3 = a2.copy()
    oracle.L.orc_collide(N, a3, objid, z, z, z, 1.0, 72.0, 1, 0)
    np.testing.assert_allclose(a3, a2, atol=1e-15)


def test_uniform_flow_is_fixed_point_and_periodic_shift(oracle):
    """A uniform-velocity Stokes equilibrium with no forcing is a fixed point
    of the full step, on any lattice shape (including 1-wide axes)."""
    for dims in ((4, 3, 5), (1, 1, 7), (2, 3, 1)):
        N = int(np.prod(dims))
        rho = np.full(N, 72.0)
        ux, uy, uz = np.full(N, 0.01), np.full(N, -0.02), np.full(N, 0.005)
        a = np.zeros(N * 15)
        oracle.L.orc_equilibrium(N, a, rho, ux, uy, uz, 72.0, 1)
        a0 = a.copy()
        objid = np.full(N, -1, dtype=np.int32)
        mask = np.zeros(N, dtype=np.int32)
        oracle.lbstep(a, mask, objid, dims, [np.zeros(N)] * 3, 0.7, 1, 0, nsteps=3)
        np.testing.assert_allclose(a, a0, atol=1e-15)


def test_propagate_moves_a_marker_along_c(oracle):
    dims = (5, 6, 7)
    N = int(np.prod(dims))
    for l in range(1, 15):
        a = np.zeros(N * 15)
        i, j, k = 2, 3, 4
        a[15 * ((i * 6 + j) * 7 + k) + l] = 1.0
        oracle.L.orc_propagate(*dims, a)
        ii, jj, kk = (i + C3[l, 0]) % 5, (j + C3[l, 1]) % 6, (k + C3[l, 2]) % 7
        assert a[15 * ((ii * 6 + jj) * 7 + kk) + l] == 1.0 and a.sum() == 1.0


def test_bounceback_conserves_mass_and_reverses_links(oracle):
    """Closed box of walls: total mass is conserved by bbl+propagate+collide,
    and a single solid node returns what its fluid neighbour sent."""
    dims = (6, 5, 7)
    N = int(np.prod(dims))
    rng = np.random.default_rng(11)
    objid = np.full(N, -1, dtype=np.int32)
    oracle.insertsph(objid, dims, 0, 1.6, 2.5, 2.0, 3.0)
    mask = oracle.makemask(objid, dims)
    a = rng.normal(0, 0.01, N * 15)
    tot = a.sum()
    oracle.lbstep(a, mask, objid, dims, [np.zeros(N)] * 3, 0.8, 0, 0, nsteps=5)
    assert abs(a.sum() - tot) < 1e-13
    # single link: node s solid, f = s + c_1 fluid
    a = np.zeros(N * 15)
    objid = np.full(N, -1, dtype=np.int32)
    s = (1 * 5 + 1) * 7 + 1
    f = (2 * 5 + 1) * 7 + 1
    objid[s] = 0
    mask = oracle.makemask(objid, dims)
    a[15 * f + OPP[1]] = 0.5   # f sends toward s along c_2
    oracle.L.orc_bbl(*dims, a, mask, None, None, None, None, None, None)
    oracle.L.orc_propagate(*dims, a)
    assert a[15 * f + 1] == 0.5 and abs(a.sum() - 0.5) < 1e-16


def test_couette_wall_force_closed_form(oracle):
    """Planar Couette between moving walls (reference example
    planar_couette.pl): wall force per node = 2 eta (v1 - v2) / h, exact."""
    h = 10
    dims = (1, 1, 2 * (h + 1))
    N = int(np.prod(dims))
    objid = np.full(N, -1, dtype=np.int32)
    objid[0] = 0
    objid[h + 1] = 1
    mask = oracle.makemask(objid, dims)
    tau, v1, v2 = 0.8, 0.11, -0.07
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, h + 1.0]])
    vel = np.array([[v1, 0, 0], [v2, 0, 0]])
    u = oracle.vspread(objid, dims, pos, vel, np.zeros((2, 3)))
    a = np.zeros(N * 15)
    b = [np.zeros(N)] * 3
    oracle.lbstep(a, mask, objid, dims, b, tau, 1, 0, u=u, nsteps=5000)
    res = oracle.measure(a, mask, objid, dims, b, tau, 1, 0, pos, u=u)
    eta = 72.0 * (tau - 0.5) / 3.0
    expect = 2 * eta * (v1 - v2) / h
    # the faster wall (v1 > v2) is dragged back by the fluid, the slower forward
    assert abs(res["frc"][0, 0] + expect) < 1e-10
    assert abs(res["frc"][1, 0] - expect) < 1e-10


def test_mp_random_walk_variance_closed_form(oracle):
    """No solids, a = 0: each axis is a random walk with variance
    (1 - alpha) t / 3 = 2 D t, D = (1 - alpha)/6 (src/mpmisc.c:106); the
    displacement-moment fields reproduce it exactly and P is conserved."""
    dims = (9, 8, 10)
    N = int(np.prod(dims))
    a = np.zeros(N * 15)
    objid = np.full(N, -1, dtype=np.int32)
    mask = np.zeros(N, dtype=np.int32)
    rho = np.full(N, 72.0)
    for alpha in (0.0, 0.3):
        for axis in range(3):
            p = np.zeros(N)
            p[(4 * 8 + 4) * 10 + 5] = 1.0
            m1, m2 = np.zeros(N), np.zeros(N)
            for _ in range(12):
                oracle.mpstep(a, mask, objid, rho, dims, p, alpha=alpha, m1=m1, m2=m2, axis=axis)
            assert abs(p.sum() - 1.0) < 1e-14
            assert abs(m1.sum()) < 1e-14
            assert abs(m2.sum() - (1 - alpha) * 12 / 3.0) < 1e-12


def test_mp_partial_adsorption_limits(oracle):
    """pads = 0 equals the reference reflecting rule; pads = 1 equals
    adall(); in between, fluid + solid tracer is conserved."""
    g = load_golden("mp.npz")
    dims = tuple(g["meta"]["dims"])
    a = np.ascontiguousarray(g["reflect_a"]).reshape(-1)
    mask, objid = g["reflect_mask"], g["reflect_objid"]
    rho = oracle.mprhoinit(a, objid)
    fl = (objid < 0).astype(np.float64)
    p0, p1 = fl.copy(), fl.copy()
    for _ in range(25):
        oracle.mpstep(a, mask, objid, rho, dims, p0, alpha=0.1, pads=0.0)
        oracle.mpstep(a, mask, objid, rho, dims, p1, alpha=0.1, pads=1.0)
    assert rel_l2(p0, g["reflect_p"]) < TOL
    assert rel_l2(p1, g["adsorb_p"]) < TOL
    p = fl.copy()
    tot = p.sum()
    for _ in range(25):
        oracle.mpstep(a, mask, objid, rho, dims, p, alpha=0.1, pads=0.3)
    assert abs(p.sum() - tot) < 1e-11 * tot
    ads = p[objid >= 0].sum()
    assert ads > 0 and ads < p1[objid >= 0].sum()


def _packs():
    import gzip
    with gzip.open(os.path.join(GOLDEN, "packs.json.gz"), "rt") as f:
        return json.load(f)


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["pack128", "pack_noncubic", "box0", "box1", "box2", "box3",
                                  "box4", "box5"])
def test_tables_at_workload_sizes(oracle, name):
    """Solid map and boundary-link mask of workload-size packs (128^3 radius-6
    pack, a non-cubic mixed-radius pack, batch-like boxes) equal the
    reference library's bit for bit (SHA-256 of the int32 tables)."""
    case = _packs()[name]
    rec = case["record"]
    dims = tuple(rec["dims"])
    N = int(np.prod(dims))
    objid = np.full(N, -1, dtype=np.int32)
    placed = []
    for k, row in enumerate(case["inserts"]):
        ns, _ = oracle.insertsph(objid, dims, k, *row)
        placed.append(1 if ns else 0)
    assert placed == case["placed_list"]
    # reference numbering: list order, newest placed object first
    kept = [k for k, p in enumerate(placed) if p][::-1]
    remap = np.full(len(placed), -1, dtype=np.int32)
    remap[kept] = np.arange(len(kept), dtype=np.int32)
    dense = np.where(objid >= 0, remap[np.maximum(objid, 0)], -1)
    mask = oracle.makemask(dense, dims)
    assert int((dense >= 0).sum()) == rec["solid"]
    assert int((mask != 0).sum()) == rec["mask_nonzero"]
    assert _sha(dense) == rec["objid_sha256"]
    assert _sha(mask) == rec["mask_sha256"]
    if "flat_mask_sha256" in rec:
        flat = np.where(dense >= 0, 0, -1)
        assert _sha(flat) == rec["flat_objid_sha256"]
        assert _sha(oracle.makemask(flat, dims)) == rec["flat_mask_sha256"]
