"""Tiny and degenerate lattices against the CPU oracle: every axis of
length 1, 2 or 3 (a node's neighbours along such an axis wrap onto the node
itself or onto one another), one solid node moving with a wall velocity,
Navier-Stokes, a non-uniform body force and a perturbed start, so that every
population and every link census of the single solid node is exercised
(src/lbkernel.c:58-170,330-454,623-877; the link mask of src/lbobjs.c:526-560).
GPU FP64 vs the oracle: populations after 9 steps, then measure()'s object
force, torque and mean velocity, and countmask against the mask."""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_parity import dense_objid, lattice, objrows

pytestmark = pytest.mark.gpu

W = np.array([16, 8, 8, 8, 8, 8, 8, 1, 1, 1, 1, 1, 1, 1, 1])


@pytest.mark.parametrize("dims", [(2, 2, 2), (1, 1, 7), (2, 3, 1), (3, 1, 1), (1, 2, 2),
                                  (3, 3, 3), (1, 1, 2)])
def test_tiny_lattice_matches_oracle(oracle, dims):
    import sunlightlb_amd as s
    lat = lattice(dims)
    L, h = lat.L, lat.h
    tau = 0.83
    lat.s.tau, lat.s.stokes_flow, lat.s.skip_solid, lat.s.objects_moving = tau, 0, 0, 1
    # one solid node at (0, 0, 0): radius 0.6 covers no other node
    sph = L.insertsph(h, 0.6, 0.0, 0.0, 0.0)
    assert sph
    sph.contents.velx, sph.contents.vely = 0.01, -0.004
    sph.contents.angz = 0.003
    L.vspread(h)
    L.zerobf(h)
    L.syncobjs(h)
    N = lat.N
    rng = np.random.default_rng(sum(dims))
    lat.sync()
    solid = lat.solid()
    assert solid.sum() == 1 and solid[0]
    for nm, sc in (("bx", 2e-4), ("by", -7e-5), ("bz", 1e-4)):
        s.view(getattr(lat.s, nm), N)[:] = np.where(solid, 0.0, sc * (1 + rng.random(N)))
    L.slb_sync_device(h)
    a0 = 1e-3 * rng.standard_normal(N * 15)
    lat.set_pops(a0)
    lat.step(9)
    L.measure(h)
    objid = dense_objid(lat)
    mask = oracle.makemask(objid, dims)
    assert (lat.mask() == mask).all()
    links = sum(int(W[l]) * int(((mask >> (l - 1)) & 1).sum()) for l in range(1, 15))
    assert L.countmask(h, b"links") == links
    rows = objrows(lat)
    u = oracle.vspread(objid, dims, rows[:, 0:3], rows[:, 3:6], rows[:, 6:9])
    bb = [lat.field("bx"), lat.field("by"), lat.field("bz")]
    ao = a0.copy()
    oracle.lbstep(ao, mask, objid, dims, bb, tau, 0, 0, u=u, nsteps=9)
    res = oracle.measure(ao, mask, objid, dims, bb, tau, 0, 0, rows[:, 0:3], u=u)
    assert rel_l2(lat.pops(), ao) < 1e-12
    if np.abs(res["frc"]).max() > 0:
        assert rel_l2(rows[:, 9:12], res["frc"]) < 1e-12
    else:
        assert np.abs(rows[:, 9:12]).max() == 0
    if np.abs(res["trq"]).max() > 0:
        assert rel_l2(rows[:, 12:15], res["trq"]) < 1e-12
    vm = np.array([lat.s.vmeanx, lat.s.vmeany, lat.s.vmeanz])
    assert rel_l2(vm, res["vmean"]) < 1e-12
    lat.free()
