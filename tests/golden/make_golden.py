"""Write the golden fixtures of tests/golden/ from the reference library.

Runs the reference SunlightLB C library compiled from its own sources by
oracle/Makefile (oracle/_ref/libslbref.so, FP64, and libslbref32.so,
-D_SINGLE) on small seeded cases and stores inputs and outputs as .npz.
Run here, where the reference sources exist:

    make -C oracle ref && python tests/golden/make_golden.py

Every lattice is non-cubic so that a transposed index cannot pass.  Body
forces use the reference's current API (zerobf + addbf on fluid nodes).
"""
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import RefLib  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def state(R, h):
    s = h.contents
    N = s.nnodes
    return dict(a=R.arr(s.a, 15 * N).astype(np.float64).reshape(N, 15).copy(),
                mask=R.arr(s.mask, N, np.int32).copy(), objid=R.objid(h))


def objrows(R, h):
    rows = []
    for o in R.objects(h):
        rows.append([o.posx, o.posy, o.posz, o.velx, o.vely, o.velz, o.angx, o.angy, o.angz,
                     o.frcx, o.frcy, o.frcz, o.trqx, o.trqy, o.trqz, o.id])
    return np.array(rows, dtype=np.float64).reshape(-1, 16)


def measure(R, h):
    R.L.measure(h)
    s = h.contents
    N = s.nnodes
    return dict(vmean=np.array([s.vmeanx, s.vmeany, s.vmeanz], dtype=np.float64),
                fx=R.arr(s.fx, N).astype(np.float64).copy(),
                fy=R.arr(s.fy, N).astype(np.float64).copy(),
                fz=R.arr(s.fz, N).astype(np.float64).copy(),
                objs=objrows(R, h))


def case_static(R, name, dims, spheres, bf, tau, stokes, skip, steps, walls=()):
    h = R.new(*dims)
    s = h.contents
    s.tau, s.stokes_flow, s.skip_solid = tau, stokes, skip
    for w in walls:
        R.L.insertwall(h, w[0].encode(), w[1])
    for (r, x, y, z) in spheres:
        R.L.insertsph(h, r, x, y, z)
    R.L.zerobf(h)
    R.L.addbf(h, *bf)
    out = {}
    done = 0
    for k in steps:
        for _ in range(k - done):
            R.L.lbstep(h)
        done = k
        st = state(R, h)
        out[f"a_{k}"] = st["a"]
    st = state(R, h)
    out["mask"], out["objid"] = st["mask"], st["objid"]
    out["countfluid"] = np.array(R.L.countfluid(h))
    out["countmask_links"] = np.array(R.L.countmask(h, b"links"))
    out["countmask_faces"] = np.array(R.L.countmask(h, b"faces"))
    m = measure(R, h)
    for k, v in m.items():
        out["meas_" + k] = v
    out["a_after_measure"] = state(R, h)["a"]
    out["discrf"] = np.array(R.L.discrf(h))
    meta = dict(dims=dims, spheres=spheres, bf=bf, tau=tau, stokes=stokes, skip=skip,
                steps=list(steps), walls=[list(w) for w in walls])
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
    print("wrote", name)


def case_moving(R):
    dims = (14, 11, 13)
    h = R.new(*dims)
    s = h.contents
    s.tau, s.stokes_flow, s.skip_solid, s.objects_moving = 0.9, 1, 0, 1
    wall = R.L.insertwall(h, b"z", 0)
    sph = R.L.insertsph(h, 3.1, 6.3, 5.4, 6.2)
    sph.contents.velx, sph.contents.vely = 0.012, -0.004
    sph.contents.angy, sph.contents.angz = 0.003, 0.002
    wall.contents.velx = -0.005
    R.L.vspread(h)
    N = s.nnodes
    u0 = np.stack([R.arr(s.ux, N), R.arr(s.uy, N), R.arr(s.uz, N)]).astype(np.float64).copy()
    for _ in range(40):
        R.L.lbstep(h)
    out = dict(a_40=state(R, h)["a"], u=u0)
    st = state(R, h)
    out["mask"], out["objid"] = st["mask"], st["objid"]
    m = measure(R, h)
    for k, v in m.items():
        out["meas_" + k] = v
    out["a_after_measure"] = state(R, h)["a"]
    meta = dict(dims=dims, tau=0.9, steps=40)
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(OUT, "moving.npz"), **out)
    print("wrote moving")


def case_mp(R):
    """MP on the flow round a sphere: reflecting, then all-adsorbing, then
    moving walls (mpmove)."""
    dims = (12, 9, 11)
    out = {}
    for variant in ("reflect", "adsorb", "moving"):
        h = R.new(*dims)
        s = h.contents
        s.tau, s.stokes_flow, s.skip_solid = 0.8, 1, 0
        if variant == "moving":
            s.objects_moving = 1
        sph = R.L.insertsph(h, 2.8, 5.1, 4.3, 5.4)
        if variant == "moving":
            sph.contents.velx = 0.01
            sph.contents.angz = 0.004
            R.L.vspread(h)
        R.L.zerobf(h)
        R.L.addbf(h, 2e-4, 5e-5, 0.0)
        for _ in range(30):
            R.L.lbstep(h)
        mp = R.L.newmp(h, b"null")
        mp.contents.alpha = 0.1
        R.L.mpinit(mp)
        R.L.mpall(mp, 1.0)
        if variant == "adsorb":
            R.L.adall(mp)
        N = s.nnodes
        st = state(R, h)
        out[f"{variant}_a"] = st["a"]
        out[f"{variant}_mask"] = st["mask"]
        out[f"{variant}_objid"] = st["objid"]
        out[f"{variant}_rho"] = R.arr(s.rho, N).astype(np.float64).copy()
        for _ in range(25):
            R.L.mpstep(mp)
        out[f"{variant}_p"] = R.arr(mp.contents.p, N).astype(np.float64).copy()
        out[f"{variant}_adtotal"] = np.array(R.L.adtotal(mp, b"null", 1))
        if variant == "moving":
            out["moving_objtab"] = objrows(R, h)[:, :9]
    out["meta"] = np.array(json.dumps(dict(dims=dims, alpha=0.1, steps=25)))
    np.savez_compressed(os.path.join(OUT, "mp.npz"), **out)
    print("wrote mp")


def case_tables(R):
    """Solid maps and boundary-link masks, bit for bit: random spheres with
    arbitrary centres and radii (some outside the box), a cylinder, a hole,
    a slab and a wall, each alone and then all together."""
    rng = np.random.default_rng(20261202)
    dims = (21, 17, 23)
    out = {}
    spheres = []
    for _ in range(40):
        r = rng.uniform(0.3, 5.0)
        c = rng.uniform(-4, 25, size=3)
        spheres.append([r, *c])
    # exact-tie radii on half-integer and integer centres
    spheres += [[2.0, 10.0, 8.0, 11.0], [3.0, 10.5, 8.5, 11.5], [np.sqrt(5.0), 4.0, 4.0, 4.0],
                [1.5, 0.0, 0.0, 0.0], [2.5, 20.5, 16.5, 22.5]]
    h = R.new(*dims)
    ns = []
    for sp in spheres:
        p = R.L.insertsph(h, *sp)
        ns.append(0 if not p else 1)
    out["sph_objid"] = R.objid(h)
    R.L.insertcyl(h, b"y", 2.6, 15.3, 4.7)
    R.L.insertcyl(h, b"zh", 9.5, 10.2, 8.1)
    R.L.insertslab(h, 1.5, 4.2, -1.0, 6.5, 17.3, 30.0)
    R.L.insertwall(h, b"x", 20)
    R.L.syncobjs(h)
    st = state(R, h)
    out["spheres"] = np.array(spheres)
    out["placed"] = np.array(ns)
    out["objid"], out["mask"] = st["objid"], st["mask"]
    out["next_id"] = np.array(h.contents.next_id)
    out["objs"] = objrows(R, h)
    R.L.flattenobj(h)
    R.L.syncobjs(h)
    out["flat_objs"] = objrows(R, h)
    R.L.invertobj(h)
    R.L.syncobjs(h)
    st = state(R, h)
    out["inv_objid"], out["inv_mask"] = st["objid"], st["mask"]
    out["inv_objs"] = objrows(R, h)
    out["meta"] = np.array(json.dumps(dict(dims=dims)))
    np.savez_compressed(os.path.join(OUT, "tables.npz"), **out)
    print("wrote tables")


def case_config1(R, R32):
    """32^3, sphere r = 4.5 at 15.5^3, F = 1e-5 along x, Stokes, tau 0.75,
    2000 steps then measure: drag and mean velocity, FP64 and FP32."""
    res = {}
    for tag, RL in (("fp64", R), ("fp32", R32)):
        h = RL.new(32, 32, 32)
        s = h.contents
        s.stokes_flow = 1
        RL.L.insertsph(h, 4.5, 15.5, 15.5, 15.5)
        RL.L.zerobf(h)
        RL.L.addbf(h, 1e-5, 0.0, 0.0)
        for _ in range(2000):
            RL.L.lbstep(h)
        m = measure(RL, h)
        a = RL.arr(s.a, 15 * s.nnodes).astype(np.float64).reshape(-1, 15)
        res[tag] = dict(frcx=float(m["objs"][0, 9]), frcy=float(m["objs"][0, 10]),
                        frcz=float(m["objs"][0, 11]), vmean=m["vmean"].tolist(),
                        pop_sum=a.sum(axis=0).tolist(), pop_l2=float(np.sqrt((a * a).sum())),
                        discrf=float(RL.L.discrf(h)))
    with open(os.path.join(OUT, "config1.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("wrote config1", res["fp64"]["frcx"], res["fp32"]["frcx"])


def case_loadvel(R):
    """lbloadvel from a binary (rho, u) file: equilibrium rebuild + collide,
    Navier-Stokes and Stokes."""
    dims = (7, 9, 8)
    rng = np.random.default_rng(7)
    N = int(np.prod(dims))
    rho = 72.0 + rng.normal(0, 0.05, N)
    u = rng.normal(0, 0.02, (N, 3))
    path = os.path.join("/tmp", "slb_golden_vel.bin")
    np.column_stack([rho, u]).astype(np.float64).tofile(path)
    out = dict(rho=rho, u=u)
    for stokes in (0, 1):
        h = R.new(*dims)
        h.contents.stokes_flow = stokes
        R.L.insertsph(h, 2.2, 3.0, 4.0, 3.5)
        R.L.lbloadvel(h, path.encode(), b"binary")
        out[f"a_stokes{stokes}"] = state(R, h)["a"]
        out[f"objid_stokes{stokes}"] = R.objid(h)
    np.savez_compressed(os.path.join(OUT, "loadvel.npz"), **out)
    print("wrote loadvel")


def case_vacf(R):
    """Tagged-tracer velocity autocorrelation (src/mpvacf.c) on the flow round
    a sphere: P and P2 after mpvacfinit, and the diffusion coefficient
    returned by mpvacfwr right after the initialisation and after 12
    (mpstep, mpstep, mpvacf) rounds, for the directions xx and xy."""
    import tempfile
    dims = (12, 10, 14)
    out = {}
    h = R.new(*dims)
    s = h.contents
    s.tau, s.stokes_flow, s.skip_solid = 0.8, 1, 0
    R.L.insertsph(h, 3.1, 5.2, 4.7, 6.9)
    R.L.zerobf(h)
    R.L.addbf(h, 2e-4, 5e-5, 0.0)
    for _ in range(40):
        R.L.lbstep(h)
    N = s.nnodes
    st = state(R, h)
    out["a"], out["mask"], out["objid"] = st["a"], st["mask"], st["objid"]
    mp = R.L.newmp(h, b"null")
    mp2 = R.L.newmp(h, b"null")
    mp.contents.alpha = mp2.contents.alpha = 0.1
    R.L.mprhoinit(mp)
    out["rho"] = R.arr(s.rho, N).astype(np.float64).copy()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "vacf.txt").encode()
        for dr in ("xx", "xy"):
            R.L.mpinit(mp)
            R.L.mpinit(mp2)
            R.L.mpvacfinit(mp, mp2, dr.encode(), 40)
            out[f"{dr}_p"] = R.arr(mp.contents.p, N).astype(np.float64).copy()
            out[f"{dr}_p2"] = R.arr(mp2.contents.p, N).astype(np.float64).copy()
            out[f"{dr}_d0"] = np.array(R.L.mpvacfwr(mp, mp2, path, 1))
            for _ in range(12):
                R.L.mpstep(mp)
                R.L.mpstep(mp2)
                R.L.mpvacf(mp, mp2)
            out[f"{dr}_d12"] = np.array(R.L.mpvacfwr(mp, mp2, path, 1))
    out["meta"] = np.array(json.dumps(dict(dims=dims, alpha=0.1, tau=0.8, steps=40, rounds=12,
                                           sphere=[3.1, 5.2, 4.7, 6.9], bf=[2e-4, 5e-5, 0.0])))
    np.savez_compressed(os.path.join(OUT, "vacf.npz"), **out)
    print("wrote vacf")


def main():
    R = RefLib(64)
    R32 = RefLib(32)
    case_tables(R)
    case_static(R, "static_stokes", (12, 10, 14), [[3.3, 5.2, 4.7, 6.9]],
                (1e-4, 2e-5, -3e-5), 0.8, 1, 0, (1, 50))
    case_static(R, "static_ns_skip", (13, 11, 10),
                [[3.0, 6.1, 5.0, 4.4], [2.2, 9.0, 5.5, 6.5]], (1e-3, 0.0, 2e-4), 0.7, 0, 1,
                (1, 30), walls=(("y", 0),))
    case_moving(R)
    case_mp(R)
    case_loadvel(R)
    case_config1(R, R32)
    case_vacf(R)


if __name__ == "__main__":
    main()
