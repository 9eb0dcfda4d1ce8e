"""Write the golden fixtures of tests/golden/ from the reference library.

Runs the reference SunlightLB C library compiled from its own sources by
oracle/Makefile (oracle/_ref/libslbref.so, FP64, and libslbref32.so,
-D_SINGLE) on small seeded cases and stores inputs and outputs as .npz.
Run here, where the reference sources exist:

    make -C oracle ref && python tests/golden/make_golden.py
    python tests/golden/make_golden.py packs      # only the workload-size tables
    python tests/golden/make_golden.py pack256    # only the 256^3 permeability pack
    python tests/golden/make_golden.py checkpoint # only the checkpoint / velocity-field cases

Every lattice is non-cubic so that a transposed index cannot pass.  Body
forces use the reference's current API (zerobf + addbf on fluid nodes).
"""
import ctypes as C
import gzip
import hashlib
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


def case_sphere_c(R):
    """The reference example driver examples/sphere.c (16^3, one sphere, Stokes,
    skip_solid, tau 0.8, 1000 steps, measure) replayed call for call on the
    reference library.  sphere.c sets the uniform body force through
    mylb->dpdx, which the current reference header lacks; here it is the
    equivalent addbf(0.01, 0, 0) on the fluid nodes after the insertion.
    Stores the five files the driver writes and its stdout."""
    import gzip
    import tempfile
    L = R.L
    vp, sp = C.c_void_p, C.c_char_p
    for n, a in {"about": [], "wrlinks": [vp], "wrmap": [vp, sp], "lbsavemap": [vp, sp, sp],
                 "wrpars": [vp], "wrmon": [vp, sp], "wrstats": [vp], "wrmeas": [vp],
                 "wrobjs": [vp], "lbsavefrc": [vp, sp, sp]}.items():
        getattr(L, n).argtypes = a
        getattr(L, n).restype = None
    libc = C.CDLL(None)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        sys.stdout.flush()
        saved = os.dup(1)
        fd = os.open("stdout.txt", os.O_WRONLY | os.O_CREAT | os.O_TRUNC)
        os.dup2(fd, 1)
        try:
            L.about()
            h = L.newlb(16, 16, 16, b"sphere_c")
            L.set_verbosity(0x04)  # OBJECT_DETAILS
            st = h.contents
            st.stokes_flow, st.objects_moving, st.skip_solid, st.tau = 1, 0, 1, 0.80
            L.wrlinks(h)
            L.insertsph(h, 3.7, 7.5, 7.5, 7.5)
            L.addbf(h, 0.01, 0.0, 0.0)
            L.wrmap(h, b"sphere_c.map")
            L.lbsavemap(h, b"sphere_c.dat", b"list")
            L.wrpars(h)
            L.wrmon(h, b"initialised")
            libc.printf(b"\nEquilibrating for %i steps\n", C.c_int(1000))
            for _ in range(1000):
                L.lbstep(h)
            L.wrstats(h)
            L.measure(h)
            L.wrmeas(h)
            L.wrobjs(h)
            L.lbsavefrc(h, b"sphere_c.frc", b"list")
            L.lbsavevel(h, b"sphere_c.vel", b"ascii")
            L.wrmon(h, b"finished")
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(fd)
            os.close(saved)
        out = {"stdout": open("stdout.txt").read(), "files": {}}
        for f in ("sphere_c.map", "sphere_c.dat", "sphere_c.frc", "sphere_c.vel", "sphere_c.mon"):
            out["files"][f] = open(f).read()
        os.chdir(cwd)
    with gzip.open(os.path.join(OUT, "sphere_c.json.gz"), "wt") as f:
        json.dump(out, f)
    print("wrote sphere_c")


IMG = dict(dims=(12, 10, 9), sphere=(3.0, 5.5, 4.5, 4.0), vel=(0.02, -0.005, 0.01), tau=0.8,
           steps=30, mp_steps=20, slice=4,
           vmap=(-0.01, 0.03, (10, 50, 50), (250, 200, 200)),
           pmap=(0.0, 2.0, (0, 0, 0), (200, 100, 50)),
           back=(255, 255, 255), obj=(0, 0, 0))


def case_images(R):
    """src/mpimage.c on a slice normal to z of a lattice the size of the image
    (where its pixel placement is right), a sphere moving through fluid at
    rest with no body force (its velocity correction then adds nothing) and
    green_lo == blue_lo (its blue channel starts from green_lo): the solid
    map, the three velocity components and the speed, and an MP field added
    to and subtracted from the map."""
    L = R.L
    vp, sp, I = C.c_void_p, C.c_char_p, C.c_int
    D = C.c_double
    for n, (r, a) in {"newlbcolour": (vp, [I, I, I]), "newlbcolourmap": (vp, [D, D, vp, vp]),
                      "newlbimage": (vp, [I, I, vp, vp]), "lbimageclear": (None, [vp]),
                      "lbimagemap": (None, [vp, vp, sp, I]),
                      "lbimagevel": (None, [vp, vp, vp, sp, sp, I]),
                      "mpimageset": (None, [vp, vp, vp, sp, sp, I])}.items():
        getattr(L, n).restype = r
        getattr(L, n).argtypes = a
    c = IMG
    h = R.new(*c["dims"])
    st = h.contents
    st.tau, st.stokes_flow, st.objects_moving = c["tau"], 1, 1
    o = L.insertsph(h, *c["sphere"])
    o.contents.velx, o.contents.vely, o.contents.velz = c["vel"]
    L.vspread(h)
    for _ in range(c["steps"]):
        L.lbstep(h)
    esi, esj, esk = c["dims"]
    back, obj = L.newlbcolour(*c["back"]), L.newlbcolour(*c["obj"])
    im = L.newlbimage(esi, esj, back, obj)
    im_c = C.cast(im, C.POINTER(C.c_int * 2)).contents
    area = esi * esj

    def pixels():
        ptrs = C.cast(im + 8, C.POINTER(C.c_void_p * 3)).contents
        return np.stack([np.ctypeslib.as_array(C.cast(ptrs[q], C.POINTER(C.c_int)), shape=(area,))
                         .copy() for q in range(3)])
    assert tuple(im_c) == (esi, esj)
    out = {"meta": json.dumps(c)}
    L.lbimagemap(im, h, b"z", c["slice"])
    out["map"] = pixels()
    lo, hi = L.newlbcolour(*c["vmap"][2]), L.newlbcolour(*c["vmap"][3])
    vmap = L.newlbcolourmap(c["vmap"][0], c["vmap"][1], lo, hi)
    for w in ("x", "y", "z", "m"):
        L.lbimagevel(im, vmap, h, w.encode(), b"z", c["slice"])
        out[f"vel_{w}"] = pixels()
    mp = L.newmp(h, b"null")
    L.mprhoinit(mp)
    L.mpinit(mp)
    L.mpall(mp, 1.0)
    for _ in range(c["mp_steps"]):
        L.mpstep(mp)
    plo, phi = L.newlbcolour(*c["pmap"][2]), L.newlbcolour(*c["pmap"][3])
    pmap = L.newlbcolourmap(c["pmap"][0], c["pmap"][1], plo, phi)
    L.lbimagemap(im, h, b"z", c["slice"])
    L.mpimageset(im, mp, pmap, b"a", b"z", c["slice"])
    out["mp_add"] = pixels()
    L.mpimageset(im, mp, pmap, b"s", b"z", c["slice"])
    out["mp_sub"] = pixels()
    N = st.nnodes
    out["a"] = R.arr(st.a, 15 * N).astype(np.float64).reshape(N, 15).copy()
    out["p"] = R.arr(mp.contents.p, N).astype(np.float64).copy()
    np.savez_compressed(os.path.join(OUT, "images.npz"), **out)
    print("wrote images")


DIAG = dict(dims=(11, 9, 10), sphere=(2.8, 5.0, 4.0, 5.5), bf=(2e-4, 0.0, 5e-5), tau=0.9,
            steps=25, mp_sphere=(3.5, 4.0, 5.0, 4.0, 1.0), mp_steps=12, iso=0.05,
            p2byte=(0.0, 0.8), adt_mult=10)


def case_diagnostics(R):
    """The diagnostic writers on a Navier-Stokes lattice with a sphere and a
    body force: lbsavestr (ascii and binary; each call advances one step),
    and on an MP field started as a sphere of tracer with adsorbing solid
    sites -- mpconwr (ascii with a step-numbered name, binary, byte-converted),
    mpspread along x, y, z, mpcheck's totals, mpmean, adtotal written to a
    file twice, the cumulative isosurface of mpiso/mpisowr, and mpsave."""
    import tempfile
    L = R.L
    vp, sp, I = C.c_void_p, C.c_char_p, C.c_int
    D = C.c_double
    for n, (r, a) in {"lbsavestr": (None, [vp, sp, sp]), "mpconwr": (None, [vp, sp, sp]),
                      "mpspread": (D, [vp, sp]), "mpmean": (D, [vp]),
                      "mpsph": (None, [vp, D, D, D, D, D]), "newmpiso": (vp, [vp, D]),
                      "mpiso": (None, [vp]), "mpisowr": (None, [vp, sp, sp]),
                      "mpsave": (None, [vp, sp, sp])}.items():
        getattr(L, n).restype = r
        getattr(L, n).argtypes = a
    c = DIAG
    h = R.new(*c["dims"])
    st = h.contents
    st.tau, st.stokes_flow = c["tau"], 0
    L.insertsph(h, *c["sphere"])
    L.zerobf(h)
    L.addbf(h, *c["bf"])
    for _ in range(c["steps"]):
        L.lbstep(h)
    out = {"meta": json.dumps(c)}
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        try:
            L.lbsavestr(h, b"str.txt", b"ascii")
            L.lbsavestr(h, b"str.bin", b"binary")
            N = st.nnodes
            out["a"] = R.arr(st.a, 15 * N).astype(np.float64).reshape(N, 15).copy()
            mp = L.newmp(h, b"mpd")
            mp.contents.alpha = 0.0  # newmp leaves alpha unset (src/mpmisc.c:38-59)
            L.mprhoinit(mp)
            L.mpinit(mp)
            L.mpsph(C.cast(mp, vp), *c["mp_sphere"])
            L.adall(mp)
            iso = L.newmpiso(C.cast(mp, vp), c["iso"])
            for _ in range(c["mp_steps"]):
                L.mpstep(mp)
                L.mpiso(iso)
            mp.contents.p2bytemin, mp.contents.p2bytemax = c["p2byte"]
            L.mpconwr(C.cast(mp, vp), b"con_%i.txt", b"ua")
            L.mpconwr(C.cast(mp, vp), b"con.bin", b"nwb")
            L.mpconwr(C.cast(mp, vp), b"con.byte", b"nwc")
            out["spread"] = np.array([L.mpspread(C.cast(mp, vp), x) for x in (b"x", b"y", b"z")])
            out["mean"] = np.array(L.mpmean(C.cast(mp, vp)))
            out["adt"] = np.array([L.adtotal(mp, b"adt.txt", c["adt_mult"]) for _ in range(2)])
            L.mpisowr(iso, b"iso.txt", b"ascii")
            L.mpsave(C.cast(mp, vp), b"mp.txt", b"ascii")
            out["p"] = R.arr(mp.contents.p, N).astype(np.float64).copy()
            step = mp.contents.step
            out["con_name"] = np.array(f"con_{step}.txt")
            for k, f in (("str_txt", "str.txt"), ("con_txt", f"con_{step}.txt"), ("adt_txt", "adt.txt"),
                         ("iso_txt", "iso.txt"), ("mp_txt", "mp.txt")):
                out[k] = np.array(open(f).read())
            for k, f in (("str_bin", "str.bin"), ("con_bin", "con.bin"), ("con_byte", "con.byte")):
                out[k] = np.frombuffer(open(f, "rb").read(), dtype=np.uint8).copy()
        finally:
            os.chdir(cwd)
    np.savez_compressed(os.path.join(OUT, "diagnostics.npz"), **out)
    print("wrote diagnostics")


def rsa_spheres(rng, dims, r_lo, r_hi, count, trials, gap=0.0):
    """Random sequential addition in a periodic box: radius U[r_lo, r_hi],
    centre uniform, accepted when the minimum-image distance to every placed
    sphere is at least the sum of the radii plus gap."""
    Ld = np.array(dims, dtype=np.float64)
    placed = []
    for _ in range(trials):
        if len(placed) == count:
            break
        r = float(rng.uniform(r_lo, r_hi))
        c = rng.uniform(0.0, 1.0, 3) * Ld
        ok = True
        for (r2, *c2) in placed:
            d = c - np.array(c2)
            d -= Ld * np.round(d / Ld)
            if np.sqrt((d * d).sum()) < r + r2 + gap:
                ok = False
                break
        if ok:
            placed.append((r, *c))
    return placed


def image_inserts(dims, spheres):
    """Each sphere and every periodic image whose bounding box meets the box,
    in insertion order, one insertsph call (one object) each."""
    out = []
    for (r, *c) in spheres:
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for e in (-1, 0, 1):
                    o = (c[0] + a * dims[0], c[1] + b * dims[1], c[2] + e * dims[2])
                    if all(-r < o[d] < dims[d] + r for d in range(3)):
                        out.append((r, *o))
    return np.array(out, dtype=np.float64).reshape(-1, 4)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def case_packs(R):
    """Solid maps and boundary-link masks at workload sizes: a 128^3 pack of
    radius-6 spheres (solid fraction about 0.25, then flattened to one
    object), a non-cubic pack of mixed radii, and boxes like the batched ones
    (edge 32..64, 2..20 spheres of radius 3..6).  The tables are stored as
    SHA-256 digests of their little-endian int32 bytes with the counts, so
    the fixture stays small while the comparison stays bit for bit."""
    rng = np.random.default_rng(20261203)
    cases = [("pack128", (128, 128, 128), rsa_spheres(rng, (128,) * 3, 6.0, 6.0, 580, 400000),
              True),
             ("pack_noncubic", (72, 60, 66), rsa_spheres(rng, (72, 60, 66), 3.0, 8.0, 60, 100000),
              True)]
    for q in range(6):
        Lb = int(rng.integers(32, 65))
        ns = int(rng.integers(2, 21))
        cases.append((f"box{q}", (Lb,) * 3, rsa_spheres(rng, (Lb,) * 3, 3.0, 6.0, ns, 20000, 1.0),
                      False))
    out = {}
    for name, dims, spheres, flatten in cases:
        ins = image_inserts(dims, spheres)
        h = R.new(*dims)
        placed = [1 if R.L.insertsph(h, *row) else 0 for row in ins]
        R.L.syncobjs(h)
        st = state(R, h)
        rec = dict(dims=list(dims), spheres=len(spheres), placed=int(sum(placed)),
                   mask_sha256=sha(st["mask"]), objid_sha256=sha(st["objid"]),
                   solid=int((st["objid"] >= 0).sum()),
                   mask_nonzero=int((st["mask"] != 0).sum()),
                   countfluid=int(R.L.countfluid(h)),
                   countmask_links=int(R.L.countmask(h, b"links")),
                   countmask_faces=int(R.L.countmask(h, b"faces")))
        if flatten:
            R.L.flattenobj(h)
            R.L.syncobjs(h)
            st = state(R, h)
            rec["flat_mask_sha256"] = sha(st["mask"])
            rec["flat_objid_sha256"] = sha(st["objid"])
        out[name] = dict(record=rec, inserts=ins.tolist(), placed_list=placed)
        print(name, {k: v for k, v in rec.items() if "sha" not in k})
    with gzip.open(os.path.join(OUT, "packs.json.gz"), "wt") as f:
        json.dump(out, f)
    print("wrote packs")


CKPT = dict(dims=(12, 10, 14), tau=0.8, sphere=(3.3, 5.2, 4.7, 6.9), wall=("y", 0),
            bf=(1e-4, 2e-5, -3e-5), sphere_vel=(0.011, 0.0, -0.002), sphere_ang=(0.0, 0.001, -0.003),
            pre_steps=28, rescale=2.0)


def ckpt_lattice(R, c):
    """The checkpoint case: Stokes flow past a sphere beside a wall, a body
    force, object velocities set (no moving walls: they only reach the
    object file)."""
    h = R.new(*c["dims"])
    st = h.contents
    st.tau, st.stokes_flow = c["tau"], 1
    R.L.insertwall(h, c["wall"][0].encode(), c["wall"][1])
    sph = R.L.insertsph(h, *c["sphere"])
    sph.contents.velx, sph.contents.vely, sph.contents.velz = c["sphere_vel"]
    sph.contents.angx, sph.contents.angy, sph.contents.angz = c["sphere_ang"]
    R.L.zerobf(h)
    R.L.addbf(h, *c["bf"])
    return h


def case_checkpoint(R):
    """Checkpoint files (src/lbstore.c:36-91 lbsave, :523-547 lbsaveobj) of
    the reference after pre_steps steps, measure() (object forces in the
    file) and one more step -- what a USR2 signal during that step writes
    (src/lbsignal.c:49-59) -- and what the reference's lbload (:102-218) and
    lbloadobj (:555-586, also with rescale) read back from them."""
    import tempfile
    L = R.L
    vp, sp = C.c_void_p, C.c_char_p
    for n, (r, a) in {"lbsaveobj": (None, [vp, sp]), "lbloadobj": (None, [vp, sp])}.items():
        getattr(L, n).restype = r
        getattr(L, n).argtypes = a
    c = CKPT
    h = ckpt_lattice(R, c)
    for _ in range(c["pre_steps"]):
        L.lbstep(h)
    L.measure(h)
    L.lbstep(h)
    out = {"meta": np.array(json.dumps(c))}
    N = h.contents.nnodes
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        try:
            L.lbsave(h, b"ck.dms")
            L.lbsaveobj(C.cast(h, vp), b"ck.obj")
            out["a_after_save"] = R.arr(h.contents.a, 15 * N).astype(np.float64).reshape(N, 15).copy()
            out["dms"] = np.frombuffer(open("ck.dms", "rb").read(), dtype=np.uint8).copy()
            out["obj"] = np.array(open("ck.obj").read())
            # read back into fresh lattices of the same geometry
            h2 = ckpt_lattice(R, c)
            L.syncobjs(h2)
            L.lbload(h2, b"ck.dms")
            out["a_loaded"] = R.arr(h2.contents.a, 15 * N).astype(np.float64).reshape(N, 15).copy()
            for o in R.objects(h2):
                o.velx = o.vely = o.velz = o.angx = o.angy = o.angz = 0.0
            L.lbloadobj(C.cast(h2, vp), b"ck.obj")
            out["loadobj_rows"] = objrows(R, h2)
            h3 = ckpt_lattice(R, c)
            h3.contents.rescale = c["rescale"]
            L.lbloadobj(C.cast(h3, vp), b"ck.obj")
            out["loadobj_rescaled_rows"] = objrows(R, h3)
            out["rescale_after"] = np.array(h3.contents.rescale)
        finally:
            os.chdir(cwd)
    np.savez_compressed(os.path.join(OUT, "checkpoint.npz"), **out)
    print("wrote checkpoint")


VF = dict(dims=(11, 13, 9), tau=0.7, sphere=(3.1, 5.5, 6.0, 4.2), steps=21, prefac=(1.0, 0.5),
          vset=[[5, 6, 4, 0.01, -0.02, 0.005], [0, 0, 0, 0.3, 0.3, 0.3], [11, 2, 2, 1.0, 1.0, 1.0],
                [6, 5, 3, -0.004, 0.0, 0.007]])


def case_vfields(R):
    """The velocity-field calls (src/lbmeasure.c:110-166 vzero/vgrab,
    src/lbobjs.c:623-643 vset) on a Navier-Stokes lattice with a sphere and
    a non-uniform body force: vgrab after collide (correction -b/2) and after
    bbl + propagate (+b/2), accumulating with different prefactors; vset on
    solid, fluid and out-of-range nodes; vzero clearing the fluid nodes."""
    L = R.L
    for n, (r, a) in {"vzero": (None, [C.c_void_p]),
                      "vgrab": (None, [C.c_void_p, C.c_double]),
                      "vset": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double])}.items():
        getattr(L, n).restype = r
        getattr(L, n).argtypes = a
    c = VF
    h = R.new(*c["dims"])
    st = h.contents
    st.tau, st.stokes_flow = c["tau"], 0
    L.insertsph(h, *c["sphere"])
    L.zerobf(h)
    N = st.nnodes
    rng = np.random.default_rng(31)
    objid = R.objid(h)
    bx, by, bz = (R.arr(getattr(st, f), N) for f in ("bx", "by", "bz"))
    for arr, sc in ((bx, 2e-4), (by, -1e-4), (bz, 5e-5)):
        arr[:] = np.where(objid < 0, sc * (1.0 + rng.random(N)), 0.0)
    out = {"meta": np.array(json.dumps(c)), "bx": bx.copy(), "by": by.copy(), "bz": bz.copy()}
    for _ in range(c["steps"]):
        L.lbstep(h)
    hv = C.cast(h, C.c_void_p)

    def fields():
        return np.stack([R.arr(getattr(st, f), N).astype(np.float64) for f in ("ux", "uy", "uz")])

    for q, (i, j, k, vx, vy, vz) in enumerate(c["vset"]):
        L.vset(hv, int(i), int(j), int(k), vx, vy, vz)
    out["u_vset"] = fields()
    L.vzero(hv)
    L.vgrab(hv, c["prefac"][0])
    out["u_grab_collide"] = fields()
    L.bbl(h)
    L.propagate(h)
    L.vgrab(hv, c["prefac"][1])
    out["u_grab_propagate"] = fields()
    L.collide(h)
    L.vzero(hv)
    out["u_zero"] = fields()
    out["a"] = R.arr(st.a, 15 * N).astype(np.float64).reshape(N, 15).copy()
    np.savez_compressed(os.path.join(OUT, "vfields.npz"), **out)
    print("wrote vfields")


def mobility2_replay(L, new, objects, vel_setter, step, measure, vspread, removeobjs, insert):
    """examples/mobility2.pl: lateral mobility of a sphere above a wall by
    relaxing to zero torque on the fly -- for each height, 20 rounds of 100
    steps + measure(), the sphere translating at vx = 0.1 with angular
    velocity ay from the secant update of the torque; vspread after each
    velocity change (the current library's moving-wall step reads it).
    Returns D/D0 per height and the last (ay, ty) per height."""
    import math
    esx, esy, esz, nsteps, a_hyd, a_nom = 48, 32, 32, 2000, 4.1, 4.3
    lb = new(esx, esy, esz)
    eta = 72.0 * (0.95 - 0.5) / 3.0
    out = []
    for height in (5.0, 6.0, 7.0, 8.0, 9.0):
        sph = insert(lb, a_nom, esx / 2 - 0.5, esy / 2 - 0.5, height + 0.5)
        vx, ntries = 0.1, 20
        num = nsteps // ntries
        ay0 = ay1 = ty0 = ty1 = 0.0
        for j in range(ntries):
            if j == 0:
                ay = 0.0
            elif j == 1:
                ay = 0.01
            else:
                ay = ay1 - ty1 * (ay1 - ay0) / (ty1 - ty0)
            vel_setter(sph, vx, ay)
            vspread(lb)
            for _ in range(num):
                step(lb)
            fx, ty = measure(lb, sph)
            if j == 0:
                ay0, ty0 = ay, ty
            elif j == 1:
                ay1, ty1 = ay, ty
            else:
                ay0, ay1, ty0, ty1 = ay1, ay, ty1, ty
        removeobjs(lb)
        m = vx / fx
        out.append(dict(height=height, a_by_h=a_hyd / height, M=m,
                        DbyD0=-6 * math.pi * eta * a_hyd * m, ay=ay, ty=ty))
    return out


def case_mobility2(R):
    L = R.L

    def new(esx, esy, esz):
        h = R.new(esx, esy, esz)
        st = h.contents
        st.objects_moving, st.vel_corrected, st.skip_solid, st.stokes_flow = 1, 1, 0, 1
        st.tau = 0.95
        return h

    def insert(h, a, x, y, z):
        L.insertwall(h, b"z", 0)
        return L.insertsph(h, a, x, y, z)

    def setv(sph, vx, ay):
        sph.contents.velx, sph.contents.angy = vx, ay

    def meas(h, sph):
        L.measure(h)
        return sph.contents.frcx, sph.contents.trqy

    L.removeobjs.argtypes = [C.c_void_p]
    L.removeobjs.restype = None
    out = mobility2_replay(L, new, None, setv, L.lbstep, meas, L.vspread,
                           lambda h: L.removeobjs(C.cast(h, C.c_void_p)), insert)
    with open(os.path.join(OUT, "mobility2.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote mobility2", [round(r["DbyD0"], 4) for r in out])


MASK64 = (1 << 64) - 1


def splitmix64(state):
    """One draw of splitmix64 (Steele, Lea, Flood 2014): returns (new state,
    output)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def sphere_nodes(dims, R, o):
    """Lattice nodes inside the box with (i-ox)^2 + (j-oy)^2 + (k-oz)^2 < R^2
    (FP64, squares summed left to right: the insertsph predicate,
    src/lbobjs.c:212-218)."""
    ax = []
    for d in range(3):
        lo = max(0, int(np.floor(o[d] - R)) - 1)
        hi = min(dims[d] - 1, int(np.ceil(o[d] + R)) + 1)
        ax.append(np.arange(lo, hi + 1, dtype=np.float64) - o[d])
    dx, dy, dz = ax
    r2 = (dx[:, None, None] * dx[:, None, None] + dy[None, :, None] * dy[None, :, None]) \
        + dz[None, None, :] * dz[None, None, :]
    return int((r2 < R * R).sum())


def rsa_pack_restated(dims, R, phi, seed, maxtrials=50_000_000):
    """Random sequential addition of radius-R spheres in a periodic box, as
    the library's slb_rsa_pack defines it: centres (u01 x, u01 y, u01 z)
    from one splitmix64 stream, u01 = (z >> 11) 2^-53; a trial is rejected
    when the minimum-image distance (dx -= L floor(dx/L + 1/2)) to a placed
    centre is below 2R (squared, summed x, y, z); each accepted sphere is
    inserted at every periodic image whose bounding box meets the box
    (a, b, e in -1..1, x slowest); stop once the inserted nodes reach
    ceil(phi N).  Returns the centres and the image rows (R, x, y, z) in
    insertion order.  A restatement for the fixture, independent of the
    library's C code, so that the library's generator is checked too."""
    L = [float(d) for d in dims]
    nc = [int(np.floor(L[d] / (2.0 * R))) for d in range(3)]
    nc = [1 if c < 3 else c for c in nc]
    cs = [L[d] / nc[d] for d in range(3)]
    cells = {}
    N = dims[0] * dims[1] * dims[2]
    target = int(np.ceil(phi * N))
    st, solid, centres, images = seed & MASK64, 0, [], []
    d2min = 4.0 * R * R

    def cell(c):
        return tuple(min(max(int(np.floor(c[d] / cs[d])), 0), nc[d] - 1) for d in range(3))

    t = 0
    while t < maxtrials and solid < target:
        t += 1
        c = []
        for d in range(3):
            st, z = splitmix64(st)
            c.append((z >> 11) * (1.0 / 9007199254740992.0) * L[d])
        ci = cell(c)
        hit = False
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for e in (-1, 0, 1):
                    key = ((ci[0] + a) % nc[0], (ci[1] + b) % nc[1], (ci[2] + e) % nc[2])
                    for q in cells.get(key, ()):
                        r2 = 0.0
                        for d in range(3):
                            dx = c[d] - centres[q][d]
                            dx -= L[d] * np.floor(dx / L[d] + 0.5)
                            r2 += dx * dx
                        if r2 < d2min:
                            hit = True
                            break
                    if hit:
                        break
                if hit:
                    break
            if hit:
                break
        if hit:
            continue
        cells.setdefault(ci, []).append(len(centres))
        centres.append(c)
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for e in (-1, 0, 1):
                    o = (c[0] + a * L[0], c[1] + b * L[1], c[2] + e * L[2])
                    if any(o[d] + R < 0.0 or o[d] - R > L[d] for d in range(3)):
                        continue
                    ns = sphere_nodes(dims, R, o)
                    if ns == 0:
                        continue
                    images.append((R, *o))
                    solid += ns
    return centres, images, t


def case_pack256(R):
    """The 256^3 permeability pack (radius 8, lattice solid fraction 0.35,
    seed 20261202): centres from the restated generator, solid map and
    boundary-link mask written by the reference library (insertsph of every
    image in order, then flattenobj), stored as SHA-256 digests."""
    dims, Rad, phi, seed = (256, 256, 256), 8.0, 0.35, 20261202
    centres, images, trials = rsa_pack_restated(dims, Rad, phi, seed)
    print(f"pack256: {len(centres)} spheres, {len(images)} images, {trials} trials")
    h = R.new(*dims)
    placed = sum(1 if R.L.insertsph(h, *row) else 0 for row in images)
    R.L.syncobjs(h)
    R.L.flattenobj(h)
    R.L.syncobjs(h)
    st = state(R, h)
    rec = dict(dims=list(dims), radius=Rad, phi=phi, seed=seed, trials=trials,
               spheres=len(centres), images=len(images), placed=placed,
               solid=int((st["objid"] >= 0).sum()),
               flat_mask_sha256=sha(st["mask"]), flat_objid_sha256=sha(st["objid"]),
               countfluid=int(R.L.countfluid(h)),
               countmask_links=int(R.L.countmask(h, b"links")),
               centres=[[float(x) for x in c] for c in centres])
    with gzip.open(os.path.join(OUT, "pack256.json.gz"), "wt") as f:
        json.dump(rec, f)
    print("wrote pack256", {k: v for k, v in rec.items() if k != "centres"})


def main():
    R = RefLib(64)
    if sys.argv[1:] == ["packs"]:
        case_packs(R)
        return
    if sys.argv[1:] == ["pack256"]:
        case_pack256(R)
        return
    if sys.argv[1:] == ["mobility2"]:
        case_mobility2(R)
        return
    if sys.argv[1:] == ["checkpoint"]:
        case_checkpoint(R)
        case_vfields(R)
        return
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
    case_sphere_c(R)
    case_images(R)
    case_diagnostics(R)
    case_checkpoint(R)
    case_vfields(R)
    case_mobility2(R)
    case_packs(R)


if __name__ == "__main__":
    main()
