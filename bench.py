#!/usr/bin/env python3
"""Throughput of the D3Q15 lattice-Boltzmann step on one MI355X.

Workload (default): the porous-media permeability case -- a periodic 256^3
box filled by random sequential addition of non-overlapping spheres of
radius 8 (seeded splitmix64, minimum-image overlap test, periodic images
inserted) to a lattice solid fraction of 0.35, flattened to one object,
uniform body force 1e-6 along x on fluid nodes, Stokes flow, tau 0.75, FP32
shifted populations.  One step = one fused bounce-back + stream + collide
pass over all 256^3 nodes (MLUPS counts every node, fluid and solid, as the
reference's ksite-updates/sec does).

Prints one JSON line.  The timed region is K steps between HIP events on
the lattice's stream with the populations already resident in HBM.
The cpu_baseline leg times the reference library (oracle/_ref, built from
the reference's own sources, -D_SINGLE) single-threaded on the same
geometry for a few steps.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may run on (the reference is serial and uses one;
    the host's count is reported beside it)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def build_pack(L, n, R, phi, seed, force):
    h = L.newlb(n, n, n, b"null")
    s = h.contents
    s.stokes_flow = 1
    s.tau = 0.75
    phi_out = C.c_double(0.0)
    t0 = time.perf_counter()
    # random sequential addition and flattenobj in one call, the sphere
    # images counted and rasterised on the GPU
    nsph = L.slb_rsa_pack_flat(h, R, phi, seed, 50_000_000, None, 0, C.byref(phi_out))
    L.zerobf(h)
    L.addbf(h, force, 0.0, 0.0)
    L.syncobjs(h)
    links = L.countmask(h, b"links")
    log(f"pack: {nsph} spheres, lattice solid fraction {phi_out.value:.4f}, "
        f"weighted links {links}, setup {time.perf_counter() - t0:.1f} s")
    return h, nsph, phi_out.value


def cpu_baseline(L, h, n, force, steps):
    """Time the reference library (FP32 build) on the same solid map."""
    from oracle import RefLib
    try:
        R = RefLib(32)
    except ImportError as e:
        log(f"cpu_baseline: reference library unavailable ({e})")
        return None
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "map.dat")
        L.lbsavemap(h, path.encode(), b"char")
        rh = R.L.newlb(n, n, n, b"null")
        rs = rh.contents
        rs.stokes_flow = 1
        rs.tau = 0.75
        R.L.insertsubs(rh, path.encode(), 0, n - 1, 0, n - 1, 0, n - 1)
    R.L.zerobf(rh)
    R.L.addbf(rh, force, 0.0, 0.0)
    R.L.lbstep(rh)  # first step syncs the tables
    t0 = time.perf_counter()
    for _ in range(steps):
        R.L.lbstep(rh)
    dt = time.perf_counter() - t0
    mlups = n ** 3 * steps / dt / 1e6
    return dict(value=round(mlups, 3), unit="MLUPS", cores=1, host_cores=host_cores(),
                kind="reference",
                sample=f"reference SunlightLB (-D_SINGLE, gcc -O3 -ffast-math, serial: 1 of "
                       f"{host_cores()} host cores) on the same {n}^3 pack, {steps} lbstep "
                       f"calls after one untimed step, {dt:.1f} s")


def ref_from_map(R, L, h, n):
    """A reference lattice holding the solid map of lattice h (n^3)."""
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "map.dat")
        L.lbsavemap(h, path.encode(), b"char")
        rh = R.L.newlb(n, n, n, b"null")
        R.L.insertsubs(rh, path.encode(), 0, n - 1, 0, n - 1, 0, n - 1)
    return rh


def cpu_mp_baseline(L, h, n, steps):
    """Time the reference's mpstep (FP32 build, one core) on the same solid
    map: uniform tracer, reflecting walls (the reference has no partial
    adsorption; the cost per node is the same), frozen flow at rest."""
    from oracle import RefLib
    try:
        R = RefLib(32)
    except ImportError as e:
        log(f"cpu_mp_baseline: reference library unavailable ({e})")
        return None
    rh = ref_from_map(R, L, h, n)
    rh.contents.stokes_flow = 1
    R.L.lbstep(rh)
    mp = R.L.newmp(rh, b"null")
    mp.contents.alpha = 0.0
    R.L.mprhoinit(mp)
    R.L.mpinit(mp)
    R.L.mpall(mp, 1.0)
    R.L.mpstep(mp)
    t0 = time.perf_counter()
    for _ in range(steps):
        R.L.mpstep(mp)
    dt = time.perf_counter() - t0
    return dict(value=round(n ** 3 * steps / dt / 1e6, 3), unit="MLUPS", cores=1,
                host_cores=host_cores(), kind="reference",
                sample=f"reference mpstep (-D_SINGLE, serial) on the same {n}^3 map, {steps} "
                       f"steps after one untimed step, {dt:.1f} s")


def _ref_moving_box(R, Lb, sph, u0, tau):
    """One box of the batch in the reference library: sphere 0 and its
    periodic images moving along x, the other spheres fixed."""
    rh = R.L.newlb(Lb, Lb, Lb, b"null")
    st = rh.contents
    st.tau, st.stokes_flow, st.objects_moving = tau, 1, 1
    for q in range(len(sph) - 1, -1, -1):
        Rq, c = sph[q][0], sph[q][1:]
        for a in (-1, 0, 1):
            for b_ in (-1, 0, 1):
                for e in (-1, 0, 1):
                    o = (c[0] + a * Lb, c[1] + b_ * Lb, c[2] + e * Lb)
                    if all(-Rq < o[d] < Lb + Rq for d in range(3)):
                        p = R.L.insertsph(rh, Rq, *o)
                        if p and q == 0:
                            p.contents.velx = u0
    R.L.vspread(rh)
    R.L.lbstep(rh)
    return rh


def _ref_boxes_worker(boxes, steps, u0, tau, barrier, out):
    """One process of the aggregate leg: set its boxes up, wait for every
    other process, then step each box in turn; reports node updates and the
    monotonic-clock interval of the stepping."""
    from oracle import RefLib
    R = RefLib(32)
    hs = [(_ref_moving_box(R, Lb, sph, u0, tau), Lb) for (Lb, sph) in boxes]
    barrier.wait()
    t0 = time.monotonic()
    for rh, _ in hs:
        for _ in range(steps):
            R.L.lbstep(rh)
    out.put((sum(Lb ** 3 for _, Lb in hs) * steps, t0, time.monotonic()))


def cpu_batch_baseline(boxes, steps, u0, tau, nbox=16, per_proc=8, max_procs=16):
    """The reference's lbstep (FP32 build) on boxes of the batch, sphere 0 and
    its images moving along x (one right-hand side per box; the other five
    cost the same).  Serial leg: one core steps the first `nbox` boxes in turn.
    Aggregate leg: the boxes are independent, so one process per host core
    (at most `max_procs`) steps `per_proc` boxes of its own, all at once; the
    aggregate is every process's node updates over the interval from the
    first start to the last finish."""
    import multiprocessing as mp
    from oracle import RefLib
    try:
        R = RefLib(32)
    except ImportError as e:
        log(f"cpu_batch_baseline: reference library unavailable ({e})")
        return None
    nodes, dt = 0, 0.0
    for (Lb, sph) in boxes[:nbox]:
        rh = _ref_moving_box(R, Lb, sph, u0, tau)
        t0 = time.perf_counter()
        for _ in range(steps):
            R.L.lbstep(rh)
        dt += time.perf_counter() - t0
        nodes += Lb ** 3
    res = dict(value=round(nodes * steps / dt / 1e6, 3), unit="MLUPS", cores=1,
               host_cores=host_cores(), kind="reference",
               sample=f"reference lbstep with moving walls (-D_SINGLE, serial) on the first "
                      f"{min(nbox, len(boxes))} boxes of the batch, one right-hand side each, "
                      f"{steps} steps per box, {dt:.1f} s")
    nproc = max(1, min(host_cores(), max_procs, len(boxes) // per_proc))
    # fresh interpreters (spawn): the parent holds a GPU context, the
    # workers touch only the host library
    ctx = mp.get_context("spawn")
    barrier, out = ctx.Barrier(nproc), ctx.Queue()
    procs = [ctx.Process(target=_ref_boxes_worker,
                         args=(boxes[w * per_proc:(w + 1) * per_proc], steps, u0, tau,
                               barrier, out)) for w in range(nproc)]
    for p in procs:
        p.start()
    try:
        got = [out.get(timeout=600) for _ in procs]
    finally:
        for p in procs:
            p.join(timeout=60)
    upd = sum(g[0] for g in got)
    span = max(g[2] for g in got) - min(g[1] for g in got)
    res["aggregate"] = dict(
        value=round(upd / span / 1e6, 3), unit="MLUPS", cores=nproc,
        sample=f"{nproc} processes at once, {per_proc} boxes each "
               f"({nproc * per_proc} distinct boxes of the batch), {steps} steps per box, "
               f"{span:.1f} s from first start to last finish")
    return res


def traffic_bytes(n, precision):
    """HBM bytes per k_step launch measured by rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE on this workload, each scaled by the factor the same counter
    reads for a plain copy of known size (profiles/k_step_traffic.json)."""
    path = os.path.join(ROOT, "profiles", "k_step_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        key = f"{n}_fp{precision}"
        return t["workloads"][key]["traffic_bytes"] if key in t["workloads"] else None
    except (OSError, KeyError, ValueError):
        return None


def converge(L, args):
    """Cold-start solve of the pack to the velocity residual tolerance in FP64
    and in FP32; permeability k = vmean eta / F (examples/rcp_perm.pl:161-167)
    from measure() at the end of each run."""
    res = {"tol": args.tol, "interval": args.interval}
    eta = 72.0 * (0.75 - 0.5) / 3.0
    for prec, window in ((64, 0), (32, 0), (64, args.aa_window), (32, args.aa_window)):
        os.environ["SLB_PRECISION"] = str(prec)
        h, _, _ = build_pack(L, args.size, args.radius, args.phi, args.seed, args.force)
        r = C.c_double(0.0)
        L.slb_residual(h)  # uploads the geometry before the clock starts
        t0 = time.perf_counter()
        if window > 0:
            steps = L.slb_run_to_residual_aa(h, args.tol, args.interval, args.maxsteps, window,
                                             C.byref(r))
        else:
            steps = L.slb_run_to_residual(h, args.tol, args.interval, args.maxsteps, C.byref(r))
        L.measure(h)
        dt = time.perf_counter() - t0
        k = h.contents.vmeanx * eta / args.force
        tag = f"fp{prec}" + (f"_anderson{window}" if window > 0 else "")
        log(f"converge {tag}: {steps} steps to residual {r.value:.3e} in {dt:.2f} s, "
            f"k = {k:.10g}")
        res[tag] = {"steps": steps, "residual": r.value, "seconds": round(dt, 3),
                    "permeability": k, "vmeanx": h.contents.vmeanx}
        L.slb_destroy(h)
    k64, k32 = res["fp64"]["permeability"], res["fp32"]["permeability"]
    res["permeability_rel_diff_fp32_vs_fp64"] = abs(k32 - k64) / abs(k64)
    if args.aa_window > 0:
        for prec in (64, 32):
            a, p_ = res[f"fp{prec}_anderson{args.aa_window}"], res[f"fp{prec}"]
            a["permeability_rel_diff_vs_fp64_plain"] = abs(a["permeability"] - k64) / abs(k64)
            a["wall_time_plain_over_anderson"] = round(p_["seconds"] / a["seconds"], 3)
            a["steps_plain_over_anderson"] = round(p_["steps"] / a["steps"], 3)
    return res


def mp_run(L, args, prec, n, R, phi, peclet, alpha, pads, baseline):
    """One MP workload run in the given lattice precision (see mp_workload)."""
    os.environ["SLB_PRECISION"] = str(prec)
    h, nsph, phi_got = build_pack(L, n, R, phi, args.seed + 1, 1e-6)
    r = C.c_double(0.0)
    steps = L.slb_run_to_residual(h, args.tol, args.interval, args.maxsteps, C.byref(r))
    L.measure(h)
    nfluid = L.countfluid(h)
    porosity = nfluid / float(n ** 3)
    U = h.contents.vmeanx / porosity
    D = (1.0 - alpha) / 6.0
    target = peclet * D / (2.0 * R)
    L.slb_scale_pops(h, target / U)
    mp = L.newmp(h, b"null")
    mp.contents.alpha = alpha
    L.mpinit(mp)
    L.mpall(mp, 1.0)
    L.slb_mp_set_adsorption(mp, pads)
    L.slb_mp_track_moments(mp, 0)
    s0 = (C.c_double * 6)()
    L.slb_mp_sums(mp, s0)
    L.slb_mp_run(mp, 10)
    L.mpinit(mp)
    L.mpall(mp, 1.0)
    L.slb_mp_track_moments(mp, 0)
    ms = L.slb_mp_time_steps(mp, args.mp_steps)
    s1 = (C.c_double * 6)()
    L.slb_mp_sums(mp, s1)
    ptot = s1[0]
    mean = s1[3] / ptot
    var = s1[4] / ptot - mean * mean
    N = n ** 3
    mlups = N * args.mp_steps / (ms * 1e-3) / 1e6
    # algorithmic bytes: the 15 link factors of a node (from its 15 pulled
    # populations of the frozen flow, precomputed once per run) and the node's
    # {P, m1, m2} read and written
    bpl = (15 + 2 * 3) * (4 if prec == 32 else 8)
    res = {
        "workload": f"{n}^3 RSA pack R {R:g} phi {phi_got:.4f} ({nsph} spheres), flow solved "
                    f"to residual {r.value:.2e} in {steps} steps, scaled to Pe {peclet:g} "
                    f"(U {target:.5g}, D {D:.5g}), alpha {alpha:g}, adsorption {pads:g}",
        "steps": args.mp_steps,
        "surviving_fraction": ptot / s0[0],
        "adsorbed_fraction": (s1[1]) / s0[0],
        "mean_displacement_x": mean,
        "variance_displacement_x": var,
        "expected_mean_no_adsorption": target * args.mp_steps,
        "ms_per_step": ms / args.mp_steps,
        "MLUPS": round(mlups, 1),
        "algorithmic_bytes_per_node": bpl,
        "achieved_GBs": round(bpl * N / (ms / args.mp_steps * 1e-3) / 1e9, 1),
    }
    if baseline:
        res["cpu_baseline"] = cpu_mp_baseline(L, h, n, args.cpu_mp_steps)
    L.slb_destroy(h)
    return res


MP_QUANTITIES = ("surviving_fraction", "adsorbed_fraction", "mean_displacement_x",
                 "variance_displacement_x")


def mp_workload(L, args, n=128, R=6.0, phi=0.25, peclet=10.0, alpha=0.0, pads=0.01):
    """Passive-scalar moment propagation on the steady flow of a sphere pack.

    Steady Stokes flow of the pack (cold start to the residual tolerance), the
    populations scaled so that Pe = U (2R) / D = peclet with U the
    interstitial mean speed (superficial vmean / porosity) and D = (1 -
    alpha)/6; uniform unit tracer on fluid nodes, partial adsorption pads at
    boundary solid nodes, displacement moments along x, mp_steps steps in the
    lattice precision with FP64 moment sums.  An FP32 run is repeated in FP64
    (flow solve included) and every reported quantity carries its relative
    difference from the FP64 run."""
    res = mp_run(L, args, args.precision, n, R, phi, peclet, alpha, pads,
                 not args.no_cpu_baseline)
    if args.precision == 32:
        r64 = mp_run(L, args, 64, n, R, phi, peclet, alpha, pads, False)
        res["fp64"] = {q: r64[q] for q in MP_QUANTITIES + ("ms_per_step",)}
        res["rel_diff_fp32_vs_fp64"] = {
            q: abs(res[q] - r64[q]) / abs(r64[q]) for q in MP_QUANTITIES}
    log(f"mp: {res}")
    return res


def make_boxes(nbox, seed):
    """Seeded boxes: cubic edge U{32..64}, 2..20 non-overlapping spheres of
    radius U[3, 6] (minimum-image gap of one lattice unit); sphere 0 is the
    one that moves."""
    import numpy as np
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(nbox):
        Lb = int(rng.integers(32, 65))
        ns = int(rng.integers(2, 21))
        sph = []
        for _t in range(20000):
            if len(sph) == ns:
                break
            R = float(rng.uniform(3.0, 6.0))
            c = rng.uniform(0.0, Lb, 3)
            ok = True
            for (R2, *c2) in sph:
                d = c - np.array(c2)
                d -= Lb * np.round(d / Lb)
                if np.sqrt((d * d).sum()) < R + R2 + 1.0:
                    ok = False
                    break
            if ok:
                sph.append((R, *c))
        boxes.append((Lb, sph))
    return boxes


def insert_periodic(L, h, Lb, R, c):
    """Insert a sphere and every periodic image whose bounding box meets the
    box, each image its own object centred on itself (correct lever arms
    for the torque, which the reference sums without minimum image)."""
    objs = []
    for a in (-1, 0, 1):
        for b_ in (-1, 0, 1):
            for e in (-1, 0, 1):
                o = (c[0] + a * Lb, c[1] + b_ * Lb, c[2] + e * Lb)
                if all(-R < o[d] < Lb + R for d in range(3)):
                    p = L.insertsph(h, R, *o)
                    if p:
                        objs.append(C.addressof(p.contents))
    return objs


def build_batch(L, prec, boxes, u0, tau):
    """The batch of the resistance workload: every box with its spheres
    (periodic images as separate objects, sphere 0 inserted last so that
    its images head the object list) and 6 right-hand sides -- sphere 0
    translating along x, y, z at u0, then rotating about x, y, z at u0/R0 --
    the other spheres fixed.  Returns the uncommitted batch and, per box,
    (object rows of sphere 0's images, object count, R0)."""
    import numpy as np
    b = L.slb_batch_new(prec)
    moving_rows = []
    for (Lb, sph) in boxes:
        h = L.newlb(Lb, Lb, Lb, b"null")
        h.contents.tau, h.contents.stokes_flow, h.contents.objects_moving = tau, 1, 1
        mover = []
        for q in range(len(sph) - 1, -1, -1):   # sphere 0 inserted last: head of the list
            objs = insert_periodic(L, h, Lb, sph[q][0], sph[q][1:])
            if q == 0:
                mover = objs
        L.syncobjs(h)
        order = []
        p = h.contents.obj_head
        while p:
            order.append(C.addressof(p.contents))
            p = p.contents.next_obj
        rows = [order.index(a) for a in mover]
        vel = rhs_velocities(len(order), rows, sph[0][0], u0)
        flat = np.ascontiguousarray(vel.reshape(-1))
        L.slb_batch_add(b, h, 6, flat.ctypes.data_as(C.POINTER(C.c_double)))
        moving_rows.append((rows, len(order), sph[0][0]))
        L.slb_destroy(h)
    return b, moving_rows


def rhs_velocities(nobj, rows, R0, u0):
    """(6, nobj, 6) object velocity rows of the 6 right-hand sides."""
    import numpy as np
    vel = np.zeros((6, nobj, 6))
    for k in range(6):
        vel[k, rows, k] = u0 if k < 3 else u0 / R0
    return vel


def resistance_matrices(L, b, moving_rows, u0):
    """6x6 resistance matrix of sphere 0 in every box after
    slb_batch_measure: column k = (force, torque) summed over sphere 0's
    images for right-hand side k, per unit velocity."""
    import numpy as np
    mats = []
    for q, (rows, nobj, R0) in enumerate(moving_rows):
        Rm = np.zeros((6, 6))
        for k in range(6):
            frc = np.zeros(3 * nobj)
            trq = np.zeros(3 * nobj)
            L.slb_batch_forces(b, q, k, frc.ctypes.data_as(C.POINTER(C.c_double)),
                               trq.ctypes.data_as(C.POINTER(C.c_double)))
            F = frc.reshape(nobj, 3)[rows].sum(0)
            T = trq.reshape(nobj, 3)[rows].sum(0)
            amp = u0 if k < 3 else u0 / R0
            Rm[:, k] = np.concatenate([F, T]) / amp
        mats.append(Rm)
    return mats


def batch_run(L, args, prec, boxes, u0, tau, tune):
    """One batched run in the given precision: returns the throughput record
    and the 6x6 resistance matrix of sphere 0 in every box."""
    os.environ["SLB_PRECISION"] = str(prec)
    t0 = time.perf_counter()
    b, moving_rows = build_batch(L, prec, boxes, u0, tau)
    L.slb_batch_commit(b)
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    if tune:
        L.slb_batch_autotune(b, None)
    tune_s = time.perf_counter() - t1
    cfg = (C.c_int * 4)()
    L.slb_batch_get_config(b, cfg)
    nodes = L.slb_batch_nodes(b)
    L.slb_batch_run(b, 32)
    ms = L.slb_batch_time_steps(b, args.batch_steps)
    L.slb_batch_measure(b)
    mlups = nodes * args.batch_steps / (ms * 1e-3) / 1e6
    bpl = 2 * 15 * (4 if prec == 32 else 8)
    mats = resistance_matrices(L, b, moving_rows, u0)
    L.slb_batch_free(b)
    res = {
        "workload": f"{len(boxes)} periodic boxes, edge U{{32..64}}, 2-20 spheres radius "
                    f"U[3,6] (seed {args.seed + 7}), sphere 0 translating/rotating (6 rhs), "
                    f"Stokes, tau {tau:g}, {'f32' if prec == 32 else 'f64'}",
        "problems": 6 * len(boxes),
        "nodes_per_step": nodes,
        "steps": args.batch_steps,
        "ms_per_step": round(ms / args.batch_steps, 4),
        "MLUPS": round(mlups, 1),
        "achieved_GBs": round(bpl * nodes / (ms / args.batch_steps * 1e-3) / 1e9, 1),
        "launch": {"block": cfg[0], "nodes_per_thread": cfg[1], "grid": cfg[2],
                   "tiles": cfg[3], "autotune_s": round(tune_s, 2)},
        "setup_s": round(setup, 1),
    }
    return res, mats


def batch_workload(L, args, u0=1e-3, tau=1.0):
    """512 independent periodic boxes, each with 6 right-hand sides (unit
    translation and rotation of sphere 0, the other spheres fixed), stepped
    together: one launch per step over all problems, autotuned launch shape,
    HIP-graph replay.  Reports throughput and the 6x6 resistance matrices of
    sphere 0 after measure(); an FP32 batch is repeated in FP64 (same launch
    shape class, not timed against it) and the matrices are compared."""
    import numpy as np
    boxes = make_boxes(args.batch_boxes, args.seed + 7)
    res, mats = batch_run(L, args, args.precision, boxes, u0, tau, True)
    eta = 72.0 * (tau - 0.5) / 3.0
    drag = [-np.mean(np.diag(Rm)[:3]) / (6 * np.pi * eta * bx[1][0][0])
            for Rm, bx in zip(mats, boxes)]
    asym = [np.abs(Rm - Rm.T).max() / np.abs(Rm).max() for Rm in mats]
    res["translational_drag_over_stokes_median"] = float(np.median(drag))
    res["resistance_asymmetry_median"] = float(np.median(asym))
    if args.precision == 32:
        r64, mats64 = batch_run(L, args, 64, boxes, u0, tau, False)
        diff = [np.abs(a - c).max() / np.abs(c).max() for a, c in zip(mats, mats64)]
        res["fp64"] = {"ms_per_step": r64["ms_per_step"], "MLUPS": r64["MLUPS"]}
        res["resistance_rel_diff_fp32_vs_fp64"] = {"max": float(np.max(diff)),
                                                   "median": float(np.median(diff))}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_batch_baseline(boxes, args.cpu_batch_steps, u0, tau)
    log(f"batch: {res}")
    return res


def warm_params(rng, count):
    """Seeded pack parameters: radius U[4, 10], lattice solid fraction
    U[0.2, 0.45]; random sequential addition up to 0.34, overlapping spheres
    (Boolean model) above, where sequential addition jams."""
    out = []
    for _ in range(count):
        R = float(rng.uniform(4.0, 10.0))
        phi = float(rng.uniform(0.2, 0.45))
        out.append((R, phi, int(phi > 0.34), int(rng.integers(1, 2 ** 62))))
    return out


def warm_maps(L, n, params, workers):
    """Solid maps of the packs, generated on host threads (slb_pack_map is
    thread-safe and ctypes releases the GIL), yielded in order."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    def one(p):
        R, phi, ov, seed = p
        m = np.empty(n ** 3, np.uint8)
        po = C.c_double(0.0)
        L.slb_pack_map(n, n, n, R, phi, seed, ov, 50_000_000, m.ctypes.data, C.byref(po))
        return m, po.value

    with ThreadPoolExecutor(max_workers=workers) as ex:
        yield from ex.map(one, params)


def warm_lattice(L, n, solid, force, precision=32):
    h = L.newlb(n, n, n, b"null")
    st = h.contents
    st.stokes_flow = 1
    st.tau = 0.75
    L.slb_insert_map(h, solid.ctypes.data)
    L.zerobf(h)
    L.addbf(h, force, 0.0, 0.0)
    L.syncobjs(h)
    L.slb_set_precision(h, precision)
    return h


def solve_loop(L, h, tol, interval, maxsteps):
    """Run to slb_residual() < tol in chunks of `interval` (as
    slb_run_to_residual); returns steps, the residual history and the final
    residual."""
    L.slb_residual_reset(h)
    L.slb_residual(h)
    steps, hist, r = 0, [], -1.0
    while steps < maxsteps:
        L.slb_run(h, interval)
        steps += interval
        r = L.slb_residual(h)
        hist.append(r)
        if 0.0 <= r < tol:
            break
    return steps, hist, r


def train_network(L, args, net, res, rng, n, F, workers):
    """Training data (cold solves of seeded packs to the residual
    tolerance, each noted as a cold residual curve and added as a sample)
    and Adam training of the network."""
    import numpy as np
    # ---- training data: cold solves of seeded packs ----
    t0 = time.perf_counter()
    cold_steps = []
    for i, (solid, phi) in enumerate(warm_maps(L, n, warm_params(rng, args.warm_train), workers)):
        h = warm_lattice(L, n, solid, F)
        steps, hist, r = solve_loop(L, h, args.tol, args.interval, args.maxsteps)
        L.slb_warm_note_cold(net, (C.c_double * len(hist))(*hist), len(hist), args.interval)
        cold_steps.append(steps)
        L.slb_warm_add_sample(net, h)
        L.slb_destroy(h)
        if (i + 1) % max(1, args.warm_train // 10) == 0:
            log(f"warm: {i + 1}/{args.warm_train} training solves, {time.perf_counter() - t0:.0f} s "
                f"(median {int(np.median(cold_steps))} steps)")
    res["data_seconds"] = round(time.perf_counter() - t0, 1)
    res["train_cold_steps_median"] = float(np.median(cold_steps))
    # ---- training ----
    hist = (C.c_double * args.warm_iters)()
    t0 = time.perf_counter()
    last = L.slb_warm_train(net, args.warm_iters, args.warm_lr, args.seed + 7, hist)
    dt = time.perf_counter() - t0
    hv = np.array(hist[:])
    k = max(1, args.warm_iters // 20)
    curve = [float(hv[i:i + k].mean()) for i in range(0, args.warm_iters, k)]
    log(f"warm: trained {args.warm_iters} iterations in {dt:.1f} s ({dt / args.warm_iters * 1e3:.2f} "
        f"ms/iter), loss {curve[0]:.3e} -> {last:.3e}")
    pr, pres = C.c_int(0), C.c_double(0.0)
    L.slb_warm_get_fallback(net, C.byref(pr), C.byref(pres))
    res.update({"train_iters": args.warm_iters, "train_seconds": round(dt, 1),
                "ms_per_iter": round(dt / args.warm_iters * 1e3, 3), "loss_curve": curve,
                "final_loss": last, "fallback_probe_steps": pr.value,
                "fallback_probe_residual": pres.value})


def warm_workload(L, args):
    """Learned warm start: generate and solve training packs on the GPU,
    train the network on them, then solve unseen packs cold and warm to the
    same residual and compare steps, wall time and permeability."""
    import numpy as np
    rng = np.random.default_rng(args.seed + 5)
    n = args.warm_size
    F = args.force
    eta = 72.0 * (0.75 - 0.5) / 3.0
    workers = max(1, min(16, os.cpu_count() or 1))
    proj = tuple(int(x) for x in args.warm_proj.split(","))
    if args.warm_load:
        net = L.slb_warm_load(args.warm_load.encode())
        if not net:
            raise RuntimeError(f"cannot load warm-start weights {args.warm_load}")
        res = {"weights": os.path.relpath(args.warm_load, ROOT), "workers": workers}
    else:
        net = L.slb_warm_new(args.seed)
        res = {"train_packs": args.warm_train, "train_size": n, "workers": workers}
        train_network(L, args, net, res, rng, n, F, workers)
    # ---- start-policy calibration: warm solves of unseen training-distribution
    # packs, without fallback, for their median residual curve (with the
    # projection when --warm-proj-calib; slb_warm_solve then projects too) ----
    calib_proj = proj if args.warm_proj_calib else (0, 0)
    L.slb_warm_set_projection(net, *calib_proj)
    if args.warm_calib > 0:
        t0 = time.perf_counter()
        crng = np.random.default_rng(args.seed + 13)
        calib = []
        for solid, phi in warm_maps(L, n, warm_params(crng, args.warm_calib), workers):
            h = warm_lattice(L, n, solid, F)
            L.slb_warm_set_mode(net, 1)
            L.slb_warm_start(net, h)
            steps, hist, r = solve_loop(L, h, args.tol, args.interval, args.maxsteps)
            L.slb_warm_note_warm(net, (C.c_double * len(hist))(*hist), len(hist), args.interval)
            calib.append(steps)
            L.slb_destroy(h)
        res["calibration"] = {"packs": args.warm_calib,
                              "seconds": round(time.perf_counter() - t0, 1),
                              "projection": list(calib_proj),
                              "warm_steps_median": float(np.median(calib)) if calib else None}
    res["start_warm_at"] = {f"{tol:.0e}": L.slb_warm_policy(net, tol, args.interval)
                            for tol in (1e-3, 1e-4, 1e-5, 1e-6)}
    log(f"warm: start policy {res['start_warm_at']}")
    if args.warm_save:
        L.slb_warm_save(net, args.warm_save.encode())
    # ---- evaluation on unseen packs ----
    sizes = [int(x) for x in args.warm_eval_sizes.split(",")]
    rows = []
    erng = np.random.default_rng(args.seed + 11)
    for size in sizes:
        params = warm_params(erng, args.warm_eval // len(sizes))
        for (solid, phi), p in zip(warm_maps(L, size, params, workers), params):
            row = {"n": size, "R": round(p[0], 3), "phi": round(phi, 4), "overlap": p[2]}
            first = len([r for r in rows if r["n"] == size]) < 4
            modes = ["cold", "warm", "warm_trace", "equilibrium_trace"]
            if proj[0] > 0:
                modes += ["proj_trace"]
            if args.aa_window > 0:
                modes += ["cold_aa", "warm_aa"]
            for mode in modes:
                if mode == "equilibrium_trace" and not first:
                    continue
                h = warm_lattice(L, size, solid, F)
                L.slb_residual(h)  # uploads the geometry before the clock starts
                hist = None
                t0 = time.perf_counter()
                if mode == "cold":
                    steps, hist, r = solve_loop(L, h, args.tol, args.interval, args.maxsteps)
                    used = 0
                elif mode == "warm":
                    L.slb_warm_set_mode(net, 1)
                    L.slb_warm_set_projection(net, *calib_proj)
                    rr, uw = C.c_double(0.0), C.c_int(0)
                    steps = L.slb_warm_solve(net, h, args.tol, args.interval, args.maxsteps,
                                             C.byref(rr), C.byref(uw))
                    r, used = rr.value, uw.value
                elif mode in ("cold_aa", "warm_aa"):
                    # Anderson-mixed solve from rest, or from the warm start
                    # with no safeguard
                    used = 0
                    if mode == "warm_aa":
                        L.slb_warm_set_mode(net, 1)
                        L.slb_warm_set_projection(net, *proj)
                        L.slb_warm_start(net, h)
                        used = 1
                    rr = C.c_double(0.0)
                    steps = L.slb_run_to_residual_aa(h, args.tol, args.interval, args.maxsteps,
                                                     args.aa_window, C.byref(rr))
                    r = rr.value
                else:
                    # the warm start without the fallback, for its residual
                    # history (proj_trace: with the divergence projection)
                    L.slb_warm_set_mode(net, 0 if mode == "equilibrium_trace" else 1)
                    L.slb_warm_set_projection(net, *(proj if mode == "proj_trace" else (0, 0)))
                    L.slb_warm_start(net, h)
                    steps, hist, r = solve_loop(L, h, args.tol, args.interval, args.maxsteps)
                    used = 1
                L.measure(h)
                dt = time.perf_counter() - t0
                kperm = h.contents.vmeanx * eta / F
                row[mode] = {"steps": steps, "seconds": round(dt, 4), "residual": r,
                             "permeability": kperm, "used_warm": used}
                if hist is not None:
                    for tol in (1e-3, 1e-4, 1e-5, 1e-6):
                        k = next((i for i, x in enumerate(hist) if 0.0 <= x < tol), None)
                        row[mode][f"steps_to_{tol:.0e}"] = (None if k is None
                                                            else (k + 1) * args.interval)
                if mode == "cold":
                    grp = (C.c_double * 3)()
                    row["pred_rel_mse"] = L.slb_warm_eval(net, h, grp)
                    row["pred_rel_mse_groups"] = list(grp)
                L.slb_destroy(h)
            for mode in ("warm", "warm_trace", "equilibrium_trace", "proj_trace", "cold_aa",
                         "warm_aa"):
                if mode in row:
                    row[mode]["k_rel_diff"] = (abs(row[mode]["permeability"] - row["cold"]["permeability"])
                                               / abs(row["cold"]["permeability"]))
            wt, ct = row["warm_trace"], row["cold"]
            log(f"warm eval n={size} R={p[0]:.2f} phi={phi:.3f}: cold {ct['steps']} steps "
                f"{ct['seconds']:.3f} s, warm {row['warm']['steps']} steps "
                f"{row['warm']['seconds']:.3f} s (kept warm: {row['warm']['used_warm']}), "
                f"to 1e-3/1e-4/1e-5 cold {ct['steps_to_1e-03']}/{ct['steps_to_1e-04']}/"
                f"{ct['steps_to_1e-05']} warm {wt['steps_to_1e-03']}/{wt['steps_to_1e-04']}/"
                f"{wt['steps_to_1e-05']}, pred err u/rho/s "
                + "/".join(f"{x:.2e}" for x in row["pred_rel_mse_groups"]))
            rows.append(row)
    L.slb_warm_free(net)
    summ = {}
    for size in sizes:
        rs = [r for r in rows if r["n"] == size]
        if not rs:
            continue
        cs = np.array([r["cold"]["seconds"] for r in rs])
        ws = np.array([r["warm"]["seconds"] for r in rs])
        cst = np.array([r["cold"]["steps"] for r in rs])
        wst = np.array([r["warm"]["steps"] for r in rs])
        dk = np.array([r["warm"]["k_rel_diff"] for r in rs])
        dkt = np.array([r["warm_trace"]["k_rel_diff"] for r in rs])
        eq = [r for r in rs if "equilibrium_trace" in r]

        def ratio(key, mode, sub=rs):
            c = [r["cold"][key] for r in sub]
            w = [r[mode][key] for r in sub]
            if any(x is None for x in c + w):
                return None
            return float(sum(c) / sum(w))

        summ[str(size)] = {
            "packs": len(rs),
            "cold_steps_mean": float(cst.mean()), "warm_steps_mean": float(wst.mean()),
            "cold_seconds_total": float(cs.sum()), "warm_seconds_total": float(ws.sum()),
            "wall_time_ratio_cold_over_warm": float(cs.sum() / ws.sum()),
            "steps_ratio_cold_over_warm": float(cst.sum() / wst.sum()),
            "started_or_fell_back_cold": int(sum(1 for r in rs if r["warm"]["used_warm"] == 0)),
            "k_rel_diff_max": float(dk.max()), "k_rel_diff_median": float(np.median(dk)),
            "no_fallback_k_rel_diff_max": float(dkt.max()),
            "no_fallback_steps_ratio": {t: ratio(f"steps_to_{t}", "warm_trace")
                                        for t in ("1e-03", "1e-04", "1e-05", "1e-06")},
            "equilibrium_steps_ratio": {t: ratio(f"steps_to_{t}", "equilibrium_trace", eq)
                                        for t in ("1e-03", "1e-04", "1e-05", "1e-06")} if eq else None,
            "projection": list(proj) if proj[0] > 0 else None,
            "projection_steps_ratio": {t: ratio(f"steps_to_{t}", "proj_trace")
                                       for t in ("1e-03", "1e-04", "1e-05", "1e-06")}
            if proj[0] > 0 else None,
            "projection_wall_ratio": (float(cs.sum() / sum(r["proj_trace"]["seconds"] for r in rs))
                                      if proj[0] > 0 else None),
            "projection_k_rel_diff_max": (float(max(r["proj_trace"]["k_rel_diff"] for r in rs))
                                          if proj[0] > 0 else None),
            "pred_rel_mse_mean": float(np.mean([r["pred_rel_mse"] for r in rs])),
        }
        if args.aa_window > 0:
            def tot(mode, key):
                return float(sum(r[mode][key] for r in rs))
            dka = np.array([abs(r[m]["permeability"] - r["cold"]["permeability"])
                            / abs(r["cold"]["permeability"]) for r in rs
                            for m in ("cold_aa", "warm_aa")])
            summ[str(size)]["anderson"] = {
                "window": args.aa_window,
                "steps_cold_over_cold_aa": tot("cold", "steps") / tot("cold_aa", "steps"),
                "wall_cold_over_cold_aa": tot("cold", "seconds") / tot("cold_aa", "seconds"),
                "steps_cold_over_warm_aa": tot("cold", "steps") / tot("warm_aa", "steps"),
                "wall_cold_over_warm_aa": tot("cold", "seconds") / tot("warm_aa", "seconds"),
                "steps_cold_aa_over_warm_aa": tot("cold_aa", "steps") / tot("warm_aa", "steps"),
                "k_rel_diff_max": float(dka.max()),
            }
    res["eval"] = summ
    res["eval_rows"] = rows
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--radius", type=float, default=8.0)
    ap.add_argument("--phi", type=float, default=0.35)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--force", type=float, default=1e-6)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-mp-steps", type=int, default=100)
    ap.add_argument("--cpu-batch-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the cold-start solves of the pack to the velocity residual "
                         "tolerance (FP64 and FP32, plain and Anderson-mixed) and the "
                         "permeability comparison")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--interval", type=int, default=100)
    ap.add_argument("--maxsteps", type=int, default=200000)
    ap.add_argument("--aa-window", type=int, default=5,
                    help="Anderson window of the accelerated cold-start solves (0: none)")
    ap.add_argument("--mp", action="store_true",
                    help="also run the passive-scalar workload: 128^3 pack (R 6, phi 0.25) "
                         "at Peclet 10, adsorption 0.01, 5000 moment-propagation steps")
    ap.add_argument("--mp-steps", type=int, default=5000)
    ap.add_argument("--batch", action="store_true",
                    help="also run the batched resistance-matrix workload (512 boxes x 6 "
                         "right-hand sides in one launch per step)")
    ap.add_argument("--batch-boxes", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--warm", action="store_true",
                    help="learned warm start: training data, training, warm vs cold evaluation")
    ap.add_argument("--warm-only", action="store_true",
                    help="run only the warm-start workload (no step-throughput line)")
    ap.add_argument("--warm-size", type=int, default=128)
    ap.add_argument("--warm-train", type=int, default=2000)
    ap.add_argument("--warm-iters", type=int, default=8000)
    ap.add_argument("--warm-lr", type=float, default=1e-3)
    ap.add_argument("--warm-eval", type=int, default=64)
    ap.add_argument("--warm-calib", type=int, default=16,
                    help="unseen packs solved warm after training for the start policy")
    ap.add_argument("--warm-eval-sizes", default="128,256")
    ap.add_argument("--warm-save", default="")
    ap.add_argument("--warm-load", default="",
                    help="evaluate the weights in this file instead of training a network")
    ap.add_argument("--warm-proj", default="200,3",
                    help="divergence projection of the warm start in the evaluation: CG "
                         "iterations per pass, passes (0,0: none)")
    ap.add_argument("--warm-proj-calib", action="store_true",
                    help="calibrate the start policy (and run slb_warm_solve) with the projection")
    ap.add_argument("--warm-weights", default=os.path.join(ROOT, "sunlightlb_amd", "weights",
                                                           "warm_unet128.bin"),
                    help="weights for the default warm-vs-cold leg (skipped when absent)")
    ap.add_argument("--warm-default-packs", type=int, default=6,
                    help="unseen 128^3 packs of the default warm-vs-cold leg")
    ap.add_argument("--no-warm", action="store_true", help="skip the default warm-vs-cold leg")
    ap.add_argument("--copy-test", action="store_true",
                    help="only run the plain device copy kernel (counter calibration)")
    args = ap.parse_args()
    # the one JSON line goes to the original stdout; everything the library
    # prints (newlb banners, lbloadvel notes) goes to stderr
    json_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    os.dup2(2, 1)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the lattice-Boltzmann path here is "
              "single-GPU (multi-GPU slab decomposition is designed, not built).",
              file=sys.stderr)
        sys.exit(2)

    os.environ["SLB_PRECISION"] = str(args.precision)
    import sunlightlb_amd as s
    L = s.lib()
    copy_bytes = 1 << 31
    copy_ms = L.slb_hbm_copy_test(copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"device copy: {copy_bytes >> 20} MiB in {copy_ms * 1e3:.1f} us, {copy_gbs:.0f} GB/s "
        f"(read + write)")
    if args.copy_test:
        return
    if args.warm_only:
        print(json.dumps({"warm": warm_workload(L, args)}), file=json_out, flush=True)
        return
    n = args.size
    h, nsph, phi = build_pack(L, n, args.radius, args.phi, args.seed, args.force)
    N = n ** 3

    kernel = L.slb_step_kernel(h).decode()
    L.slb_run(h, args.warmup)
    ms = L.slb_time_steps(h, args.steps)
    ms_per_step = ms / args.steps
    mlups = N * args.steps / (ms * 1e-3) / 1e6
    bytes_per_lup = 2 * 15 * (4 if args.precision == 32 else 8)
    achieved = bytes_per_lup * N / (ms_per_step * 1e-3) / 1e9
    log(f"gpu: {ms_per_step * 1e3:.1f} us/step, {mlups:.1f} MLUPS, {achieved:.0f} GB/s "
        f"algorithmic ({bytes_per_lup} B/LUP)")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(L, h, n, args.force, args.cpu_steps)
    conv = None
    if not args.no_converge:
        L.slb_destroy(h)
        conv = converge(L, args)

    dtype = "f32" if args.precision == 32 else "f64"
    out = {
        "metric": "MLUPS (D3Q15 fused bounce-back+stream+collide, all nodes)",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": f"porous-media permeability: periodic {n}^3 RSA pack of radius-"
                        f"{args.radius:g} spheres to lattice solid fraction {args.phi:g} "
                        f"(seed {args.seed}), body force {args.force:g} x, Stokes, tau 0.75",
            "nodes": N, "spheres": nsph, "solid_fraction": round(phi, 5),
            "populations": f"{dtype} shifted (a = f - w rho0/72), tiles of 1024 nodes x 15 runs, "
                           "pull, ping-pong",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic_bytes(n, args.precision),
            # the kernel the library launches for this lattice's steps
            "kernel": kernel + {"k_step_lds": " (fused pull step, k pulls staged in LDS)",
                                "k_step_ldsr": " (fused pull step, k pulls staged in LDS, "
                                               "whole rows per block)",
                                "k_step": " (fused pull step)"}.get(kernel, ""),
            "algorithmic_bytes_per_node": bytes_per_lup,
            "device_copy_GBs": round(copy_gbs, 1),
        },
        "cpu_baseline": cpu,
    }
    if conv is not None:
        if cpu is not None:
            # the reference's FP32 step time on this pack times the GPU FP32
            # run's step count: extrapolated, not run
            conv["cpu_fp32_seconds_to_residual_extrapolated"] = round(
                conv["fp32"]["steps"] * N / (cpu["value"] * 1e6), 1)
        out["converge"] = conv
    if args.mp:
        out["mp"] = mp_workload(L, args)
    if args.batch:
        out["batch"] = batch_workload(L, args)
    if args.warm:
        out["warm"] = warm_workload(L, args)
    elif not args.no_warm and os.path.exists(args.warm_weights):
        # the learned warm start on the committed weights: unseen 128^3 packs
        # solved cold and warm to the residual tolerance
        args.warm_load, args.warm_calib = args.warm_weights, 0
        args.warm_eval, args.warm_eval_sizes = args.warm_default_packs, "128"
        out["warm"] = warm_workload(L, args)
        out["warm"].pop("eval_rows", None)
    print(json.dumps(out), file=json_out, flush=True)


if __name__ == "__main__":
    main()
