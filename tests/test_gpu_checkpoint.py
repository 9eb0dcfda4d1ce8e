"""Checkpoint and velocity-field calls against the reference library's own
files and fields (tests/golden/checkpoint.npz, vfields.npz, written by
make_golden.py from oracle/_ref):

* lbsave's .dms (src/lbstore.c:36-91: bbl + propagate, 10 moments per node
  as float, collide) and lbsaveobj's .obj (:523-547) of the same state;
* the USR2 signal path (sighook, then the end of the next collide:
  src/lbsignal.c:40-88, src/lbkernel.c:873-875) writing those two files,
  and USR1 refreshing the monitor file;
* lbload (:102-218) and lbloadobj (:555-586, with rescale) of the
  reference's files;
* vset (src/lbobjs.c:623-643), vzero and vgrab (src/lbmeasure.c:110-166,
  corrected by -b/2 after collide and +b/2 after propagate);
  the reference library is built with its own -O3 -ffast-math, so values
  computed by the same expression may differ in the last bit;
* the velocity residual recomputed in numpy from two population snapshots.

The .dms holds the moments rounded to float.  The GPU's FP64 moments equal
the reference's to about 1e-16 relative, so the floats agree bit for bit
except where a value lies within that distance of a float rounding
boundary (one unit in the last place apart) or is zero up to roundoff
(a stress component that cancels by symmetry, whose sign and low bits
depend on summation order); compare_dms states the bound.
"""
import ctypes as C
import os
import signal
import time

import numpy as np
import pytest

from conftest import load_golden, rel_l2

pytestmark = pytest.mark.gpu


def lib():
    import sunlightlb_amd as s
    return s.lib()


def ckpt_lattice(c, name="null"):
    import sunlightlb_amd as s
    lat = s.Lattice(*c["dims"], name=name, precision=64)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow = c["tau"], 1
    L.insertwall(h, c["wall"][0].encode(), int(c["wall"][1]))
    sph = L.insertsph(h, *c["sphere"])
    sph.contents.velx, sph.contents.vely, sph.contents.velz = c["sphere_vel"]
    sph.contents.angx, sph.contents.angy, sph.contents.angz = c["sphere_ang"]
    L.zerobf(h)
    L.addbf(h, *c["bf"])
    return lat


def objrows(lat):
    return np.array([[o.posx, o.posy, o.posz, o.velx, o.vely, o.velz, o.angx, o.angy, o.angz,
                      o.frcx, o.frcy, o.frcz, o.trqx, o.trqy, o.trqz, o.id]
                     for o in lat.objects()])


def compare_dms(got, ref):
    """Same length and layout (10 floats per node); every float equal to the
    reference's, or one float ulp from it, or -- for a moment that is zero
    up to roundoff on that node (a stress component cancelling by symmetry)
    -- within 1e-12 of that moment's largest magnitude, where summation
    order decides the sign and low bits.  Returns the count of floats not
    bit-identical."""
    assert len(got) == len(ref) and len(ref) % 40 == 0
    g = np.frombuffer(bytes(got), dtype=np.float32).reshape(-1, 10)
    r = np.frombuffer(bytes(ref), dtype=np.float32).reshape(-1, 10)
    same = g.view(np.uint32) == r.view(np.uint32)
    d = np.abs(g.astype(np.float64) - r.astype(np.float64))
    ulp = np.spacing(np.maximum(np.abs(g), np.abs(r))).astype(np.float64)
    scale = np.abs(r).astype(np.float64).max(axis=0)
    ok = same | (d <= ulp) | (d <= 1e-12 * scale[None, :])
    bad = np.argwhere(~ok)
    assert len(bad) == 0, f"{len(bad)} floats differ, first {[(int(n), int(c), float(g[n, c]), float(r[n, c])) for n, c in bad[:5]]}"
    ndiff = int((~same).sum())
    assert ndiff <= g.size // 100, f"{ndiff} of {g.size} floats not bit-identical"
    return ndiff


def run_to_checkpoint(lat, c):
    for _ in range(c["pre_steps"]):
        lat.L.lbstep(lat.h)
    lat.L.measure(lat.h)


def test_lbsave_and_lbsaveobj_match_reference_files(tmp_path):
    g = load_golden("checkpoint.npz")
    c = g["meta"]
    lat = ckpt_lattice(c)
    run_to_checkpoint(lat, c)
    lat.L.lbstep(lat.h)
    dms, obj = str(tmp_path / "ck.dms").encode(), str(tmp_path / "ck.obj").encode()
    lat.L.lbsave(lat.h, dms)
    lat.L.lbsaveobj(lat.h, obj)
    nd = compare_dms(open(dms, "rb").read(), g["dms"].tobytes())
    print(f"lbsave: {nd} floats not bit-identical to the reference's")
    assert open(obj).read() == str(g["obj"])
    assert rel_l2(lat.pops(), g["a_after_save"]) < 1e-12
    lat.free()


def test_usr2_signal_writes_the_checkpoint(tmp_path):
    """sighook(SIGUSR2) during a step: at the end of that step's collide the
    library saves <name>.dms and <name>.obj, as lbsave + lbsaveobj would,
    and clears the signal flag (without re-entering the handler from
    lbsave's own collide)."""
    g = load_golden("checkpoint.npz")
    c = g["meta"]
    name = str(tmp_path / "sig")
    lat = ckpt_lattice(c, name=name)
    L = lat.L
    run_to_checkpoint(lat, c)
    flag = C.c_int.in_dll(L, "slb_signal_flag")
    flag.value = 0
    L.sighook(signal.SIGUSR2)
    assert flag.value == 2
    L.lbstep(lat.h)
    assert flag.value == 0
    nd = compare_dms(open(name + ".dms", "rb").read(), g["dms"].tobytes())
    print(f"USR2 checkpoint: {nd} floats not bit-identical to the reference's")
    assert open(name + ".obj").read() == str(g["obj"])
    assert rel_l2(lat.pops(), g["a_after_save"]) < 1e-12
    # a real signal delivered to the process takes the same path
    os.remove(name + ".dms")
    os.kill(os.getpid(), signal.SIGUSR2)
    for _ in range(100):                 # delivered to some thread of the process
        if flag.value == 2:
            break
        time.sleep(0.01)
    assert flag.value == 2
    L.lbstep(lat.h)
    assert flag.value == 0 and os.path.exists(name + ".dms")
    lat.free()


def test_usr1_signal_refreshes_the_monitor(tmp_path):
    g = load_golden("checkpoint.npz")
    c = g["meta"]
    name = str(tmp_path / "mon")
    lat = ckpt_lattice(c, name=name)
    L = lat.L
    lat.s.monitoring = 1
    L.lbstep(lat.h)
    assert not os.path.exists(name + ".mon")
    L.sighook(signal.SIGUSR1)
    L.lbstep(lat.h)
    head = open(name + ".mon").read().split("\n")[0].split()
    assert head[0] == str(os.getpid()) and head[1] == name and head[2] == "running"
    assert C.c_int.in_dll(L, "slb_signal_flag").value == 0
    lat.free()


def test_lbload_and_lbloadobj_read_reference_files(tmp_path):
    g = load_golden("checkpoint.npz")
    c = g["meta"]
    dms, obj = tmp_path / "ref.dms", tmp_path / "ref.obj"
    dms.write_bytes(g["dms"].tobytes())
    obj.write_text(str(g["obj"]))
    lat = ckpt_lattice(c)
    L = lat.L
    L.syncobjs(lat.h)
    L.lbload(lat.h, str(dms).encode())
    assert rel_l2(lat.pops(), g["a_loaded"]) < 1e-12
    for o in lat.objects():
        o.velx = o.vely = o.velz = o.angx = o.angy = o.angz = 0.0
    L.lbloadobj(lat.h, str(obj).encode())
    rows = objrows(lat)
    assert np.array_equal(rows[:, :9], g["loadobj_rows"][:, :9])
    lat.free()
    lat = ckpt_lattice(c)
    lat.s.rescale = float(c["rescale"])
    lat.L.lbloadobj(lat.h, str(obj).encode())
    assert np.array_equal(objrows(lat)[:, :9], g["loadobj_rescaled_rows"][:, :9])
    assert lat.s.rescale == 1.0
    lat.free()


def test_vset_vzero_vgrab_match_reference():
    import sunlightlb_amd as s
    g = load_golden("vfields.npz")
    c = g["meta"]
    lat = s.Lattice(*c["dims"], precision=64)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow = c["tau"], 0
    L.insertsph(h, *c["sphere"])
    L.zerobf(h)
    lat.sync()
    for f in ("bx", "by", "bz"):
        s.view(getattr(lat.s, f), lat.N)[:] = g[f]
    L.slb_sync_device(h)
    for _ in range(c["steps"]):
        L.lbstep(h)

    def fields():
        return np.stack([lat.field(f) for f in ("ux", "uy", "uz")])

    for i, j, k, vx, vy, vz in c["vset"]:
        L.vset(h, int(i), int(j), int(k), vx, vy, vz)
    # rho0 v / 12: the reference library is built with -ffast-math, which may
    # turn the division into a multiplication by 1/12 (one rounding apart)
    np.testing.assert_allclose(fields(), g["u_vset"], rtol=2e-16, atol=0)
    assert ((fields() != 0) == (g["u_vset"] != 0)).all()
    L.vzero(h)
    L.vgrab(h, c["prefac"][0])
    assert rel_l2(fields(), g["u_grab_collide"]) < 1e-12
    L.bbl(h)
    L.propagate(h)
    L.vgrab(h, c["prefac"][1])
    assert rel_l2(fields(), g["u_grab_propagate"]) < 1e-12
    L.collide(h)
    L.vzero(h)
    u = fields()
    solid = lat.solid()
    np.testing.assert_allclose(u[:, solid], g["u_zero"][:, solid], rtol=2e-16, atol=0)
    assert not u[:, ~solid].any()
    assert rel_l2(lat.pops(), g["a"]) < 1e-12
    lat.free()


C15 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1],
                [1, -1, -1], [-1, -1, -1]], dtype=np.float64)


@pytest.mark.parametrize("prec", [64, 32])
def test_residual_recomputed_from_snapshots(prec):
    """slb_residual = ||u(t) - u(t-100)|| / ||u(t)|| over fluid nodes with
    u = (j - b/2) / rho, j = sum c_l a_l the momentum of the post-collide
    populations (which carry the full body force b) and rho = rho0 + sum a_l
    (vgrab after collide, src/lbmeasure.c:141-166), recomputed in numpy from
    the two population snapshots and the per-node body force (the D3Q15
    velocities of src/lbmisc.c:89-104)."""
    import sunlightlb_amd as s
    dims = (24, 20, 28)
    lat = s.Lattice(*dims, precision=prec)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow = 0.75, 1
    for sp in ((5.0, 6.0, 7.0, 8.0), (4.0, 17.0, 12.0, 20.0), (3.5, 12.0, 2.0, 25.0)):
        L.insertsph(h, *sp)
    L.zerobf(h)
    lat.sync()
    rng = np.random.default_rng(3)
    solid = lat.solid()
    b = np.zeros((lat.N, 3))
    for q, (f, sc) in enumerate((("bx", 1e-5), ("by", 2e-6), ("bz", -3e-6))):
        b[:, q] = np.where(solid, 0.0, sc * (1.0 + rng.random(lat.N)))
        s.view(getattr(lat.s, f), lat.N)[:] = b[:, q]
    L.slb_sync_device(h)
    L.slb_run(h, 150)
    L.slb_residual_reset(h)
    assert L.slb_residual(h) == -1.0
    a1 = lat.pops()
    L.slb_run(h, 100)
    r = L.slb_residual(h)
    a2 = lat.pops()
    fl = ~solid

    def vel(a):
        rho = 72.0 + a.sum(1)
        j = a @ C15
        return ((j - 0.5 * b) / rho[:, None])[fl]

    u1, u2 = vel(a1), vel(a2)
    r_np = np.sqrt(((u2 - u1) ** 2).sum()) / np.sqrt((u2 ** 2).sum())
    print(f"residual fp{prec}: library {r:.15e} numpy {r_np:.15e}")
    assert r > 0 and abs(r - r_np) <= 1e-12 * r_np
    lat.free()
