"""GPU parity of the image entry points (src/mpimage.c) through the C ABI.

Against the reference library (tests/golden/images.npz): a slice normal to
z of a lattice the size of the image, a sphere moving through fluid at rest
(no body force) and a colour map with green_lo == blue_lo -- the cases where
the reference's pixel placement and colour code are right -- the solid map,
the three velocity components and the speed, and an MP field added to and
subtracted from the map.  The lattice state itself matches the reference to
1e-12 first; the pixel colours then match exactly (a colour truncates a value
that agrees to ~1e-13; at most one pixel per image is allowed to sit on an
integer boundary and differ by one).

Slices normal to x and y, and images larger than the slice (centred, the
rest left at the background), are checked against the same colour rule
applied in numpy to this library's own populations."""
import numpy as np
import pytest

from conftest import load_golden, rel_l2
from imagelib import pixels

pytestmark = pytest.mark.gpu


def close_images(got, ref, what):
    d = np.abs(got.astype(int) - ref.astype(int))
    assert d.max() <= 1 and (d.max(axis=0) > 0).sum() <= 1, (what, np.argwhere(d > 0)[:5])


def setup(g):
    import sunlightlb_amd as s
    m = g["meta"]
    lat = s.Lattice(*m["dims"], precision=64)
    L, h = lat.L, lat.h
    lat.s.tau, lat.s.stokes_flow, lat.s.objects_moving = m["tau"], 1, 1
    o = L.insertsph(h, *m["sphere"])
    o.contents.velx, o.contents.vely, o.contents.velz = m["vel"]
    L.vspread(h)
    lat.step(m["steps"])
    assert rel_l2(lat.pops(), g["a"]) < 1e-12
    return lat


def test_images_match_reference_on_z_slices():
    g = load_golden("images.npz")
    m = g["meta"]
    lat = setup(g)
    L, h = lat.L, lat.h
    esi, esj, _ = m["dims"]
    im = L.newlbimage(esi, esj, L.newlbcolour(*m["back"]), L.newlbcolour(*m["obj"]))
    L.lbimagemap(im, h, b"z", m["slice"])
    assert np.array_equal(pixels(im), g["map"])
    vmap = L.newlbcolourmap(m["vmap"][0], m["vmap"][1], L.newlbcolour(*m["vmap"][2]),
                            L.newlbcolour(*m["vmap"][3]))
    for w in ("x", "y", "z", "m"):
        L.lbimagevel(im, vmap, h, w.encode(), b"z", m["slice"])
        close_images(pixels(im), g[f"vel_{w}"], w)
    mp = L.newmp(h, b"null")
    L.mprhoinit(mp)
    L.mpinit(mp)
    L.mpall(mp, 1.0)
    for _ in range(m["mp_steps"]):
        L.mpstep(mp)
    lat.L.slb_mp_sync_host(mp)
    import sunlightlb_amd as s
    assert rel_l2(s.view(mp.contents.p, lat.N), g["p"]) < 1e-12
    pmap = L.newlbcolourmap(m["pmap"][0], m["pmap"][1], L.newlbcolour(*m["pmap"][2]),
                            L.newlbcolour(*m["pmap"][3]))
    L.lbimagemap(im, h, b"z", m["slice"])
    L.mpimageset(im, mp, pmap, b"a", b"z", m["slice"])
    close_images(pixels(im), g["mp_add"], "mp add")
    L.mpimageset(im, mp, pmap, b"s", b"z", m["slice"])
    close_images(pixels(im), g["mp_sub"], "mp subtract")
    lat.free()


def expected_vel(a, solid, dims, axis, sl, W, H, what, vmap, back, obj, fac):
    """The colour rule on this library's populations: u = j / rho (no body
    force), slice (j,k) / (i,k) / (i,j) centred in a W x H image."""
    c = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                  [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1],
                  [1, -1, -1], [-1, -1, -1]])
    A = a.reshape(*dims, 15)
    S = solid.reshape(dims)
    A, S = np.take(A, sl, axis=axis), np.take(S, sl, axis=axis)
    rho = 72.0 + np.zeros(A.shape[:2])
    for l in range(15):
        rho = rho + A[..., l]
    u = np.einsum("abl,lq->abq", A, c) / rho[..., None]
    v = {"x": u[..., 0], "y": u[..., 1], "z": u[..., 2]}.get(what, np.sqrt((u ** 2).sum(-1)))
    lo, hi = np.array(vmap[2]), np.array(vmap[3])
    t = (1.0 / (vmap[1] - vmap[0])) * (v - vmap[0])
    col = np.clip(np.trunc(lo + t[..., None] * (hi - lo)).astype(int), 0, 255)
    col[S] = obj
    img = np.zeros((W, H, 3), int) + np.array(back)
    na, nb = A.shape[:2]
    ho, vo = (W - na) // 2, (H - nb) // 2
    img[ho:ho + na, vo:vo + nb] = col
    return img.reshape(W * H, 3).T


@pytest.mark.parametrize("axis,name", [(0, b"x"), (1, b"y")])
def test_x_and_y_slices_centred_in_a_larger_image(axis, name):
    g = load_golden("images.npz")
    m = g["meta"]
    lat = setup(g)
    L, h = lat.L, lat.h
    W, H = 15, 14
    im = L.newlbimage(W, H, L.newlbcolour(*m["back"]), L.newlbcolour(*m["obj"]))
    vmap = L.newlbcolourmap(m["vmap"][0], m["vmap"][1], L.newlbcolour(*m["vmap"][2]),
                            L.newlbcolour(*m["vmap"][3]))
    a, solid = lat.pops(), lat.solid()
    sl = 5
    for w in ("x", "z", "m"):
        L.lbimageclear(im)
        L.lbimagevel(im, vmap, h, w.encode(), name, sl)
        exp = expected_vel(a, solid, m["dims"], axis, sl, W, H, w, m["vmap"], m["back"],
                           m["obj"], 0.0)
        close_images(pixels(im), exp, (name, w))
    L.lbimageclear(im)
    L.lbimagemap(im, h, name, sl)
    px = pixels(im)
    S = np.take(solid.reshape(m["dims"]), sl, axis=axis)
    na, nb = S.shape
    ho, vo = (W - na) // 2, (H - nb) // 2
    red = px[0].reshape(W, H)
    assert (red[ho:ho + na, vo:vo + nb] == np.where(S, m["obj"][0], m["back"][0])).all()
    inside = np.zeros((W, H), bool)
    inside[ho:ho + na, vo:vo + nb] = True
    assert (red[~inside] == m["back"][0]).all()
    lat.free()
