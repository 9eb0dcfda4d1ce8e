"""Object tables at a size where the host passes run on several threads
(552,960 nodes), checked against numpy restatements of the reference's
rules: insertsph's FP64 predicate (i-ox)^2 + (j-oy)^2 + (k-oz)^2 < r^2 with
the last insert winning (src/lbobjs.c:204-232), insertslab's open box
(:170-199), insertcyl's (i-a)^2 + (k-b)^2 < r^2 along y (:241-287),
makemask's link bits (:526-560), countfluid / countmask
(:648-714), flattenobj / invertobj centroids (:362-428) and zerobf / addbf
on fluid nodes (:593-619)."""
import numpy as np
import pytest

from test_gpu_parity import dense_objid, lattice, objrows

pytestmark = pytest.mark.gpu

DIMS = (80, 72, 96)
C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1],
              [1, -1, -1], [-1, -1, -1]])
W = np.array([16, 8, 8, 8, 8, 8, 8, 1, 1, 1, 1, 1, 1, 1, 1])


def raster(spheres, slab, cyl, dims):
    """Dense object index (list order: the newest object is 0), -1 fluid."""
    i, j, k = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in dims], indexing="ij")
    owner = np.full(dims, -1, dtype=np.int64)
    serial = 0
    for (r, ox, oy, oz) in spheres:
        inside = (i - ox) * (i - ox) + (j - oy) * (j - oy) + (k - oz) * (k - oz) < r * r
        if inside.any():
            owner[inside] = serial
            serial += 1
    x1, x2, y1, y2, z1, z2 = slab
    inside = (i > x1) & (i < x2) & (j > y1) & (j < y2) & (k > z1) & (k < z2)
    owner[inside] = serial
    serial += 1
    r, a, b = cyl                   # along y: (i - a)^2 + (k - b)^2 < r^2
    inside = (i - a) * (i - a) + (k - b) * (k - b) < r * r
    owner[inside] = serial
    serial += 1
    return np.where(owner < 0, -1, serial - 1 - owner).reshape(-1).astype(np.int32), serial


def makemask(solid, dims):
    s = solid.reshape(dims)
    m = np.zeros(dims, dtype=np.int64)
    for l in range(1, 15):
        nb = np.roll(s, shift=tuple(-C[l]), axis=(0, 1, 2))     # neighbour at +c_l
        m |= np.where(s & ~nb, 1 << (l - 1), 0)
    return m.reshape(-1)


def centroid(sel, dims):
    idx = np.nonzero(sel.reshape(dims))
    ns = len(idx[0])
    return [float(np.sum(a.astype(np.int64))) / ns for a in idx], ns


def test_tables_on_threads_match_restatement():
    rng = np.random.default_rng(11)
    spheres = [(float(rng.uniform(2.5, 9.0)), *(float(rng.uniform(-4, n + 4)) for n in DIMS))
               for _ in range(40)]
    slab = (10.5, 30.2, -1.0, 20.5, 40.3, 70.0)
    cyl = (7.3, 52.6, 17.9)
    lat = lattice(DIMS)
    L, h = lat.L, lat.h
    assert lat.N > 1 << 18
    for sp in spheres:
        L.insertsph(h, *sp)
    L.insertslab(h, *slab)
    L.insertcyl(h, b"y", *cyl)
    L.syncobjs(h)
    want, nobj = raster(spheres, slab, cyl, DIMS)
    got = dense_objid(lat)
    assert (got == want).all()
    solid = want >= 0
    mask = makemask(solid, DIMS)
    assert (lat.mask() == mask).all()
    assert L.countfluid(h) == int((~solid).sum())
    links = sum(int(W[l]) * int(((mask >> (l - 1)) & 1).sum()) for l in range(1, 15))
    faces = sum(int(((mask >> (l - 1)) & 1).sum()) for l in range(1, 7))
    assert L.countmask(h, b"links") == links
    assert L.countmask(h, b"faces") == faces

    L.zerobf(h)
    L.addbf(h, 1e-6, -2e-6, 0.0)
    L.addbf(h, 3e-7, 0.0, 5e-7)
    for name, v in (("bx", 1e-6 + 3e-7), ("by", -2e-6), ("bz", 5e-7)):
        f = lat.field(name)
        assert (f[~solid] == v).all() and (f[solid] == 0.0).all()

    L.flattenobj(h)
    L.syncobjs(h)
    (cx, cy, cz), ns = centroid(solid, DIMS)
    rows = objrows(lat)
    assert rows.shape[0] == 1
    assert (rows[0, 0], rows[0, 1], rows[0, 2]) == (cx, cy, cz)
    assert (dense_objid(lat) == np.where(solid, 0, -1)).all()
    assert L.countfluid(h) == lat.N - ns

    L.invertobj(h)
    L.syncobjs(h)
    (cx, cy, cz), nf = centroid(~solid, DIMS)
    rows = objrows(lat)
    assert (rows[0, 0], rows[0, 1], rows[0, 2]) == (cx, cy, cz)
    assert (dense_objid(lat) == np.where(solid, -1, 0)).all()
    assert (lat.mask() == makemask(~solid, DIMS)).all()
    assert L.countfluid(h) == lat.N - nf
    lat.free()
