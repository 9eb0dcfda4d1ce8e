"""slb_warm_features_noalloc (jump-flooding features with the seed buffers
in the library's static device scratch) against slb_warm_features (the
same kernels on buffers it allocates): bit-identical features and the same
normalising length, on two calls in a row -- the second on a smaller
lattice, so stale scratch contents from the first call would show -- and
a lattice beyond the scratch refused with -1 and the output untouched."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def lib():
    import sunlightlb_amd as s
    return s.lib()


def pack(dims, spheres):
    import sunlightlb_amd as s
    lat = s.Lattice(*dims, precision=32)
    lat.s.stokes_flow = 1
    for r, x, y, z in spheres:
        lat.L.insertsph(lat.h, r, x, y, z)
    lat.L.zerobf(lat.h)
    lat.L.addbf(lat.h, 1e-5, 0.0, 0.0)
    return lat


def features(L, lat, fn, C=4):
    """Run one of the feature entry points into a fresh device tensor
    (prefilled with a marker so an untouched output shows)."""
    n = lat.N * C
    p = L.slbn_malloc(2 * n)
    marker = np.full(n, 0x7fc1, np.uint16)
    L.slbn_h2d(p, marker.ctypes.data, 2 * n)
    ell = fn(lat.h, p, C)
    out = np.empty(n, np.uint16)
    L.slbn_d2h(out.ctypes.data, p, 2 * n)
    L.slbn_free(p)
    return ell, out


CASES = [
    ((48, 40, 56), [(6.5, 10.0, 12.0, 14.0), (5.0, 30.0, 8.0, 40.0), (7.2, 40.0, 30.0, 20.0),
                    (4.0, 0.5, 39.0, 55.0)]),
    ((16, 24, 20), [(3.5, 4.0, 5.0, 6.0), (2.5, 11.0, 3.0, 17.0), (4.0, 12.5, 12.0, 8.0)]),
]


def test_noalloc_features_equal_allocating_entry_point():
    L = lib()
    for dims, spheres in CASES:          # the second call is on a smaller lattice
        print(f"no-alloc features: dims {dims}", flush=True)
        lat = pack(dims, spheres)
        ell_a, x_a = features(L, lat, L.slb_warm_features)
        ell_b, x_b = features(L, lat, L.slb_warm_features_noalloc)
        lat.free()
        assert ell_a > 0.0 and ell_b == ell_a, (dims, ell_a, ell_b)
        assert np.array_equal(x_a, x_b), dims
        assert not np.any(x_b.reshape(-1, 4)[:, 2:] == 0x7fc1)   # every channel written


def test_noalloc_features_refuse_a_lattice_beyond_the_scratch():
    """A lattice of more than 2^21 nodes is refused before any device work:
    -1 returned, a message on stderr, the output tensor untouched."""
    code = r"""
import numpy as np, sys
sys.path.insert(0, %r)
import sunlightlb_amd as s
L = s.lib()
lat = s.Lattice(130, 128, 128, precision=32)
lat.L.insertsph(lat.h, 10.0, 60.0, 60.0, 60.0)
n = lat.N * 2
p = L.slbn_malloc(2 * n)
m = np.full(n, 0x7fc1, np.uint16)
L.slbn_h2d(p, m.ctypes.data, 2 * n)
ell = L.slb_warm_features_noalloc(lat.h, p, 2)
out = np.empty(n, np.uint16)
L.slbn_d2h(out.ctypes.data, p, 2 * n)
print("RESULT", ell, bool((out == 0x7fc1).all()))
""" % ROOT
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    line = [l for l in p.stdout.splitlines() if l.startswith("RESULT")][0]
    assert line.split()[1:] == ["-1.0", "True"], line
    assert "exceeds the static scratch" in p.stderr
