"""Host-only image entry points (src/mpimage.c): colours are clamped to
0..255, a colour map keeps its end colours, a new image clears to its
background, and the pixel getters refuse coordinates outside the image
(including i == width and j == height, which the reference reads one past
the end).  No lattice and no GPU call."""
import ctypes as C

import numpy as np

from imagelib import pixels, size


def lib():
    import sunlightlb_amd as s
    return s.lib()


def test_colours_maps_and_image_clear():
    L = lib()
    c = L.newlbcolour(300, -5, 128)
    rgb = C.cast(c, C.POINTER(C.c_int * 3)).contents
    assert tuple(rgb) == (255, 0, 128)
    L.lbcolourset(c, 12, 256, -1)
    assert tuple(rgb) == (12, 255, 0)
    d = L.newlbcolour(0, 0, 0)
    L.lbcolourcopy(d, c)
    assert tuple(C.cast(d, C.POINTER(C.c_int * 3)).contents) == (12, 255, 0)
    r, g, b = C.c_int(-7), C.c_int(999), C.c_int(42)
    L.lbcolourclamp(C.byref(r), C.byref(g), C.byref(b))
    assert (r.value, g.value, b.value) == (0, 255, 42)
    back, obj = L.newlbcolour(1, 2, 3), L.newlbcolour(200, 100, 50)
    m = L.newlbcolourmap(-1.0, 2.0, back, obj)
    lo = C.cast(m, C.POINTER(C.c_int * 6)).contents
    assert tuple(lo) == (1, 2, 3, 200, 100, 50)
    im = L.newlbimage(7, 5, back, obj)
    assert size(im) == (7, 5)
    L.lbimageclear(im)
    px = pixels(im)
    assert (px[0] == 1).all() and (px[1] == 2).all() and (px[2] == 3).all()
    assert (L.lbgetred(im, 6, 4), L.lbgetgreen(im, 6, 4), L.lbgetblue(im, 6, 4)) == (1, 2, 3)
    for i, j in ((7, 0), (0, 5), (-1, 0), (0, -1)):
        assert L.lbgetred(im, i, j) == 0
        assert L.lbgetgreen(im, i, j) == 0
    got = L.newlbcolour(9, 9, 9)
    L.lbcolourget(im, got, 7, 5)   # out of range: unchanged
    assert tuple(C.cast(got, C.POINTER(C.c_int * 3)).contents) == (9, 9, 9)
    L.lbcolourget(im, got, 3, 2)
    assert tuple(C.cast(got, C.POINTER(C.c_int * 3)).contents) == (1, 2, 3)
