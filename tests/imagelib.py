"""Helpers shared by the image tests: pixel access of an LBImage handle."""
import ctypes as C

import numpy as np


class _Img(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int),
                ("red", C.POINTER(C.c_int)), ("green", C.POINTER(C.c_int)),
                ("blue", C.POINTER(C.c_int)),
                ("back", C.c_int * 3), ("obj", C.c_int * 3)]


def pixels(im):
    """(3, width * height) copy of an image's channels, pixel (a, b) at
    height * a + b."""
    s = C.cast(im, C.POINTER(_Img)).contents
    n = s.width * s.height
    return np.stack([np.ctypeslib.as_array(ch, shape=(n,)).copy()
                     for ch in (s.red, s.green, s.blue)])


def size(im):
    s = C.cast(im, C.POINTER(_Img)).contents
    return s.width, s.height
