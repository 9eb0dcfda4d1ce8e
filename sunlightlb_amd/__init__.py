"""Python binding of the MI355X SunlightLB library (libsunlightlb.so).

The library exports the SunlightLB C API (include/sunlightlb.h, the reference
interface src/sunlightlb.h) and the GPU extensions (include/sunlightlb_gpu.h).
This module mirrors the reference's scripting binding (src/SunlightLB.i wraps
the same header): the struct layouts as ctypes Structures, every function
with its C signature, and numpy views of the host mirrors.

There is no CPU fallback: importing works anywhere the shared library is
built, and every compute call runs on the GPU (the library exits with an
error when no HIP device is visible).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsunlightlb.so")

# verbosity bits and cycle states (src/sunlightlb.h:39-49)
TRACE_ROUTINE_EXITS = 0x02
OBJECT_DETAILS = 0x04
MP_OBJECT_DETAILS = 0x04
FORCE_DETAILS = 0x08
POST_BBL, POST_PROPAGATE, POST_COLLIDE = 42, 47, 49

Real = C.c_double
RealP = C.POINTER(Real)
IntP = C.POINTER(C.c_int)


class LBObj_type(C.Structure):
    pass


LBObj = C.POINTER(LBObj_type)
LBObj_type._fields_ = [("id", C.c_int)] + [
    (n, Real) for n in ("posx", "posy", "posz", "velx", "vely", "velz", "angx", "angy", "angz",
                        "frcx", "frcy", "frcz", "trqx", "trqy", "trqz")
] + [("next_obj", LBObj), ("base_obj", LBObj)]


class LBSys_type(C.Structure):
    _fields_ = [
        ("w", IntP), ("c", IntP), ("a", RealP),
        ("ux", RealP), ("uy", RealP), ("uz", RealP),
        ("fx", RealP), ("fy", RealP), ("fz", RealP),
        ("bx", RealP), ("by", RealP), ("bz", RealP),
        ("esi", C.c_int), ("esj", C.c_int), ("esk", C.c_int), ("nnodes", C.c_int),
        ("step", C.c_int), ("status", C.c_int), ("changed_objects", C.c_int),
        ("objects_moving", C.c_int), ("skip_solid", C.c_int), ("stokes_flow", C.c_int),
        ("vel_corrected", C.c_int), ("monitoring", C.c_int),
        ("tau", Real), ("rho0", Real),
        ("vmeanx", Real), ("vmeany", Real), ("vmeanz", Real),
        ("psi", RealP), ("rho", RealP), ("name", C.c_char_p),
        ("rescale", Real), ("load_average", C.c_int),
        ("obj", C.POINTER(LBObj)), ("obj_head", LBObj), ("mask", IntP), ("next_id", C.c_int),
        ("dpdx", Real), ("dpdy", Real), ("dpdz", Real),
    ]


LBSys = C.POINTER(LBSys_type)


class MPSys_type(C.Structure):
    _fields_ = [
        ("lb", LBSys), ("p", RealP), ("newp", RealP),
        ("step", C.c_int), ("objects_moving", C.c_int), ("to_be_initialised", C.c_int),
        ("alpha", Real), ("isoslice", Real), ("poffset", Real),
        ("p2bytemin", Real), ("p2bytemax", Real), ("name", C.c_char_p),
    ]


MPSys = C.POINTER(MPSys_type)
S = C.c_char_p
D = C.c_double
I = C.c_int
VP = C.c_void_p
LongP = C.POINTER(C.c_long)

# name: (restype, argtypes) for every entry point of include/sunlightlb.h and
# include/sunlightlb_gpu.h
SIGNATURES = {
    "newlb": (LBSys, [I, I, I, S]),
    "set_verbosity": (None, [I]), "unset_verbosity": (None, [I]), "report_verbosity": (None, []),
    "lbstep": (None, [LBSys]), "bbl": (None, [LBSys]), "bblforce": (None, [LBSys]),
    "collide": (None, [LBSys]), "propagate": (None, [LBSys]),
    "lookupobj": (LBObj, [LBSys, I]), "addobj": (LBObj, [LBSys]), "zeroobj": (None, [LBObj]),
    "insertsubs": (LBObj, [LBSys, S, I, I, I, I, I, I]),
    "insertlist": (LBObj, [LBSys, S]),
    "insertslab": (LBObj, [LBSys, D, D, D, D, D, D]),
    "insertsph": (LBObj, [LBSys, D, D, D, D]),
    "insertcyl": (LBObj, [LBSys, S, D, D, D]),
    "insertwall": (LBObj, [LBSys, S, I]),
    "flattenobj": (LBObj, [LBSys]), "invertobj": (LBObj, [LBSys]),
    "removeobjs": (None, [LBSys]), "wrobjs": (None, [LBSys]), "wrmap": (None, [LBSys, S]),
    "makemask": (None, [LBSys]), "vspread": (None, [LBSys]), "zerobf": (None, [LBSys]),
    "addbf": (None, [LBSys, D, D, D]), "vset": (None, [LBSys, I, I, I, D, D, D]),
    "countmask": (I, [LBSys, S]), "countfluid": (I, [LBSys]), "syncobjs": (None, [LBSys]),
    "lbsave": (None, [LBSys, S]), "lbload": (None, [LBSys, S]),
    "lbsavevel": (None, [LBSys, S, S]), "lbloadvel": (None, [LBSys, S, S]),
    "lbsavestr": (None, [LBSys, S, S]), "lbsaveobj": (None, [LBSys, S]),
    "lbloadobj": (None, [LBSys, S]), "lbsavefrc": (None, [LBSys, S, S]),
    "lbsavemap": (None, [LBSys, S, S]),
    "error": (None, [S]), "about": (None, []), "wrpars": (None, [LBSys]),
    "wrstatus": (None, [LBSys, C.c_void_p]), "wrlinks": (None, [LBSys]),
    "wrstats": (None, [LBSys]), "fwrstats": (None, [C.c_void_p, I, I]),
    "startclock": (None, []), "mycat": (C.c_void_p, [S, S]), "mycopy": (C.c_void_p, [S]),
    "myscmp": (I, [S, S]),
    "accvmean": (None, [LBSys, D]), "vzero": (None, [LBSys]), "vgrab": (None, [LBSys, D]),
    "vsave": (None, [LBSys, S]), "measure": (None, [LBSys]), "zeroforce": (None, [LBSys]),
    "normforce": (None, [LBSys, D]), "distribf": (None, [LBSys]), "wrmeas": (None, [LBSys]),
    "discrf": (D, [LBSys]), "wrmon": (None, [LBSys, S]),
    "mpstep": (None, [MPSys]), "newmp": (MPSys, [LBSys, S]),
    "newmpiso": (C.c_void_p, [MPSys, D]), "mpiso": (None, [C.c_void_p]),
    "mpisowr": (None, [C.c_void_p, S, S]),
    "mpinit": (None, [MPSys]), "mpwrpars": (None, [MPSys]), "mprhoinit": (None, [MPSys]),
    "mpcheck": (None, [MPSys]), "mpmean": (D, [MPSys]), "mpmax": (D, [MPSys]),
    "mpmin": (D, [MPSys]), "mprescale": (None, [MPSys, D]), "mpcopy": (None, [MPSys, MPSys]),
    "mpsave": (None, [MPSys, S, S]), "mpload": (None, [MPSys, S, S]),
    "mpconwr": (None, [MPSys, S, S]), "mpspread": (D, [MPSys, S]),
    "mpall": (None, [MPSys, D]), "mpslab": (None, [MPSys, D, D, D, D, D, D, D]),
    "mpwall": (None, [MPSys, S, I, D]), "mpsph": (None, [MPSys, D, D, D, D, D]),
    "mpcyl": (None, [MPSys, S, D, D, D, D]),
    "mpsubs": (None, [MPSys, S, I, I, I, I, I, I, D]), "mplist": (None, [MPSys, S, D]),
    "mpvacfinit": (None, [MPSys, MPSys, S, I]), "mpvacfinit2": (None, [MPSys, MPSys]),
    "mpvacf": (None, [MPSys, MPSys]), "mpvacfwr": (D, [MPSys, MPSys, S, I]),
    "adall": (None, [MPSys]), "adnone": (None, [MPSys]),
    "adslab": (None, [MPSys, D, D, D, D, D, D]), "adsph": (None, [MPSys, D, D, D, D]),
    "adcyl": (None, [MPSys, S, D, D, D]), "adtotal": (D, [MPSys, S, I]),
    "adzero": (None, [MPSys]),
    "newlbcolour": (C.c_void_p, [I, I, I]), "lbcolourset": (None, [C.c_void_p, I, I, I]),
    "lbcolourcopy": (None, [C.c_void_p, C.c_void_p]),
    "newlbcolourmap": (C.c_void_p, [D, D, C.c_void_p, C.c_void_p]),
    "lbcolourmapset": (None, [C.c_void_p, D, D, C.c_void_p, C.c_void_p]),
    "newlbimage": (C.c_void_p, [I, I, C.c_void_p, C.c_void_p]),
    "lbimageclear": (None, [C.c_void_p]),
    "lbimagemap": (None, [C.c_void_p, LBSys, S, I]),
    "lbimagevel": (None, [C.c_void_p, C.c_void_p, LBSys, S, S, I]),
    "lbcolourget": (None, [C.c_void_p, C.c_void_p, I, I]),
    "lbgetred": (I, [C.c_void_p, I, I]), "lbgetgreen": (I, [C.c_void_p, I, I]),
    "lbgetblue": (I, [C.c_void_p, I, I]),
    "lbcolourclamp": (None, [IntP, IntP, IntP]),
    "mpimageset": (None, [C.c_void_p, MPSys, C.c_void_p, S, S, I]),
    "sighandle": (None, [LBSys]), "sighook": (None, [I]), "siginit": (None, []),
    # sunlightlb_gpu.h
    "slb_set_precision": (I, [LBSys, I]), "slb_get_precision": (I, [LBSys]),
    "slb_sync_host": (None, [LBSys]), "slb_sync_device": (None, [LBSys]),
    "slb_mp_sync_host": (None, [MPSys]), "slb_mp_sync_device": (None, [MPSys]),
    "slb_destroy": (None, [LBSys]),
    "slb_run": (None, [LBSys, I]), "slb_residual": (D, [LBSys]),
    "slb_residual_reset": (None, [LBSys]),
    "slb_run_to_residual": (I, [LBSys, D, I, I, RealP]),
    "slb_run_to_residual_aa": (I, [LBSys, D, I, I, I, RealP]),
    "slb_time_steps": (D, [LBSys, I]), "slb_scale_pops": (None, [LBSys, D]),
    "slb_sum_pops": (None, [LBSys, RealP]),
    "slb_insert_spheres": (I, [LBSys, I, RealP]),
    "slb_rsa_pack": (I, [LBSys, D, D, C.c_ulonglong, C.c_long, RealP, I, RealP]),
    "slb_rsa_pack_flat": (I, [LBSys, D, D, C.c_ulonglong, C.c_long, RealP, I, RealP]),
    "slb_pack_map": (I, [I, I, I, D, D, C.c_ulonglong, I, C.c_long, VP, RealP]),
    "slb_insert_map": (LBObj, [LBSys, VP]),
    "slb_mp_set_adsorption": (None, [MPSys, D]), "slb_mp_track_moments": (None, [MPSys, I]),
    "slb_mp_sums": (None, [MPSys, RealP]), "slb_mp_run": (None, [MPSys, I]),
    "slb_step_kernel": (C.c_char_p, [LBSys]),
    "slb_mp_time_steps": (D, [MPSys, I]),
    "slb_device_bytes": (C.c_size_t, [LBSys]), "slb_device_arch": (C.c_char_p, []),
    "slb_hbm_copy_test": (D, [C.c_size_t, I]),
    "slb_batch_new": (C.c_void_p, [I]),
    "slb_batch_add": (I, [C.c_void_p, LBSys, I, RealP]),
    "slb_batch_commit": (I, [C.c_void_p]),
    "slb_batch_autotune": (I, [C.c_void_p, S]),
    "slb_batch_set_config": (None, [C.c_void_p, I, I, I]),
    "slb_batch_get_config": (None, [C.c_void_p, IntP]),
    "slb_batch_run": (None, [C.c_void_p, I]),
    "slb_batch_time_steps": (D, [C.c_void_p, I]),
    "slb_batch_measure": (None, [C.c_void_p]),
    "slb_batch_forces": (I, [C.c_void_p, I, I, RealP, RealP]),
    "slb_batch_get_pops": (I, [C.c_void_p, I, I, RealP]),
    "slb_batch_nodes": (C.c_long, [C.c_void_p]),
    "slb_batch_free": (None, [C.c_void_p]),
    # learned warm start
    "slb_warm_new": (C.c_void_p, [C.c_ulonglong]),
    "slb_warm_free": (None, [C.c_void_p]),
    "slb_warm_save": (I, [C.c_void_p, S]), "slb_warm_load": (C.c_void_p, [S]),
    "slb_warm_add_sample": (I, [C.c_void_p, LBSys]),
    "slb_warm_nsamples": (I, [C.c_void_p]),
    "slb_warm_sample_stats": (None, [C.c_void_p, I, RealP]),
    "slb_warm_set_mode": (None, [C.c_void_p, I]), "slb_warm_get_mode": (I, [C.c_void_p]),
    "slb_warm_features": (D, [LBSys, VP, I]),
    "slb_warm_features_noalloc": (D, [LBSys, VP, I]),
    "slb_warm_set_projection": (None, [VP, I, I]),
    "slb_warm_get_projection": (None, [VP, IntP, IntP]),
    "slb_warm_project": (D, [LBSys, VP, I, I, I]), "slb_warm_targets": (None, [LBSys, VP, I, D]),
    "slb_warm_moments": (None, [LBSys, VP, I, D]), "slb_warm_inject": (None, [LBSys, VP, I, I, D]),
    "slb_warm_train": (D, [C.c_void_p, I, D, C.c_ulonglong, RealP]),
    "slb_warm_eval": (D, [C.c_void_p, LBSys, RealP]),
    "slb_warm_predict": (None, [C.c_void_p, LBSys, VP]),
    "slb_warm_start": (None, [C.c_void_p, LBSys]),
    "slb_warm_set_fallback": (None, [C.c_void_p, I, D]),
    "slb_warm_get_fallback": (None, [C.c_void_p, IntP, RealP]),
    "slb_warm_note_cold": (None, [C.c_void_p, RealP, I, I]),
    "slb_warm_note_warm": (None, [C.c_void_p, RealP, I, I]),
    "slb_warm_policy": (I, [C.c_void_p, D, I]),
    "slb_warm_solve": (I, [C.c_void_p, LBSys, D, I, I, RealP, IntP]),
    "slb_warm_nparams": (C.c_long, [C.c_void_p]),
    "slb_warm_layout": (None, [IntP, IntP, IntP, LongP, LongP]),
    "slb_warm_get_params": (None, [C.c_void_p, VP]),
    "slb_warm_set_params": (None, [C.c_void_p, VP]),
    "slb_warm_fwd_bwd": (D, [C.c_void_p, IntP, VP, VP, VP, VP, VP]),
    # network kernels (device pointers)
    "slbn_kpad": (I, [I]), "slbn_set_stream": (None, [VP]),
    "slbn_conv": (None, [VP, I, VP, VP, VP, VP, I, I, I, I, I, I]),
    "slbn_conv_wgrad": (None, [VP, I, VP, I, VP, I, I, I]),
    "slbn_conv_wgrad_small": (None, [VP, I, VP, I, VP, I, I, I]),
    "slbn_wgrad_blocked_ok": (I, [I, I, I, I, I]),
    "slbn_wgrad_lds16_ok": (I, [I, I, I, I, I]),
    "slbn_wtrans": (None, [VP, I, I, VP]),
    "slbn_relu_bwd": (None, [VP, VP, VP, C.c_long]),
    "slbn_bias_grad": (None, [VP, I, I, VP]),
    "slbn_relu_bwd_bias": (None, [VP, VP, VP, I, I, VP]),
    "slbn_pool": (None, [VP, VP, I, I, I, I]), "slbn_pool_bwd": (None, [VP, VP, I, I, I, I]),
    "slbn_up_add": (None, [VP, VP, VP, I, I, I, I]), "slbn_up_bwd": (None, [VP, VP, I, I, I, I]),
    "slbn_add_into": (None, [VP, VP, C.c_long]),
    "slbn_mse": (None, [VP, I, VP, I, I, I, VP, VP, VP]),
    "slbn_adam": (None, [VP, VP, VP, VP, VP, C.c_long, C.c_float, C.c_float, C.c_float,
                         C.c_float, I]),
    "slbn_f2bf": (None, [VP, VP, C.c_long]), "slbn_bf2f": (None, [VP, VP, C.c_long]),
    "slbn_expand": (None, [VP, I, VP, I, C.c_long]),
    "slbn_malloc": (VP, [C.c_size_t]), "slbn_free": (None, [VP]),
    "slbn_zero": (None, [VP, C.c_size_t]),
    "slbn_h2d": (None, [VP, VP, C.c_size_t]), "slbn_d2h": (None, [VP, VP, C.c_size_t]),
    "slbn_sync": (None, []),
}

GLOBALS = ("slb_verbosity", "slb_signal_flag")

_lib = None


def lib():
    """Load libsunlightlb.so (in-tree build) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(
            f"{LIBPATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
    L = C.CDLL(LIBPATH, mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def b(s):
    return s.encode() if isinstance(s, str) else s


def view(ptr, n, dtype=np.float64):
    """numpy view (no copy) of n elements of a host mirror."""
    if not ptr:
        return None
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


class Lattice:
    """Thin convenience wrapper around an LBSys handle."""

    def __init__(self, esi, esj, esk, name="null", precision=None):
        L = lib()
        self.L = L
        self.h = L.newlb(esi, esj, esk, b(name))
        self.s = self.h.contents
        if precision is not None:
            if L.slb_set_precision(self.h, int(precision)) != 0:
                raise ValueError("precision must be 32 or 64")

    @classmethod
    def wrap(cls, h):
        """A wrapper around an existing LBSys handle (not owned: free()
        destroys it)."""
        self = cls.__new__(cls)
        self.L, self.h, self.s = lib(), h, h.contents
        return self

    @property
    def shape(self):
        return (self.s.esi, self.s.esj, self.s.esk)

    @property
    def N(self):
        return self.s.nnodes

    def sync(self):
        self.L.slb_sync_host(self.h)

    def pops(self):
        """Host copy (N, 15) of the shifted populations (AoS like lb->a)."""
        self.sync()
        return view(self.s.a, 15 * self.N).reshape(self.N, 15).copy()

    def set_pops(self, a):
        self.sync()
        view(self.s.a, 15 * self.N)[:] = np.asarray(a, dtype=np.float64).reshape(-1)
        self.L.slb_sync_device(self.h)

    def mask(self):
        self.sync()
        return view(self.s.mask, self.N, np.int32).copy()

    def solid(self):
        o = np.ctypeslib.as_array(C.cast(self.s.obj, C.POINTER(C.c_uint64)), shape=(self.N,))
        return o != 0

    def field(self, name):
        self.sync()
        return view(getattr(self.s, name), self.N).copy()

    def objects(self):
        out = []
        p = self.s.obj_head
        while p:
            out.append(p.contents)
            p = p.contents.next_obj
        return out

    def free(self):
        if self.h:
            self.L.slb_destroy(self.h)
            self.h = None

    def step(self, n=1):
        if n == 1:
            self.L.lbstep(self.h)
        else:
            self.L.slb_run(self.h, n)

    def sum_pops(self):
        out = (C.c_double * 15)()
        self.L.slb_sum_pops(self.h, out)
        return np.array(out[:])
