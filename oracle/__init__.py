"""CPU oracles for the SunlightLB lattice-Boltzmann step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the product path.

Two oracles:
  Oracle  -- liboracle.so, this project's clean-room C restatement of the
             reference algorithm on plain arrays (slb_oracle.c).
  RefLib  -- oracle/_ref/libslbref.so (FP64) and libslbref32.so (-D_SINGLE),
             the reference's own C sources compiled where they lie by
             oracle/Makefile.  Present only where the reference sources were
             available at build time; the committed golden fixtures in
             tests/golden/ carry its outputs everywhere else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = {64: os.path.join(HERE, "_ref", "libslbref.so"),
          32: os.path.join(HERE, "_ref", "libslbref32.so")}

dP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
iP = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _optd(x):
    return None if x is None else np.ascontiguousarray(x, dtype=np.float64)


class _OptD:
    """argtype accepting None or a float64 array."""

    @classmethod
    def from_param(cls, v):
        if v is None:
            return None
        return v.ctypes.data_as(C.POINTER(C.c_double))


C3 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
               [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1],
               [1, -1, -1], [-1, -1, -1]], dtype=np.int64)
W = np.array([16, 8, 8, 8, 8, 8, 8, 1, 1, 1, 1, 1, 1, 1, 1], dtype=np.float64)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7])


class Oracle:
    """ctypes front end of liboracle.so (FP64, AoS a[15n+l])."""

    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        O = _OptD
        sig = {
            "orc_insertsph": (C.c_int, [C.c_int] * 3 + [iP, C.c_int] + [C.c_double] * 4
                              + [C.POINTER(C.c_int)]),
            "orc_makemask": (None, [C.c_int] * 3 + [iP, iP]),
            "orc_countfluid": (C.c_int, [C.c_long, iP]),
            "orc_countmask": (C.c_int, [C.c_long, iP, C.c_int]),
            "orc_bbl": (None, [C.c_int] * 3 + [dP, iP, O, O, O, O, O, O]),
            "orc_propagate": (None, [C.c_int] * 3 + [dP]),
            "orc_collide": (None, [C.c_long, dP, iP, dP, dP, dP, C.c_double, C.c_double,
                                   C.c_int, C.c_int]),
            "orc_lbstep": (None, [C.c_int] * 3 + [dP, iP, iP, dP, dP, dP, O, O, O,
                                                  C.c_double, C.c_double, C.c_int, C.c_int,
                                                  C.c_int, C.c_long]),
            "orc_vspread": (None, [C.c_int] * 3 + [iP, dP, dP, dP, C.c_double, dP, dP, dP]),
            "orc_accvmean": (None, [C.c_long, dP, iP, dP, dP, dP, O, C.c_double, C.c_double,
                                    dP]),
            "orc_distribf": (None, [C.c_int] * 3 + [iP, C.c_int, dP, dP, dP, dP, dP, dP]),
            "orc_measure": (None, [C.c_int] * 3 + [dP, iP, iP, dP, dP, dP, O, O, O,
                                                   C.c_double, C.c_double, C.c_int, C.c_int,
                                                   C.c_int, C.c_int, C.c_int, dP, dP, dP, dP,
                                                   dP, dP, dP]),
            "orc_vfield": (None, [C.c_long, dP, iP, dP, dP, dP, C.c_double, C.c_double,
                                  dP, dP, dP]),
            "orc_equilibrium": (None, [C.c_long, dP, dP, dP, dP, dP, C.c_double, C.c_int]),
            "orc_moments": (None, [C.c_long, dP, dP]),
            "orc_mprhoinit": (None, [C.c_long, dP, iP, C.c_double, dP]),
            "orc_mpstep": (None, [C.c_int] * 3 + [dP, iP, iP, dP, C.c_double, C.c_double,
                                                  C.c_int, O, O, O, C.c_double, C.c_int,
                                                  dP, O, O, C.c_int]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        self.L = L

    # --- geometry
    def insertsph(self, objid, dims, oid, rad, ox, oy, oz):
        ov = C.c_int(0)
        ns = self.L.orc_insertsph(*dims, objid, oid, rad, ox, oy, oz, C.byref(ov))
        return ns, ov.value

    def makemask(self, objid, dims):
        m = np.zeros(int(np.prod(dims)), dtype=np.int32)
        self.L.orc_makemask(*dims, np.ascontiguousarray(objid, dtype=np.int32), m)
        return m

    # --- stepping
    def lbstep(self, a, mask, objid, dims, b, tau, stokes, skip_solid, u=None, nsteps=1,
               rho0=72.0):
        moving = u is not None
        ux, uy, uz = (None, None, None) if u is None else (_optd(u[0]), _optd(u[1]), _optd(u[2]))
        self.L.orc_lbstep(*dims, a, mask, objid, _optd(b[0]), _optd(b[1]), _optd(b[2]),
                          ux, uy, uz, tau, rho0, int(stokes), int(skip_solid), int(moving),
                          int(nsteps))
        return a

    def measure(self, a, mask, objid, dims, b, tau, stokes, skip_solid, objpos, u=None,
                vel_corrected=1, rho0=72.0):
        N = int(np.prod(dims))
        nobj = len(objpos)
        fx, fy, fz = np.zeros(N), np.zeros(N), np.zeros(N)
        vm = np.zeros(3)
        frc, trq = np.zeros(3 * max(nobj, 1)), np.zeros(3 * max(nobj, 1))
        moving = u is not None
        ux, uy, uz = (None, None, None) if u is None else (_optd(u[0]), _optd(u[1]), _optd(u[2]))
        self.L.orc_measure(*dims, a, mask, objid, _optd(b[0]), _optd(b[1]), _optd(b[2]),
                           ux, uy, uz, tau, rho0, int(stokes), int(skip_solid), int(moving),
                           int(vel_corrected), nobj,
                           np.ascontiguousarray(objpos, dtype=np.float64).reshape(-1)
                           if nobj else np.zeros(3), fx, fy, fz, vm, frc, trq)
        return dict(vmean=vm, fx=fx, fy=fy, fz=fz, frc=frc[:3 * nobj].reshape(nobj, 3),
                    trq=trq[:3 * nobj].reshape(nobj, 3))

    def vspread(self, objid, dims, objpos, objvel, objang, rho0=72.0):
        N = int(np.prod(dims))
        ux, uy, uz = np.zeros(N), np.zeros(N), np.zeros(N)
        f = lambda v: np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        self.L.orc_vspread(*dims, objid, f(objpos), f(objvel), f(objang), rho0, ux, uy, uz)
        return ux, uy, uz

    def mprhoinit(self, a, objid, rho0=72.0):
        N = len(objid)
        rho = np.zeros(N)
        self.L.orc_mprhoinit(N, a, objid, rho0, rho)
        return rho

    def mpstep(self, a, mask, objid, rho, dims, p, alpha=0.0, pads=0.0, m1=None, m2=None,
               axis=0, moving=False, objtab=None, rho0=72.0):
        nmom = 3 if m1 is not None else 1
        if moving:
            t = np.asarray(objtab, dtype=np.float64).reshape(-1, 9)
            pos, vel, ang = (np.ascontiguousarray(t[:, 0:3]).reshape(-1),
                             np.ascontiguousarray(t[:, 3:6]).reshape(-1),
                             np.ascontiguousarray(t[:, 6:9]).reshape(-1))
        else:
            pos = vel = ang = None
        self.L.orc_mpstep(*dims, a, mask, objid, rho, rho0, alpha, int(moving), pos, vel, ang,
                          pads, nmom, p, m1, m2, axis)


# --------------------------------------------------------------------------
# The reference library itself (oracle/_ref), driven through its C API.

def _ref_structs(real):
    RP = C.POINTER(real)

    class Obj(C.Structure):
        pass

    ObjP = C.POINTER(Obj)
    Obj._fields_ = [("id", C.c_int)] + [(n, real) for n in (
        "posx", "posy", "posz", "velx", "vely", "velz", "angx", "angy", "angz",
        "frcx", "frcy", "frcz", "trqx", "trqy", "trqz")] + [("next_obj", ObjP),
                                                             ("base_obj", ObjP)]

    class Sys(C.Structure):
        _fields_ = [
            ("w", C.POINTER(C.c_int)), ("c", C.POINTER(C.c_int)), ("a", RP),
            ("ux", RP), ("uy", RP), ("uz", RP), ("fx", RP), ("fy", RP), ("fz", RP),
            ("bx", RP), ("by", RP), ("bz", RP),
            ("esi", C.c_int), ("esj", C.c_int), ("esk", C.c_int), ("nnodes", C.c_int),
            ("step", C.c_int), ("status", C.c_int), ("changed_objects", C.c_int),
            ("objects_moving", C.c_int), ("skip_solid", C.c_int), ("stokes_flow", C.c_int),
            ("vel_corrected", C.c_int), ("monitoring", C.c_int),
            ("tau", real), ("rho0", real), ("vmeanx", real), ("vmeany", real),
            ("vmeanz", real), ("psi", RP), ("rho", RP), ("name", C.c_char_p),
            ("rescale", real), ("load_average", C.c_int),
            ("obj", C.POINTER(ObjP)), ("obj_head", ObjP), ("mask", C.POINTER(C.c_int)),
            ("next_id", C.c_int)]

    class MP(C.Structure):
        _fields_ = [("lb", C.POINTER(Sys)), ("p", RP), ("newp", RP), ("step", C.c_int),
                    ("objects_moving", C.c_int), ("to_be_initialised", C.c_int),
                    ("alpha", real), ("isoslice", real), ("poffset", real),
                    ("p2bytemin", real), ("p2bytemax", real), ("name", C.c_char_p)]

    return Obj, Sys, MP


class RefLib:
    """The reference library (oracle/_ref) behind the same C API."""

    def __init__(self, bits=64):
        path = REF_SO[bits]
        if not os.path.exists(path):
            raise ImportError(f"{path} missing (built by `make -C oracle ref` where the "
                              "reference sources exist)")
        real = C.c_double if bits == 64 else C.c_float
        self.bits = bits
        self.real = real
        self.np_real = np.float64 if bits == 64 else np.float32
        self.Obj, self.Sys, self.MP = _ref_structs(real)
        L = C.CDLL(path)
        SysP, ObjP, MPP = C.POINTER(self.Sys), C.POINTER(self.Obj), C.POINTER(self.MP)
        R, S, I = real, C.c_char_p, C.c_int
        sig = {
            "newlb": (SysP, [I, I, I, S]), "lbstep": (None, [SysP]), "bbl": (None, [SysP]),
            "bblforce": (None, [SysP]), "propagate": (None, [SysP]), "collide": (None, [SysP]),
            "insertsph": (ObjP, [SysP, R, R, R, R]), "insertwall": (ObjP, [SysP, S, I]),
            "insertslab": (ObjP, [SysP, R, R, R, R, R, R]),
            "insertcyl": (ObjP, [SysP, S, R, R, R]), "insertlist": (ObjP, [SysP, S]),
            "insertsubs": (ObjP, [SysP, S, I, I, I, I, I, I]),
            "flattenobj": (ObjP, [SysP]), "invertobj": (ObjP, [SysP]),
            "removeobjs": (None, [SysP]), "makemask": (None, [SysP]),
            "syncobjs": (None, [SysP]), "vspread": (None, [SysP]), "zerobf": (None, [SysP]),
            "addbf": (None, [SysP, C.c_double, C.c_double, C.c_double]),
            "measure": (None, [SysP]), "accvmean": (None, [SysP, R]),
            "countfluid": (I, [SysP]), "countmask": (I, [SysP, S]), "discrf": (R, [SysP]),
            "newmp": (MPP, [SysP, S]), "mpinit": (None, [MPP]), "mpstep": (None, [MPP]),
            "mpall": (None, [MPP, R]),
            "mpvacfinit": (None, [MPP, MPP, S, C.c_int]), "mpvacf": (None, [MPP, MPP]),
            "mpvacfwr": (R, [MPP, MPP, S, C.c_int]), "adall": (None, [MPP]), "mprhoinit": (None, [MPP]),
            "adtotal": (R, [MPP, S, I]), "lbloadvel": (None, [SysP, S, S]),
            "lbsavevel": (None, [SysP, S, S]), "lbsave": (None, [SysP, S]),
            "lbload": (None, [SysP, S]), "set_verbosity": (None, [I]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        self.L = L

    def arr(self, ptr, n, dtype=None):
        dt = dtype or self.np_real
        return np.ctypeslib.as_array(ptr, shape=(n,)).view(dt)

    def new(self, esi, esj, esk):
        return self.L.newlb(esi, esj, esk, b"null")

    def objid(self, h):
        """Dense object index per node in list order (head = 0), -1 fluid."""
        s = h.contents
        order = {}
        p, k = s.obj_head, 0
        while p:
            order[C.addressof(p.contents)] = k
            k += 1
            p = p.contents.next_obj
        objs = np.ctypeslib.as_array(C.cast(s.obj, C.POINTER(C.c_uint64)), shape=(s.nnodes,))
        out = np.full(s.nnodes, -1, dtype=np.int32)
        for addr, k in order.items():
            out[objs == addr] = k
        return out

    def objects(self, h):
        out = []
        p = h.contents.obj_head
        while p:
            out.append(p.contents)
            p = p.contents.next_obj
        return out
