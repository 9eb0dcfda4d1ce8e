"""The drop-in boundary, checked without a GPU: the library loads, exports
every function the public headers declare, its structs have the reference
layout, and a reference driver compiles and links against it unchanged."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "sunlightlb_amd", "libsunlightlb.so")
REF = "/root/reference"


def declared_functions(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = []
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, re.M):
        if m.group(0).lstrip().startswith(("typedef", "#")):
            continue
        names.append(m.group(1))
    return names


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if len(l.split()) >= 3}


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    syms = exported_symbols(LIB)
    for h in ("sunlightlb.h", "sunlightlb_gpu.h", "sunlightlb_nn.h"):
        decl = declared_functions(os.path.join(INC, h))
        assert len(decl) > 20 or h != "sunlightlb.h"
        missing = [d for d in decl if d not in syms]
        assert not missing, f"{h}: not exported: {missing}"
    for g in ("slb_verbosity", "slb_signal_flag"):
        assert g in syms


def test_python_binding_declares_every_symbol():
    import sunlightlb_amd
    decl = set(declared_functions(os.path.join(INC, "sunlightlb.h")))
    decl |= set(declared_functions(os.path.join(INC, "sunlightlb_gpu.h")))
    decl |= set(declared_functions(os.path.join(INC, "sunlightlb_nn.h")))
    assert decl <= set(sunlightlb_amd.SIGNATURES), decl - set(sunlightlb_amd.SIGNATURES)
    L = sunlightlb_amd.lib()  # loading needs no GPU
    assert L.myscmp(b"null", b"nullfile") == 1 and L.myscmp(b"abc", b"abd") == 0


OFFSETS_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "sunlightlb.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  P(LBObj_type, id); P(LBObj_type, posx); P(LBObj_type, trqz); P(LBObj_type, next_obj);
  P(LBObj_type, base_obj);
  P(LBSys_type, w); P(LBSys_type, c); P(LBSys_type, a); P(LBSys_type, bz); P(LBSys_type, esi);
  P(LBSys_type, monitoring); P(LBSys_type, tau); P(LBSys_type, rho0); P(LBSys_type, vmeanz);
  P(LBSys_type, psi); P(LBSys_type, rho); P(LBSys_type, name); P(LBSys_type, rescale);
  P(LBSys_type, load_average); P(LBSys_type, obj); P(LBSys_type, obj_head);
  P(LBSys_type, mask); P(LBSys_type, next_id);
  P(MPSys_type, lb); P(MPSys_type, p); P(MPSys_type, newp); P(MPSys_type, step);
  P(MPSys_type, alpha); P(MPSys_type, p2bytemax); P(MPSys_type, name);
  P(MPIso_type, count); P(LBColourMap_type, v_hi); P(LBImage_type, object_blue);
  printf("sizeof.MPSys %zu\n", sizeof(MPSys_type));
  printf("sizeof.LBObj %zu\n", sizeof(LBObj_type));
  return 0;
}
"""


def _offsets(incdir, tmp, extra=()):
    src = os.path.join(tmp, "off.c")
    exe = os.path.join(tmp, "off")
    with open(src, "w") as f:
        f.write(OFFSETS_C)
    subprocess.run(["gcc", "-I", incdir, *extra, src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return dict(l.rsplit(" ", 1) for l in out.strip().splitlines())


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference absent")
@pytest.mark.parametrize("extra", [(), ("-D_SINGLE",)])
def test_struct_layout_matches_reference_header(tmp_path, extra):
    mine = _offsets(INC, str(tmp_path), extra)
    ref = _offsets(os.path.join(REF, "src"), str(tmp_path), extra)
    assert mine == ref


def test_dpd_fields_are_appended():
    import sunlightlb_amd as s
    names = [f[0] for f in s.LBSys_type._fields_]
    assert names[-3:] == ["dpdx", "dpdy", "dpdz"]
    assert s.LBSys_type.dpdx.offset > s.LBSys_type.next_id.offset


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "examples", "sphere.c")),
                    reason="reference absent")
def test_reference_sphere_driver_links_unchanged(tmp_path):
    exe = str(tmp_path / "sphere_c")
    r = subprocess.run(["gcc", "-O2", "-Wall", "-I", INC,
                        os.path.join(REF, "examples", "sphere.c"), "-L",
                        os.path.dirname(LIB), "-lsunlightlb", "-lm", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(exe)


def test_own_c_driver_compiles(tmp_path):
    drv = os.path.join(ROOT, "examples", "config1.c")
    exe = str(tmp_path / "config1")
    r = subprocess.run(["gcc", "-O2", "-Wall", "-I", INC, drv, "-L", os.path.dirname(LIB),
                        "-lsunlightlb", "-lm", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
