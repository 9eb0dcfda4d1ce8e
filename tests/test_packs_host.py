"""The host pack generator (slb_pack_map: thread-safe, no lattice, no GPU)
against the reference library: the solid map of the 256^3 permeability pack
(radius 8, lattice solid fraction 0.35, seed 20261202) equals, bit for bit,
the one the reference's insertsph wrote from the same sphere centres
(tests/golden/pack256.json.gz, SHA-256 of the int32 object index after
flattenobj); the sphere count and solid fraction agree."""
import ctypes as C
import gzip
import hashlib
import json
import os

import numpy as np

from conftest import GOLDEN


def test_pack_map_matches_reference_digest():
    import sunlightlb_amd as s
    L = s.lib()
    with gzip.open(os.path.join(GOLDEN, "pack256.json.gz"), "rt") as f:
        rec = json.load(f)
    n = 256
    m = np.empty(n ** 3, np.uint8)
    phi = C.c_double(0.0)
    placed = L.slb_pack_map(n, n, n, rec["radius"], rec["phi"], rec["seed"], 0, 50_000_000,
                            m.ctypes.data, C.byref(phi))
    assert placed == rec["spheres"]
    assert int(m.sum()) == rec["solid"]
    assert abs(phi.value - rec["solid"] / n ** 3) < 1e-15
    objid = np.where(m > 0, 0, -1).astype("<i4")
    assert hashlib.sha256(objid.tobytes()).hexdigest() == rec["flat_objid_sha256"]
