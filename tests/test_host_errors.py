"""Error paths of the host API that need no GPU: a warm-start weights file
with a wrong magic, a truncated one, or a missing one is refused (NULL), and
bench.py refuses more than one GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def lib():
    import sunlightlb_amd as s
    return s.lib()


def test_warm_weights_file_errors(tmp_path):
    L = lib()
    assert not L.slb_warm_load(str(tmp_path / "missing.bin").encode())
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTANET!" + bytes(64))
    assert not L.slb_warm_load(str(bad).encode())
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(b"SLBWNET4" + bytes(10))
    assert not L.slb_warm_load(str(trunc).encode())


def _weights_file(path, ncurve, nwcurve):
    """A complete weights file in the layout of slb_warm_save (magic, layer
    table, fallback threshold, both curves, zero parameters) with the given
    curve counts."""
    import ctypes as C
    import struct
    L = lib()
    nl = C.c_int(0)
    cin = (C.c_int * 32)()
    cout = (C.c_int * 32)()
    bofs = (C.c_long * 32)()
    L.slb_warm_layout(C.byref(nl), cin, cout, None, bofs)
    nparam = bofs[nl.value - 1] + 64
    hdr = [nl.value] + [v for q in range(nl.value) for v in (cin[q], cout[q])] + [0]
    blob = b"SLBWNET4" + struct.pack(f"<{len(hdr)}i", *hdr) + struct.pack("<d", 0.0)
    blob += struct.pack("<ii", ncurve, 100) + bytes(8 * 256)
    blob += struct.pack("<ii", nwcurve, 100) + bytes(8 * 256)
    blob += struct.pack("<q", nparam) + bytes(4 * nparam)
    path.write_bytes(blob)


def test_warm_weights_file_with_bad_curve_counts(tmp_path):
    """A curve count beyond the fixed curve table is refused before any
    parameter is read or any device memory is touched (a subprocess, so
    that a regression which went on to the device cannot end this one)."""
    def load(nc, nw):
        f = tmp_path / f"curve_{nc}_{nw}.bin"
        _weights_file(f, nc, nw)
        code = ("import sunlightlb_amd as s; L = s.lib(); "
                f"print('loaded' if L.slb_warm_load({str(f)!r}.encode()) else 'refused')")
        return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                              timeout=120, cwd=ROOT)

    for nc, nw in ((100000, 10), (10, 100000), (-1, 10), (10, -5)):
        p = load(nc, nw)
        assert p.returncode == 0 and p.stdout.strip() == "refused", (nc, nw, p.stdout, p.stderr)
    # the same file with counts inside the table is not refused: it goes on
    # to the device (which a host without a GPU reports as an error)
    p = load(256, 12)
    assert "refused" not in p.stdout, p.stdout


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)


def test_newlb_rejects_bad_dimensions():
    """Non-positive dimensions, or more nodes than the int node count holds,
    end the process with the library's error message before any device work
    (the check precedes device set-up, so it runs without a GPU)."""
    for dims in ("0, 4, 4", "4, -1, 4", "2048, 1024, 1024"):
        code = ("import sunlightlb_amd as s; L = s.lib(); "
                f"L.newlb({dims}, b'bad'); print('returned')")
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           timeout=120, cwd=ROOT)
        assert p.returncode == 1, (dims, p.stdout, p.stderr)
        assert "lattice dimensions" in p.stderr
        assert "returned" not in p.stdout


def test_string_helpers_need_no_device():
    """mycat, mycopy and myscmp (src/lbmisc.c) are host code: concatenation and
    copy into fresh memory, and myscmp(s, t) true when t starts with s (the
    file-name tests "null", "stdout", "stderr" of the writers rely on it)."""
    import ctypes as C
    L = lib()
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    p = L.mycat(b"sphere", b".dat")
    q = L.mycopy(b"null")
    try:
        assert C.string_at(p) == b"sphere.dat"
        assert C.string_at(q) == b"null"
    finally:
        libc.free(p)
        libc.free(q)
    assert L.myscmp(b"null", b"null") == 1
    assert L.myscmp(b"std", b"stdout") == 1
    assert L.myscmp(b"stdout", b"std") == 0
    assert L.myscmp(b"null", b"nul") == 0
