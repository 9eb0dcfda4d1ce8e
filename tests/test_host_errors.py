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
