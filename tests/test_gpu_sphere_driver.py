"""The reference's C example driver (examples/sphere.c), compiled unchanged
against include/sunlightlb.h and linked to this library (oracle/Makefile,
target `drivers`), run on the GPU: the files it writes and its stdout must
match the reference library's run of the same calls
(tests/golden/sphere_c.json.gz, written by tests/golden/make_golden.py).

sphere.c sets the body force through mylb->dpdx, a field the current
reference header lacks; the fixture gives the reference the equivalent
addbf(0.01, 0, 0) on the fluid nodes.  Bars: the solid map and solid-node
list identical; every number in the force list, the velocity dump and the
printed measurements equal to the reference's within one unit of the last
printed digit (the FP64 GPU state agrees with the reference to ~1e-13, so
a printed digit differs only where a value sits on a rounding boundary);
stdout line for line from newlb on (about() identifies each library), apart
from process ids and wall-clock timings."""
import gzip
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "sphere_c")
GOLDEN = os.path.join(ROOT, "tests", "golden", "sphere_c.json.gz")

NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?$")
VOLATILE = re.compile(r"process id|elapsed|updates/sec|^Clock started")


def last_digit(tok):
    """One unit of the last printed digit of a numeric token."""
    mant, _, exp = tok.lower().partition("e")
    dec = len(mant.split(".")[1]) if "." in mant else 0
    return 10.0 ** (int(exp or 0) - dec)


def same_tokens(a, b):
    ta, tb = a.split(), b.split()
    if len(ta) != len(tb):
        return False
    for x, y in zip(ta, tb):
        if x == y:
            continue
        if not (NUM.match(x) and NUM.match(y)):
            return False
        if abs(float(x) - float(y)) > 1.01 * max(last_digit(x), last_digit(y)):
            return False
    return True


def compare_text(got, ref, what, skip=None):
    gl = [l for l in got.splitlines() if not (skip and skip.search(l))]
    rl = [l for l in ref.splitlines() if not (skip and skip.search(l))]
    assert len(gl) == len(rl), f"{what}: {len(gl)} lines, reference {len(rl)}"
    bad = [(i, g, r) for i, (g, r) in enumerate(zip(gl, rl)) if not same_tokens(g, r)]
    assert not bad, f"{what}: {len(bad)} lines differ, first: {bad[:3]}"


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/sphere_c is built from the reference example by "
                    "`make -C oracle drivers` where the reference sources exist")
    d = tmp_path_factory.mktemp("sphere_c")
    p = subprocess.run([BIN], cwd=d, capture_output=True, text=True, timeout=180)
    assert p.returncode == 0, p.stderr[-2000:]
    with gzip.open(GOLDEN, "rt") as f:
        ref = json.load(f)
    return d, p.stdout, ref


def test_files_match_reference(run):
    d, _, ref = run
    for name in ("sphere_c.map", "sphere_c.dat"):
        assert open(os.path.join(d, name)).read() == ref["files"][name], name
    for name in ("sphere_c.frc", "sphere_c.vel"):
        compare_text(open(os.path.join(d, name)).read(), ref["files"][name], name)
    compare_text(open(os.path.join(d, "sphere_c.mon")).read(), ref["files"]["sphere_c.mon"],
                 "sphere_c.mon", skip=re.compile(r"elapsed|updates/sec|:\d\d:\d\d"))


def after_banner(text):
    """stdout from newlb's first line on: about() prints each library's own
    identification, not part of the compared format."""
    i = text.find("LB system ")
    assert i >= 0
    return text[i:]


def test_stdout_matches_reference(run):
    _, out, ref = run
    compare_text(after_banner(out), after_banner(ref["stdout"]), "stdout", skip=VOLATILE)
