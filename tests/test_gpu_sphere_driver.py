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

from textcmp import compare_text

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "sphere_c")
GOLDEN = os.path.join(ROOT, "tests", "golden", "sphere_c.json.gz")

VOLATILE = re.compile(r"process id|elapsed|updates/sec|^Clock started")


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
