"""Comparison of printed outputs: numbers equal within one unit of the last
printed digit, other tokens identical."""
import re

NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?$")


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
