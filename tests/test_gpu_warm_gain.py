"""What the trained warm-start network gains over a cold start, on the GPU.

The network is trained by the library's own GPU kernels on converged Stokes
flows of 160 seeded 64^3 sphere packs (radius 4-10, solid fraction
0.2-0.45); on held-out packs it must

  * reduce the training loss by at least 3x;
  * reach a residual of 1e-4 in fewer steps than a cold start (summed over
    the held-out packs, at least 1.2x fewer), i.e. it starts much closer to
    the steady state;
  * with the cold-start fallback, still reach 1e-6 and give the cold-start
    permeability within 2e-5 relative.

At 1e-6 the warm start alone is not faster on these packs: the prediction's
~10% density and divergence errors leave a slowly relaxing mass imbalance
between weakly connected pores (a tail near 1e-5 that a cold start does not
have); the fallback bounds that loss.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 64
F = 1e-6
ETA = 72.0 * (0.75 - 0.5) / 3.0


def lib():
    import sunlightlb_amd as s
    return s.lib()


def pack(L, R, phi, seed, overlap):
    m = np.empty(N ** 3, np.uint8)
    po = C.c_double(0.0)
    L.slb_pack_map(N, N, N, R, phi, seed, overlap, 5_000_000, m.ctypes.data, C.byref(po))
    h = L.newlb(N, N, N, b"null")
    h.contents.stokes_flow = 1
    h.contents.tau = 0.75
    L.slb_insert_map(h, m.ctypes.data)
    L.zerobf(h)
    L.addbf(h, F, 0.0, 0.0)
    L.syncobjs(h)
    L.slb_set_precision(h, 32)
    return h


def solve(L, h, tol, maxsteps=60000):
    L.slb_residual_reset(h)
    L.slb_residual(h)
    steps, hist = 0, []
    while steps < maxsteps:
        L.slb_run(h, 100)
        steps += 100
        r = L.slb_residual(h)
        hist.append(r)
        if 0.0 <= r < tol:
            break
    return steps, hist


def first_below(hist, tol):
    return next((i + 1) * 100 for i, x in enumerate(hist) if 0.0 <= x < tol)


def params(rng, n):
    out = []
    for _ in range(n):
        R = float(rng.uniform(4.0, 10.0))
        phi = float(rng.uniform(0.2, 0.45))
        out.append((R, phi, int(rng.integers(1, 2 ** 62)), int(phi > 0.34)))
    return out


def test_trained_network_starts_closer_than_cold():
    L = lib()
    rng = np.random.default_rng(2026)
    net = L.slb_warm_new(11)
    try:
        for R, phi, seed, ov in params(rng, 160):
            h = pack(L, R, phi, seed, ov)
            _, hist = solve(L, h, 1e-6)
            L.slb_warm_note_cold(net, (C.c_double * len(hist))(*hist), len(hist), 100)
            L.slb_warm_add_sample(net, h)
            L.slb_destroy(h)
        iters = 8000
        hist = (C.c_double * iters)()
        last = L.slb_warm_train(net, iters, 1e-3, 3, hist)
        first = float(np.mean(hist[:100]))
        print(f"training loss {first:.3f} -> {last:.3f}", flush=True)
        assert last < first / 3

        cold4 = warm4 = 0
        for R, phi, seed, ov in params(np.random.default_rng(77), 6):
            h = pack(L, R, phi, seed, ov)
            cs, ch = solve(L, h, 1e-6)
            L.measure(h)
            kc = h.contents.vmeanx * ETA / F
            L.slb_destroy(h)

            h = pack(L, R, phi, seed, ov)
            L.slb_warm_start(net, h)
            _, wh = solve(L, h, 1e-4)
            L.slb_destroy(h)

            h = pack(L, R, phi, seed, ov)
            r, used = C.c_double(0.0), C.c_int(0)
            ws = L.slb_warm_solve(net, h, 1e-6, 100, 60000, C.byref(r), C.byref(used))
            L.measure(h)
            kw = h.contents.vmeanx * ETA / F
            L.slb_destroy(h)
            c4, w4 = first_below(ch, 1e-4), first_below(wh, 1e-4)
            cold4 += c4
            warm4 += w4
            print(f"R={R:.2f} phi={phi:.3f}: to 1e-4 cold {c4} warm {w4}; to 1e-6 cold {cs} "
                  f"warm+fallback {ws} (kept warm {used.value}), dk/k {abs(kw - kc) / kc:.1e}",
                  flush=True)
            assert 0.0 <= r.value < 1e-6
            assert abs(kw - kc) / kc < 2e-5
        print(f"steps to 1e-4, summed: cold {cold4}, warm {warm4}")
        assert cold4 >= 1.2 * warm4
    finally:
        L.slb_warm_free(net)
