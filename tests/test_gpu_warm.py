"""GPU parity of the warm-start network kernels and lattice hooks, called
through the C ABI (libsunlightlb.so), against the CPU oracle (oracle/nn.py,
itself pinned against torch's CPU conv3d / autograd).

Bars (written per test): FP32 convolution outputs and weight gradients
within 1e-5 of the oracle relative to the largest value (FP32 accumulation
order only); BF16 outputs within one BF16 rounding step; jump-flooding
features bit-exact; whole-network gradients within 2e-2 relative L2 per
layer (BF16 activations whose roundings can flip); equilibrium injection
within 1e-6 relative L2 of lbloadvel fed the same predicted velocity (FP64
lattice, FP32 network output); moment injection exact on its own moments."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from conftest import rel_l2
from oracle import nn as onn

pytestmark = pytest.mark.gpu


def lib():
    import sunlightlb_amd as s
    return s.lib()


class Dev:
    """A device buffer owned by the test (slbn_malloc / slbn_free)."""

    def __init__(self, nbytes=None, host=None):
        self.L = lib()
        if host is not None:
            host = np.ascontiguousarray(host)
            nbytes = host.nbytes
        self.nbytes = int(nbytes)
        self.p = self.L.slbn_malloc(self.nbytes)
        if host is not None:
            self.L.slbn_h2d(self.p, host.ctypes.data, self.nbytes)
        else:
            self.L.slbn_zero(self.p, self.nbytes)

    def get(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        self.L.slbn_d2h(out.ctypes.data, self.p, out.nbytes)
        return out

    def free(self):
        if self.p:
            self.L.slbn_free(self.p)
            self.p = None


def bf_random(rng, shape, scale=1.0):
    bits = onn.bf16_bits(rng.standard_normal(shape) * scale)
    return bits, onn.from_bf16_bits(bits)


def weights(rng, cout, cin):
    """W[co][Kp] BF16 bits with zero K padding, and w[co, t, ci] float64."""
    kp = onn.kpad(cin)
    bits, vals = bf_random(rng, (cout, 27 * cin), np.sqrt(2.0 / (27 * cin)))
    W = np.zeros((cout, kp), np.uint16)
    W[:, :27 * cin] = bits
    return W, vals.reshape(cout, 27, cin)


def within_bf16_step(got_bits, ref, atol=1e-5):
    """BF16 results within one rounding step of the float64 reference, plus
    atol * max|ref| for the FP32 accumulation order (values near a ReLU
    cut or a rounding boundary); returns a bool, printing the worst case."""
    got = onn.from_bf16_bits(got_bits)
    bound = np.abs(ref) * 2.0 ** -7 + atol * np.abs(ref).max() + 1e-30
    bad = np.abs(got - ref) > bound
    if bad.any():
        i = np.argmax(np.abs(got - ref) - bound)
        print(f"  {bad.sum()} of {bad.size} outside: got {got.flat[i]!r} ref {ref.flat[i]!r}")
    return not bad.any()


CONV_CASES = [((16, 16, 16), 16, 16, 1), ((8, 16, 24), 32, 64, 1), ((16, 8, 8), 64, 32, 0),
              ((8, 8, 16), 16, 48, 1), ((5, 7, 9), 8, 16, 0), ((4, 6, 8), 16, 8, 1)]


@pytest.mark.parametrize("shape,cin,cout,relu", CONV_CASES)
def test_conv_forward_matches_oracle(shape, cin, cout, relu):
    print(f"conv fwd {shape} {cin}->{cout} relu={relu}", flush=True)
    L = lib()
    rng = np.random.default_rng(10 + cin + cout)
    N = int(np.prod(shape))
    xb, x = bf_random(rng, shape + (cin,))
    W, w = weights(rng, cout, cin)
    bias = (rng.standard_normal(cout) * 0.1).astype(np.float32)
    dX, dW, dB = Dev(host=xb), Dev(host=W), Dev(host=bias)
    dY, dYf = Dev(N * cout * 2), Dev(N * cout * 4)
    L.slbn_conv(dX.p, cin, dW.p, dB.p, dY.p, dYf.p, cout, relu, 0, *shape)
    L.slbn_sync()
    ref = onn.conv3d(x, w, bias.astype(np.float64))
    if relu:
        ref = np.maximum(ref, 0.0)
    yf = dYf.get(np.float32, shape + (cout,)).astype(np.float64)
    yb = dY.get(np.uint16, shape + (cout,))
    for d in (dX, dW, dB, dY, dYf):
        d.free()
    err = np.abs(yf - ref).max() / np.abs(ref).max()
    print(f"  max rel err {err:.2e}")
    assert err < 1e-5
    assert within_bf16_step(yb, ref)


@pytest.mark.parametrize("shape,cin,cout", [((16, 16, 16), 16, 16), ((8, 16, 24), 32, 64),
                                            ((16, 8, 8), 64, 32), ((8, 8, 8), 16, 64)])
def test_conv_input_gradient_matches_oracle(shape, cin, cout):
    print(f"conv bwd-data {shape} {cin}->{cout}", flush=True)
    L = lib()
    rng = np.random.default_rng(20 + cin + cout)
    N = int(np.prod(shape))
    gb, g = bf_random(rng, shape + (cout,))
    W, w = weights(rng, cout, cin)
    dG, dW = Dev(host=gb), Dev(host=W)
    dWT = Dev(cin * onn.kpad(cout) * 2)
    dY, dYf = Dev(N * cin * 2), Dev(N * cin * 4)
    L.slbn_wtrans(dW.p, cout, cin, dWT.p)
    L.slbn_conv(dG.p, cout, dWT.p, None, dY.p, dYf.p, cin, 0, 1, *shape)
    L.slbn_sync()
    ref = onn.conv3d_bwd_data(g, w)
    yf = dYf.get(np.float32, shape + (cin,)).astype(np.float64)
    yb = dY.get(np.uint16, shape + (cin,))
    for d in (dG, dW, dWT, dY, dYf):
        d.free()
    err = np.abs(yf - ref).max() / np.abs(ref).max()
    print(f"  max rel err {err:.2e}")
    assert err < 1e-5
    assert within_bf16_step(yb, ref)


@pytest.mark.parametrize("shape,cin,cout,small", [((16, 16, 16), 16, 16, 0), ((8, 16, 24), 32, 64, 0),
                                                  ((32, 32, 32), 64, 64, 0), ((16, 8, 8), 64, 32, 0),
                                                  ((8, 8, 12), 16, 32, 0), ((4, 6, 6), 64, 16, 0),
                                                  ((8, 16, 16), 32, 16, 0),
                                                  ((6, 10, 24), 16, 16, 0), ((3, 2, 8), 16, 16, 0),
                                                  ((5, 6, 128), 16, 16, 0), ((4, 8, 40), 16, 16, 0),
                                                  ((4, 5, 16), 16, 16, 0), ((2, 4, 136), 16, 16, 0),
                                                  ((8, 8, 12), 8, 16, 1), ((6, 8, 10), 16, 8, 1)])
def test_conv_weight_gradient_matches_oracle(shape, cin, cout, small):
    L = lib()
    print(f"conv wgrad {shape} {cin}->{cout} small={small} "
          f"lds16={L.slbn_wgrad_lds16_ok(cin, cout, *shape)} "
          f"blocked={L.slbn_wgrad_blocked_ok(cin, cout, *shape)}", flush=True)
    rng = np.random.default_rng(30 + cin + cout)
    xb, x = bf_random(rng, shape + (cin,))
    gb, g = bf_random(rng, shape + (cout,))
    kp = onn.kpad(cin)
    dX, dG, dW = Dev(host=xb), Dev(host=gb), Dev(cout * kp * 4)
    f = L.slbn_conv_wgrad_small if small else L.slbn_conv_wgrad
    f(dX.p, cin, dG.p, cout, dW.p, *shape)
    L.slbn_sync()
    got = dW.get(np.float32, (cout, kp)).astype(np.float64)
    for d in (dX, dG, dW):
        d.free()
    ref = onn.conv3d_bwd_weight(x, g).reshape(cout, 27 * cin)
    assert np.all(got[:, 27 * cin:] == 0.0)
    err = np.abs(got[:, :27 * cin] - ref).max() / np.abs(ref).max()
    print(f"  max rel err {err:.2e}")
    assert err < 1e-5


def test_pool_upsample_and_elementwise_kernels():
    L = lib()
    rng = np.random.default_rng(40)
    shape, C = (8, 12, 16), 32
    N = int(np.prod(shape))
    cshape = tuple(s // 2 for s in shape)
    Nc = N // 8
    xb, x = bf_random(rng, shape + (C,))
    cb, c = bf_random(rng, cshape + (C,))
    dX, dC = Dev(host=xb), Dev(host=cb)
    out_c, out_f = Dev(Nc * C * 2), Dev(N * C * 2)
    # pooling and its gradient
    L.slbn_pool(dX.p, out_c.p, C, *shape)
    L.slbn_sync()
    assert within_bf16_step(out_c.get(np.uint16, cshape + (C,)), onn.pool(x))
    L.slbn_pool_bwd(dC.p, out_f.p, C, *shape)
    L.slbn_sync()
    assert within_bf16_step(out_f.get(np.uint16, shape + (C,)), onn.pool_bwd(c))
    # upsample + skip, upsample alone, and the upsample gradient
    L.slbn_up_add(dC.p, dX.p, out_f.p, C, *shape)
    L.slbn_sync()
    assert within_bf16_step(out_f.get(np.uint16, shape + (C,)), onn.up(c) + x)
    L.slbn_up_add(dC.p, None, out_f.p, C, *shape)
    L.slbn_sync()
    assert np.array_equal(out_f.get(np.uint16, shape + (C,)),
                          onn.bf16_bits(onn.up(c)))
    L.slbn_up_bwd(dX.p, out_c.p, C, *shape)
    L.slbn_sync()
    assert within_bf16_step(out_c.get(np.uint16, cshape + (C,)), onn.up_bwd(x))
    # ReLU gradient, add, bias gradient
    yb, y = bf_random(rng, shape + (C,))
    gb, g = bf_random(rng, shape + (C,))
    dYr, dGr, dP = Dev(host=yb), Dev(host=gb), Dev(N * C * 2)
    L.slbn_relu_bwd(dYr.p, dGr.p, dP.p, N * C)
    L.slbn_sync()
    assert np.array_equal(dP.get(np.uint16, (N, C)).reshape(-1),
                          np.where(y.reshape(-1) > 0, gb.reshape(-1), 0).astype(np.uint16))
    L.slbn_add_into(dYr.p, dGr.p, N * C)
    L.slbn_sync()
    assert within_bf16_step(dYr.get(np.uint16, (N * C,)), (y + g).reshape(-1))
    dB = Dev(C * 4)
    L.slbn_bias_grad(dGr.p, C, N, dB.p)
    L.slbn_sync()
    db = dB.get(np.float32, (C,))
    ref = g.reshape(N, C).sum(0)
    assert np.abs(db - ref).max() < 1e-5 * np.abs(g).sum(axis=(0, 1, 2)).max()
    # expand a packed 2-channel sample to 16 channels
    sb, _ = bf_random(rng, (N, 2))
    dS, dE = Dev(host=sb), Dev(N * 16 * 2)
    L.slbn_expand(dS.p, 2, dE.p, 16, N)
    L.slbn_sync()
    e = dE.get(np.uint16, (N, 16))
    assert np.array_equal(e[:, :2], sb) and not e[:, 2:].any()
    for d in (dX, dC, out_c, out_f, dYr, dGr, dP, dB, dS, dE):
        d.free()


@pytest.mark.parametrize("C,N", [(16, 5000), (32, 3001), (64, 777)])
def test_fused_relu_backward_and_bias_gradient(C, N):
    """slbn_relu_bwd_bias writes the same dP as slbn_relu_bwd (bit for bit) and
    adds the bias gradient slbn_bias_grad computes from it (FP32, float
    atomics over blocks: to rounding), and both match the direct formula."""
    L = lib()
    rng = np.random.default_rng(50 + C)
    yb, y = bf_random(rng, (N, C))
    gb, g = bf_random(rng, (N, C))
    dY, dG = Dev(host=yb), Dev(host=gb)
    dP1, dP2, db1, db2 = Dev(N * C * 2), Dev(N * C * 2), Dev(C * 4), Dev(C * 4)
    L.slbn_relu_bwd(dY.p, dG.p, dP1.p, N * C)
    L.slbn_bias_grad(dP1.p, C, N, db1.p)
    L.slbn_relu_bwd_bias(dY.p, dG.p, dP2.p, C, N, db2.p)
    L.slbn_sync()
    p1, p2 = dP1.get(np.uint16, (N, C)), dP2.get(np.uint16, (N, C))
    b1, b2 = db1.get(np.float32, (C,)), db2.get(np.float32, (C,))
    for d in (dY, dG, dP1, dP2, db1, db2):
        d.free()
    assert np.array_equal(p1, p2)
    assert np.array_equal(p2, np.where(y > 0, gb, 0).astype(np.uint16))
    scale = np.abs(g).sum(0).max()
    assert np.abs(b2 - b1).max() <= 1e-6 * scale
    ref = np.where(y > 0, g, 0.0).sum(0)
    assert np.abs(b2 - ref).max() < 1e-5 * scale


def test_loss_and_adam_kernels():
    L = lib()
    rng = np.random.default_rng(41)
    N, Cc, Ct = 3000, 16, 10
    yf = rng.standard_normal((N, Cc)).astype(np.float32)
    tb, t = bf_random(rng, (N, Ct))
    wc = (rng.random(Ct) * 1e-3).astype(np.float32)
    dYf, dT, dWc = Dev(host=yf), Dev(host=tb), Dev(host=wc)
    dG, dL = Dev(N * Cc * 2), Dev(4)
    L.slbn_mse(dYf.p, Cc, dT.p, Ct, Ct, N, dWc.p, dG.p, dL.p)
    L.slbn_sync()
    loss, g = onn.mse_grad(yf.astype(np.float64), t, wc.astype(np.float64))
    assert abs(dL.get(np.float32, (1,))[0] - loss) < 1e-5 * loss
    assert within_bf16_step(dG.get(np.uint16, (N, Cc)), g)
    # Adam: FP32 master weights and the refreshed BF16 copy
    n = 5000
    w = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    dw, dm, dv = Dev(host=w), Dev(host=m), Dev(host=v)
    dwb = Dev(n * 2)
    ref_w, ref_m, ref_v = w.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    for step in range(1, 4):
        gr = rng.standard_normal(n).astype(np.float32)
        dg = Dev(host=gr)
        L.slbn_adam(dw.p, dwb.p, dg.p, dm.p, dv.p, n, 1e-3, 0.9, 0.999, 1e-8, step)
        L.slbn_sync()
        dg.free()
        ref_w, ref_m, ref_v = onn.adam(ref_w, gr.astype(np.float64), ref_m, ref_v, 1e-3, 0.9,
                                       0.999, 1e-8, step)
    got = dw.get(np.float32, (n,))
    assert np.abs(got - ref_w).max() < 1e-6
    assert np.array_equal(dwb.get(np.uint16, (n,)), onn.bf16_bits(got))
    for d in (dYf, dT, dWc, dG, dL, dw, dm, dv, dwb):
        d.free()


def test_network_forward_backward_matches_oracle():
    """One forward+backward of the whole network on the GPU against the
    oracle with the same BF16 rounding points; gradients per layer."""
    print("network fwd/bwd 16x16x24", flush=True)
    L = lib()
    rng = np.random.default_rng(50)
    net = L.slb_warm_new(7)
    try:
        n = L.slb_warm_nparams(net)
        arena = np.empty(n, np.float32)
        L.slb_warm_get_params(net, arena.ctypes.data)
        wofs, bofs, total = onn.layout()
        assert total == n
        # non-zero biases exercise the bias path
        for Lr in range(12):
            arena[bofs[Lr]:bofs[Lr] + onn.LCOUT[Lr]] = rng.standard_normal(onn.LCOUT[Lr]) * 0.05
        L.slb_warm_set_params(net, arena.ctypes.data)
        shape = (16, 16, 24)
        N = int(np.prod(shape))
        x = np.zeros(shape + (onn.NCH,))
        x[..., 0] = rng.random(shape) < 0.3
        x[..., 1] = rng.standard_normal(shape)
        xb = onn.bf16_bits(x)
        tb, t = bf_random(rng, shape + (onn.NTARG,))
        wc = onn.target_weights(t).astype(np.float32)
        dX, dT = Dev(host=xb), Dev(host=tb)
        dims = (C.c_int * 3)(*shape)
        outf = np.empty((N, onn.NCH), np.float32)
        grad = np.empty(n, np.float32)
        se = L.slb_warm_fwd_bwd(net, dims, dX.p, dT.p, wc.ctypes.data, outf.ctypes.data,
                                grad.ctypes.data)
        dX.free()
        dT.free()
    finally:
        L.slb_warm_free(net)
    params = onn.unpack(arena.astype(np.float64))
    out, cache = onn.forward(params, onn.from_bf16_bits(xb), bf=True)
    loss, dout = onn.mse_grad(out, t, wc.astype(np.float64), bf=True)
    grads = onn.backward(cache, dout)
    e_out = rel_l2(outf.reshape(shape + (onn.NCH,)), out)
    print(f"  output rel L2 {e_out:.2e}, loss {se:.6e} vs {loss:.6e}")
    assert e_out < 5e-3
    assert abs(se - loss) < 1e-2 * loss
    g_ref = onn.pack(grads)
    for Lr in range(12):
        a, b = wofs[Lr], bofs[Lr] + onn.LCOUT[Lr]
        e = rel_l2(grad[a:b], g_ref[a:b])
        print(f"  layer {Lr} grad rel L2 {e:.2e}")
        assert e < 2e-2, Lr
    # padding columns and unused bias slots stay exactly zero
    used = np.zeros(n, bool)
    for Lr in range(12):
        kp = onn.kpad(onn.LCIN[Lr])
        blk = np.zeros((onn.LCOUT[Lr], kp), bool)
        blk[:, :27 * onn.LCIN[Lr]] = True
        used[wofs[Lr]:wofs[Lr] + blk.size] = blk.reshape(-1)
        used[bofs[Lr]:bofs[Lr] + onn.LCOUT[Lr]] = True
    assert not grad[~used].any()


def pack_lattice(n, spheres, force=1e-5, prec=64):
    import sunlightlb_amd as s
    lat = s.Lattice(n[0], n[1], n[2], precision=prec)
    lat.s.stokes_flow = 1
    lat.s.tau = 0.75
    for r, x, y, z in spheres:
        lat.L.insertsph(lat.h, r, x, y, z)
    lat.L.zerobf(lat.h)
    lat.L.addbf(lat.h, force, 0.0, 0.0)
    return lat


SPHERES = [(3.5, 4.0, 5.0, 6.0), (2.5, 11.0, 3.0, 17.0), (4.0, 12.5, 12.0, 8.0), (3.0, 0.0, 0.0, 0.0)]


def test_features_match_oracle_jump_flooding():
    L = lib()
    lat = pack_lattice((16, 16, 24), SPHERES)
    N = lat.N
    dF = Dev(N * 2 * 2)
    ell = L.slb_warm_features(lat.h, dF.p, 2)
    got = dF.get(np.uint16, lat.shape + (2,))
    dF.free()
    solid = lat.solid().reshape(lat.shape)
    f, ell_ref = onn.features(solid)
    lat.free()
    assert abs(ell - ell_ref) < 1e-12 * ell_ref
    assert np.array_equal(got, onn.bf16_bits(f))


def test_targets_are_the_moments_of_the_state():
    """Targets against vgrab's velocity (reference-pinned) and the moments
    of the populations; scaled by the documented group factors."""
    L = lib()
    lat = pack_lattice((16, 16, 24), SPHERES)
    lat.step(150)
    N, F = lat.N, 1e-5
    eta = 72.0 * 0.25 / 3.0
    dT = Dev(N * 16 * 2)
    ell = 2.5
    L.slb_warm_targets(lat.h, dT.p, 16, ell)
    t = onn.from_bf16_bits(dT.get(np.uint16, (N, 16)))
    dT.free()
    L.vzero(lat.h)
    L.vgrab(lat.h, 1.0)
    u = np.stack([lat.field(f) for f in ("ux", "uy", "uz")], -1)
    a = lat.pops()
    solid = lat.solid()
    lat.free()
    cv = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                   [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1],
                   [1, -1, -1], [-1, -1, -1]], float)
    d = a.sum(1)
    Pi = np.einsum("nl,la,lb->nab", a, cv, cv)
    s6 = np.stack([Pi[:, 0, 0] - d / 3, Pi[:, 1, 1] - d / 3, Pi[:, 2, 2] - d / 3, Pi[:, 0, 1],
                   Pi[:, 0, 2], Pi[:, 1, 2]], -1)
    ref = np.concatenate([u * eta / (F * ell ** 2), d[:, None] / (8 * F * ell), s6 / (F * ell)], -1)
    ref[solid] = 0.0
    assert not t[:, 10:].any()
    ok = np.abs(t[:, :10] - ref) <= np.abs(ref) * 2.0 ** -7 + 1e-6 * np.abs(ref).max(0)
    assert ok.mean() > 0.999 and not t[solid].any()


def test_equilibrium_injection_matches_lbloadvel():
    """SLB_WARM_EQUILIBRIUM builds exactly the state lbloadvel builds from
    the predicted velocity (file written with rho0 and u - b/(2 rho0))."""
    L = lib()
    F = 1e-5
    net = L.slb_warm_new(3)
    try:
        L.slb_warm_set_mode(net, 0)
        A = pack_lattice((16, 16, 24), SPHERES, F)
        m = np.empty((A.N, 10), np.float32)
        L.slb_warm_predict(net, A.h, m.ctypes.data)
        L.slb_warm_start(net, A.h)
        pa = A.pops()
        solid = A.solid()
        A.free()
    finally:
        L.slb_warm_free(net)
    assert np.abs(m[~solid, :3]).max() > 0 and not m[solid].any()
    B = pack_lattice((16, 16, 24), SPHERES, F)
    v = np.zeros((B.N, 4))
    v[:, 0] = 72.0
    v[~solid, 1:] = m[~solid, :3].astype(np.float64)
    v[~solid, 1] -= F / (2 * 72.0)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "vel.bin")
        v.tofile(p)
        L.lbloadvel(B.h, p.encode(), b"binary")
    pb = B.pops()
    B.free()
    e = rel_l2(pa, pb)
    print(f"inject vs lbloadvel rel L2 {e:.2e}")
    assert e < 1e-6


def test_moment_injection_round_trip():
    """SLB_WARM_MOMENTS injection of a state's own FP32 moments reproduces
    those moments, and a converged state rebuilt from its moments (ghost
    moments dropped) is still converged: the residual over the next 100
    steps stays below 1e-6."""
    L = lib()
    F = 1e-5
    A = pack_lattice((16, 16, 24), SPHERES, F)
    steps = L.slb_run_to_residual(A.h, 1e-10, 100, 30000, None)
    print(f"converged in {steps} steps", flush=True)
    assert steps < 30000
    N = A.N
    dM = Dev(N * 12 * 4)
    ell = 1.7
    L.slb_warm_moments(A.h, dM.p, 12, ell)
    m = dM.get(np.float32, (N, 12))
    assert not m[:, 10:].any()
    L.measure(A.h)
    vm = A.s.vmeanx
    # fresh lattice, same geometry: inject, compare moments, run on
    B = pack_lattice((16, 16, 24), SPHERES, F)
    L.slb_warm_inject(B.h, dM.p, 12, 1, ell)
    dM2 = Dev(N * 12 * 4)
    L.slb_warm_moments(B.h, dM2.p, 12, ell)
    m2 = dM2.get(np.float32, (N, 12))
    for d in (dM, dM2):
        d.free()
    e = np.abs(m2 - m).max(0)[:10] / np.abs(m).max(0)[:10]
    print(f"moment round trip max rel err per channel {e}")
    assert (e < 1e-5).all()
    L.slb_residual_reset(B.h)
    L.slb_residual(B.h)
    L.slb_run(B.h, 100)
    r = L.slb_residual(B.h)
    L.measure(B.h)
    dv = abs(B.s.vmeanx - vm) / abs(vm)
    print(f"residual after rebuild {r:.2e}, vmean rel change {dv:.2e}")
    A.free()
    B.free()
    assert r < 1e-6 and dv < 1e-6


def test_start_policy_weights_file_and_cold_start():
    """The start policy compares the median warm and cold residual curves at
    the tolerance; it survives a save/load round trip; set_fallback clears
    it; and a solve the policy starts cold takes the steps of a cold run."""
    L = lib()
    net = L.slb_warm_new(5)
    D = C.c_double
    try:
        assert L.slb_warm_policy(net, 1e-6, 100) == 1  # no curves: warm
        # cold: res_k = 10^-(k+1)/2 -> below 1e-4 after 900 steps, 1e-6 after 1300
        cold = [10.0 ** (-(k + 1) / 2.0) for k in range(30)]
        # warm: res_k = 1e-3 0.6^k -> below 1e-4 after 600 steps, 1e-6 after 1500
        warm = [1e-3 * 0.6 ** k for k in range(30)]
        for shift in (1.0, 0.5, 2.0):  # three runs: the median is the middle one
            L.slb_warm_note_cold(net, (D * 30)(*[x * shift for x in cold]), 30, 100)
            L.slb_warm_note_warm(net, (D * 30)(*[x * shift for x in warm]), 30, 100)
        assert L.slb_warm_policy(net, 1e-4, 100) == 1
        assert L.slb_warm_policy(net, 1e-6, 100) == 0
        assert L.slb_warm_policy(net, 1e-6, 50) == 1  # curves of another interval
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "net.bin").encode()
            assert L.slb_warm_save(net, path) == 0
            net2 = L.slb_warm_load(path)
            assert net2
            try:
                assert L.slb_warm_policy(net2, 1e-4, 100) == 1
                assert L.slb_warm_policy(net2, 1e-6, 100) == 0
            finally:
                L.slb_warm_free(net2)
        # a solve the policy starts cold: same steps and state as a cold run
        runs = []
        for use_net in (True, False):
            lat = pack_lattice((16, 12, 20), SPHERES, prec=64)
            if use_net:
                lat.L.slb_run(lat.h, 7)  # a state that the cold start must discard
                r, used = D(0.0), C.c_int(-1)
                steps = L.slb_warm_solve(net, lat.h, 1e-6, 100, 20000, C.byref(r), C.byref(used))
                assert used.value == 0
            else:
                r = D(0.0)
                steps = lat.L.slb_run_to_residual(lat.h, 1e-6, 100, 20000, C.byref(r))
            runs.append((steps, r.value, lat.pops().copy()))
            lat.free()
        assert runs[0][0] == runs[1][0]
        assert runs[0][1] == runs[1][1]
        assert np.array_equal(runs[0][2], runs[1][2])
        L.slb_warm_set_fallback(net, 0, 0.0)
        assert L.slb_warm_policy(net, 1e-6, 100) == 1
    finally:
        L.slb_warm_free(net)


@pytest.mark.parametrize("shape,cin,cout,transposed", [
    ((4, 8, 64), 16, 16, 0), ((3, 4, 128), 16, 32, 1), ((4, 8, 64), 16, 48, 0),
    ((2, 4, 64), 16, 64, 1), ((4, 8, 32), 32, 32, 0), ((3, 4, 64), 32, 64, 1),
    ((4, 4, 16), 64, 64, 0), ((2, 8, 32), 64, 16, 1),
    ((2, 4, 128), 16, 32, 1), ((6, 4, 64), 16, 16, 1), ((8, 8, 64), 16, 24, 0),
    ((2, 4, 32), 32, 16, 0), ((2, 4, 32), 32, 64, 1)])
def test_conv_lds_kernel_is_bit_identical_and_matches_oracle(shape, cin, cout, transposed):
    """16, 32 or 64 input channels on rows of a multiple of 1024/cin nodes take
    the LDS-staged kernel (two planes per block for an even z extent and 16
    inputs with at most 32 outputs, 32 -> 32, or 64 inputs; one plane
    otherwise); SLBN_CONV_LDS=0
    forces the general one: outputs
    bit-identical (same K order and MFMA sequence), and within FP32 rounding
    of the oracle."""
    L = lib()
    rng = np.random.default_rng(77 + cin + cout + transposed)
    N = int(np.prod(shape))
    xb, x = bf_random(rng, shape + (cin,))
    W, w = weights(rng, cout, cin)
    bias = (rng.standard_normal(cout) * 0.1).astype(np.float32)
    dX, dW, dB = Dev(host=xb), Dev(host=W), Dev(host=bias)
    outs = []
    for env in ("1", "0"):
        os.environ["SLBN_CONV_LDS"] = env
        try:
            dY, dYf = Dev(N * cout * 2), Dev(N * cout * 4)
            L.slbn_conv(dX.p, cin, dW.p, dB.p, dY.p, dYf.p, cout, 1, transposed, *shape)
            L.slbn_sync()
            outs.append((dY.get(np.uint16, shape + (cout,)), dYf.get(np.float32, shape + (cout,))))
            dY.free()
            dYf.free()
        finally:
            os.environ.pop("SLBN_CONV_LDS", None)
    for d in (dX, dW, dB):
        d.free()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))
    if not transposed:
        ref = np.maximum(onn.conv3d(x, w, bias.astype(np.float64)), 0.0)
        err = np.abs(outs[0][1].astype(np.float64) - ref).max() / np.abs(ref).max()
        assert err < 1e-5


@pytest.mark.parametrize("passes", [1, 3])
def test_projection_matches_oracle(passes):
    """slb_warm_project (CG on the fluid-graph Laplacian, FP64, then the
    node correction; FP32 tensor) against the oracle's exact solve of each
    pass, on a pack with a smooth divergent field plus noise; the channels
    beyond 0..2 and the solid nodes are left untouched."""
    import sunlightlb_amd as s
    L = lib()
    dims = (12, 14, 16)
    lat = s.Lattice(*dims, precision=32)
    for sp in ((3.5, 5.0, 6.0, 7.0), (3.0, 10.0, 2.0, 12.0), (2.5, 1.0, 11.0, 3.0)):
        L.insertsph(lat.h, *sp)
    L.syncobjs(lat.h)
    solid = lat.solid().reshape(dims)
    rng = np.random.default_rng(8)
    i, j, k = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    u = np.stack([np.cos(2 * np.pi * (i / 12 + 2 * k / 16)), np.sin(2 * np.pi * j / 14),
                  0.5 * np.cos(2 * np.pi * (i / 12 - j / 14))], -1)
    u += 0.1 * rng.standard_normal(u.shape)
    u[solid] = 0.0
    C4 = 4
    m = np.zeros(dims + (C4,), np.float32)
    m[..., :3] = u
    m[..., 3] = 7.0
    u = m[..., :3].astype(np.float64)          # the FP32 values the GPU starts from
    d = Dev(host=m)
    res = L.slb_warm_project(lat.h, d.p, C4, 600, passes)
    got = d.get(np.float32, dims + (C4,))
    d.free()
    lat.free()
    ref, _ = onn.project(u, solid, passes=passes)
    print(f"projection passes {passes}: CG residual {res:.2e}, "
          f"rel L2 vs oracle {np.linalg.norm(got[..., :3] - ref) / np.linalg.norm(ref):.2e}")
    assert res < 1e-8
    assert np.linalg.norm(got[..., :3] - ref) < 1e-5 * np.linalg.norm(ref)
    assert (got[..., 3] == 7.0).all()
    assert (got[solid][:, :3] == 0.0).all()
    d0 = np.linalg.norm(onn.face_divergence(u, solid))
    assert np.linalg.norm(onn.face_divergence(got[..., :3].astype(np.float64), solid)) < 0.3 * d0
