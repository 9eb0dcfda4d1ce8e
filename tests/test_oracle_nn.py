"""The warm-start network oracle (oracle/nn.py) pinned on the CPU against
torch's own conv3d / avg_pool3d / nearest upsampling / autograd / Adam as
library routines, and its jump-flooding distance against a brute-force
periodic distance transform."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import torch.nn.functional as F

from oracle import nn as onn

torch.set_default_dtype(torch.float64)


def to_t(x):
    """channels-last numpy [D,H,W,C] -> torch [1,C,D,H,W]"""
    return torch.tensor(np.ascontiguousarray(x)).permute(3, 0, 1, 2).unsqueeze(0)


def from_t(t):
    return t.squeeze(0).permute(1, 2, 3, 0).detach().numpy()


def w_t(w):
    cout, _, cin = w.shape
    return torch.tensor(w.reshape(cout, 3, 3, 3, cin)).permute(0, 4, 1, 2, 3).contiguous()


def tconv(xt, wt, b=None):
    return F.conv3d(F.pad(xt, (1, 1, 1, 1, 1, 1), mode="circular"), wt, b)


@pytest.mark.parametrize("shape,cin,cout", [((4, 5, 6), 3, 2), ((8, 8, 8), 16, 32),
                                            ((2, 6, 4), 8, 16), ((1, 4, 7), 4, 4)])
def test_conv3d_matches_torch_circular(shape, cin, cout):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape + (cin,))
    w = rng.standard_normal((cout, 27, cin))
    b = rng.standard_normal(cout)
    y = onn.conv3d(x, w, b)
    yt = from_t(tconv(to_t(x), w_t(w), torch.tensor(b)))
    assert np.abs(y - yt).max() < 1e-10 * np.abs(yt).max()


@pytest.mark.parametrize("shape,cin,cout", [((4, 5, 6), 3, 2), ((8, 8, 8), 16, 16)])
def test_conv3d_backward_matches_autograd(shape, cin, cout):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape + (cin,))
    w = rng.standard_normal((cout, 27, cin))
    dy = rng.standard_normal(shape + (cout,))
    xt = to_t(x).requires_grad_(True)
    wt = w_t(w).requires_grad_(True)
    tconv(xt, wt).backward(to_t(dy))
    dx = onn.conv3d_bwd_data(dy, w)
    dw = onn.conv3d_bwd_weight(x, dy)
    assert np.abs(dx - from_t(xt.grad)).max() < 1e-10 * np.abs(dx).max()
    dwt = wt.grad.permute(0, 2, 3, 4, 1).reshape(cout, 27, cin).numpy()
    assert np.abs(dw - dwt).max() < 1e-10 * np.abs(dw).max()


def test_pool_and_upsample_match_torch():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 6, 8, 5))
    assert np.allclose(onn.pool(x), from_t(F.avg_pool3d(to_t(x), 2)), rtol=0, atol=1e-13)
    c = rng.standard_normal((2, 3, 4, 5))
    assert np.array_equal(onn.up(c), from_t(F.interpolate(to_t(c), scale_factor=2, mode="nearest")))
    # adjoints: <pool x, y> = <x, pool_bwd y>, <up c, z> = <c, up_bwd z>
    y = rng.standard_normal((2, 3, 4, 5))
    assert abs((onn.pool(x) * y).sum() - (x * onn.pool_bwd(y)).sum()) < 1e-12
    z = rng.standard_normal((4, 6, 8, 5))
    assert abs((onn.up(c) * z).sum() - (c * onn.up_bwd(z)).sum()) < 1e-12


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.float32([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.4e38])])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(onn.bf16(x), ref)
    assert np.array_equal(onn.from_bf16_bits(onn.bf16_bits(x)), ref)


class TorchNet(torch.nn.Module):
    """The network of slb_warm.c written with torch layers (circular
    padding), used only to pin the numpy oracle's forward and backward."""

    def __init__(self, params):
        super().__init__()
        self.w = torch.nn.ParameterList([torch.nn.Parameter(w_t(w)) for w, _ in params])
        self.b = torch.nn.ParameterList([torch.nn.Parameter(torch.tensor(b)) for _, b in params])

    def c(self, L, x):
        y = tconv(x, self.w[L], self.b[L])
        return y if L == onn.OUT else torch.relu(y)

    def forward(self, x):
        up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
        h = self.c(onn.E1B, self.c(onn.E1A, x))
        h2 = self.c(onn.E2B, self.c(onn.E2A, F.avg_pool3d(h, 2)))
        h3 = self.c(onn.E3B, self.c(onn.E3A, F.avg_pool3d(h2, 2)))
        h4 = self.c(onn.E4B, self.c(onn.E4A, F.avg_pool3d(h3, 2)))
        u2 = up(self.c(onn.D3, up(h4) + h3)) + h2
        u1 = up(self.c(onn.D2, u2)) + h
        return self.c(onn.OUT, self.c(onn.D1, u1))


def random_params(rng, scale=1.0):
    ps = []
    for L in range(12):
        cin, cout = onn.LCIN[L], onn.LCOUT[L]
        ps.append((rng.standard_normal((cout, 27, cin)) * scale * np.sqrt(2.0 / (27 * cin)),
                   rng.standard_normal(cout) * 0.1))
    return ps


def test_network_forward_backward_match_autograd():
    rng = np.random.default_rng(5)
    params = random_params(rng)
    x = np.zeros((16, 16, 24, onn.NCH))
    x[..., :2] = rng.standard_normal((16, 16, 24, 2))
    t = rng.standard_normal((16, 16, 24, onn.NTARG)) * (np.array(onn.GROUP) + 1.0)
    wc = onn.target_weights(t)
    out, cache = onn.forward(params, x)
    loss, dout = onn.mse_grad(out, t, wc)
    grads = onn.backward(cache, dout)

    net = TorchNet(params)
    yt = net(to_t(x))
    lt = (((yt[:, :onn.NTARG] - to_t(t)) ** 2) * torch.tensor(wc).view(1, -1, 1, 1, 1)).sum()
    lt.backward()
    assert np.abs(out - from_t(yt)).max() < 1e-9 * np.abs(out).max()
    assert abs(loss - lt.item()) < 1e-10 * abs(lt.item())
    for L in range(12):
        dw, db = grads[L]
        dwt = net.w[L].grad.permute(0, 2, 3, 4, 1).reshape(dw.shape).numpy()
        assert np.abs(dw - dwt).max() <= 1e-9 * np.abs(dwt).max(), L
        assert np.abs(db - net.b[L].grad.numpy()).max() <= 1e-9 * np.abs(db).max() + 1e-300, L


def test_target_weights_equalise_groups():
    rng = np.random.default_rng(8)
    t = rng.standard_normal((4, 4, 4, onn.NTARG)) * np.array([1, 2, 3, 10, .1, .2, .3, .4, .5, .6])
    wc = onn.target_weights(t)
    # a prediction of zero has relative error 1 in every group: loss 1
    loss, _ = onn.mse_grad(np.zeros((4, 4, 4, onn.NCH)), t, wc)
    assert abs(loss - 1.0) < 1e-12
    # scaling one group's targets and prediction together leaves the loss
    p = t + 0.1 * rng.standard_normal(t.shape)
    pad = lambda a: np.concatenate([a, np.zeros(a.shape[:3] + (onn.NCH - onn.NTARG,))], -1)
    l1, _ = onn.mse_grad(pad(p), t, wc)
    k = np.ones(onn.NTARG); k[3] = 7.0
    l2, _ = onn.mse_grad(pad(p * k), t * k, onn.target_weights(t * k))
    assert abs(l1 - l2) < 1e-12 * l1


def test_parameter_arena_round_trip():
    rng = np.random.default_rng(6)
    params = random_params(rng)
    v = onn.pack(params)
    back = onn.unpack(v)
    for (w, b), (w2, b2) in zip(params, back):
        assert np.array_equal(w, w2) and np.array_equal(b, b2)
    wofs, bofs, n = onn.layout()
    assert n == v.size and all(o % 64 == 0 for o in wofs + bofs)


def test_adam_matches_torch():
    rng = np.random.default_rng(7)
    w0 = rng.standard_normal(50)
    p = torch.nn.Parameter(torch.tensor(w0))
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.standard_normal(50)
        p.grad = torch.tensor(g)
        opt.step()
        w, m, v = onn.adam(w, g, m, v, 1e-2, 0.9, 0.999, 1e-8, step)
    assert np.abs(w - p.detach().numpy()).max() < 1e-12


@pytest.mark.parametrize("shape,frac,seed", [((8, 8, 8), 0.05, 1), ((6, 10, 12), 0.2, 2),
                                             ((16, 16, 16), 0.01, 3), ((4, 4, 4), 0.0, 4)])
def test_jump_flooding_distance(shape, frac, seed):
    rng = np.random.default_rng(seed)
    seeds = rng.random(shape) < frac
    near = onn.jfa(seeds)
    if not seeds.any():
        assert (near == -1).all()
        return
    got = np.sqrt(onn._pdist2(shape, np.arange(seeds.size), near)).reshape(shape)
    exact = onn.exact_distance(seeds)
    # every answer is a seed, never closer than the exact distance, and
    # jump flooding with the extra stride-1 pass is exact on (almost) all
    assert seeds.reshape(-1)[near].all()
    assert (got >= exact - 1e-12).all()
    assert (np.abs(got - exact) < 1e-12).mean() > 0.995
    assert (got - exact).max() < 1.0


def test_features_signs_and_scale():
    solid = np.zeros((8, 8, 8), bool)
    solid[3:5, 3:5, 3:5] = True
    f, ell = onn.features(solid)
    assert np.array_equal(f[..., 0] == 1.0, solid)
    assert (f[..., 1][solid] < 0).all() and (f[..., 1][~solid] > 0).all()
    # ell is the mean fluid distance: brute force over the fluid nodes
    d = onn.exact_distance(solid)
    assert abs(ell - d[~solid].mean()) < 1e-6 * ell
    # node (0, 4, 4): nearest solid is (3, 4, 4), distance 3
    assert f[0, 4, 4, 1] == np.float32(3.0) * np.float32(1.0 / ell)
    # solid node (3, 3, 3): nearest fluid is a face neighbour
    assert f[3, 3, 3, 1] == -np.float32(1.0 / ell)


# ---------------------------------------------------------------- projection

def _pack(shape, spheres):
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    solid = np.zeros(shape, bool)
    for r, x, y, z in spheres:
        solid |= (i - x) ** 2 + (j - y) ** 2 + (k - z) ** 2 < r * r
    return solid


def _smooth_field(shape, seed, modes=4):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    u = np.zeros(shape + (3,))
    for _ in range(modes):
        kv = rng.integers(0, 3, 3)
        ph = rng.uniform(0, 2 * np.pi)
        arg = 2 * np.pi * (kv[0] * i / shape[0] + kv[1] * j / shape[1] + kv[2] * k / shape[2]) + ph
        u += rng.normal(size=3) * np.cos(arg)[..., None]
    return u


def test_projection_oracle_laplacian_and_solution():
    """The fluid-graph Laplacian is symmetric, annihilates constants and is
    positive semi-definite; the solved potential satisfies A phi = -div F to
    roundoff; a field without face divergence (uniform flow, shear flow in
    an all-fluid periodic box) is left unchanged."""
    solid = _pack((9, 10, 11), [(2.5, 4.0, 5.0, 5.0), (2.0, 1.0, 1.0, 9.0)])
    A = onn.fluid_laplacian(solid)
    assert abs(A - A.T).max() == 0
    assert np.abs(A @ np.ones(solid.size)).max() == 0
    ev = np.linalg.eigvalsh(A.toarray())
    assert ev.min() > -1e-10
    u = _smooth_field(solid.shape, 1)
    u[solid] = 0.0
    out, phi = onn.project(u, solid)
    rhs = -onn.face_divergence(u, solid).reshape(-1)
    assert np.abs(A @ phi.reshape(-1) - rhs).max() < 1e-10 * np.abs(rhs).max()
    free = np.zeros((8, 6, 7), bool)
    uni = np.zeros((8, 6, 7, 3))
    uni[..., 0], uni[..., 1], uni[..., 2] = 0.3, -0.1, 0.05
    assert np.abs(onn.face_divergence(uni, free)).max() < 1e-15
    assert np.allclose(onn.project(uni, free)[0], uni, atol=1e-14)
    shear = np.zeros((8, 6, 7, 3))
    shear[..., 0] = np.sin(2 * np.pi * np.arange(6) / 6)[None, :, None]
    assert np.allclose(onn.project(shear, free)[0], shear, atol=1e-13)


def _coarse(d, b=4):
    s = d.shape
    return d[:s[0] // b * b, :s[1] // b * b, :s[2] // b * b].reshape(
        s[0] // b, b, s[1] // b, b, s[2] // b, b).sum((1, 3, 5))


def test_projection_removes_the_divergence_of_smooth_fields():
    """On a sphere pack, one pass cuts a strongly divergent smooth field's
    face divergence several-fold, three passes its 4^3-block sums (mass
    imbalance between regions) more than twenty-fold; the same correction
    with the sign flipped (a plausible slip) makes it larger."""
    solid = _pack((12, 14, 16), [(3.5, 5.0, 6.0, 7.0), (3.0, 10.0, 2.0, 12.0), (2.5, 1.0, 11.0, 3.0)])
    u = _smooth_field(solid.shape, 2)
    u[solid] = 0.0
    div0 = onn.face_divergence(u, solid)
    out, phi = onn.project(u, solid)
    assert np.linalg.norm(onn.face_divergence(out, solid)) < 0.3 * np.linalg.norm(div0)
    wrong = 2 * u - out                 # the correction added instead of subtracted
    assert np.linalg.norm(onn.face_divergence(wrong, solid)) > np.linalg.norm(div0)
    out3, _ = onn.project(u, solid, passes=3)
    c0 = np.linalg.norm(_coarse(div0))
    c3 = np.linalg.norm(_coarse(onn.face_divergence(out3, solid)))
    assert c3 < 0.05 * c0, (c0, c3)
