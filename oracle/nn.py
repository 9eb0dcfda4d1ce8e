"""CPU oracle of the learned warm start's network and lattice features.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nowhere in the product
path) as the checker of the gfx950 network kernels.  The reference has no
learned component; these functions restate the network the library defines
(sunlightlb_amd/csrc/slb_warm.c) in numpy, and are themselves pinned in the
CPU tests against torch's conv3d / avg_pool3d / autograd on the CPU, and the
jump-flooding distance against a brute-force periodic distance transform.

Tensors are channels-last, x[i, j, k, c] (k fastest, matching the lattice
node index esk*(esj*i+j)+k).  Convolutions are 3x3x3 with periodic padding;
tap t = 9*(di+1) + 3*(dj+1) + (dk+1) reads x[i+di, j+dj, k+dk].  Weights
w[co, t, ci].  With bf16=True every tensor the GPU stores in BF16 is rounded
to BF16 (round to nearest even) at the same point, and sums run in float64.
"""
import numpy as np

# layer table of the network (slb_warm.c)
E1A, E1B, E2A, E2B, E3A, E3B, E4A, E4B, D3, D2, D1, OUT = range(12)
LCIN = (16, 16, 16, 32, 32, 64, 64, 64, 64, 32, 16, 16)
LCOUT = (16, 16, 32, 32, 64, 64, 64, 64, 32, 16, 16, 16)
LLEVEL = (0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0, 0)
NCH = 16    # input/output channels
NFEAT = 2   # used input channels
NTARG = 10  # used output channels: velocity (3), density (1), stress (6)
GROUP = (0, 0, 0, 1, 2, 2, 2, 2, 2, 2)


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    """BF16 bit patterns (uint16) of x, rounded to nearest even."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(
        np.float64)


def _shift(x, d):
    """y[i, j, k] = x[i+d0, j+d1, k+d2] (periodic)."""
    return np.roll(x, shift=(-d[0], -d[1], -d[2]), axis=(0, 1, 2))


def conv3d(x, w, b=None):
    """y[n, co] = b[co] + sum_{t, ci} w[co, t, ci] x[n + off(t), ci]."""
    D, H, W, cin = x.shape
    y = np.zeros((D, H, W, w.shape[0]))
    for t, d in enumerate(taps()):
        y += _shift(x, d) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3d_bwd_data(dy, w):
    """dx[n, ci] = sum_{t, co} dy[n - off(t), co] w[co, t, ci]."""
    dx = np.zeros(dy.shape[:3] + (w.shape[2],))
    for t, d in enumerate(taps()):
        dx += _shift(dy, (-d[0], -d[1], -d[2])) @ w[:, t, :]
    return dx


def conv3d_bwd_weight(x, dy):
    """dw[co, t, ci] = sum_n dy[n, co] x[n + off(t), ci]."""
    cin, cout = x.shape[3], dy.shape[3]
    dyf = dy.reshape(-1, cout)
    dw = np.zeros((cout, 27, cin))
    for t, d in enumerate(taps()):
        dw[:, t, :] = dyf.T @ _shift(x, d).reshape(-1, cin)
    return dw


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def pool_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_bwd(dy):
    D, H, W, C = dy.shape
    return dy.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def kpad(cin):
    return ((27 * cin + 31) // 32) * 32


def layout():
    """Offsets of W[co][Kp] and the 64 bias slots of each layer in the
    parameter arena (slb_warm_layout)."""
    wofs, bofs, o = [], [], 0
    for L in range(12):
        wofs.append(o)
        o += LCOUT[L] * kpad(LCIN[L])
        bofs.append(o)
        o += 64
    return wofs, bofs, o


def unpack(arena):
    """Per layer (w[co, t, ci], b[co]) from the parameter arena."""
    wofs, bofs, n = layout()
    assert arena.size == n
    out = []
    for L in range(12):
        cin, cout, kp = LCIN[L], LCOUT[L], kpad(LCIN[L])
        W = arena[wofs[L]:wofs[L] + cout * kp].reshape(cout, kp)[:, :27 * cin]
        out.append((W.reshape(cout, 27, cin).astype(np.float64),
                    arena[bofs[L]:bofs[L] + cout].astype(np.float64)))
    return out


def pack(grads, n=None):
    """Parameter-arena vector from per-layer (dw, db)."""
    wofs, bofs, total = layout()
    v = np.zeros(total)
    for L, (dw, db) in enumerate(grads):
        cin, cout, kp = LCIN[L], LCOUT[L], kpad(LCIN[L])
        M = np.zeros((cout, kp))
        M[:, :27 * cin] = dw.reshape(cout, 27 * cin)
        v[wofs[L]:wofs[L] + cout * kp] = M.reshape(-1)
        v[bofs[L]:bofs[L] + cout] = db
    return v


def _r(x, on):
    return bf16(x) if on else x


def forward(params, x, bf=False):
    """Network forward; returns (output float64 [D,H,W,16], cache)."""
    wb = [(_r(w, bf), b) for (w, b) in params]
    a = {}
    inp = {}
    x = _r(x, bf)

    def conv(L, xin):
        inp[L] = xin
        y = conv3d(xin, *wb[L])
        if L != OUT:
            y = _r(np.maximum(y, 0.0), bf)
        a[L] = y
        return y

    h = conv(E1B, conv(E1A, x))
    p1 = _r(pool(h), bf)
    h2 = conv(E2B, conv(E2A, p1))
    p2 = _r(pool(h2), bf)
    h3 = conv(E3B, conv(E3A, p2))
    p3 = _r(pool(h3), bf)
    h4 = conv(E4B, conv(E4A, p3))
    u3 = _r(up(h4) + h3, bf)
    u2 = _r(up(conv(D3, u3)) + h2, bf)
    u1 = _r(up(conv(D2, u2)) + h, bf)
    out = conv(OUT, conv(D1, u1))
    return out, {"a": a, "in": inp, "w": wb, "bf": bf}


def mse_grad(out, t, wc, bf=False):
    """Weighted squared error sum_n sum_c wc[c] (out - t)^2 over the first
    len(wc) channels and its gradient 2 wc[c] (out - t) (zero beyond)."""
    wc = np.asarray(wc, dtype=np.float64)
    k = wc.size
    d = out[..., :k] - t[..., :k]
    g = np.zeros_like(out)
    g[..., :k] = 2.0 * d * wc
    return float((wc * d * d).sum()), _r(g, bf)


def target_weights(t):
    """Loss weights making each target group's relative mean-square error
    count equally: wc[c] = 1 / (3 nodes |group| ms[group])."""
    t = t.reshape(-1, NTARG)
    n = t.shape[0]
    ms = np.zeros(3)
    cnt = np.zeros(3)
    for c in range(NTARG):
        ms[GROUP[c]] += (t[:, c] ** 2).sum()
        cnt[GROUP[c]] += 1
    ms /= n * cnt
    return np.array([1.0 / (3.0 * n * cnt[GROUP[c]] * (ms[GROUP[c]] or 1.0)) for c in range(NTARG)])


def backward(cache, dout):
    """Parameter gradients [(dw, db)] of the network for output gradient dout."""
    a, inp, wb, bf = cache["a"], cache["in"], cache["w"], cache["bf"]
    grads = [None] * 12

    def cb(L, gy, need_dx=True):
        dp = gy if L == OUT else _r(gy * (a[L] > 0), bf)
        grads[L] = (conv3d_bwd_weight(inp[L], dp), dp.sum(axis=(0, 1, 2)))
        return _r(conv3d_bwd_data(dp, wb[L][0]), bf) if need_dx else None

    g_u1 = cb(D1, cb(OUT, dout))
    g_u2 = cb(D2, _r(up_bwd(g_u1), bf))
    g_u3 = cb(D3, _r(up_bwd(g_u2), bf))
    g = cb(E4A, cb(E4B, _r(up_bwd(g_u3), bf)))
    g = _r(_r(pool_bwd(g), bf) + g_u3, bf)
    g = cb(E3A, cb(E3B, g))
    g = _r(_r(pool_bwd(g), bf) + g_u2, bf)
    g = cb(E2A, cb(E2B, g))
    g = _r(_r(pool_bwd(g), bf) + g_u1, bf)
    cb(E1A, cb(E1B, g), need_dx=False)
    return grads


def adam(w, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return w, m, v


# ---------------------------------------------------------------------- #
# Lattice features: periodic jump flooding (the GPU algorithm, same
# candidate order and tie rule) and the exact distance it approximates.

def _pdist2(shape, a, b):
    ai = np.stack(np.unravel_index(a, shape), -1)
    bi = np.stack(np.unravel_index(b, shape), -1)
    d = np.abs(ai - bi)
    d = np.minimum(d, np.array(shape) - d)
    return (d * d).sum(-1)


def jfa(seed_mask):
    """Nearest-seed node index of every node (periodic, squared Euclidean
    distance, -1 if there is no seed): passes with strides s0, s0/2, .., 1
    and one more stride-1 pass; each pass keeps the node's current seed
    unless one of its 27 candidates (dz, dy, dx) in {-1,0,1}^3 order at
    offsets s*(dz,dy,dx) is strictly closer."""
    shape = seed_mask.shape
    N = seed_mask.size
    idx = np.arange(N)
    cur = np.where(seed_mask.reshape(-1), idx, -1)
    maxd = max(shape)
    s0 = 1
    while 2 * s0 < maxd:
        s0 *= 2
    strides = []
    s = s0
    while s >= 1:
        strides.append(s)
        s //= 2
    strides.append(1)
    I, J, K = np.unravel_index(idx, shape)
    for s in strides:
        best = cur.copy()
        bd = np.where(best >= 0, _pdist2(shape, idx, np.maximum(best, 0)), np.iinfo(np.int64).max)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    nb = np.ravel_multi_index(((I + dz * s) % shape[0], (J + dy * s) % shape[1],
                                               (K + dx * s) % shape[2]), shape)
                    c = cur[nb]
                    ok = c >= 0
                    d = np.where(ok, _pdist2(shape, idx, np.maximum(c, 0)), np.iinfo(np.int64).max)
                    better = ok & (d < bd)
                    best = np.where(better, c, best)
                    bd = np.where(better, d, bd)
        cur = best
    return cur


def features(solid):
    """(solid flag, signed distance / ell) per node as float32 before BF16
    rounding, and ell: + distance to the nearest solid node on fluid nodes,
    - the distance to the nearest fluid node on solid nodes (64 without
    any); ell = the mean over fluid nodes of the distance to the solid
    (float32 distances summed in float64; 1 without fluid or solid)."""
    shape = solid.shape
    idx = np.arange(solid.size)
    ns = jfa(solid)
    nf = jfa(~solid)
    s = solid.reshape(-1)
    dist_s = np.where(ns >= 0, np.sqrt(_pdist2(shape, idx, np.maximum(ns, 0)).astype(np.float32)),
                      np.float32(64.0))
    dist_f = np.where(nf >= 0, np.sqrt(_pdist2(shape, idx, np.maximum(nf, 0)).astype(np.float32)),
                      np.float32(64.0))
    fl = ~s
    tot = dist_s[fl].astype(np.float64).sum()
    ell = tot / fl.sum() if fl.any() and tot > 0 else 1.0
    scale = np.float32(1.0 / ell)
    sd = np.where(s, -dist_f, dist_s).astype(np.float32) * scale
    return np.stack([s.astype(np.float32), sd.astype(np.float32)], -1).reshape(shape + (2,)), ell


def exact_distance(seed_mask):
    """Brute-force periodic Euclidean distance to the nearest seed node."""
    shape = seed_mask.shape
    seeds = np.flatnonzero(seed_mask.reshape(-1))
    idx = np.arange(seed_mask.size)
    best = np.full(seed_mask.size, np.inf)
    for s in seeds:
        best = np.minimum(best, _pdist2(shape, idx, np.full_like(idx, s)))
    return np.sqrt(best).reshape(shape)


# --------------------------------------------------------------------------
# Divergence projection of a predicted velocity field (slb_warm_project).

def face_divergence(u, solid):
    """div F(x) = sum_a [F(x, x+e_a) - F(x-e_a, x)] on fluid nodes (0 on
    solid), face flux F(x, x+e_a) = (u_a(x) + u_a(x+e_a))/2 when both nodes
    are fluid and 0 otherwise, periodic.  u[i, j, k, a], solid[i, j, k]."""
    fl = ~solid
    div = np.zeros(solid.shape)
    for a in range(3):
        nxt = np.roll(fl, -1, axis=a)
        ua = u[..., a]
        F = np.where(fl & nxt, 0.5 * (ua + np.roll(ua, -1, axis=a)), 0.0)   # F(x, x+e_a)
        div += F - np.roll(F, 1, axis=a)
    return np.where(fl, div, 0.0)


def fluid_laplacian(solid):
    """Sparse Laplacian of the graph of fluid nodes joined across open faces
    (A phi(x) = sum_open (phi(x) - phi(nb)), zero rows on solid nodes)."""
    import scipy.sparse as sp
    shape = solid.shape
    N = solid.size
    idx = np.arange(N).reshape(shape)
    fl = ~solid
    rows, cols, vals = [], [], []
    diag = np.zeros(N)
    for a in range(3):
        for sgn in (-1, 1):
            nb = np.roll(idx, -sgn, axis=a)
            open_ = (fl & np.roll(fl, -sgn, axis=a)).reshape(-1)
            src = idx.reshape(-1)[open_]
            dst = nb.reshape(-1)[open_]
            rows.append(src)
            cols.append(dst)
            vals.append(-np.ones(len(src)))
            np.add.at(diag, src, 1.0)
    rows.append(np.arange(N))
    cols.append(np.arange(N))
    vals.append(diag)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))


def project(u, solid, passes=1):
    """The projection slb_warm_project defines, each pass solved to machine
    precision (least squares on the singular Laplacian): phi with
    A phi = -div F, then on each fluid node and axis a, u_a -= mean over the
    open faces along a of phi's difference across the face (phi(x+e_a) -
    phi(x) forward, phi(x) - phi(x-e_a) backward); repeated `passes` times.
    Returns (u', phi of the last pass)."""
    out, phi = u, None
    for _ in range(passes):
        out, phi = _project_pass(out, solid)
    return out, phi


def _project_pass(u, solid):
    import scipy.sparse.linalg as sla
    A = fluid_laplacian(solid)
    rhs = -face_divergence(u, solid).reshape(-1)
    phi = sla.lsqr(A, rhs, atol=1e-15, btol=1e-15, iter_lim=20000)[0].reshape(solid.shape)
    fl = ~solid
    out = u.copy()
    for a in range(3):
        fwd = np.roll(fl, -1, axis=a) & fl
        bwd = np.roll(fl, 1, axis=a) & fl
        d = np.where(fwd, np.roll(phi, -1, axis=a) - phi, 0.0) + \
            np.where(bwd, phi - np.roll(phi, 1, axis=a), 0.0)
        c = fwd.astype(int) + bwd.astype(int)
        out[..., a] = np.where(fl & (c > 0), u[..., a] - d / np.maximum(c, 1), u[..., a])
    return out, phi
