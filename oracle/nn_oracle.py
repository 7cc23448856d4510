"""CPU oracle of the learned initial guess (test infrastructure only).

Restates, in numpy FP64, the encoder-decoder that liquidhairinteration_amd/csrc/k_nn.hip
runs on MFMA: forward pass, masked MSE loss and the full backward pass.  It has no
counterpart in the reference, whose CGSolver<3>::Init starts every solve from x = 0
(PressureSolver.cpp:666-680); the pressure problem the network approximates is the
one of PressureSolver.cpp:544-698.

With ``bf16=True`` the oracle rounds to BF16 (round to nearest even) at the same
points as the GPU: the input, every weight, every activation after ReLU / pooling /
upsampling, and every activation gradient; sums are FP64 (the GPU accumulates in
FP32), so the two agree to accumulation-order effects.  With ``bf16=False`` it is the
exact FP64 network, which the tests pin against torch.nn.functional (float64) and
against finite differences.

Only tests/, the smoke check and the CPU baseline of bench.py import this module.

Layout: tensors are channels-last numpy arrays [x, y, z, C] (z fastest, the Array3
order).  Parameters: per layer W[cout, 27, cin] (tap = (dx+1)*9 + (dy+1)*3 + (dz+1))
and bias[cout], flattened in the order W, bias, layer after layer (E0 E1 E2 E3 D2 D1
D0 OUT) -- the layout of lh_warmstart_get_params.
"""
import numpy as np

CIN = (2, 16, 32, 64, 160, 96, 48, 16)
COUT = (16, 32, 64, 96, 64, 32, 16, 1)
NAMES = ("E0", "E1", "E2", "E3", "D2", "D1", "D0", "OUT")
TAPS = [((t // 9) - 1, ((t // 3) % 3) - 1, (t % 3) - 1) for t in range(27)]


def num_params():
    return sum(co * 27 * ci + co for ci, co in zip(CIN, COUT))


def unflatten(theta):
    """flat parameter vector -> list of (W[cout,27,cin], b[cout]) in float64"""
    theta = np.asarray(theta, dtype=np.float64)
    out, q = [], 0
    for ci, co in zip(CIN, COUT):
        w = theta[q:q + co * 27 * ci].reshape(co, 27, ci)
        q += co * 27 * ci
        b = theta[q:q + co]
        q += co
        out.append((w, b))
    assert q == theta.size
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def bf16(a):
    """round to BF16 (nearest even), via FP32 as on the GPU; returned as float64"""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _shift(x, o):
    """s[v] = x[v + o] with zeros outside (x: [d0,d1,d2,C])"""
    d = x.shape[:3]
    s = np.zeros_like(x)
    src = tuple(slice(max(0, oa), min(da, da + oa)) for oa, da in zip(o, d))
    dst = tuple(slice(max(0, -oa), min(da, da - oa)) for oa, da in zip(o, d))
    s[dst] = x[src]
    return s


def conv(x, w, b=None):
    """out[v, co] = b[co] + sum_tap sum_ci w[co, tap, ci] x[v + off(tap), ci] (zero padding)"""
    d = x.shape[:3]
    out = np.zeros(d + (w.shape[0],))
    for t, o in enumerate(TAPS):
        out += _shift(x, o) @ w[:, t, :].T
    if b is not None:
        out += b
    return out


def conv_bwd_data(g, w):
    """gradient of conv() w.r.t. x for the output gradient g"""
    d = g.shape[:3]
    out = np.zeros(d + (w.shape[2],))
    for t, o in enumerate(TAPS):
        out += _shift(g, tuple(-oa for oa in o)) @ w[:, t, :]
    return out


def conv_bwd_weight(g, x):
    """gradients of conv() w.r.t. w and b"""
    co, ci = g.shape[3], x.shape[3]
    gw = np.zeros((co, 27, ci))
    g2 = g.reshape(-1, co)
    for t, o in enumerate(TAPS):
        gw[:, t, :] = g2.T @ _shift(x, o).reshape(-1, ci)
    return gw, g2.sum(axis=0)


def pool(x):
    d0, d1, d2, c = x.shape
    return x.reshape(d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, c).sum(axis=(1, 3, 5)) * 0.125


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def up_bwd(g):
    d0, d1, d2, c = g.shape
    return g.reshape(d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, c).sum(axis=(1, 3, 5))


def forward(theta, x_in, rnd=True):
    """x_in: [d0,d1,d2,2] -> (y [d0,d1,d2], cache)"""
    R = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    L = unflatten(theta)
    W = [(R(w), b) for w, b in L]
    relu = lambda a: np.maximum(a, 0.0)
    x0 = R(x_in)
    a0 = R(relu(conv(x0, *W[0])))
    p0 = R(pool(a0))
    a1 = R(relu(conv(p0, *W[1])))
    p1 = R(pool(a1))
    a2 = R(relu(conv(p1, *W[2])))
    p2 = R(pool(a2))
    a3 = R(relu(conv(p2, *W[3])))
    c2 = np.concatenate([up(a3), a2], axis=3)
    d2 = R(relu(conv(c2, *W[4])))
    c1 = np.concatenate([up(d2), a1], axis=3)
    d1 = R(relu(conv(c1, *W[5])))
    c0 = np.concatenate([up(d1), a0], axis=3)
    d0 = R(relu(conv(c0, *W[6])))
    y = conv(d0, *W[7])[..., 0]
    cache = dict(W=W, x0=x0, a0=a0, p0=p0, a1=a1, p1=p1, a2=a2, p2=p2, a3=a3, c2=c2, d2=d2, c1=c1, d1=d1,
                 c0=c0, d0=d0)
    return y, cache


def loss_grad(theta, x_in, target, mask, rnd=True):
    """masked MSE loss sum m (y - t)^2 / sum m and its gradient w.r.t. theta (flat)"""
    R = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    y, k = forward(theta, x_in, rnd)
    W = k["W"]
    m = np.asarray(mask, dtype=np.float64)
    n = m.sum()
    dlt = (y - target) * m
    loss = float((dlt * dlt).sum() / n) if n > 0 else 0.0
    gy = (2.0 * dlt / n if n > 0 else 0.0 * dlt)[..., None]
    pos = lambda a: (a > 0).astype(np.float64)
    grads = [None] * 8
    grads[7] = conv_bwd_weight(gy, k["d0"])
    g_d0 = R(pos(k["d0"]) * conv_bwd_data(gy, W[7][0]))
    grads[6] = conv_bwd_weight(g_d0, k["c0"])
    g_c0 = R(pos(k["c0"]) * conv_bwd_data(g_d0, W[6][0]))
    g_d1 = R(up_bwd(g_c0[..., :32]))
    grads[5] = conv_bwd_weight(g_d1, k["c1"])
    g_c1 = R(pos(k["c1"]) * conv_bwd_data(g_d1, W[5][0]))
    g_d2 = R(up_bwd(g_c1[..., :64]))
    grads[4] = conv_bwd_weight(g_d2, k["c2"])
    g_c2 = R(pos(k["c2"]) * conv_bwd_data(g_d2, W[4][0]))
    g_a3 = R(up_bwd(g_c2[..., :96]))
    grads[3] = conv_bwd_weight(g_a3, k["p2"])
    g_p2 = R(conv_bwd_data(g_a3, W[3][0]))
    g_a2 = R(g_c2[..., 96:] + pos(k["a2"]) * up(g_p2) * 0.125)
    grads[2] = conv_bwd_weight(g_a2, k["p1"])
    g_p1 = R(conv_bwd_data(g_a2, W[2][0]))
    g_a1 = R(g_c1[..., 64:] + pos(k["a1"]) * up(g_p1) * 0.125)
    grads[1] = conv_bwd_weight(g_a1, k["p0"])
    g_p0 = R(conv_bwd_data(g_a1, W[1][0]))
    g_a0 = R(g_c0[..., 32:] + pos(k["a0"]) * up(g_p0) * 0.125)
    grads[0] = conv_bwd_weight(g_a0, k["x0"])
    return loss, flatten(grads), y


def init_params(seed=0):
    """He-normal weights, zero biases (the distribution of lh_warmstart_init; not its stream)"""
    rng = np.random.default_rng(seed)
    layers = []
    for i, (ci, co) in enumerate(zip(CIN, COUT)):
        std = np.sqrt((1.0 if i == 7 else 2.0) / (27.0 * ci))
        layers.append((rng.normal(0.0, std, (co, 27, ci)), np.zeros(co)))
    return flatten(layers)
