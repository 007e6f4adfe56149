"""FP64 oracle of the learned denoiser -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ only, as the checker of the device network; the product
never imports it.  The denoiser has no counterpart in the reference: this
restates the network the library's header documents (a three-level 3-D
U-Net, 16 / 32 / 64 channels, 3^3 convolutions with zero padding, 2^3 mean
pooling, nearest 2x upsampling, concatenated skips, ReLU), in numpy FP64:

  * forward(params, x)           the correction o and every activation,
  * loss(params, x, t)           mean((x + m o - t)^2), m = [|x| > support],
  * gradient(params, x, t)       its gradient by reverse-mode differentiation,
  * fd_gradient(params, x, t, idx, h)  central finite differences.

Pinned by tests/test_unet_oracle.py: the convolution against
scipy.ndimage.correlate channel by channel, the pooling and upsampling
against their closed forms and as each other's adjoint (up^T = 8 pool), and
the reverse-mode gradient of every layer against the central differences.

Tensors are [z][y][x][c] (x the grid's fastest index); a parameter vector
holds, layer by layer, w[tap][cin][cout] then b[cout], tap = 9 dz + 3 dy + dx
for the offsets (dx, dy, dz) - 1.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, relu) of layers 0..10
LAYERS = ((1, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True),
          (64, 64, True), (64, 32, True), (64, 32, True), (32, 16, True), (32, 16, True),
          (16, 1, False))


def layer_sizes():
    return [27 * ci * co + co for ci, co, _ in LAYERS]


def param_count() -> int:
    return sum(layer_sizes())


def unpack(params):
    """[(w (27, cin, cout), b (cout,))] views of a parameter vector."""
    p = np.asarray(params, dtype=np.float64)
    if p.size != param_count():
        raise ValueError(f"{p.size} parameters, the network has {param_count()}")
    out, o = [], 0
    for ci, co, _ in LAYERS:
        w = p[o:o + 27 * ci * co].reshape(27, ci, co)
        o += 27 * ci * co
        out.append((w, p[o:o + co]))
        o += co
    return out


def bf16(a):
    """Round-to-nearest-even to bfloat16 (kept as float64)."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# --- the layers ----------------------------------------------------------------
def conv3(x, w, b=None):
    """out[v, co] = b[co] + sum_tap sum_ci w[tap, ci, co] x[v + off(tap), ci]
    (zero padding); x (D, D, D, cin)."""
    D = x.shape[0]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros(x.shape[:3] + (w.shape[2],))
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        out += xp[dz:dz + D, dy:dy + D, dx:dx + D, :] @ w[tap]
    if b is not None:
        out += b
    return out


def conv3_back(x, w, dout):
    """Gradients of conv3 with respect to (x, w, b), given d out."""
    D = x.shape[0]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    flat = dout.reshape(-1, dout.shape[-1])
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        xs = xp[dz:dz + D, dy:dy + D, dx:dx + D, :].reshape(-1, x.shape[-1])
        dw[tap] = xs.T @ flat
        dxp[dz:dz + D, dy:dy + D, dx:dx + D, :] += dout @ w[tap].T
    return dxp[1:-1, 1:-1, 1:-1, :], dw, flat.sum(axis=0)


def pool(x):
    """2^3 mean pooling."""
    D, c = x.shape[0], x.shape[-1]
    return x.reshape(D // 2, 2, D // 2, 2, D // 2, 2, c).mean(axis=(1, 3, 5))


def pool_back(dout):
    return np.repeat(np.repeat(np.repeat(dout, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    """nearest 2x upsampling."""
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_back(dout):
    D, c = dout.shape[0], dout.shape[-1]
    return dout.reshape(D // 2, 2, D // 2, 2, D // 2, 2, c).sum(axis=(1, 3, 5))


# --- the network -----------------------------------------------------------------
def forward(params, x, act_bf16=False):
    """x (D, D, D) with D a multiple of 4 -> (o (D, D, D), cache).  With
    act_bf16 the input, every ReLU output and every pooled tensor are rounded
    to bfloat16 (the matrix-core inference's storage)."""
    L = unpack(params)
    q = bf16 if act_bf16 else (lambda a: a)
    relu = lambda z: q(np.maximum(z, 0.0))
    c = {"x": q(x)[..., None]}
    c["a0"] = relu(conv3(c["x"], *L[0]))
    c["a1"] = relu(conv3(c["a0"], *L[1]))
    c["p1"] = q(pool(c["a1"]))
    c["a2"] = relu(conv3(c["p1"], *L[2]))
    c["a3"] = relu(conv3(c["a2"], *L[3]))
    c["p2"] = q(pool(c["a3"]))
    c["a4"] = relu(conv3(c["p2"], *L[4]))
    c["a5"] = relu(conv3(c["a4"], *L[5]))
    c["a6"] = relu(conv3(c["a5"], *L[6]))
    c["cat2"] = np.concatenate([up(c["a6"]), c["a3"]], axis=-1)
    c["a7"] = relu(conv3(c["cat2"], *L[7]))
    c["a8"] = relu(conv3(c["a7"], *L[8]))
    c["cat1"] = np.concatenate([up(c["a8"]), c["a1"]], axis=-1)
    c["a9"] = relu(conv3(c["cat1"], *L[9]))
    o = conv3(c["a9"], *L[10])[..., 0]
    return o, c


SUPPORT = 1e-4   # the library's default support threshold


def support_mask(x, support=SUPPORT):
    """m = [|x| > support] in float32, as the device tests it."""
    return np.abs(np.asarray(x, dtype=np.float32)) > np.float32(support)


def loss(params, x, t, support=SUPPORT):
    """mean((x + m o - t)^2): no correction outside the support."""
    o, _ = forward(params, x)
    return float(np.mean((x + support_mask(x, support) * o - t) ** 2))


PLANTS = ("SKIP_SWAP",)


def gradient(params, x, t, plant=None, support=SUPPORT):
    """(loss, d loss / d params) by reverse-mode differentiation.  `plant`
    names a deliberate mistake, for the tests that show the checks catch
    it: "SKIP_SWAP" takes the level-1 skip's gradient from the upsampled
    half of the concatenation."""
    if plant is not None and plant not in PLANTS:
        raise ValueError(f"unknown plant {plant!r}")
    L = unpack(params)
    o, c = forward(params, x)
    m = support_mask(x, support)
    r = x + m * o - t
    nv = r.size
    grads = [None] * len(LAYERS)

    def back(l, inp, dout, act):
        if act is not None:
            dout = dout * (act > 0)
        dx, dw, db = conv3_back(inp, L[l][0], dout)
        grads[l] = np.concatenate([dw.ravel(), db])
        return dx

    d = (2.0 * m * r / nv)[..., None]
    d = back(10, c["a9"], d, None)
    d = back(9, c["cat1"], d, c["a9"])
    da8 = up_back(d[..., :16])
    da1_skip = d[..., :16] if plant == "SKIP_SWAP" else d[..., 16:]
    d = back(8, c["a7"], da8, c["a8"])
    d = back(7, c["cat2"], d, c["a7"])
    da6 = up_back(d[..., :32])
    da3_skip = d[..., 32:]
    d = back(6, c["a5"], da6, c["a6"])
    d = back(5, c["a4"], d, c["a5"])
    d = back(4, c["p2"], d, c["a4"])
    d = back(3, c["a2"], da3_skip + pool_back(d), c["a3"])
    d = back(2, c["p1"], d, c["a2"])
    d = back(1, c["a0"], da1_skip + pool_back(d), c["a1"])
    back(0, c["x"], d, c["a0"])
    return float(np.mean(r ** 2)), np.concatenate(grads)


def fd_gradient(params, x, t, idx, h=1e-6, support=SUPPORT):
    """Central differences of the loss for the parameters `idx`."""
    p = np.array(params, dtype=np.float64)
    out = np.empty(len(idx))
    for k, i in enumerate(idx):
        keep = p[i]
        p[i] = keep + h
        lp = loss(p, x, t, support)
        p[i] = keep - h
        lm = loss(p, x, t, support)
        p[i] = keep
        out[k] = (lp - lm) / (2 * h)
    return out


def random_params(seed, scale=1.0):
    """Weights of He scale (all layers, the last one too) and small biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co, _ in LAYERS:
        parts.append(rng.standard_normal(27 * ci * co) * scale * np.sqrt(2.0 / (27 * ci)))
        parts.append(rng.standard_normal(co) * 0.05)
    return np.concatenate(parts)
