"""numpy restatement of the learned warm-start network (TEST INFRASTRUCTURE ONLY).

The checker for odx_mlp.hip's forward and backward kernels.  It mirrors the
GPU's precision choices: every MFMA operand is rounded to bf16 (round to
nearest even) and accumulated in float32; elementwise work is float32.
The network: 16 -> 128 -> 128 -> 3 with tanh-form GELU, inputs and targets
normalised by (x - mean) / scale.  Parameter vector layout:
W1[128][16] b1[128] W2[128][128] b2[128] W3[3][128] b3[3].

Setting EXACT = True switches to float64 without bf16 rounding (used to pin
the analytic gradients against finite differences).
"""
import numpy as np

NIN, NH, NOUT = 16, 128, 3
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NOUT, NH)), ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)

EXACT = False


def _f():
    return np.float64 if EXACT else np.float32


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], _f()).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], _f()).ravel() for name, _ in SIZES])


def bf16(a):
    """Round float32 values to bfloat16 (nearest even) and back to float32."""
    if EXACT:
        return np.asarray(a, np.float64)
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def gelu(x):
    """tanh-form GELU and its derivative."""
    F = _f()
    x = np.asarray(x, F)
    c, a = F(0.7978845608028654), F(0.044715)
    u = c * (x + a * x * x * x)
    th = np.tanh(u).astype(F)
    d = F(0.5) * (1 + th) + F(0.5) * x * (1 - th * th) * c * (1 + 3 * a * x * x)
    return (F(0.5) * x * (1 + th)).astype(F), d.astype(F)


def forward(p, X, norm):
    F = _f()
    W = unpack(p)
    norm = np.asarray(norm, F)
    mu, sig = norm[0:16], norm[16:32]
    xn = ((np.asarray(X, F) - mu) / sig).astype(F)
    z1 = bf16(xn) @ bf16(W["W1"]).T + W["b1"]
    h1, d1 = gelu(z1)
    a1, g1 = bf16(h1), bf16(d1)
    z2 = a1 @ bf16(W["W2"]).T + W["b2"]
    h2, d2 = gelu(z2)
    a2, g2 = bf16(h2), bf16(d2)
    y = (a2 @ bf16(W["W3"]).T + W["b3"]).astype(F)
    return y, dict(xn=xn, a1=a1, g1=g1, a2=a2, g2=g2)


def predict(p, X, norm):
    """Impulses in force units: s * (base + y * out_scale + out_mean), with
    s = X[:, 15] and base = X[:, 7:10] (matched previous impulses / s)."""
    F = _f()
    norm = np.asarray(norm, F)
    y, _ = forward(p, X, norm)
    X = np.asarray(X, F)
    return X[:, 15:16] * (X[:, 7:10] + y * norm[35:38] + norm[32:35])


def gradients(p, X, T, norm):
    """Gradient of mean_s sum_o (y - t_norm)^2 and the loss."""
    F = _f()
    W = unpack(p)
    norm = np.asarray(norm, F)
    N = len(X)
    y, c = forward(p, X, norm)
    tn = ((np.asarray(T, F) - norm[32:35]) / norm[35:38]).astype(F)
    e = y - tn
    loss = float(np.sum(e.astype(np.float64) ** 2) / N)
    dz3 = (2.0 * e / N).astype(F)
    g = {}
    g["W3"] = dz3.T @ c["a2"]
    g["b3"] = dz3.sum(0)
    dh2 = dz3 @ W["W3"]
    dz2 = bf16(dh2 * c["g2"])
    g["W2"] = dz2.T @ c["a1"]
    g["b2"] = dz2.sum(0)
    dh1 = dz2 @ bf16(W["W2"])
    dz1 = bf16(dh1 * c["g1"])
    g["W1"] = dz1.T @ bf16(c["xn"])
    g["b1"] = dz1.sum(0)
    return pack(g), loss
