"""CPU reference of the learned-bracket network (test infrastructure).

Restates the library's forward pass, loss and backward pass in numpy, rounding every MFMA
operand to BF16 (round-to-nearest-even) exactly where the device kernels do -- split
hi + lo pairs in the forward products, hi parts in the backward ones -- and accumulating
in float64, so device outputs and gradients can be checked against it.  With
``bf16=False`` it is the plain float64 network, which the finite-difference test pins.
"""
import numpy as np

NIN, NHID, NOUT = 4, 64, 2
SHAPES = [("W1", (NHID, NIN)), ("b1", (NHID,)), ("W2", (NHID, NHID)), ("b2", (NHID,)),
          ("W3", (NHID, NHID)), ("b3", (NHID,)), ("W4", (NOUT, NHID)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    out, o = {}, 0
    for name, sh in SHAPES:
        k = int(np.prod(sh))
        out[name] = np.asarray(w[o:o + k], np.float64).reshape(sh)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, (u >> 16) << 16, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _split(a):
    """hi + lo BF16 pair of float32 values: hi = bf16(a), lo = bf16(a - hi)."""
    a32 = np.asarray(a, np.float32)
    hi = bf16(a32)
    lo = bf16((a32.astype(np.float64) - hi).astype(np.float32))
    return hi, lo


def _mm(a, W, bf):
    """a @ W.T as the device forms it: split BF16 (hi.hi + lo.hi + hi.lo) or exact."""
    if not bf:
        return np.asarray(a, np.float64) @ np.asarray(W, np.float64).T
    ah, al = _split(a)
    wh, wl = _split(W)
    return ah @ wh.T + al @ wh.T + ah @ wl.T


def forward(w, x, bf=True):
    """x: (n, 4) standardized features.  Returns outputs (n, 2) and the activations
    (h0 = the BF16 hi part of the features, as the weight gradient of layer 1 uses it)."""
    p = unpack(w)
    h0 = bf16(x) if bf else np.asarray(x, np.float64)
    z1 = _mm(x, p["W1"], bf) + p["b1"]
    h1 = np.maximum(z1, 0)
    z2 = _mm(h1, p["W2"], bf) + p["b2"]
    h2 = np.maximum(z2, 0)
    z3 = _mm(h2, p["W3"], bf) + p["b3"]
    h3 = np.maximum(z3, 0)
    o = _mm(h3, p["W4"], bf) + p["b4"]
    return o, (h0, h1, h2, h3)


def prior_z0(x, mean, std):
    """z0 = max(1 + q - b^2/2, 1) from standardized features, in fp32 as the device forms it."""
    x = np.asarray(x, np.float32)
    mean, std = np.asarray(mean, np.float32), np.asarray(std, np.float32)
    qv = np.exp(x[:, 0] * std[0] + mean[0]) - np.float32(1e-12)
    bsq = np.exp(x[:, 2] * std[2] + mean[2]) - np.float32(1e-12)
    return np.maximum(np.float32(1) + qv - np.float32(0.5) * bsq, np.float32(1)).astype(np.float64)


def targets(o1, mu, z0):
    """t = log(mu z0) and the fixed width target log(|1 - exp(t - o1)| + 1e-12)."""
    t1 = np.log(np.clip(mu, 1e-30, 1.0) * z0)
    t2 = np.log(np.abs(1 - np.exp(np.clip(t1 - o1, -30, 30))) + 1e-12)
    return t1, t2


def loss_and_grad(w, x, mu, lam, mean=(0, 0, 0, 0), std=(1, 1, 1, 1), bf=True):
    """Loss (o1 - t)^2 + lam (o2 - log(|1 - exp(t - o1)| + 1e-12))^2 with t = log(mu z0),
    mean over n; the second target is held fixed (no gradient through it)."""
    n = x.shape[0]
    p = unpack(w)
    q = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    o, (h0, h1, h2, h3) = forward(w, x, bf)
    t1, t2 = targets(o[:, 0], mu, prior_z0(x, mean, std))
    e1, e2 = o[:, 0] - t1, o[:, 1] - t2
    L = np.sum(e1 * e1 + lam * e2 * e2) / n
    dO = np.stack([2 * e1 / n, 2 * lam * e2 / n], axis=1)
    W4, W3, W2 = q(p["W4"]), q(p["W3"]), q(p["W2"])
    d3 = (q(dO) @ W4) * (h3 > 0)
    d2 = (q(d3) @ W3) * (h2 > 0)
    d1 = (q(d2) @ W2) * (h1 > 0)
    g = {"W4": q(dO).T @ q(h3), "b4": dO.sum(0),
         "W3": q(d3).T @ q(h2), "b3": d3.sum(0),
         "W2": q(d2).T @ q(h1), "b2": d2.sum(0),
         "W1": q(d1).T @ h0, "b1": d1.sum(0)}
    return L, pack(g)


def init(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, sh in SHAPES:
        if name.startswith("W"):
            d[name] = rng.uniform(-1, 1, sh) * np.sqrt(6.0 / sh[1])
        else:
            d[name] = rng.normal(0, 0.1, sh)
    return pack(d)
