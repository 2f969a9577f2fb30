"""ORACLE (test infrastructure only) -- FP64 numpy statement of the learned warm start:
the message-passing network over the tet face-adjacency graph, its per-tet features, the three
training losses (squared error, residual, A-norm energy), the hand-derived backward pass and the Adam update.

The network has no counterpart in the reference (it solves every PCG from dlambda = 0,
cons/mgpbd_gpu.py:426 and cons/mgpcg_gpu.py:271-272); its definition is the build's own:

    h0  = relu(x W_in + b_in)                               x: n x 8 per-tet features
    m_l = mean_{j in N(i)} h_{l-1, j}                       N: face neighbours (<= 4)
    h_l = relu(h_{l-1} W_l[:64] + m_l W_l[64:] + b_l)       l = 1, 2, 3 (hidden 64)
    y   = h3 w_out + b_out;     dlambda = lam_prev + s y    s = feature scale

``emulate=True`` rounds to BF16 where the GPU stores BF16 (activations h_l, m_l, the
message-passing weights, the backward's masked gradients dZ and neighbour-message gradients)
and keeps everything else in FP64, so the GPU (FP32 accumulation) differs from it only by
accumulation order and rare BF16 rounding-boundary flips. ``emulate=False`` is the exact
FP64 network, the one the finite-difference and autograd checks pin.

Only ``tests/`` use this module; the product never imports it.
"""

from __future__ import annotations

import numpy as np

H, F, NL = 64, 8, 3


def offsets():
    """Flattened parameter layout (FP32 vector on the GPU): W_in[8][64], b_in[64],
    {W_l[128][64], b_l[64]} x 3, w_out[64], b_out."""
    o = {"W_in": (0, (F, H)), "b_in": (F * H, (H,))}
    p = F * H + H
    for l in range(NL):
        o[f"W{l}"] = (p, (2 * H, H))
        o[f"b{l}"] = (p + 2 * H * H, (H,))
        p += 2 * H * H + H
    o["w_out"] = (p, (H,))
    o["b_out"] = (p + H, (1,))
    return o, p + H + 1


def n_params() -> int:
    return offsets()[1]


def unpack(params):
    o, _ = offsets()
    P = np.asarray(params, np.float64)
    return {k: P[a:a + int(np.prod(s))].reshape(s) for k, (a, s) in o.items()}


def pack(d):
    o, n = offsets()
    P = np.zeros(n, np.float64)
    for k, (a, s) in o.items():
        P[a:a + int(np.prod(s))] = np.asarray(d[k], np.float64).reshape(-1)
    return P


def bf16(a):
    """Round-to-nearest-even to bfloat16 (via float32), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def _q(emulate):
    return bf16 if emulate else (lambda a: np.asarray(a, np.float64))


def degree(nbr):
    return (np.asarray(nbr) >= 0).sum(axis=1)


def mean_neighbours(h, nbr):
    nbr = np.asarray(nbr)
    d = degree(nbr)
    s = np.zeros_like(h)
    for q in range(nbr.shape[1]):
        j = nbr[:, q]
        ok = j >= 0
        s[ok] += h[j[ok]]
    inv = np.where(d > 0, 1.0 / np.maximum(d, 1), 0.0)
    return s * inv[:, None]


def mean_neighbours_T(dm, nbr):
    """Exact transpose of ``mean_neighbours`` (no symmetry assumed)."""
    nbr = np.asarray(nbr)
    d = degree(nbr)
    inv = np.where(d > 0, 1.0 / np.maximum(d, 1), 0.0)
    out = np.zeros_like(dm)
    scaled = dm * inv[:, None]
    for q in range(nbr.shape[1]):
        j = nbr[:, q]
        ok = j >= 0
        np.add.at(out, j[ok], scaled[ok])
    return out


def feature_stats(b, dinv, lam_prev, C, alpha, vol):
    """FP64 statistics of the features, cast to float32 as the GPU stores them:
    [s, sC, mean log alpha, std log alpha, mean log A_ii, std log A_ii, mean volume, 0]."""
    b, dinv, C, alpha, vol = (np.asarray(a, np.float32).astype(np.float64)
                              for a in (b, dinv, C, alpha, vol))
    lp = np.zeros_like(b) if lam_prev is None else np.asarray(lam_prev, np.float32).astype(np.float64)
    n = max(len(b), 1)
    jac = b * dinv
    la = np.log(alpha)
    ld = np.where(dinv > 0, -np.log(np.where(dinv > 0, dinv, 1.0)), 0.0)
    s = max(np.sqrt((jac * jac).sum() / n), np.sqrt((lp * lp).sum() / n), 1e-30)
    ma, md = la.sum() / n, ld.sum() / n
    st = [s, max(np.sqrt((C * C).sum() / n), 1e-30), ma,
          max(np.sqrt(max((la * la).sum() / n - ma * ma, 0.0)), 1e-6), md,
          max(np.sqrt(max((ld * ld).sum() / n - md * md, 0.0)), 1e-6),
          max(vol.sum() / n, 1e-30), 0.0]
    return np.asarray(st, np.float32)


def features(b, dinv, lam_prev, C, alpha, vol, nbr, stats=None):
    """Per-tet features (float32, as the GPU computes them from float32 statistics):
    [b/A_ii / s, lam_prev / s, C / sC, std. log alpha, std. log A_ii, vol / mean - 1, deg / 4, 1]."""
    st = feature_stats(b, dinv, lam_prev, C, alpha, vol) if stats is None else np.asarray(stats, np.float32)
    f32 = np.float32
    b, dinv, C, alpha, vol = (np.asarray(a, f32) for a in (b, dinv, C, alpha, vol))
    n = len(b)
    lp = np.zeros(n, f32) if lam_prev is None else np.asarray(lam_prev, f32)
    x = np.empty((n, F), f32)
    la = np.log(alpha)
    ld = np.where(dinv > 0, -np.log(np.where(dinv > 0, dinv, f32(1))), st[4]).astype(f32)
    x[:, 0] = (b * dinv) / st[0]
    x[:, 1] = lp / st[0]
    x[:, 2] = C / st[1]
    x[:, 3] = (la - st[2]) / st[3]
    x[:, 4] = (ld - st[4]) / st[5]
    x[:, 5] = vol / st[6] - f32(1)
    x[:, 6] = f32(0.25) * degree(nbr).astype(f32)
    x[:, 7] = 1
    return x, st


def forward(params, x, nbr, stats, lam_prev=None, emulate=False):
    """Returns (y, dlambda, cache)."""
    q = _q(emulate)
    p = unpack(params)
    x = np.asarray(x, np.float64)
    h = [q(np.maximum(x @ p["W_in"] + p["b_in"], 0.0))]
    m = [None]
    for l in range(NL):
        W = p[f"W{l}"]
        ml = q(mean_neighbours(h[l], nbr))
        z = h[l] @ q(W[:H]) + ml @ q(W[H:]) + p[f"b{l}"]
        h.append(q(np.maximum(z, 0.0)))
        m.append(ml)
    y = h[NL] @ p["w_out"] + p["b_out"][0]
    s = float(np.asarray(stats, np.float32)[0])
    lp = 0.0 if lam_prev is None else np.asarray(lam_prev, np.float64)
    return y, lp + s * y, {"h": h, "m": m, "x": x}


def loss_mse(y, target, lam_prev, stats):
    s = float(np.asarray(stats, np.float32)[0])
    lp = 0.0 if lam_prev is None else np.asarray(lam_prev, np.float64)
    e = y - (np.asarray(target, np.float64) - lp) / s
    return float((e * e).mean()), 2.0 * e / len(e)


def loss_residual(A, b, dlam, stats):
    """||A dlam - b||^2 / ||b||^2 and its gradient with respect to y (dlam = lam_prev + s y)."""
    s = float(np.asarray(stats, np.float32)[0])
    b = np.asarray(b, np.float64)
    r = A @ np.asarray(dlam, np.float64) - b
    bb = float(b @ b)
    return float(r @ r) / bb, (2.0 * s / bb) * (A.T @ r)


def loss_energy(A, target, dlam, stats):
    """(dlam - t)^T A (dlam - t) / t^T A t and its gradient with respect to y."""
    s = float(np.asarray(stats, np.float32)[0])
    t = np.asarray(target, np.float64)
    e = np.asarray(dlam, np.float64) - t
    Ae = A @ e
    tAt = float(t @ (A @ t))
    return float(e @ Ae) / tAt, (2.0 * s / tAt) * Ae


def backward(params, cache, dy, nbr, emulate=False):
    """d loss / d params for d loss / d y = dy (hand-derived; pinned by finite differences)."""
    q = _q(emulate)
    p = unpack(params)
    h, m, x = cache["h"], cache["m"], cache["x"]
    dy = np.asarray(dy, np.float64)
    g = {"w_out": h[NL].T @ dy, "b_out": np.array([dy.sum()])}
    dh = np.outer(dy, p["w_out"])
    for l in range(NL - 1, -1, -1):
        W = p[f"W{l}"]
        dz = q(dh * (h[l + 1] > 0))
        g[f"W{l}"] = np.vstack([h[l].T @ dz, m[l + 1].T @ dz])
        g[f"b{l}"] = dz.sum(axis=0)
        dm = q(dz @ q(W[H:]).T)
        dh = dz @ q(W[:H]).T + mean_neighbours_T(dm, nbr)
    dz0 = q(dh * (h[0] > 0))
    g["W_in"] = x.T @ dz0
    g["b_in"] = dz0.sum(axis=0)
    return pack(g)


def adam(params, grad, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015, Algorithm 1)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """He-normal weights, zero biases, zero decoder (so the untrained network predicts
    dlambda = lam_prev: the plain warm start)."""
    rng = np.random.default_rng(seed)
    d = {"W_in": rng.normal(0, np.sqrt(2.0 / F), (F, H)), "b_in": np.zeros(H)}
    for l in range(NL):
        d[f"W{l}"] = rng.normal(0, np.sqrt(2.0 / (2 * H)), (2 * H, H))
        d[f"b{l}"] = np.zeros(H)
    d["w_out"] = np.zeros(H)
    d["b_out"] = np.zeros(1)
    return pack(d).astype(np.float32)
