"""Pins for the learned-warm-start oracle (oracle/gnn_oracle.py): its hand-derived backward
against central finite differences and against torch autograd (float64), the neighbour-mean
transpose against the adjoint identity, BF16 rounding against torch's bfloat16 conversion,
Adam against torch.optim.Adam, and the features against the invariances they are built to
have. A wrong sign, a transposed weight, a dropped ReLU mask or a wrong neighbour index in the
backward fails the finite-difference and autograd checks."""

import numpy as np
import pytest
import scipy.sparse as sp

from controllable_pbd_3d_amd.mesh import box_mesh
from oracle import gnn_oracle as go
from oracle import structure

torch = pytest.importorskip("torch")


def mesh_graph(n=3, seed=0):
    m = box_mesh(n, 0.1, jitter=0.05, seed=seed)
    t = m.tets.reshape(-1).astype(np.int32)
    nt = len(t) // 4
    adj = structure.build_tet_face_adjacency(t, nt)
    nbr = -np.ones((nt, 4), np.int32)
    for i in range(nt):
        a = sorted(adj[i])
        nbr[i, :len(a)] = a
    return m, nbr


def random_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    P = go.init_params(seed).astype(np.float64)
    d = go.unpack(P)
    d["w_out"] = rng.normal(0, 0.3, go.H)
    d["b_out"] = np.array([0.1])
    for l in range(go.NL):
        d[f"b{l}"] = rng.normal(0, 0.1, go.H)
    d["b_in"] = rng.normal(0, 0.1, go.H)
    return go.pack(d) * scale


def random_inputs(nt, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (nt, go.F))
    x[:, 7] = 1.0
    lp = rng.normal(0, 1, nt)
    stats = np.array([0.7, 1, 0, 1, 0, 1, 1, 0], np.float32)
    return x, lp, stats


def spd_on_graph(nbr, seed=0):
    rng = np.random.default_rng(seed)
    n = len(nbr)
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in nbr[i]:
            if j > i:
                v = rng.uniform(-0.5, 0.5)
                rows += [i, j]; cols += [j, i]; vals += [v, v]
    B = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    d = np.abs(B).sum(axis=1).A1 + rng.uniform(0.5, 2.0, n)
    return (B + sp.diags(d)).tocsr()


def total_loss(P, x, nbr, stats, lp, kind, target=None, A=None, b=None):
    y, dlam, _ = go.forward(P, x, nbr, stats, lp)
    if kind == "mse":
        return go.loss_mse(y, target, lp, stats)[0]
    if kind == "energy":
        return go.loss_energy(A, target, dlam, stats)[0]
    return go.loss_residual(A, b, dlam, stats)[0]


@pytest.mark.parametrize("kind", ["mse", "residual", "energy"])
def test_backward_matches_finite_differences(kind):
    _, nbr = mesh_graph(2)
    nt = len(nbr)
    x, lp, stats = random_inputs(nt, 1)
    P = random_params(2)
    rng = np.random.default_rng(3)
    target = lp + rng.normal(0, 0.5, nt)
    A = spd_on_graph(nbr, 4)
    b = rng.normal(0, 1, nt)
    y, dlam, cache = go.forward(P, x, nbr, stats, lp)
    if kind == "mse":
        _, dy = go.loss_mse(y, target, lp, stats)
    elif kind == "energy":
        _, dy = go.loss_energy(A, target, dlam, stats)
    else:
        _, dy = go.loss_residual(A, b, dlam, stats)
    g = go.backward(P, cache, dy, nbr)
    o, n = go.offsets()
    # every parameter group, a few entries each (the decoder fully)
    idx = list(range(o["w_out"][0], n))
    for k in ("W_in", "b_in", "W0", "b0", "W1", "b1", "W2", "b2"):
        a, shp = o[k]
        idx += list(rng.choice(int(np.prod(shp)), 6, replace=False) + a)
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (total_loss(Pp, x, nbr, stats, lp, kind, target, A, b) -
              total_loss(Pm, x, nbr, stats, lp, kind, target, A, b)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def torch_forward(P, x, nbr, stats, lp):
    """The network in torch float64 for autograd (independent derivative)."""
    o, _ = go.offsets()
    T = torch.tensor(P, dtype=torch.float64, requires_grad=True)

    def g(k):
        a, s = o[k]
        return T[a:a + int(np.prod(s))].reshape(s)

    X = torch.tensor(x, dtype=torch.float64)
    nb = torch.tensor(nbr, dtype=torch.long)
    valid = (nb >= 0).to(torch.float64)
    deg = valid.sum(1).clamp(min=1)
    h = torch.relu(X @ g("W_in") + g("b_in"))
    for l in range(go.NL):
        gathered = h[nb.clamp(min=0)] * valid[..., None]
        m = gathered.sum(1) / deg[:, None]
        W = g(f"W{l}")
        h = torch.relu(h @ W[:go.H] + m @ W[go.H:] + g(f"b{l}"))
    y = h @ g("w_out") + g("b_out")[0]
    return T, y


def test_backward_matches_autograd():
    _, nbr = mesh_graph(3)
    nt = len(nbr)
    x, lp, stats = random_inputs(nt, 5)
    P = random_params(6)
    target = lp + np.random.default_rng(7).normal(0, 0.5, nt)
    T, yt = torch_forward(P, x, nbr, stats, lp)
    s = float(stats[0])
    t = torch.tensor((target - lp) / s)
    loss = ((yt - t) ** 2).mean()
    loss.backward()
    y, _, cache = go.forward(P, x, nbr, stats, lp)
    L, dy = go.loss_mse(y, target, lp, stats)
    assert abs(L - loss.item()) < 1e-12 * max(1, L)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    g = go.backward(P, cache, dy, nbr)
    np.testing.assert_allclose(g, T.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_neighbour_mean_transpose_is_adjoint():
    _, nbr = mesh_graph(3)
    nbr = nbr.copy()
    nbr[5, 2:] = -1          # a ragged row: the transpose must not assume symmetry
    rng = np.random.default_rng(0)
    h = rng.normal(size=(len(nbr), 7))
    g = rng.normal(size=(len(nbr), 7))
    lhs = (go.mean_neighbours(h, nbr) * g).sum()
    rhs = (h * go.mean_neighbours_T(g, nbr)).sum()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # an isolated tet gets a zero message
    nbr[7] = -1
    assert np.all(go.mean_neighbours(h, nbr)[7] == 0)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 1, 5000) * 10.0 ** rng.integers(-30, 30, 5000),
                        [0.0, -0.0, 1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 65504.0]])
    v = v.astype(np.float32)
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(go.bf16(v), ref)
    assert go.bf16(np.float32(1 + 2 ** -8)) == 1.0                 # tie to even
    assert go.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6


def test_adam_matches_torch():
    rng = np.random.default_rng(1)
    P0 = rng.normal(size=50)
    p = torch.tensor(P0.copy(), requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.99), eps=1e-8)
    P, m, v = P0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.normal(size=50)
        p.grad = torch.tensor(g)
        opt.step()
        P, m, v = go.adam(P, g, m, v, step, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8)
    np.testing.assert_allclose(P, p.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_feature_invariances():
    m, nbr = mesh_graph(3)
    nt = len(nbr)
    rng = np.random.default_rng(2)
    b = rng.normal(0, 1e-3, nt)
    dinv = 10.0 ** rng.uniform(-3, -1, nt)
    lp = rng.normal(0, 1e-2, nt)
    C = np.abs(rng.normal(0, 1e-3, nt))
    alpha = 10.0 ** rng.uniform(-6, -4, nt)
    vol = rng.uniform(1e-4, 2e-4, nt)
    x, st = go.features(b, dinv, lp, C, alpha, vol, nbr)
    # b and lam_prev share one scale: scaling both leaves their features unchanged
    x2, st2 = go.features(8 * b, dinv, 8 * lp, C, alpha, vol, nbr)
    np.testing.assert_allclose(x2[:, :2], x[:, :2], rtol=1e-6)
    assert st2[0] == pytest.approx(8 * st[0], rel=1e-6)
    # standardised columns
    for c in (3, 4):
        assert abs(x[:, c].mean()) < 1e-4 and abs(x[:, c].std() - 1) < 1e-4
    assert abs(x[:, 5].mean()) < 1e-5
    # rms of the Jacobi column is 1 when it dominates the scale
    if np.sqrt(np.mean((b * dinv) ** 2)) > np.sqrt(np.mean(lp ** 2)):
        assert np.sqrt(np.mean(x[:, 0].astype(np.float64) ** 2)) == pytest.approx(1, rel=1e-5)
    np.testing.assert_array_equal(x[:, 6] * 4, (nbr >= 0).sum(1))
    assert np.all(x[:, 7] == 1)
    # lam_prev missing is lam_prev = 0
    x0, _ = go.features(b, dinv, None, C, alpha, vol, nbr)
    x00, _ = go.features(b, dinv, np.zeros(nt), C, alpha, vol, nbr)
    np.testing.assert_array_equal(x0, x00)


def test_zero_decoder_predicts_previous_lambda():
    _, nbr = mesh_graph(2)
    x, lp, stats = random_inputs(len(nbr), 3)
    P = go.init_params(0)
    y, dlam, _ = go.forward(P, x, nbr, stats, lp, emulate=True)
    assert np.all(y == 0) and np.array_equal(dlam, lp)
