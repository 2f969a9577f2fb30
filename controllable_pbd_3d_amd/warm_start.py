"""Learned warm start for the first PCG of every substep (extension; no reference
counterpart -- the reference zeroes lambda per substep, cons/mgpbd_gpu.py:402, and starts every
PCG from dlambda = 0, cons/mgpbd_gpu.py:426, cons/mgpcg_gpu.py:271-272).

A three-layer message-passing network over the tet face-adjacency graph predicts dlambda of the
first Newton iteration from per-tet features (b / A_ii, the previous substep's lambda, C,
alpha~, A_ii, rest volume, degree). Inference and training both run on libmgpbd's kernels:
feature and neighbour-mean gather kernels, BF16 MFMA layers with FP32 accumulation, their
backward kernels and Adam. ``MGPBDSolverGPU(warm_start='learned', warm_start_model=...)``
hands the prediction to ``mgpbd_pcg_solve_guess`` together with the previous lambda; the
device keeps whichever start has the smaller residual, and every start is iterated to the
same tolerance.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import dptr
from .runtime import get_context, to_device

HIDDEN, FEATURES, LAYERS = 64, 8, 3


def _offsets():
    o = {"W_in": (0, (FEATURES, HIDDEN)), "b_in": (FEATURES * HIDDEN, (HIDDEN,))}
    p = FEATURES * HIDDEN + HIDDEN
    for l in range(LAYERS):
        o[f"W{l}"] = (p, (2 * HIDDEN, HIDDEN))
        o[f"b{l}"] = (p + 2 * HIDDEN * HIDDEN, (HIDDEN,))
        p += 2 * HIDDEN * HIDDEN + HIDDEN
    o["w_out"] = (p, (HIDDEN,))
    o["b_out"] = (p + HIDDEN, (1,))
    return o, p + HIDDEN + 1


def initial_parameters(seed: int = 0) -> np.ndarray:
    """He-normal message-passing and encoder weights, zero biases and a zero decoder: the
    untrained network predicts dlambda = previous lambda (the plain warm start)."""
    o, n = _offsets()
    rng = np.random.default_rng(seed)
    P = np.zeros(n, np.float32)
    a, s = o["W_in"]
    P[a:a + s[0] * s[1]] = rng.normal(0, np.sqrt(2.0 / FEATURES), s).reshape(-1)
    for l in range(LAYERS):
        a, s = o[f"W{l}"]
        P[a:a + s[0] * s[1]] = rng.normal(0, np.sqrt(2.0 / (2 * HIDDEN)), s).reshape(-1)
    return P


def neighbour_table(t_i, n_tets: int, per_tet: int = 1) -> np.ndarray:
    """(n_tets * per_tet, 4) int32 face neighbours, -1 padded (libmgpbd's face adjacency).
    With two constraints per tet (ARAP 2t, volume 2t+1) each constraint's neighbours are the
    same kind of constraint on the face-adjacent tets."""
    from .tet_aggregation import face_adjacency_arrays
    adj, _ = face_adjacency_arrays(np.asarray(t_i, np.int32).reshape(-1), n_tets)
    adj = adj.reshape(n_tets, 4)
    if per_tet == 1:
        return adj
    out = np.repeat(adj, per_tet, axis=0).astype(np.int64)
    kind = np.tile(np.arange(per_tet), n_tets)[:, None]
    out = np.where(out >= 0, per_tet * out + kind, -1)
    return out.astype(np.int32)


class LearnedWarmStart:
    """Network state (FP32 parameters, Adam moments, BF16 weight copies) and workspaces for
    graphs of up to ``n`` tets on one GPU."""

    def __init__(self, nbr, *, params=None, seed: int = 0, device=None, lr: float = 1e-3,
                 betas=(0.9, 0.999), eps: float = 1e-8):
        self.ctx = get_context(device)
        dev = self.device = self.ctx.device
        lib = self.lib = self.ctx.lib
        self.n_params = int(lib.mgpbd_gnn_param_count())
        if self.n_params != _offsets()[1]:
            raise _lib.MGPBDError("libmgpbd network layout differs from this module's")
        self.nbr = to_device(nbr, torch.int32, dev, (-1, 4))
        self.n = int(self.nbr.shape[0])
        P = initial_parameters(seed) if params is None else np.asarray(params, np.float32)
        if P.shape != (self.n_params,):
            raise ValueError(f"params must have {self.n_params} values")
        f32 = dict(dtype=torch.float32, device=dev)
        self.params = torch.from_numpy(P.copy()).to(dev)
        self.grad = torch.zeros(self.n_params, **f32)
        self.m = torch.zeros(self.n_params, **f32)
        self.v = torch.zeros(self.n_params, **f32)
        self.step_count = 0
        self.lr, self.betas, self.eps = float(lr), tuple(betas), float(eps)
        nw = 2 * LAYERS * HIDDEN * HIDDEN
        self.wt = torch.zeros(nw, dtype=torch.int16, device=dev)
        self.w = torch.zeros(nw, dtype=torch.int16, device=dev)
        self.ctx("mgpbd_gnn_pack", dptr(self.params), dptr(self.wt), dptr(self.w))
        # features and inference: two ping-ponged activation buffers; the training workspace
        # (every layer's activations, gradients, partials) is allocated on first training use
        # and dropped by release_training_memory()
        wi = int(lib.mgpbd_gnn_infer_workspace_bytes(self.n))
        self.ws_infer = torch.zeros(wi, dtype=torch.uint8, device=dev)
        self._train = None
        self.x = torch.zeros(self.n, FEATURES, **f32)
        self.stats = torch.zeros(8, **f32)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)

    @property
    def ws(self):
        """The training workspace (allocated on first use)."""
        if self._train is None:
            f32 = dict(dtype=torch.float32, device=self.device)
            nbytes = int(self.lib.mgpbd_gnn_workspace_bytes(self.n))
            self._train = {"ws": torch.zeros(nbytes, dtype=torch.uint8, device=self.device),
                           "dy": torch.zeros(self.n, **f32), "dlam": torch.zeros(self.n, **f32),
                           "tmp": torch.zeros(2 * self.n, **f32)}
        return self._train["ws"]

    def _train_buf(self, key):
        self.ws   # allocates the training buffers on first use
        return self._train[key]

    @property
    def dy(self):
        return self._train_buf("dy")

    @property
    def dlam(self):
        return self._train_buf("dlam")

    @property
    def tmp(self):
        return self._train_buf("tmp")

    def release_training_memory(self) -> None:
        """Frees the training workspace (activations of every layer, gradients, weight-gradient
        partials); inference keeps its own small one. Training again reallocates it."""
        self._train = None

    @classmethod
    def for_mesh(cls, t_i, n_tets: int, per_tet: int = 1, **kw) -> "LearnedWarmStart":
        """Network over the constraints of a tet mesh: one per tet, or two (ARAP and volume)
        with ``per_tet=2``."""
        return cls(neighbour_table(t_i, n_tets, per_tet), **kw)

    # ------------------------------------------------------------------ inference
    def features(self, b, dinv, lam_prev, C, alpha, vol, x=None, stats=None, ctx=None):
        x = self.x if x is None else x
        stats = self.stats if stats is None else stats
        ctx = self.ctx if ctx is None else ctx
        ctx("mgpbd_gnn_features", self.n, dptr(b), dptr(dinv), dptr(lam_prev), dptr(C),
            dptr(alpha), dptr(vol), dptr(self.nbr), dptr(x), dptr(stats), dptr(self.ws_infer))
        return x, stats

    def forward(self, x, stats, lam_prev, out=None):
        self.ctx("mgpbd_gnn_forward", self.n, dptr(self.nbr), dptr(x), dptr(stats),
                 dptr(lam_prev), dptr(self.params), dptr(self.wt), dptr(self.ws), dptr(out))
        return out

    def infer(self, x, stats, lam_prev, out, ctx=None):
        """The forward's prediction through the fused inference kernels (bit-identical;
        keeps no activations for a backward pass). ctx: another context of the same device
        (e.g. bound to a second stream), default the model's."""
        ctx = self.ctx if ctx is None else ctx
        ctx("mgpbd_gnn_infer", self.n, dptr(self.nbr), dptr(x), dptr(stats),
            dptr(lam_prev), dptr(self.params), dptr(self.wt), dptr(self.ws_infer), dptr(out))
        return out

    def predict(self, b, dinv, lam_prev, C, alpha, vol, out, ctx=None):
        """out = predicted dlambda of the first Newton iteration (device tensors, n floats);
        ctx: the context (stream) to run on, default the model's."""
        x, st = self.features(b, dinv, lam_prev, C, alpha, vol, ctx=ctx)
        return self.infer(x, st, lam_prev, out, ctx=ctx)

    # ------------------------------------------------------------------ training
    def make_sample(self, b, dinv, lam_prev, C, alpha, vol, target, matrix=None) -> dict:
        """One training pair from a solved first Newton iteration: features, scale, previous
        lambda, converged dlambda and (for the energy and residual losses) the right-hand side
        and a copy of the level-0 values of A (``matrix``: the GPUSparseMatrix of that solve)."""
        x = torch.empty(self.n, FEATURES, dtype=torch.float32, device=self.device)
        st = torch.empty(8, dtype=torch.float32, device=self.device)
        self.features(b, dinv, lam_prev, C, alpha, vol, x, st)
        s = {"x": x, "stats": st, "lam_prev": lam_prev.clone(), "target": target.clone()}
        if matrix is not None:
            s.update(b=b.clone(), val=matrix.data.clone(), slice_ptr=matrix.slice_ptr,
                     sell_idx=matrix.sell_idx)
        return s

    def _loss(self, sample: dict, loss: str):
        if loss == "mse":
            self.ctx("mgpbd_gnn_loss_mse", self.n, dptr(sample["stats"]),
                     dptr(sample["lam_prev"]), dptr(sample["target"]), dptr(self.ws),
                     dptr(self.dy), dptr(self.loss))
        elif loss == "residual":
            self.ctx("mgpbd_gnn_loss_residual", self.n, dptr(sample["slice_ptr"]),
                     dptr(sample["sell_idx"]), dptr(sample["val"]), dptr(sample["b"]),
                     dptr(self.dlam), dptr(sample["stats"]), dptr(self.tmp), dptr(self.ws),
                     dptr(self.dy), dptr(self.loss))
        elif loss == "energy":
            # t.A t is constant per pair: computed at the pair's first loss, then reused
            ready = "tat" in sample
            if not ready:
                sample["tat"] = torch.zeros(1, dtype=torch.float64, device=self.dlam.device)
            self.ctx("mgpbd_gnn_loss_energy", self.n, dptr(sample["slice_ptr"]),
                     dptr(sample["sell_idx"]), dptr(sample["val"]), dptr(sample["target"]),
                     dptr(self.dlam), dptr(sample["stats"]), dptr(self.tmp), dptr(self.ws),
                     dptr(self.dy), dptr(self.loss), dptr(sample["tat"]), int(ready))
        else:
            raise ValueError("loss must be 'energy', 'residual' or 'mse'")

    def loss_and_grad(self, sample: dict, loss: str = "energy"):
        """Forward, loss and backward for one sample; the gradient lands in ``self.grad`` and
        the loss in ``self.loss`` (device). 'energy' (default) is the A-norm error of the start
        relative to the converged dlambda, the quantity that bounds CG's iteration count;
        'residual' is ||A dlambda - b||^2 / ||b||^2; 'mse' the plain squared error."""
        self.forward(sample["x"], sample["stats"], sample["lam_prev"], self.dlam)
        self._loss(sample, loss)
        self.ctx("mgpbd_gnn_backward", self.n, dptr(self.nbr), dptr(sample["x"]),
                 dptr(self.params), dptr(self.w), dptr(self.dy), dptr(self.ws), dptr(self.grad))
        return self.loss

    def adam_step(self):
        self.step_count += 1
        b1, b2 = self.betas
        self.ctx("mgpbd_gnn_adam", dptr(self.params), dptr(self.grad), dptr(self.m),
                 dptr(self.v), self.lr, b1, b2, self.eps, self.step_count, dptr(self.wt),
                 dptr(self.w))

    def train(self, samples, epochs: int = 1, loss: str = "energy", seed: int = 0,
              lr_decay: float = 1.0):
        """Adam over the samples, one step per sample, shuffled per epoch; returns the mean
        loss of every epoch (one host sync per epoch)."""
        rng = np.random.default_rng(seed)
        hist = []
        acc = torch.zeros(1, dtype=torch.float64, device=self.device)
        for _ in range(epochs):
            acc.zero_()
            for i in rng.permutation(len(samples)):
                acc += self.loss_and_grad(samples[i], loss)
                self.adam_step()
            hist.append(float(acc.item()) / max(len(samples), 1))
            self.lr *= lr_decay
        return hist

    def evaluate(self, samples, loss: str = "energy"):
        """Mean loss over samples without updating the parameters."""
        tot = 0.0
        for s in samples:
            self.forward(s["x"], s["stats"], s["lam_prev"], self.dlam)
            self._loss(s, loss)
            tot += float(self.loss.item())
        return tot / max(len(samples), 1)

    # ------------------------------------------------------------------ state
    def state_dict(self) -> dict:
        return {"params": self.params.cpu().numpy().copy()}

    def load_state_dict(self, d: dict) -> None:
        P = np.asarray(d["params"], np.float32)
        if P.shape != (self.n_params,):
            raise ValueError(f"params must have {self.n_params} values")
        self.params.copy_(torch.from_numpy(P))
        self.ctx("mgpbd_gnn_pack", dptr(self.params), dptr(self.wt), dptr(self.w))

    def save(self, path: str) -> None:
        np.savez(path, params=self.params.cpu().numpy(), hidden=HIDDEN, layers=LAYERS,
                 features=FEATURES)

    @staticmethod
    def load_parameters(path: str) -> np.ndarray:
        with np.load(path, allow_pickle=False) as z:
            if int(z["hidden"]) != HIDDEN or int(z["layers"]) != LAYERS:
                raise ValueError("checkpoint network shape differs")
            return z["params"].astype(np.float32)
