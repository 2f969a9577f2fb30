"""Device sparse matrices and BLAS-1, same interface as the reference's
``cons/sparse_gpu.py`` (``GPUSparseMatrix`` :78-144, ``gpu_dot`` ... :17-69).

Storage is SELL-64 (64-row slices = one wavefront, entries k-major inside a slice) so every
column/value load of the SpMV is one coalesced 256-byte segment; the logical CSR (host
``indptr`` / ``indices``, row order and in-row order of the reference) is kept alongside and
``csr_data()`` returns values in CSR order.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._lib import check, dptr, np_ptr
from .runtime import get_context


def sell_layout(indptr: np.ndarray, indices: np.ndarray, n_rows: int):
    """(slice_ptr int32[n_slices+1], sell_idx int32[padded], sell_pos int32[nnz])."""
    lib = _lib.load()
    indptr = np.ascontiguousarray(indptr, np.int32)
    indices = np.ascontiguousarray(indices, np.int32)
    ns = (n_rows + 63) // 64
    slice_ptr = np.zeros(ns + 1, np.int32)
    check(lib.mgpbd_sell_layout(n_rows, np_ptr(indptr), np_ptr(slice_ptr)), "mgpbd_sell_layout")
    sell_idx = np.empty(max(int(slice_ptr[-1]), 1), np.int32)
    check(lib.mgpbd_sell_columns(n_rows, np_ptr(indptr), np_ptr(indices), np_ptr(slice_ptr),
                                 np_ptr(sell_idx)), "mgpbd_sell_columns")
    sell_pos = np.empty(max(len(indices), 1), np.int32)
    check(lib.mgpbd_sell_map(n_rows, np_ptr(indptr), np_ptr(slice_ptr), np_ptr(sell_pos)),
          "mgpbd_sell_map")
    return slice_ptr, sell_idx, sell_pos[:len(indices)]


class GPUSparseMatrix:
    """CSR matrix resident on the GPU (cons/sparse_gpu.py:78-144)."""

    def __init__(self, n_rows: int, nnz: int, device=None):
        self.ctx = get_context(device)
        self.device = self.ctx.device
        self.n_rows = int(n_rows)
        self.nnz = int(nnz)
        self.indptr = None        # host int32 CSR
        self.indices = None
        self.diag_pos = None      # device int32: SELL position of each row's diagonal, -1 if none
        self.data = None          # device float32 values (SELL layout)
        self.diag_inv = torch.zeros(self.n_rows, dtype=torch.float32, device=self.device)

    # -- structure -------------------------------------------------------------------------
    def set_pattern(self, indptr: np.ndarray, indices: np.ndarray, diag_csr_pos=None):
        """Uploads the CSR pattern (host arrays) in SELL-64 form."""
        indptr = np.ascontiguousarray(indptr, np.int32)
        indices = np.ascontiguousarray(indices, np.int32)
        assert indptr.shape[0] == self.n_rows + 1, "indptr size"
        assert int(indptr[-1]) == self.nnz == indices.shape[0], "nnz mismatch"
        self.indptr, self.indices = indptr, indices
        slice_ptr, sell_idx, sell_pos = sell_layout(indptr, indices, self.n_rows)
        self.sell_pos_host = sell_pos
        dev = self.device
        self.slice_ptr = torch.from_numpy(slice_ptr).to(dev)
        self.sell_idx = torch.from_numpy(sell_idx).to(dev)
        self.sell_pos = torch.from_numpy(sell_pos.astype(np.int64)).to(dev)
        self.padded = int(slice_ptr[-1])
        self.data = torch.zeros(max(self.padded, 1), dtype=torch.float32, device=dev)
        self.row_len = torch.from_numpy(np.diff(indptr).astype(np.int32)).to(dev)
        if diag_csr_pos is None:   # scan each row for its diagonal (sparse_gpu.py:107-121)
            rows = np.repeat(np.arange(self.n_rows, dtype=np.int64), np.diff(indptr))
            hit = np.nonzero(indices == rows)[0]
            diag_csr_pos = -np.ones(self.n_rows, np.int64)
            # first diagonal entry of each row, as the reference's scan breaks at the first
            first = np.unique(rows[hit], return_index=True)
            diag_csr_pos[first[0]] = hit[first[1]]
        dp = np.where(np.asarray(diag_csr_pos) >= 0,
                      sell_pos[np.maximum(np.asarray(diag_csr_pos), 0)] if self.nnz else -1, -1)
        self.diag_pos = torch.from_numpy(np.asarray(dp, np.int32)).to(dev)
        self.is_diagonal = bool(self.nnz == self.n_rows and
                                np.all(indices == np.arange(self.n_rows)) and
                                np.all(np.diff(indptr) == 1))

    def from_scipy(self, A: sp.csr_matrix):
        """Pattern and values from a scipy CSR (cons/sparse_gpu.py:93-101)."""
        A = A.tocsr()
        assert A.shape[0] == self.n_rows
        assert A.nnz == self.nnz
        self.set_pattern(A.indptr, A.indices)
        self.update_data(np.asarray(A.data, dtype=np.float32))
        self.cache_diag_inv()

    def update_data(self, data_np):
        """New values in CSR order (cons/sparse_gpu.py:103-105). Accepts numpy or a device
        tensor in CSR order."""
        d = torch.as_tensor(data_np, dtype=torch.float32).to(self.device).reshape(-1)
        assert d.numel() == self.nnz
        self.data[self.sell_pos] = d

    def csr_data(self) -> np.ndarray:
        """Values in CSR order (host)."""
        return self.data[self.sell_pos].cpu().numpy()

    def to_scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.csr_data(), self.indices.copy(), self.indptr.copy()),
                             shape=(self.n_rows, self.n_rows))

    # -- kernels ---------------------------------------------------------------------------
    def cache_diag_inv(self):
        """1/diag (0 where |diag| <= 1e-30) (cons/sparse_gpu.py:107-121)."""
        self.ctx("mgpbd_diag_inv", self.n_rows, dptr(self.diag_pos), dptr(self.data),
                 dptr(self.diag_inv))

    def spmv(self, x: torch.Tensor, y: torch.Tensor):
        """y = A x (cons/sparse_gpu.py:123-132)."""
        self.ctx("mgpbd_spmv", self.n_rows, dptr(self.slice_ptr), dptr(self.sell_idx),
                 dptr(self.data), dptr(x), dptr(y))

    def jacobi_smooth(self, b, x, x_new, omega: float):
        """x_new = x + omega D^-1 (b - A x) (cons/sparse_gpu.py:134-144)."""
        self.ctx("mgpbd_jacobi", self.n_rows, dptr(self.slice_ptr), dptr(self.sell_idx),
                 dptr(self.data), dptr(self.diag_inv), dptr(b), dptr(x), dptr(x_new),
                 float(omega))


# -- BLAS-1 (cons/sparse_gpu.py:17-69); FP64 deterministic reductions ----------------------
def _ctx(t: torch.Tensor):
    return get_context(t.device.index)


def gpu_dot(a, b, n: int) -> float:
    ctx = _ctx(a)
    ctx("mgpbd_dot", int(n), dptr(a), dptr(b), dptr(ctx.scalars))
    return float(ctx.scalars[0].item())


def gpu_norm_sq(x, n: int) -> float:
    return gpu_dot(x, x, n)


def gpu_axpy(alpha: float, x, y, n: int):
    """y += alpha x"""
    _ctx(x)("mgpbd_axpy", int(n), float(alpha), dptr(x), dptr(y))


def gpu_xpay(alpha: float, x, y, n: int):
    """y = x + alpha y"""
    _ctx(x)("mgpbd_xpay", int(n), float(alpha), dptr(x), dptr(y))


def gpu_copy(src, dst, n: int):
    _ctx(src)("mgpbd_copy", int(n), dptr(src), dptr(dst))


def gpu_fill_zero(x, n: int):
    _ctx(x)("mgpbd_fill", int(n), 0.0, dptr(x))


def gpu_scale(x, alpha: float, n: int):
    _ctx(x)("mgpbd_scale", int(n), float(alpha), dptr(x))


def gpu_negate_and_add(b, r, n: int):
    """r = b - r"""
    _ctx(b)("mgpbd_negate_and_add", int(n), dptr(b), dptr(r))
