"""The allocation-free constraint pass (mgpbd_constraints_noalloc) against the context entry
point mgpbd_constraints, through the C ABI: same kernels, so C, gradC, rhs and ||rhs||^2 must
agree bit for bit, on two calls in a row (the second on a smaller mesh, so the shared
file-scope scratch still holds the first call's partials)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from tests.scenarios import deformed, small_block   # noqa: E402


def _inputs(n, seed):
    from controllable_pbd_3d_amd.runtime import get_context
    mesh = small_block(n, seed=seed)
    dev = torch.device("cuda", 0)
    ctx = get_context(0)
    t_i = torch.from_numpy(mesh.tets.copy()).to(dev)
    x_ref = torch.from_numpy(mesh.verts.copy()).to(dev)
    nt = mesh.n_tets
    Dm = torch.zeros(nt, 9, dtype=torch.float32, device=dev)
    vol = torch.zeros(nt, dtype=torch.float32, device=dev)
    alpha = torch.zeros(nt, dtype=torch.float32, device=dev)
    ctx("mgpbd_rest_shape", nt, t_i.data_ptr(), x_ref.data_ptr(), None, 1e6, 1.0 / 300.0,
        Dm.data_ptr(), vol.data_ptr(), alpha.data_ptr())
    x = torch.from_numpy(deformed(mesh, seed=seed)).to(dev)
    g = torch.Generator().manual_seed(seed)
    lam = (torch.rand(nt, generator=g) * 1e-3).to(dev)
    return ctx, mesh, t_i, x, Dm, alpha, lam


def _outputs(nt, dev):
    f = dict(dtype=torch.float32, device=dev)
    return (torch.full((nt,), 7.0, **f), torch.full((nt, 12), 7.0, **f),
            torch.full((nt,), 7.0, **f), torch.zeros(1, dtype=torch.float64, device=dev))


def _call_both(n, seed):
    from controllable_pbd_3d_amd import _lib
    ctx, mesh, t_i, x, Dm, alpha, lam = _inputs(n, seed)
    nt, dev = mesh.n_tets, x.device
    C0, G0, b0, n0 = _outputs(nt, dev)
    ctx("mgpbd_constraints", nt, t_i.data_ptr(), x.data_ptr(), Dm.data_ptr(), alpha.data_ptr(),
        lam.data_ptr(), C0.data_ptr(), G0.data_ptr(), b0.data_ptr(), n0.data_ptr())
    C1, G1, b1, n1 = _outputs(nt, dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    print(f"[no_alloc] box{n} n_tets={nt} seed={seed}", flush=True)
    _lib.check(_lib.load().mgpbd_constraints_noalloc(
        stream, nt, t_i.data_ptr(), x.data_ptr(), Dm.data_ptr(), alpha.data_ptr(),
        lam.data_ptr(), C1.data_ptr(), G1.data_ptr(), b1.data_ptr(), n1.data_ptr()),
        "mgpbd_constraints_noalloc")
    torch.cuda.synchronize()
    assert torch.equal(C0, C1)
    assert torch.equal(G0, G1)
    assert torch.equal(b0, b1)
    assert n0.item() == n1.item()
    assert np.isfinite(n1.item()) and n1.item() > 0
    # the norm is the FP64 sum of the rhs squares
    ref = float((b1.double() ** 2).sum().item())
    assert abs(n1.item() - ref) <= 1e-12 * ref


def test_noalloc_constraints_equal_context_entry_point_twice():
    _call_both(9, seed=1)     # 4,374 tets: several blocks of partials
    _call_both(3, seed=2)     # 162 tets: one block; stale partials of the first call remain


def test_noalloc_rejects_bad_sizes():
    from controllable_pbd_3d_amd import _lib
    lib = _lib.load()
    assert lib.mgpbd_constraints_noalloc(None, -1, None, None, None, None, None, None, None,
                                         None, None) == -1
    assert b"n_tets must be in" in lib.mgpbd_last_error()
    assert lib.mgpbd_constraints_noalloc(None, (1 << 28) + 1, None, None, None, None, None,
                                         None, None, None, None) == -1
    assert lib.mgpbd_constraints_noalloc(None, 10, None, None, None, None, None, None, None,
                                         None, None) == -1
    assert b"null array" in lib.mgpbd_last_error()
