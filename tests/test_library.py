"""The C ABI (CPU-side checks): libmgpbd.so loads, exports every function include/mgpbd.h
declares, the ctypes table matches the header, host entry points reject bad input with a
status and a message, and the GPU classes fail loudly without a GPU."""

import os
import re

import numpy as np
import pytest

from controllable_pbd_3d_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "mgpbd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mgpbd_[A-Za-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) > 40
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert lib.mgpbd_abi_version() == 1


def test_substep_structures_match_the_header():
    """The ctypes restatements of mgpbd_substep_desc / mgpbd_newton_stat have the C sizes,
    and every field of the descriptor is distinct (none dropped or doubled)."""
    import ctypes
    lib = _lib.load()
    sizes = (ctypes.c_int32 * 2)()
    assert lib.mgpbd_substep_struct_sizes(sizes) == 0
    assert (sizes[0], sizes[1]) == (ctypes.sizeof(_lib.SubstepDesc),
                                    ctypes.sizeof(_lib.NewtonStat))
    names = [f[0] for f in _lib.SubstepDesc._fields_]
    assert len(names) == len(set(names))
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = src[src.index("typedef struct mgpbd_substep_desc"):src.index("} mgpbd_substep_desc;")]
    c_fields = re.findall(r"\*?\s*(\w+)(?:\[\d+\])?\s*[;,]", body.split("{", 1)[1])
    assert [("lam" if f == "lambda" else f) for f in c_fields] == names


def test_host_setup_rejects_bad_input():
    lib = _lib.load()
    t = np.array([0, 1, 2, 7], np.int32)          # vertex id 7 with n_verts = 4
    assert not lib.mgpbd_pattern_build(1, 4, _lib.np_ptr(t))
    assert b"out of range" in lib.mgpbd_last_error()
    adj = np.empty(4, np.int32)
    deg = np.empty(1, np.int32)
    assert lib.mgpbd_tet_aggregate(1, _lib.np_ptr(adj), _lib.np_ptr(deg), 0,
                                   _lib.np_ptr(adj), _lib.np_ptr(deg)) == -1
    with pytest.raises(_lib.MGPBDError, match="mgpbd_tet_aggregate"):
        _lib.check(-1, "mgpbd_tet_aggregate")


def test_device_entry_points_reject_null_context():
    lib = _lib.load()
    assert lib.mgpbd_spmv(None, 10, None, None, None, None, None) == -1
    assert b"null context" in lib.mgpbd_last_error()
    assert lib.mgpbd_pcg_solve(None, None, None, None, 1, 1e-6, 0, None, None) == -1


def test_solver_requires_a_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from controllable_pbd_3d_amd import MGPBDSolverGPU
    from controllable_pbd_3d_amd.mesh import box_mesh
    m = box_mesh(2, 0.1)
    with pytest.raises(_lib.MGPBDError, match="needs a HIP GPU"):
        MGPBDSolverGPU(m.verts, m.verts, m.invm, m.tets, m.tet_mass, (0, -9.8, 0), 1 / 300)
