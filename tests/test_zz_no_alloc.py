"""sds_solve_noalloc (the solve with its workspace in the library's fixed per-GPU scratch array)
against sds_solve on the same factor and arguments: bitwise the same solutions and refinement
steps, over a sequence of calls that share the scratch array (a large call, then smaller ones
on other handles, then the large one again), and the size and FGMRES refusals."""
import ctypes

import numpy as np
import pytest

import sparsedirectsolver as sds
from generators import kkt_matrix
from sparsedirectsolver import _abi
from sparsedirectsolver import matrices as M

pytestmark = pytest.mark.gpu

SCRATCH = 128 << 20  # SDS_NOALLOC_SCRATCH_MIB of include/sds_hip.h


def _spd(shape, fp=None):
    A = M.laplacian3d(*shape)
    rp, cv, nz = M.csr_arrays(M.lower_csr(A))
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "SPD", "L")
    sds.setparam(s, "user_perm", sds.grid_nd_perm(*shape))
    sds.setparam(s, "amalgamation", (512, 0.25, 8))
    if fp:
        sds.setparam(s, "factor_precision", fp)
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    return s


def _ldlt():
    A = kkt_matrix(200, 80, 1e-8, np.random.default_rng(3))
    rp, cv, nz = M.csr_arrays(M.lower_csr(A))
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "S", "L")
    sds.setparam(s, "amalgamation", (64, 0.4, 8))
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    return s


def _lu():
    A = M.convection_diffusion3d(7, peclet=10.0).tocsr()
    A.sort_indices()
    rp, cv, nz = M.csr_arrays(A)
    s = sds.DirectSolver((rp, cv, nz, A.shape[0]), "G", "F")
    sds.execute("analysis", s)
    sds.execute("factorization", s)
    return s


def _call(entry, s, B, nrhs, steps, mode=0, phase=1008):
    """One call of `entry` (sds_solve or sds_solve_noalloc) on the solver's handle; B and the
    returned X are nrhs x n (column-major n x nrhs)."""
    import torch
    X = torch.zeros_like(B)
    done = ctypes.c_int32(-1)
    rc = entry(s._h, phase, mode, ctypes.c_void_p(B.data_ptr()), s.n, ctypes.c_void_p(X.data_ptr()), s.n, nrhs,
               ctypes.c_void_p(s.nzval.data_ptr()), steps, 0.0, ctypes.byref(done))
    torch.cuda.synchronize()
    assert rc == _abi.SDS_SUCCESS, s._lib.sds_last_error(s._h)
    return X.cpu().numpy(), done.value


def _compare(s, nrhs, steps, seed, mode=0, phase=1008):
    import torch
    lib = s._lib
    B = torch.tensor(np.random.default_rng(seed).standard_normal((nrhs, s.n)), dtype=torch.float64, device="cuda")
    Xa, da = _call(lib.sds_solve, s, B, nrhs, steps, mode, phase)
    Xb, db = _call(lib.sds_solve_noalloc, s, B, nrhs, steps, mode, phase)
    assert da == db == (steps if phase & 512 else 0)
    assert np.array_equal(Xa, Xb), np.abs(Xa - Xb).max()
    assert np.isfinite(Xa).all() and np.abs(Xa).max() > 0


def test_noalloc_matches_solve_over_a_sequence_of_calls():
    big = _spd((20, 20, 20))
    print(f"noalloc: SPD n={big.n} nrhs 8 + 2 refinement steps", flush=True)
    _compare(big, 8, 2, seed=1)
    small = _ldlt()
    print(f"noalloc: LDL^T n={small.n} nrhs 2 (smaller, after the larger call)", flush=True)
    _compare(small, 2, 1, seed=2)
    lu = _lu()
    print(f"noalloc: LU n={lu.n} transposed, nrhs 1", flush=True)
    _compare(lu, 1, 1, seed=3, mode=1)
    _compare(lu, 3, 0, seed=4, phase=16 | 32 | 64 | 128 | 256)  # every sub-phase but refinement
    f32 = _spd((9, 8, 7), fp="Float32")
    print(f"noalloc: FP32 factor n={f32.n} nrhs 4 + 2 refinement steps", flush=True)
    _compare(f32, 4, 2, seed=5)
    print("noalloc: SPD again, nrhs 5", flush=True)
    _compare(big, 5, 1, seed=6)


def test_noalloc_refuses_what_it_cannot_hold():
    import torch
    s = _spd((42, 42, 40))
    nrhs = 64
    assert (32 * s.n * nrhs) > SCRATCH  # 2 s + 16 bytes per entry at FP64, before the update vectors
    print(f"noalloc refusal: n={s.n} nrhs {nrhs}", flush=True)
    lib = s._lib
    B = torch.zeros((nrhs, s.n), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    rc = lib.sds_solve_noalloc(s._h, 1008, 0, ctypes.c_void_p(B.data_ptr()), s.n, ctypes.c_void_p(X.data_ptr()), s.n,
                               nrhs, ctypes.c_void_p(s.nzval.data_ptr()), 0, 0.0, None)
    assert rc == _abi.SDS_INVALID_VALUE
    assert b"scratch" in lib.sds_last_error(s._h)
    # a size that fits still solves on the same handle
    _compare(s, 2, 1, seed=7)
    assert lib.sds_set_ir_mode(s._h, 1) == _abi.SDS_SUCCESS
    B1 = B[:2].contiguous()
    rc = lib.sds_solve_noalloc(s._h, 1008, 0, ctypes.c_void_p(B1.data_ptr()), s.n, ctypes.c_void_p(X.data_ptr()),
                               s.n, 2, ctypes.c_void_p(s.nzval.data_ptr()), 3, 0.0, None)
    assert rc == _abi.SDS_NOT_SUPPORTED
    assert lib.sds_set_ir_mode(s._h, 0) == _abi.SDS_SUCCESS
