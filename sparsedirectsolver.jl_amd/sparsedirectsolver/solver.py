"""The solver handle and its phases, mirroring SparseDirectSolver.jl's API.

``DirectSolver``, ``execute``, ``setparam``, ``getparam`` and ``update`` keep the
names, phase strings, parameter names and error behaviour of the reference
(src/solver.jl:187-966, src/options.jl:155-657) so that a caller of the
reference (or of CUDSS.jl) finds the same surface. The host analysis runs in
``libsds_analysis.so``; every numeric and solve phase runs on the GPU through
the C ABI of ``libsds_hip.so`` (include/sds_hip.h). There is no CPU path: without
the HIP library the phases raise.

Device arrays are torch tensors on ``cuda`` (plumbing for HIP memory and streams).
"""
import ctypes

import numpy as np

from . import _abi
from ._native import hip_lib
from .analysis import (MATCHING_JOBS, compute_matching, entry_scaling, full_maps, matching_pairs, merge_symbolic,
                       symbolic_analysis)
from .enums import PHASES, PIVOT_TYPES, STRUCTURES, VIEWS
from .errors import FactorizationError, InterruptedError_, InvalidValueError, NotSupportedError

CONFIG_PARAMETERS = ("reordering_alg", "factorization_alg", "solve_alg", "matching_alg", "solve_mode", "ir_n_steps",
                     "ir_tol", "pivot_type", "pivot_threshold", "pivot_epsilon", "max_lu_nnz", "hybrid_memory_mode",
                     "hybrid_device_memory_limit", "use_cuda_register_memory", "host_nthreads",
                     "hybrid_execute_mode", "pivot_epsilon_alg", "nd_nlevels", "ubatch_size", "ubatch_index",
                     "use_superpanels", "device_count", "device_indices", "schur_mode", "deterministic_mode",
                     "nd_ubfactor", "nd_nseps", "nd_seed")
# refinement modes (src/types.jl:306): plain steps, or FGMRES preconditioned by the factorization
IR_MODES = {"ir": 0, "fgmres": 1}
EXTRA_CONFIG = ("ir_mode", "factor_precision", "amalgamation", "schedule", "pivot_pairs", "pivot_pair_tolerance")
USER_DATA = ("user_perm", "user_host_interrupt", "pivot_sign", "user_schur_indices", "user_nd_partition_tree",
             "ubatch_mask")
DATA_PARAMETERS = ("info", "lu_nnz", "npivots", "inertia", "perm_reorder_row", "perm_reorder_col", "perm_row",
                   "perm_col", "diag", "user_perm", "hybrid_device_memory_min", "comm_device", "comm_host",
                   "memory_estimates", "perm_matching", "scale_row", "scale_col", "nsuperpanels",
                   "user_schur_indices", "schur_shape", "schur_matrix", "user_nd_partition_tree",
                   "nd_partition_tree", "user_host_interrupt", "ir_n_steps", "ubatch_mask", "flops", "pivot_sign",
                   "pivot_stats")

_EPS_ALGS = {"default": 0, "algo1": 1, "algo2": 2}

STAGE_NONE, STAGE_REORDERED, STAGE_ANALYZED, STAGE_FACTORIZED = range(4)


def _defaults():
    return {
        "reordering_alg": "default", "factorization_alg": "default", "solve_alg": "default",
        "matching_alg": "default", "solve_mode": 0, "ir_n_steps": 0, "ir_tol": 0.0, "pivot_type": "A",
        "pivot_threshold": 0.01, "pivot_epsilon": None, "pivot_epsilon_alg": "default", "max_lu_nnz": -1,
        "hybrid_memory_mode": 0, "hybrid_device_memory_limit": 0, "use_cuda_register_memory": 1,
        "host_nthreads": 0, "hybrid_execute_mode": 0, "nd_nlevels": 10, "ubatch_size": 0, "ubatch_index": -1,
        "use_superpanels": 1, "device_count": 1, "device_indices": None, "schur_mode": 0,
        "deterministic_mode": 0, "nd_ubfactor": -1, "nd_nseps": -1, "nd_seed": -1,
        "ir_mode": "ir", "factor_precision": None, "amalgamation": (32, 0.25, 8), "schedule": "auto",
        "pivot_pairs": "default", "pivot_pair_tolerance": 1e-6,
        "user_perm": None, "user_host_interrupt": None, "pivot_sign": None, "user_schur_indices": None,
        "user_nd_partition_tree": None, "ubatch_mask": None,
    }


def _check(lib, h, rc):
    if rc == _abi.SDS_SUCCESS:
        return
    msg = lib.sds_last_error(h).decode()
    if rc == _abi.SDS_INVALID_VALUE:
        raise InvalidValueError(msg)
    if rc == _abi.SDS_NOT_SUPPORTED:
        raise NotSupportedError(msg)
    if rc == _abi.SDS_FACTORIZATION_ERROR:
        raise FactorizationError(0, msg)
    if rc == _abi.SDS_INTERRUPTED:
        raise InterruptedError_(msg)
    raise RuntimeError(f"libsds_hip: {msg}")


def _torch():
    import torch
    return torch


class DirectSolver:
    """Solver handle for one sparse matrix (src/solver.jl:187-244).

    ``DirectSolver(A, structure, view)`` with ``A`` a scipy CSR matrix (host), or
    ``DirectSolver((rowptr, colval, nzval, n), structure, view, index="O")`` with
    the CSR arrays; ``nzval`` may be a torch tensor on ``cuda`` (used in place)."""

    def __init__(self, A, structure, view="F", index="O"):
        torch = _torch()
        if structure not in STRUCTURES:
            raise InvalidValueError(f'structure must be "G", "S", "H", "SPD" or "HPD", got {structure!r}')
        if view not in VIEWS:
            raise InvalidValueError(f"view must be 'L', 'U' or 'F', got {view!r}")
        if isinstance(A, tuple):
            rowptr, colval, nzval, n = A
            base = 1 if index == "O" else 0
        else:
            A = A.tocsr()
            A.sort_indices()
            if A.shape[0] != A.shape[1]:
                raise InvalidValueError(f"the matrix must be square, got {A.shape[0]} × {A.shape[1]}")
            rowptr, colval, nzval, n = A.indptr + 1, A.indices + 1, A.data, A.shape[0]
            base, index = 1, "O"
        self.n = int(n)
        self.rowptr = np.asarray(rowptr, dtype=np.int64)
        self.colval = np.asarray(colval, dtype=np.int64)
        self.index = index
        # uniform batch (src/solver.jl:135-150): nbatch matrices with this pattern, values member
        # after member (a vector of nbatch nnz entries or an nnz x nbatch matrix); the members are
        # factored as the blocks of one block-diagonal matrix whose elimination forest holds
        # nbatch copies of the member's analysis, so each level is one set of launches for all
        nnz = self.colval.shape[0]
        if getattr(nzval, "ndim", 1) == 2:
            if nzval.shape[0] != nnz:
                raise InvalidValueError(f"nzval is {tuple(nzval.shape)}, the pattern has {nnz} entries")
            nzval = nzval.t().reshape(-1) if isinstance(nzval, torch.Tensor) else np.asarray(nzval).T.reshape(-1)
        count = int(nzval.numel() if isinstance(nzval, torch.Tensor) else np.size(nzval))
        if nnz == 0 or count % nnz != 0 or count == 0:
            raise InvalidValueError(f"nzval has {count} entries, the pattern has {nnz}")
        self.nbatch = count // nnz
        self.n_member = self.n
        if self.nbatch > 1:
            nb, b = self.nbatch, (1 if index == "O" else 0)
            rp0 = self.rowptr - b
            self.rowptr = np.concatenate([rp0[:-1] + k * nnz for k in range(nb)] + [[nb * nnz]]) + b
            self.colval = np.concatenate([self.colval + k * self.n for k in range(nb)])
            self.n = self.n * nb
            self._member_pattern = (rp0 + b, self.colval[:nnz])
        self._base = base
        self.structure = structure
        self.view = view
        self.options = _defaults()
        self.stage = STAGE_NONE
        self.info = 0
        self.ir_steps = -1
        self.symbolic = None
        self._h = None
        self._lib = None
        self._keep = []
        self._analysis = None
        self._set_values(nzval)

    def _set_values(self, nzval):
        torch = _torch()
        if getattr(nzval, "ndim", 1) == 2:  # nnz x nbatch: member after member
            nzval = nzval.t().reshape(-1) if isinstance(nzval, torch.Tensor) else np.asarray(nzval).T.reshape(-1)
        if isinstance(nzval, torch.Tensor):
            if nzval.dtype != torch.float64:
                raise NotSupportedError("matrix values must be Float64 (factor_precision selects an FP32 factor)")
            self.nzval = nzval if nzval.is_cuda else nzval.cuda()
        else:
            self.nzval = torch.as_tensor(np.ascontiguousarray(nzval, dtype=np.float64)).cuda()
        if self.nzval.numel() != self.colval.shape[0]:
            raise InvalidValueError(f"nzval has {self.nzval.numel()} entries, the pattern has {self.colval.shape[0]}")

    def __repr__(self):
        st = ("created", "reordered", "analyzed", "factorized")[self.stage]
        return f'DirectSolver(n = {self.n}, nnz = {self.colval.shape[0]}, structure = "{self.structure}", ' \
               f"view = '{self.view}', {st})"

    def __del__(self):
        if self._h is not None and self._lib is not None:
            try:
                self._lib.sds_destroy(self._h)
            except Exception:
                pass
            self._h = None

    # -- phases ---------------------------------------------------------------------------------------
    def _reorder(self):
        o = self.options
        if self.structure not in ("SPD", "HPD", "S", "H", "G"):
            raise NotSupportedError(f'structure "{self.structure}" is not available in this build')
        schur = None
        if o["schur_mode"] != 0:
            # Schur complement mode (src/solver.jl:400-416): the flagged rows and columns last, one
            # dense root supernode that the factorization assembles without factoring
            if self.nbatch > 1 or self._analysis is not None:
                raise NotSupportedError("schur_mode = 1 is not supported for batches")
            schur = o["user_schur_indices"]
            if schur is None:
                raise InvalidValueError("schur_mode = 1 needs user_schur_indices (n flags 0/1)")
            schur = np.ascontiguousarray(schur, dtype=np.int8)
            if schur.shape != (self.n,) or np.any((schur != 0) & (schur != 1)) or not schur.any():
                raise InvalidValueError("user_schur_indices must hold n flags 0/1, at least one of them 1")
        # matching and scaling (src/solver.jl:543-579), from the values at analysis
        self.matching = None
        if o["matching_alg"] not in ("default", "none"):
            job = MATCHING_JOBS.get(o["matching_alg"])
            if job is None:
                raise InvalidValueError(f"matching_alg = {o['matching_alg']!r} is unknown")
            if schur is not None:
                raise NotSupportedError(f"matching_alg = {o['matching_alg']!r} is not supported with schur_mode = 1")
            if self.nbatch > 1 or self._analysis is not None:
                raise NotSupportedError("matching is not supported for batches in this build")
            self.matching = compute_matching(self.rowptr, self.colval, self._host_values(), self.n, self.structure,
                                             self.view, self.index, job)
        if self._analysis is not None:
            # an analysis assembled elsewhere (the block-diagonal forest of an uneven batch)
            self.symbolic = self._analysis(o)
            self.stage = STAGE_REORDERED
            return
        if o["ubatch_size"] not in (0, self.nbatch):
            raise InvalidValueError(f"ubatch_size = {o['ubatch_size']}, but the matrix values hold {self.nbatch} "
                                    "batch members")
        ordering = o["reordering_alg"] if o["reordering_alg"] in ("default", "algo1", "algo2", "algo3", "algo4",
                                                                  "algo5") else None
        if ordering is None and o["user_perm"] is None:
            raise NotSupportedError(f"reordering_alg = {o['reordering_alg']!r} is not available in this build")
        if self.nbatch > 1:
            rp, cv = self._member_pattern
            S1 = symbolic_analysis(rp, cv, self.n_member, self.structure, self.view, self.index, perm=o["user_perm"],
                                   ordering=ordering or "default", amalgamation=o["amalgamation"],
                                   use_superpanels=o["use_superpanels"])
            self.symbolic = merge_symbolic([S1] * self.nbatch)
            self._ns_member = S1.ns
        else:
            m = self.matching
            # 2x2 pivot pairs (src/symbolic/pairs.jl:27-35): "S"/"H", the AMD ordering, no user_perm
            pp = o["pivot_pairs"]
            pairs_on = (self.structure in ("S", "H") and o["user_perm"] is None and ordering != "algo5"
                        and pp != "none" and schur is None)
            colval, values, fixed = self.colval, None, None
            if m is not None and self.structure == "G":
                # the analysis runs on M = (Dr A Dc)[:, q]: column c becomes column q^-1(c)
                qinv = np.empty(self.n, dtype=np.int64)
                qinv[m.perm - 1] = np.arange(1, self.n + 1)
                base = 0 if self.index in ("Z", 0) else 1
                colval = qinv[np.asarray(self.colval, dtype=np.int64) - base] - 1 + base
            elif m is not None and m.job == 5 and pairs_on:
                # pairs from the cycles of the scaled symmetric matching (src/solver.jl:560-566)
                fixed = matching_pairs(m, self.rowptr, self.colval, self._host_values(), self.n, self.structure,
                                       self.view, self.index, pp, o["pivot_pair_tolerance"])
            elif pairs_on:
                values = self._host_values()
            self.symbolic = symbolic_analysis(self.rowptr, colval, self.n, self.structure, self.view,
                                              self.index, perm=o["user_perm"], ordering=ordering or "default",
                                              amalgamation=o["amalgamation"], use_superpanels=o["use_superpanels"],
                                              schur=schur, values=values, pivot_pairs=pp if pairs_on else "none",
                                              pair_tolerance=o["pivot_pair_tolerance"],
                                              pivot_threshold=o["pivot_threshold"], fixed_pairs=fixed)
            if m is not None and self.structure == "G":
                # the refinement residual and the adjoint scatter read A itself, in its own numbering
                fr, fc, fs = full_maps(self.rowptr, self.colval, self.n, self.structure, self.view, self.index)
                self.symbolic.maps.update(full_rowptr=fr, full_colval=fc, full_src=fs)
        self.stage = STAGE_REORDERED

    def _host_values(self):
        """The first batch member's values on the host: the analysis reads them once (pivot pairs,
        matching; src/solver.jl:545-572)."""
        return self.nzval[:len(self.colval)].detach().cpu().numpy().astype(np.float64)

    def _symbolic_factorization(self):
        if self.stage < STAGE_REORDERED:
            raise FactorizationError(0, 'phase "symbolic_factorization": needs "reordering"')
        lib = hip_lib()
        if self._h is None:
            h = ctypes.c_void_p()
            torch = _torch()
            stream = torch.cuda.current_stream().cuda_stream
            _check(lib, None, lib.sds_create(ctypes.byref(h), ctypes.c_void_p(stream)))
            self._h, self._lib = h, lib
        S = self.symbolic
        keep = {k: np.ascontiguousarray(S.maps[k], dtype=np.int64) for k in S.maps}
        d = _abi.SymbolicDesc()
        d.n, d.nnz, d.ns = S.n, S.nnz, S.ns
        d.structure = STRUCTURES[self.structure]
        d.index_bits = 64
        for k in ("perm", "super_ptr", "snparent", "rowptr", "rowval", "front_ptr", "front_nrows", "front_ncols",
                  "child_ptr", "child_list", "relind_ptr", "relind", "amap", "amap_ptr", "amap_src"):
            setattr(d, k, keep[k].ctypes.data)
        d.full_nrows = keep["full_rowptr"].shape[0] - 1
        d.full_nnz = keep["full_colval"].shape[0]
        d.full_rowptr, d.full_colval, d.full_src = (keep["full_rowptr"].ctypes.data, keep["full_colval"].ctypes.data,
                                                    keep["full_src"].ctypes.data)
        _check(lib, self._h, lib.sds_upload_symbolic(self._h, ctypes.byref(d)))
        _check(lib, self._h, lib.sds_set_schur(self._h, int(getattr(S, "nschur", 0))))
        m = getattr(self, "matching", None)
        if m is not None and (not m.symmetric or m.job == 5):
            # jobs 1-4 of the symmetric structures have no scaling: nothing to apply
            w = np.ascontiguousarray(entry_scaling(m, self.rowptr, self.colval, self.n, self.index))
            rs = np.ascontiguousarray(m.rscale, dtype=np.float64)
            cs = np.ascontiguousarray(m.cscale, dtype=np.float64)
            q = None if m.symmetric else np.ascontiguousarray(m.perm, dtype=np.int64)
            _check(lib, self._h, lib.sds_set_matching(self._h, w.ctypes.data, rs.ctypes.data, cs.ctypes.data,
                                                      q.ctypes.data if q is not None else None))
        fp = self.options["factor_precision"]
        dtype = _abi.SDS_R32 if fp in ("Float32", "float32", np.float32) else _abi.SDS_R64
        _check(lib, self._h, lib.sds_alloc_numeric(self._h, dtype, 1))
        self._dtype = dtype
        self.stage = STAGE_ANALYZED

    def _factorize(self, phase):
        if self.stage < STAGE_ANALYZED:
            raise FactorizationError(0, f'phase "{phase}": needs "analysis"')
        if phase == "refactorization" and self.stage < STAGE_FACTORIZED:
            raise FactorizationError(0, 'phase "refactorization": needs a previous "factorization"')
        p, keep = self._factor_params()
        info = ctypes.c_int64(0)
        vals = self.nzval
        members = self._members()
        if self.nbatch > 1 and len(members) < self.nbatch:
            # the inactive members keep their factors: they are factored again from the values of
            # their last factorization, which gives the same factor bit for bit
            if getattr(self, "_fact_vals", None) is None:
                raise InvalidValueError("ubatch_index / ubatch_mask need a factorization of every member first")
            nnz = self.nzval.numel() // self.nbatch
            vals = self._fact_vals.clone()
            for k in members:
                vals[k * nnz:(k + 1) * nnz] = self.nzval[k * nnz:(k + 1) * nnz]
        self.stage = STAGE_ANALYZED
        _check(self._lib, self._h, self._lib.sds_factorize(self._h, ctypes.c_void_p(vals.data_ptr()),
                                                            ctypes.byref(p), ctypes.byref(info)))
        del keep
        self.info = int(info.value)
        if self.nbatch > 1:
            self._fact_vals = vals.clone()
        self.stage = STAGE_FACTORIZED
        if self.nbatch > 1:
            self.info = self._member_info()

    def _member_stats(self):
        """Per-member sums of the fronts' (npos, nneg, nzero, nperturbed, n2x2) and each member's
        info: 0 or its original column (1-based) of its first failed pivot."""
        S = self.symbolic
        st = np.zeros(6 * S.ns, dtype=np.int64)
        _check(self._lib, self._h, self._lib.sds_get_stats(self._h, st.ctypes.data, None))
        st = st.reshape(S.ns, 6)
        ns1, out = self._ns_member, []
        for k in range(self.nbatch):
            part = st[k * ns1:(k + 1) * ns1]
            fi = part[:, 5]
            bad = np.nonzero(fi > 0)[0]
            info = 0
            if bad.size:
                col = int(np.min(S.super_ptr[k * ns1 + bad] - 1 + fi[bad]))  # 1-based factor column
                info = int(S.perm[col - 1]) - k * self.n_member
            out.append((part[:, :5].sum(axis=0), info))
        return out

    def _member_info(self):
        return [info for _, info in self._member_stats()]

    def _members(self):
        """The batch members the next numeric or solve phase processes (ubatch_index, ubatch_mask;
        src/solver.jl:694-707), 0-based."""
        o = self.options
        if self.nbatch == 1:
            return [0]
        idx, mask = int(o["ubatch_index"]), o["ubatch_mask"]
        if idx < -1 or idx >= self.nbatch:
            raise InvalidValueError(f"ubatch_index must be -1 or a member index below {self.nbatch}")
        members = list(range(self.nbatch)) if idx == -1 else [idx]
        if mask is not None:
            mask = np.asarray(mask)
            if mask.shape != (self.nbatch,) or np.any((mask != 0) & (mask != 1)):
                raise InvalidValueError(f"ubatch_mask must hold {self.nbatch} flags in {{0, 1}}")
            members = [k for k in members if mask[k]]
        if not members:
            raise InvalidValueError("ubatch_index and ubatch_mask select no batch member")
        return members

    def _factor_params(self):
        """(sds_factor_params, objects to keep alive during the call)"""
        o = self.options
        p = _abi.FactorParams()
        p.pivot_type = PIVOT_TYPES[o["pivot_type"]]
        p.pivot_epsilon_alg = _EPS_ALGS[o["pivot_epsilon_alg"]]
        p.pivot_threshold = float(o["pivot_threshold"])
        p.pivot_epsilon = -1.0 if o["pivot_epsilon"] is None else float(o["pivot_epsilon"])
        flag = o["user_host_interrupt"]
        if flag is not None and not isinstance(flag, (ctypes.c_bool, ctypes.c_uint8)):
            raise InvalidValueError("user_host_interrupt must be a one-byte flag (ctypes.c_bool or c_uint8) or None")
        p.interrupt = ctypes.addressof(flag) if flag is not None else None
        sign = o["pivot_sign"]
        if sign is not None:
            sign = np.ascontiguousarray(sign, dtype=np.int8)
            if sign.shape != (self.n_member,) or np.any(np.abs(sign) > 1):
                raise InvalidValueError("pivot_sign must hold n entries in {-1, 0, 1}")
            sign = np.ascontiguousarray(np.tile(sign, self.nbatch))  # the same signs for every member
            p.pivot_sign = sign.ctypes.data
        return p, (sign, flag)

    def _solve(self, phase, X, B):
        if self.stage < STAGE_FACTORIZED:
            raise FactorizationError(0, f'phase "{phase}": needs a "factorization"')
        if phase in ("solve_fwd_schur", "solve_bwd_schur") and not getattr(self.symbolic, "nschur", 0):
            raise FactorizationError(0, f'phase "{phase}" needs an analysis with schur_mode = 1')
        if self.nbatch > 1:
            return self._solve_batch(phase, X, B)
        return self._solve_one(phase, X, B)

    def _solve_batch(self, phase, X, B):
        """Uniform batch: right-hand sides and solutions hold nrhs columns per member, member after
        member (an n x (nrhs nbatch) matrix, an n nrhs nbatch vector or an n x nrhs x nbatch array,
        src/solver.jl:142-147); the inactive members' solutions are left untouched."""
        n, nb = self.n_member, self.nbatch

        def merged(T, what):
            if T.dim() == 1 and T.numel() % (n * nb) == 0:
                r = T.numel() // (n * nb)
                return T.reshape(nb, r, n).permute(0, 2, 1).reshape(nb * n, r), r
            if T.dim() == 2 and T.shape[0] == n and T.shape[1] % nb == 0:
                r = T.shape[1] // nb
                return T.reshape(n, nb, r).permute(1, 0, 2).reshape(nb * n, r), r
            if T.dim() == 3 and T.shape[0] == n and T.shape[2] == nb:
                return T.permute(2, 0, 1).reshape(nb * n, T.shape[1]), T.shape[1]
            raise InvalidValueError(f"{what} is {tuple(T.shape)}: a batch of {nb} members of order {n} needs an "
                                    "n x (nrhs nbatch) matrix, an n nrhs nbatch vector or an n x nrhs x nbatch array")
        Bm, r = merged(B, "B")
        Xm, rx = merged(X, "X")
        if rx != r:
            raise InvalidValueError(f"X holds {rx} columns per member, B {r}")
        Xm = Xm.contiguous().clone()
        self._solve_one(phase, Xm, Bm.contiguous())
        for k in self._members():
            xk = Xm[k * n:(k + 1) * n]
            if X.dim() == 1:
                X.view(nb, r, n)[k].copy_(xk.t())
            elif X.dim() == 2:
                X.view(n, nb, r)[:, k].copy_(xk)
            else:
                X[:, :, k].copy_(xk)

    def _solve_one(self, phase, X, B):
        torch = _torch()
        bits = PHASES[phase]
        vec = B.dim() == 1
        nrhs = 1 if vec else B.shape[1]
        if B.shape[0] != self.n or X.shape[0] != self.n or (not vec and X.shape[1] != nrhs):
            raise InvalidValueError(f"A is {self.n}×{self.n}, X is {tuple(X.shape)}, B is {tuple(B.shape)}")
        # column-major n x nrhs views for the C ABI
        Bc = B.reshape(self.n, nrhs).t().contiguous().to(torch.float64)
        Xc = X.reshape(self.n, nrhs).t().contiguous().to(torch.float64)
        o = self.options
        steps = ctypes.c_int32(0)
        nsteps = int(o["ir_n_steps"]) if bits & PHASES["solve_refinement"] else 0
        _check(self._lib, self._h, self._lib.sds_set_ir_mode(self._h, IR_MODES[o["ir_mode"]]))
        _check(self._lib, self._h, self._lib.sds_solve(
            self._h, bits, int(o["solve_mode"]), ctypes.c_void_p(Bc.data_ptr()), self.n,
            ctypes.c_void_p(Xc.data_ptr()), self.n, nrhs, ctypes.c_void_p(self.nzval.data_ptr()), nsteps,
            float(o["ir_tol"]), ctypes.byref(steps)))
        if bits & PHASES["solve_refinement"]:
            self.ir_steps = int(steps.value)
        X.copy_(Xc.t().reshape(X.shape))


def execute(phase, solver, X=None, B=None, asynchronous=True):
    """Run `phase` on `solver` (src/solver.jl:436-460)."""
    if phase not in PHASES:
        raise ValueError(f'unknown phase "{phase}"; expected one of {", ".join(repr(p) for p in PHASES)}')
    if phase == "reordering":
        solver._reorder()
    elif phase == "symbolic_factorization":
        solver._symbolic_factorization()
    elif phase == "analysis":
        solver._reorder()
        solver._symbolic_factorization()
    elif phase in ("factorization", "refactorization"):
        solver._factorize(phase)
    else:
        solver._solve(phase, X, B)
    if not asynchronous:
        _torch().cuda.synchronize()


def graph_capture(solver, nrhs=1, phase="solve"):
    """Capture one refactorization + `phase` of the solver into a HIP graph (sds_graph_capture).

    The graph reads the solver's current values tensor and two device buffers the solver keeps
    (n x nrhs right-hand sides and solutions); graph_replay copies B in, replays and copies X
    out. The values are read at each replay: update them in place (solver.nzval) between
    replays, or call update() and capture again."""
    torch = _torch()
    if solver.stage < STAGE_ANALYZED:
        raise FactorizationError(0, "graph_capture needs the analysis")
    n = solver.n
    solver._gB = torch.zeros((nrhs, n), dtype=torch.float64, device=solver.nzval.device)
    solver._gX = torch.zeros((nrhs, n), dtype=torch.float64, device=solver.nzval.device)
    p, keep = solver._factor_params()
    o = solver.options
    nsteps = int(o["ir_n_steps"]) if PHASES[phase] & PHASES["solve_refinement"] else 0
    _check(solver._lib, solver._h, solver._lib.sds_set_ir_mode(solver._h, IR_MODES[o["ir_mode"]]))
    _check(solver._lib, solver._h, solver._lib.sds_graph_capture(
        solver._h, ctypes.c_void_p(solver.nzval.data_ptr()), ctypes.byref(p), PHASES[phase], int(o["solve_mode"]),
        ctypes.c_void_p(solver._gB.data_ptr()), n, ctypes.c_void_p(solver._gX.data_ptr()), n, nrhs, nsteps))
    del keep
    solver._gnrhs = nrhs


def graph_replay(solver, X=None, B=None, check_info=False):
    """Replay the captured graph: B (n x nrhs or n) is copied into the graph's buffer first and
    the solution into X afterwards (both optional). check_info: wait and return info."""
    n, nrhs = solver.n, solver._gnrhs
    if B is not None:
        solver._gB.copy_(B.reshape(n, nrhs).t())
    _check(solver._lib, solver._h, solver._lib.sds_graph_launch(solver._h))
    solver.stage = STAGE_FACTORIZED
    if X is not None:
        X.copy_(solver._gX.t().reshape(X.shape))
    if check_info:
        info = ctypes.c_int64(0)
        _check(solver._lib, solver._h, solver._lib.sds_graph_info(solver._h, ctypes.byref(info)))
        solver.info = int(info.value)
        return solver.info
    return None


def update(solver, nzval):
    """Point the solver at new values with the same pattern (src/solver.jl:309-321)."""
    solver._set_values(nzval)


def autotune(solver, reps=2):
    """Choose the front kernels' tile shapes for this analysis by timing `reps` factorizations of
    the current values per candidate (libsds_hip's sds_autotune; no counterpart in the reference).
    Returns {"extend_add_rows": r, "cb_k_chunk": kc, "trailing_split": t, "step_two_per_cu": o,
    "small_front_kernel": k, "subtree_kib": b, "panel_block": p}; the solver then needs a
    "factorization" again."""
    if solver.stage < STAGE_ANALYZED:
        raise FactorizationError(0, 'autotune needs "analysis"')
    p, keep = solver._factor_params()
    tiles = (ctypes.c_int32 * 8)()
    _check(solver._lib, solver._h, solver._lib.sds_autotune(solver._h, ctypes.c_void_p(solver.nzval.data_ptr()),
                                                             ctypes.byref(p), int(reps), tiles))
    del keep
    solver.stage = STAGE_ANALYZED
    return {"extend_add_rows": int(tiles[0]), "cb_k_chunk": int(tiles[1]), "trailing_split": int(tiles[2]),
            "step_two_per_cu": int(tiles[3]), "small_front_kernel": int(tiles[4]), "subtree_kib": int(tiles[5]),
            "panel_block": int(tiles[6]), "pivot_panels": int(tiles[7])}


def set_tiles(solver, tiles):
    """Set the front kernels' tile shapes as `autotune` returns them (a dict with the keys
    "extend_add_rows", "cb_k_chunk", "trailing_split", "step_two_per_cu" and optionally
    "small_front_kernel", "subtree_kib", "panel_block" and "pivot_panels", default 0;
    libsds_hip's sds_set_tiles) and reallocate the numeric storage for them: the solver then needs a
    "factorization" again."""
    if solver.stage < STAGE_ANALYZED:
        raise FactorizationError(0, 'set_tiles needs "analysis"')
    t = (ctypes.c_int32 * 8)(int(tiles["extend_add_rows"]), int(tiles["cb_k_chunk"]), int(tiles["trailing_split"]),
                             int(tiles["step_two_per_cu"]), int(tiles.get("small_front_kernel", 0)),
                             int(tiles.get("subtree_kib", 0)), int(tiles.get("panel_block", 0)),
                             int(tiles.get("pivot_panels", 0)))
    _check(solver._lib, solver._h, solver._lib.sds_set_tiles(solver._h, t))
    fp = solver.options["factor_precision"]
    dtype = _abi.SDS_R32 if fp in ("Float32", "float32", np.float32) else _abi.SDS_R64
    _check(solver._lib, solver._h, solver._lib.sds_alloc_numeric(solver._h, dtype, 1))
    solver.stage = STAGE_ANALYZED


def setparam(solver, name, value):
    """Set a configuration or user data parameter (src/options.jl:610)."""
    if name not in CONFIG_PARAMETERS + EXTRA_CONFIG + USER_DATA:
        if name in DATA_PARAMETERS:
            raise InvalidValueError(f'"{name}" is a data parameter computed by the solver; it cannot be set')
        raise ValueError(f'unknown parameter "{name}"')
    o = solver.options
    if name == "pivot_type":
        if value not in PIVOT_TYPES:
            raise InvalidValueError(f"pivot_type must be one of {', '.join(PIVOT_TYPES)}, got {value!r}")
        if value in ("C", "R"):
            raise NotSupportedError("global pivoting ('C', 'R') is not supported")
    elif name in ("ir_n_steps",) and (not isinstance(value, (int, np.integer)) or value < 0):
        raise InvalidValueError("ir_n_steps must be a nonnegative integer")
    elif name == "ir_mode" and value not in IR_MODES:
        raise InvalidValueError(f"ir_mode must be one of {', '.join(IR_MODES)}, got {value!r}")
    elif name in ("ir_tol", "pivot_threshold") and not (float(value) >= 0):
        raise InvalidValueError(f"{name} must be nonnegative")
    elif name == "pivot_epsilon" and value is not None and not (float(value) >= 0):
        raise InvalidValueError("pivot_epsilon must be nonnegative")
    elif name == "pivot_epsilon_alg" and value not in _EPS_ALGS:
        raise InvalidValueError(f"pivot_epsilon_alg must be one of {', '.join(_EPS_ALGS)}")
    elif name == "solve_mode" and value not in (0, 1, 2):
        raise InvalidValueError("solve_mode must be 0, 1 or 2")
    elif name == "user_perm" and value is not None:
        v = np.asarray(value, dtype=np.int64)
        if v.shape != (solver.n_member,) or not np.array_equal(np.sort(v), np.arange(1, solver.n_member + 1)):
            raise InvalidValueError("user_perm must be a permutation of 1:n")
        value = v
    elif name == "amalgamation":
        mw, zf, mn = value
        if int(mw) < 1 or not (0 <= float(zf)) or int(mn) < 0:
            raise InvalidValueError("amalgamation must be (max_width ≥ 1, zero_fraction ≥ 0, min_width ≥ 0)")
        value = (int(mw), float(zf), int(mn))
    elif name == "factor_precision" and value not in (None, "Float64", "Float32", np.float32, np.float64):
        raise InvalidValueError("factor_precision must be Float64 or Float32")
    o[name] = value


def getparam(solver, name):
    """Read a data or configuration parameter (src/options.jl:648, src/solver.jl getparam)."""
    if name in solver.options and name not in ("ir_n_steps", "user_perm"):
        return solver.options[name]
    if name == "user_perm":
        return solver.options["user_perm"]
    if name == "ir_n_steps":
        return solver.ir_steps
    if name == "info":
        return solver.info
    S = solver.symbolic
    if name in ("perm_reorder_row", "perm_reorder_col", "perm_row", "perm_col"):
        if S is None:
            raise FactorizationError(0, f'"{name}" needs the "reordering" phase')
        return S.perm.copy()
    if name == "nsuperpanels":
        return S.ns
    if name == "flops":
        return S.flops
    if name == "lu_nnz":
        return S.nnz_stored
    if name in ("schur_shape", "schur_matrix") and (S is None or not getattr(S, "nschur", 0)):
        raise InvalidValueError(f'the data parameter "{name}" needs an analysis with schur_mode = 1')
    if name == "schur_shape":
        # (rows, columns, nonzeros of the pattern): the analysis keeps S dense, one triangle with the
        # diagonal for the symmetric structures (src/solver.jl:1067)
        ns = S.nschur
        return (ns, ns, ns * ns if solver.structure == "G" else ns * (ns + 1) // 2)
    if solver.stage < STAGE_ANALYZED:
        raise FactorizationError(0, f'"{name}" needs the analysis')
    lib, h = solver._lib, solver._h
    if name == "memory_estimates":
        est = np.zeros(16, dtype=np.int64)
        _check(lib, h, lib.sds_memory_estimates(h, est.ctypes.data))
        return est
    if solver.stage < STAGE_FACTORIZED:
        raise FactorizationError(0, f'"{name}" needs a "factorization"')
    if name in ("inertia", "npivots", "pivot_stats") and solver.nbatch > 1:
        # uniform batch: one entry per member (src/solver.jl:1070-1074)
        out = []
        for t, _ in solver._member_stats():
            out.append((int(t[0]), int(t[1])) if name == "inertia" else int(t[3]) if name == "npivots" else
                       dict(npos=int(t[0]), nneg=int(t[1]), nzero=int(t[2]), nperturbed=int(t[3]), n2x2=int(t[4])))
        return out
    if name in ("inertia", "npivots", "pivot_stats"):
        tot = np.zeros(6, dtype=np.int64)
        _check(lib, h, lib.sds_get_stats(h, None, tot.ctypes.data))
        if name == "inertia":
            return (int(tot[0]), int(tot[1]))
        if name == "npivots":
            return int(tot[3])
        return dict(npos=int(tot[0]), nneg=int(tot[1]), nzero=int(tot[2]), nperturbed=int(tot[3]), n2x2=int(tot[4]))
    if name == "schur_matrix":
        # S = A22 - A21 A11^-1 A12 of the flagged rows and columns in increasing order, a new dense
        # ns x ns device matrix (src/solver.jl:1068)
        torch = _torch()
        ns = S.nschur
        buf = torch.empty((ns, ns), dtype=torch.float64, device="cuda")
        _check(lib, h, lib.sds_schur_export(h, ctypes.c_void_p(buf.data_ptr()), ns))
        return buf.t().contiguous()  # the export is column-major
    if name in ("perm_matching", "scale_row", "scale_col"):
        m = getattr(solver, "matching", None)
        if m is None:
            raise InvalidValueError(f'"{name}" needs a matching (matching_alg other than "default" at the analysis)')
        return {"perm_matching": m.perm, "scale_row": m.rscale, "scale_col": m.cscale}[name].copy()
    if name == "diag":
        d = np.zeros(2 * solver.n, dtype=np.float64 if solver._dtype == _abi.SDS_R64 else np.float32)
        _check(lib, h, lib.sds_get_diag(h, d.ctypes.data))
        return d[:solver.n].astype(np.float64)
    raise ValueError(f'unknown parameter "{name}"')


# test and debugging glue (the reference's extract.jl host copies)
def numeric_factor(solver, which=0):
    """Host copy of the factor panels."""
    S = solver.symbolic
    a = np.zeros(S.factor_len, dtype=np.float64 if solver._dtype == _abi.SDS_R64 else np.float32)
    _check(solver._lib, solver._h, solver._lib.sds_get_factor(solver._h, which, a.ctypes.data))
    return a


def numeric_pivots(solver):
    """(piv, kind): the local pivot order as 1-based factor columns and the pivot kinds."""
    piv = np.zeros(solver.n, dtype=np.int32)
    kind = np.zeros(solver.n, dtype=np.int8)
    _check(solver._lib, solver._h, solver._lib.sds_get_pivots(solver._h, piv.ctypes.data, kind.ctypes.data))
    return piv, kind


def numeric_diag(solver):
    """D as stored: 2n entries (the diagonal, then the 2x2 subdiagonals at the first position)."""
    d = np.zeros(2 * solver.n, dtype=np.float64 if solver._dtype == _abi.SDS_R64 else np.float32)
    _check(solver._lib, solver._h, solver._lib.sds_get_diag(solver._h, d.ctypes.data))
    return d


def numeric_stats(solver):
    st = np.zeros(6 * solver.symbolic.ns, dtype=np.int64)
    tot = np.zeros(6, dtype=np.int64)
    _check(solver._lib, solver._h, solver._lib.sds_get_stats(solver._h, st.ctypes.data, tot.ctypes.data))
    return st, tot


def profile(solver, enable=True):
    """Start (clearing the totals) or stop per-kernel-class timing on the solver's stream."""
    _check(solver._lib, solver._h, solver._lib.sds_profile(solver._h, 1 if enable else 0))


def profile_read(solver):
    """{kind: (ms, launches, flops, bytes)} accumulated since profile(solver, True)."""
    out = {}
    for k, name in enumerate(_abi.PROFILE_KINDS):
        ms, n, fl, by = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _check(solver._lib, solver._h, solver._lib.sds_profile_read(solver._h, k, ctypes.byref(ms), ctypes.byref(n),
                                                                     ctypes.byref(fl), ctypes.byref(by)))
        if n.value:
            out[name] = (ms.value, n.value, fl.value, by.value)
    return out
