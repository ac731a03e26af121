"""Uneven batches: many independent systems of different sizes factored and solved together.

The members form one block-diagonal matrix whose elimination forest is the union of the members'
trees (the reference's plan for non-uniform batches, PLAN.md:635-639; its uniform batch layouts
are src/numeric/batch.jl). Each member is analysed on its own (ordering, etree, supernodes and
amalgamation exactly as if it were solved alone; the amalgamation's zero budget is per matrix)
and the analyses are concatenated into the forest's. Every tree level of the forest holds the
fronts of all members at that height, so each level of the factorization and of the sweeps is
one set of launches for the whole batch, whatever the members' sizes; a member's factor and
solution are bitwise those of solving it by itself.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.sparse as sp

from .analysis import merge_symbolic, symbolic_analysis
from .solver import DirectSolver, execute, getparam, setparam


def _torch():
    import torch
    return torch


class UnevenBatch:
    """``UnevenBatch(members, structure, view="L", perms=None)``: ``members`` are scipy CSR
    matrices (each stored as `view` says), ``perms`` optional per-member 1-based fill-reducing
    orderings (default: the analysis' own ordering of the block-diagonal matrix, which orders
    each member independently)."""

    def __init__(self, members, structure, view="L", perms=None):
        self.sizes = np.array([m.shape[0] for m in members], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        A = sp.block_diag([m.tocsr() for m in members], format="csr")
        A.sort_indices()
        self.A = A
        self.n = int(A.shape[0])
        self.solver = DirectSolver(A, structure, view)
        self.members = [m.tocsr() for m in members]
        for m in self.members:
            m.sort_indices()
        self.perms = perms
        self.solver._analysis = self._analyze

    def _analyze(self, o):
        ordering = {"default": "default", "algo3": "algo3", "algo5": "algo5"}.get(o["reordering_alg"], "default")

        def one(k):
            m = self.members[k]
            return symbolic_analysis(m.indptr + 1, m.indices + 1, m.shape[0], self.solver.structure,
                                     self.solver.view, "O", perm=None if self.perms is None else self.perms[k],
                                     ordering=ordering, amalgamation=o["amalgamation"],
                                     use_superpanels=o["use_superpanels"])
        with ThreadPoolExecutor(max_workers=8) as ex:
            parts = list(ex.map(one, range(len(self.members))))
        return merge_symbolic(parts)

    def __len__(self):
        return len(self.sizes)

    def setparam(self, name, value):
        setparam(self.solver, name, value)

    def getparam(self, name):
        return getparam(self.solver, name)

    def execute(self, phase, X=None, B=None):
        execute(phase, self.solver, X, B)

    def stack(self, parts):
        """One (sum n) x nrhs device tensor from the members' right-hand sides."""
        torch = _torch()
        return torch.cat([torch.as_tensor(p, dtype=torch.float64).to("cuda") for p in parts], dim=0)

    def split(self, X):
        """The members' slices of a stacked (sum n) x nrhs tensor (views)."""
        return [X[self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.sizes))]
