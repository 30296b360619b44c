"""Batched independent Euler cases over the fvk_batch_* C ABI.

All cases share the numerics (fvk_params) and are stepped by the same kernel
launches; each has its own mesh, boundary conditions and initial state, and
its own residual history.  `run_to_convergence` freezes every case once its
density residual has dropped by the requested factor from its first-step
value, as the batched-airfoil workload specifies.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import abi
from .solver import _check, lib


def _setup(L):
    if getattr(L, "_batch_ready", False):
        return
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.fvk_batch_create.argtypes = [C.c_int32, vp, vp, vp, C.POINTER(abi.ParamsC), vp, C.POINTER(vp)]
    L.fvk_batch_destroy.argtypes = [vp]
    L.fvk_batch_destroy.restype = None
    L.fvk_batch_step.argtypes = [vp, C.c_int32, dp]
    L.fvk_batch_set_active.argtypes = [vp, vp]
    L.fvk_batch_set_graph.argtypes = [vp, C.c_int32]
    L.fvk_batch_get_field.argtypes = [vp, C.c_int32, dp]
    L.fvk_batch_solver.argtypes = [vp]
    L.fvk_batch_solver.restype = vp
    L._batch_ready = True


class FVKBatch:
    def __init__(self, meshes: Sequence, bcs: Sequence[Sequence[abi.BCC]], params: abi.ParamsC,
                 ics: Sequence[abi.InitialConditionC]):
        self.L = lib()
        _setup(self.L)
        self.n_cases = len(meshes)
        self.meshes = list(meshes)
        self._descs = [abi.MeshDesc(m) for m in meshes]
        descs = (abi.MeshDescC * self.n_cases)(*[d.desc for d in self._descs])
        flat = [b for bl in bcs for b in bl]
        bc_arr = (abi.BCC * max(1, len(flat)))(*flat)
        nb = (C.c_int32 * self.n_cases)(*[len(bl) for bl in bcs])
        ic_arr = (abi.InitialConditionC * self.n_cases)(*ics)
        self._keep = (descs, bc_arr, nb, ic_arr)
        self.h = C.c_void_p()
        _check(self.L.fvk_batch_create(self.n_cases, C.cast(descs, C.c_void_p), C.cast(bc_arr, C.c_void_p),
                                       C.cast(nb, C.c_void_p), C.byref(params), C.cast(ic_arr, C.c_void_p),
                                       C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.fvk_batch_destroy(self.h)
            self.h = C.c_void_p()

    def step(self, n: int) -> np.ndarray:
        """n steps; returns the per-case residual norms [n, n_cases, 4]."""
        r = np.empty((n, self.n_cases, 4))
        _check(self.L.fvk_batch_step(self.h, n, abi.as_dptr(r)))
        return r

    def synchronize(self) -> None:
        """Nothing to wait for: a batch step returns with its stream drained."""

    def set_active(self, active) -> None:
        a = np.ascontiguousarray(active, dtype=np.int32)
        _check(self.L.fvk_batch_set_active(self.h, a.ctypes.data))

    def set_graph(self, on: bool) -> None:
        _check(self.L.fvk_batch_set_graph(self.h, 1 if on else 0))

    def field(self, k: int) -> np.ndarray:
        U = np.empty((self.meshes[k].n_cells, 4))
        _check(self.L.fvk_batch_get_field(self.h, k, abi.as_dptr(U)))
        return U

    def solver_handle(self) -> C.c_void_p:
        return C.c_void_p(self.L.fvk_batch_solver(self.h))

    def run_to_convergence(self, drop: float = 1e-6, max_steps: int = 20000, chunk: int = 100,
                           progress=None) -> Dict[str, np.ndarray]:
        """Steps until every case's density residual is below drop x its
        first-step value (or max_steps).  A converged case is frozen at the
        end of the chunk in which it converged.  Returns per case the step
        of convergence (-1: not converged) and the wall time at which it was
        frozen."""
        conv_step = np.full(self.n_cases, -1, dtype=np.int64)
        conv_time = np.full(self.n_cases, np.nan)
        active = np.ones(self.n_cases, dtype=np.int32)
        first = None
        done = 0
        t0 = time.perf_counter()
        while done < max_steps and active.any():
            n = min(chunk, max_steps - done)
            r = self.step(n)
            if first is None:
                first = r[0, :, 0].copy()
            now = time.perf_counter() - t0
            for k in np.nonzero(active)[0]:
                hit = np.nonzero(r[:, k, 0] <= drop * first[k])[0]
                if hit.size:
                    conv_step[k] = done + int(hit[0]) + 1
                    conv_time[k] = now
                    active[k] = 0
            done += n
            self.set_active(active)
            if progress:
                progress(done, int(active.sum()), now)
        return {"steps": conv_step, "time": conv_time, "total_steps": done,
                "wall": time.perf_counter() - t0}
