"""Batches of independent fixed-mesh cases advanced together on one GPU
(cfd_batch in libcfd_amd.so): a sweep over meshes of uneven size and over
Mach number, angle of attack and Reynolds number, one hipGraph replay per
step for every case.  The reference runs one case per process; each case of
a batch gives the same results, bit for bit, as a single-case Session of the
same precision and patch shape."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FIELDS, SCALARS, BcLists, Consts, check, dptr, iptr, lib
from .api import bc_lists, consts_of


class BatchCase(C.Structure):
    _fields_ = [("nelem", C.c_int), ("npoin", C.c_int), ("inpoel", C.POINTER(C.c_int)),
                ("X", C.POINTER(C.c_double)), ("Y", C.POINTER(C.c_double)), ("c", Consts),
                ("bc", BcLists)]


SCALAR_OF = {"DTMIN": "CFD_S_DTMIN", "TIME": "CFD_S_TIME", "DTMIN1": "CFD_S_DTMIN1",
             "BANDERA": "CFD_S_BANDERA", "ITER": "CFD_S_ITER"}
WIDTH = {"U": 4, "U1": 4, "RHS": 4, "DNX": 3, "DNY": 3}


class Batch:
    """cases: mesh.Case objects with fixed meshes.  precision 1 = FP32
    element arithmetic (FP64 node sums and update), 0 = FP64."""

    def __init__(self, cases, precision: int = 1, itlocal: int = 0, patch_elems: int = 0,
                 patch_block: int = 0):
        self.cases = list(cases)
        self._keep = []
        arr = (BatchCase * len(self.cases))()
        for k, cs in enumerate(self.cases):
            if cs.moving:
                raise ValueError("a batch holds fixed-mesh cases")
            m = cs.mesh
            inp = np.ascontiguousarray(m.inpoel, np.int32).reshape(-1)
            X = np.ascontiguousarray(m.x, np.float64)
            Y = np.ascontiguousarray(m.y, np.float64)
            self._keep += [inp, X, Y]
            arr[k].nelem, arr[k].npoin = m.nelem, m.npoin
            arr[k].inpoel, arr[k].X, arr[k].Y = iptr(inp), dptr(X), dptr(Y)
            arr[k].c = consts_of(cs)
            arr[k].bc = bc_lists(cs.lists, self._keep)
        self._h = C.c_void_p()
        check(lib().cfd_batch_create(C.byref(self._h), len(self.cases), arr, int(precision),
                                     int(itlocal), int(patch_elems), int(patch_block)))

    def close(self):
        if self._h:
            check(lib().cfd_batch_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_freestream(self):
        check(lib().cfd_batch_init_freestream(self._h))

    def use_graph(self, on: bool = True):
        check(lib().cfd_batch_set_graph(self._h, int(on)))

    def step(self, n: int = 1):
        check(lib().cfd_batch_step(self._h, int(n)))

    def sync(self):
        check(lib().cfd_batch_sync(self._h))

    def get(self, icase: int, name: str) -> np.ndarray:
        fid = FIELDS["CFD_F_" + name]
        n = C.c_longlong()
        check(lib().cfd_batch_case_size(self._h, int(icase), fid, C.byref(n)))
        a = np.empty(n.value)
        check(lib().cfd_batch_get_field(self._h, int(icase), fid, dptr(a)))
        w = WIDTH.get(name, 1)
        return a.reshape(-1, w) if w > 1 else a

    def scalars(self, icase: int) -> dict:
        a = np.zeros(SCALARS["CFD_NSCALARS"])
        check(lib().cfd_batch_get_scalars(self._h, int(icase), dptr(a)))
        return {k: float(a[SCALARS[v]]) for k, v in SCALAR_OF.items()}

    def info(self) -> dict:
        npn, ne = C.c_longlong(), C.c_longlong()
        npatch, pe, blk = C.c_int(), C.c_int(), C.c_int()
        check(lib().cfd_batch_info(self._h, C.byref(npn), C.byref(ne), C.byref(npatch),
                                   C.byref(pe), C.byref(blk)))
        return {"npoin": npn.value, "nelem": ne.value, "npatch": npatch.value,
                "patch_elems": pe.value, "patch_block": blk.value}

    def autotune(self, candidates, steps: int = 5):
        """Times `steps` steps of each (patch_elems, patch_block) candidate,
        keeps the fastest and restarts every case from the free stream;
        returns (best candidate, {candidate: ms per step or None if it does
        not fit})."""
        cand = list(candidates)
        pe = (C.c_int * len(cand))(*[c[0] for c in cand])
        bl = (C.c_int * len(cand))(*[c[1] for c in cand])
        ms = (C.c_double * len(cand))()
        best = C.c_int()
        check(lib().cfd_batch_autotune(self._h, len(cand), pe, bl, int(steps), ms, C.byref(best)))
        times = {c: (ms[i] if ms[i] >= 0 else None) for i, c in enumerate(cand)}
        return cand[best.value], times
