"""Device-resident solver session (the main program's time loop,
ns2DComp.ALE.f90:138-281) on the gfx950 library.

State is exchanged with the same names the oracle and the reference harness
use (U, T, VEL_X, ..., LAP_SPARSE; DTMIN, TIME, ...), so states compare field
by field and a step can be teacher-forced from any of them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FIELDS, KTIMERS, SCALARS, CaseDesc, check, dptr, iptr, lib
from .api import MeshContext, bc_lists, consts_of

# state-dict key -> field id
FIELD_OF = {k[len("CFD_F_"):]: v for k, v in FIELDS.items() if k != "CFD_NFIELDS"}
SCALAR_OF = {"DTMIN": "CFD_S_DTMIN", "TIME": "CFD_S_TIME", "DTMIN1": "CFD_S_DTMIN1",
             "BANDERA": "CFD_S_BANDERA", "ITER": "CFD_S_ITER", "HMIN": "CFD_S_HMIN",
             "FX1": "CFD_S_FX", "FY1": "CFD_S_FY", "RM1": "CFD_S_RM", "DISN2": "CFD_S_DISN2",
             "PCG_ITX": "CFD_S_PCG_ITX", "PCG_ITY": "CFD_S_PCG_ITY", "PCG_RZX": "CFD_S_PCG_RZX",
             "PCG_RZY": "CFD_S_PCG_RZY", "WS_LEARNED": "CFD_S_WS_LEARNED",
             "WS_SOLVES": "CFD_S_WS_SOLVES", "WS_LOSS": "CFD_S_WS_LOSS",
             "PCG_ITSUM": "CFD_S_PCG_ITSUM", "FVX1": "CFD_S_FVX", "FVY1": "CFD_S_FVY",
             "PCG_UNCONV": "CFD_S_PCG_UNCONV"}
NSCALARS = SCALARS["CFD_NSCALARS"]
WIDTH = {"U": 4, "U1": 4, "RHS": 4, "DNX": 3, "DNY": 3}


class Session:
    def __init__(self, cs, mesh: MeshContext | None = None, motion: int | None = None,
                 motion_amp: float | None = None, motion_omega: float | None = None,
                 pcg_tol: float = 0.0,
                 pcg_maxit: int = 0, gcl: int = -1, precision: int = 0, init: bool = True,
                 pcg_rel: int = 0, warm_start: int = 0, patch_elems: int = 0,
                 patch_block: int = 0, renumber: int = 0):
        self.cs = cs
        m = cs.mesh
        self.mesh = mesh if mesh is not None else MeshContext(m.inpoel, m.npoin)
        L = cs.lists
        self._keep = []
        iset = np.ascontiguousarray(L["iset"], np.int32).reshape(-1)
        ilaux = np.ascontiguousarray(L["ilaux"], np.int32)
        X = np.ascontiguousarray(m.x, np.float64)
        Y = np.ascontiguousarray(m.y, np.float64)
        self._keep += [iset, ilaux, X, Y]
        d = CaseDesc()
        d.c = consts_of(cs)
        d.nrk = 4
        d.moving = int(cs.moving)
        d.itlocal = int(cs.itlocal)
        d.xref, d.yref = cs.xref, cs.yref
        d.nsetedges = L["iset"].shape[0]
        d.iset = iptr(iset)
        d.nmove, d.nfix_move = int(L["nmove"]), int(L["nfix_move"])
        d.ilaux = iptr(ilaux) if ilaux.size else None
        d.bc = bc_lists(L, self._keep)
        d.X, d.Y = dptr(X), dptr(Y)
        # the pitching law: the case's unless given here
        d.motion = int(cs.motion if motion is None else motion)
        d.motion_amp = float(cs.motion_amp if motion_amp is None else motion_amp)
        d.motion_omega = float(cs.motion_omega if motion_omega is None else motion_omega)
        d.pcg_tol, d.pcg_maxit, d.gcl, d.precision = pcg_tol, pcg_maxit, gcl, precision
        d.pcg_rel = pcg_rel
        d.warm_start = warm_start
        d.patch_elems, d.patch_block = patch_elems, patch_block
        d.renumber = renumber
        self._h = C.c_void_p()
        check(lib().cfd_session_create(C.byref(self._h), self.mesh.handle, C.byref(d)))
        self.npoin, self.nelem, self.nnz = m.npoin, m.nelem, self.mesh.nnz
        if init:
            self.init_freestream()

    def close(self):
        if self._h:
            check(lib().cfd_session_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_freestream(self):
        check(lib().cfd_session_init_freestream(self._h))

    def step(self, n: int = 1):
        check(lib().cfd_session_step(self._h, int(n)))

    def sync(self):
        check(lib().cfd_session_sync(self._h))

    # ---- fields / scalars -------------------------------------------------------
    def _count(self, fid):
        n = C.c_longlong()
        check(lib().cfd_session_field_size(self._h, fid, C.byref(n)))
        return n.value

    def get(self, name: str) -> np.ndarray:
        fid = FIELD_OF[name]
        a = np.empty(self._count(fid))
        check(lib().cfd_session_get_field(self._h, fid, dptr(a)))
        w = WIDTH.get(name, 1)
        return a.reshape(-1, w) if w > 1 else a

    def set(self, name: str, value) -> None:
        fid = FIELD_OF[name]
        a = np.ascontiguousarray(value, np.float64).reshape(-1)
        if a.size != self._count(fid):
            raise ValueError(f"{name}: expected {self._count(fid)} values, got {a.size}")
        check(lib().cfd_session_set_field(self._h, fid, dptr(a)))

    def scalars(self) -> dict:
        a = np.zeros(NSCALARS)
        check(lib().cfd_session_get_scalars(self._h, dptr(a)))
        return {k: float(a[SCALARS[v]]) for k, v in SCALAR_OF.items()}

    def set_scalars(self, **kw) -> None:
        a = np.zeros(NSCALARS)
        check(lib().cfd_session_get_scalars(self._h, dptr(a)))
        for k, v in kw.items():
            a[SCALARS[SCALAR_OF[k]]] = v
        check(lib().cfd_session_set_scalars(self._h, dptr(a)))

    def state(self) -> dict:
        st = {k: self.get(k) for k in FIELD_OF}
        st.update(self.scalars())
        st.update(npoin=self.npoin, nelem=self.nelem, nnz=self.nnz)
        return st

    def set_state(self, st: dict, fields=None) -> None:
        """Loads a state (oracle / reference names); scalars DTMIN, TIME,
        DTMIN1, BANDERA, ITER and the previous pitch angle DISN2 if given."""
        for k in (fields or FIELD_OF):
            if k in st:
                self.set(k, st[k])
        sc = {k: st[k] for k in ("DTMIN", "TIME", "DTMIN1", "BANDERA", "ITER", "DISN2") if k in st}
        if sc:
            self.set_scalars(**sc)

    # ---- restart files (ns2DComp.ALE.f90:394-435, 898-917) -------------------------
    def write_restart(self, path: str) -> None:
        check(lib().cfd_session_write_restart(self._h, str(path).encode()))

    def read_restart(self, path: str, keep_clock: bool = False) -> None:
        """U and T from the file (U1 and the primitives follow); the step
        control restarts; TIME and ITER from the file only with keep_clock."""
        check(lib().cfd_session_read_restart(self._h, str(path).encode(), int(keep_clock)))

    # ---- timing --------------------------------------------------------------------
    def timing(self, enable: bool = True, slots=None) -> None:
        """HIP-event timing of every slot (enable), or only of `slots`
        (names such as "ELEMENT"); resets the accumulators."""
        if slots is None:
            check(lib().cfd_session_timing(self._h, int(enable)))
            return
        mask = 0
        for s in slots:
            mask |= 1 << KTIMERS["CFD_K_" + s]
        check(lib().cfd_session_timing_mask(self._h, C.c_uint(mask if enable else 0)))

    def kernel_time(self, slot: str):
        t, n = C.c_double(), C.c_longlong()
        check(lib().cfd_session_kernel_time(self._h, KTIMERS["CFD_K_" + slot], C.byref(t),
                                            C.byref(n)))
        return t.value, n.value
