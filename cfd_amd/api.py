"""Host-side mirror of the reference routines on the path, backed by the
gfx950 library (no CPU fallback).

Names, argument meaning and accumulate/overwrite semantics follow the
reference (file:line of each routine in its docstring).  Arrays are numpy,
node-major like the Fortran ones: U has shape (npoin, 4) = U(4,npoin),
dNx (nelem, 3) = dNx(3,nelem), inpoel (nelem, 3) int32 with 1-based ids.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BcLists, Consts, RkElems, RkNodes, check, dptr, iptr, lib


def consts_of(cs) -> Consts:
    return Consts(gama=cs.gama, fr=cs.fr, fmu=cs.fmu, fk=cs.fk, fcv=cs.fcv, t_inf=cs.t_inf,
                  rho_inf=cs.rho_inf, u_inf=cs.u_inf, v_inf=cs.v_inf, cte=cs.cte, fsafe=cs.fsafe)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def bc_lists(lists: dict, keep: list) -> BcLists:
    """cfd_bc_lists view of a case's loader lists (arrays kept alive in keep)."""
    arrs = {k: _i32(lists[k]) if k.startswith("i") or k == "wall" else _f64(lists[k])
            for k in ("ifixv_node", "rfixv_valuex", "rfixv_valuey", "ifixrho_node",
                      "rfixrho_value", "ifixt_node", "rfixt_value", "wall")}
    keep.extend(arrs.values())
    return BcLists(nfixv=arrs["ifixv_node"].size, ifixv_node=iptr(arrs["ifixv_node"]),
                   rfixv_valuex=dptr(arrs["rfixv_valuex"]), rfixv_valuey=dptr(arrs["rfixv_valuey"]),
                   nfixrho=arrs["ifixrho_node"].size, ifixrho_node=iptr(arrs["ifixrho_node"]),
                   rfixrho_value=dptr(arrs["rfixrho_value"]), nfixt=arrs["ifixt_node"].size,
                   ifixt_node=iptr(arrs["ifixt_node"]), rfixt_value=dptr(arrs["rfixt_value"]),
                   nwall=arrs["wall"].shape[0], wall=iptr(arrs["wall"].reshape(-1)))


class MeshContext:
    """Neighbour tables and scratch for one mesh (cfd_mesh)."""

    def __init__(self, inpoel: np.ndarray, npoin: int):
        self.inpoel = _i32(inpoel).reshape(-1, 3)
        self.nelem = self.inpoel.shape[0]
        self.npoin = int(npoin)
        self._h = C.c_void_p()
        check(lib().cfd_mesh_create(C.byref(self._h), self.nelem, self.npoin, iptr(self.inpoel)))
        a, b = C.c_int(), C.c_int()
        check(lib().cfd_mesh_sizes(self._h, C.byref(a), C.byref(b)))
        self.npsup, self.nnz = a.value, b.value

    def close(self):
        if self._h:
            check(lib().cfd_mesh_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- PointNeighbor / Mlaplace tables --------------------------------------
    def getEsup(self):
        """pointNeighbor.f90:5-43 -> (esup1, esup2)"""
        e1 = np.zeros(3 * self.nelem, np.int32)
        e2 = np.zeros(self.npoin + 1, np.int32)
        check(lib().cfd_get_esup(self._h, iptr(e1), iptr(e2)))
        return e1, e2

    def getPsup(self):
        """pointNeighbor.f90:45-91 -> (psup1, psup2)"""
        p1 = np.zeros(max(self.npsup, 1), np.int32)
        p2 = np.zeros(self.npoin + 1, np.int32)
        check(lib().cfd_get_psup(self._h, iptr(p1), iptr(p2)))
        return p1[: self.npsup], p2

    def lap_structure(self):
        """mLaplace.f90:60-94 -> (lap_rowptr, lap_idx)"""
        r = np.zeros(self.npoin + 1, np.int32)
        i = np.zeros(self.nnz, np.int32)
        check(lib().cfd_lap_structure(self._h, iptr(r), iptr(i)))
        return r, i

    # ---- geometry -----------------------------------------------------------------
    def laplace(self, X, Y, dNx, dNy):
        """mLaplace.f90:7-58 -> (lap_sparse, lap_diag)"""
        ls = np.zeros(self.nnz)
        ld = np.zeros(self.npoin)
        check(lib().cfd_laplace(self._h, dptr(_f64(X)), dptr(_f64(Y)), dptr(_f64(dNx)),
                                dptr(_f64(dNy)), dptr(ls), dptr(ld)))
        return ls, ld

    def deriv(self, X, Y):
        """subrutinas.f90:88-126 -> dict(area, dNx, dNy, HH, HHX, HHY, hmin)"""
        ne = self.nelem
        o = {k: np.zeros(ne) for k in ("area", "HH", "HHX", "HHY")}
        o["dNx"] = np.zeros((ne, 3))
        o["dNy"] = np.zeros((ne, 3))
        h = C.c_double()
        check(lib().cfd_deriv(self._h, dptr(_f64(X)), dptr(_f64(Y)), dptr(o["area"]), dptr(o["dNx"]),
                              dptr(o["dNy"]), dptr(o["HH"]), dptr(o["HHX"]), dptr(o["HHY"]),
                              C.byref(h)))
        o["hmin"] = h.value
        return o

    def MASAS(self, area):
        """subrutinas.f90:128-153 -> M"""
        M = np.zeros(self.npoin)
        check(lib().cfd_masas(self._h, dptr(_f64(area)), dptr(M)))
        return M

    def deltat(self, consts: Consts, area, T, VEL_X, VEL_Y, W_X, W_Y):
        """subrutinas.f90:155-218 -> (DT, dtmin)"""
        DT = np.zeros(self.nelem)
        d = C.c_double()
        check(lib().cfd_deltat(self._h, C.byref(consts), dptr(_f64(area)), dptr(_f64(T)),
                               dptr(_f64(VEL_X)), dptr(_f64(VEL_Y)), dptr(_f64(W_X)),
                               dptr(_f64(W_Y)), dptr(DT), C.byref(d)))
        return DT, d.value

    def ESTAB(self, consts: Consts, U, T, GAMM, VEL_X, VEL_Y, W_X, W_Y, dNx, dNy, dtmin):
        """subrutinas.f90:332-446 -> (SHOC, T_SUGN1, T_SUGN2, T_SUGN3)"""
        out = [np.zeros(self.nelem) for _ in range(4)]
        check(lib().cfd_estab(self._h, C.byref(consts), dptr(_f64(U)), dptr(_f64(T)),
                              dptr(_f64(GAMM)), dptr(_f64(VEL_X)), dptr(_f64(VEL_Y)),
                              dptr(_f64(W_X)), dptr(_f64(W_Y)), dptr(_f64(dNx)), dptr(_f64(dNy)),
                              C.c_double(dtmin), *[dptr(o) for o in out]))
        return tuple(out)

    def calcRHS(self, consts: Consts, rhs, U, theta, T, dNx, dNy, area, shoc, dtl, t1, t2, t3):
        """calcRHS.f90:4-154; ACCUMULATES into rhs (numpy (npoin,4), updated in place)"""
        assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
        th = None if theta is None else _f64(theta)
        check(lib().cfd_calcrhs(self._h, C.byref(consts), dptr(rhs), dptr(_f64(U)), dptr(th),
                                dptr(_f64(T)), dptr(_f64(dNx)), dptr(_f64(dNy)), dptr(_f64(area)),
                                dptr(_f64(shoc)), dptr(_f64(dtl)), dptr(_f64(t1)), dptr(_f64(t2)),
                                dptr(_f64(t3))))
        return rhs

    def FUENTE(self, rhs, U, W_X, W_Y, dNx, dNy, area, dtl):
        """subrutinas.f90:1036-1094; ACCUMULATES into rhs"""
        assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
        check(lib().cfd_fuente(self._h, dptr(rhs), dptr(_f64(U)), dptr(_f64(W_X)), dptr(_f64(W_Y)),
                               dptr(_f64(dNx)), dptr(_f64(dNy)), dptr(_f64(area)), dptr(_f64(dtl))))
        return rhs

    def normales(self, lists: dict, X, Y):
        """subrutinas.f90:7-64 (normals kept in the context for RK)"""
        keep = []
        bc = bc_lists(lists, keep)
        check(lib().cfd_normales(self._h, C.byref(bc), dptr(_f64(X)), dptr(_f64(Y))))

    def normalvel(self, VEL_X, VEL_Y, W_X, W_Y):
        """subrutinas.f90:66-85 with the normals of the last normales():
        returns the projected (VEL_X, VEL_Y)"""
        vx, vy = _f64(VEL_X).copy(), _f64(VEL_Y).copy()
        check(lib().cfd_normalvel(self._h, dptr(vx), dptr(vy), dptr(_f64(W_X)), dptr(_f64(W_Y))))
        return vx, vy

    def smoothing(self, X, Y, fixed, dX=None, dY=None, order: int = 0):
        """smoothing.f90:21-75 (smoothing_mod::smoothing) on copies of X, Y:
        returns (X, Y, sweeps, H_MIN_GLOBAL).  fixed: nodes held (nonzero).
        dX, dY: the reference's optional arguments (setTol_move after the
        mesh check; the mode then stays on for this context).  order 0: the
        reference's node order (bit for bit, at most 65536 nodes); 1: colour
        by colour, in parallel."""
        x, y = _f64(X).copy(), _f64(Y).copy()
        fx = _i32(np.asarray(fixed) != 0)
        dx = None if dX is None else _f64(dX)
        dy = None if dY is None else _f64(dY)
        n, h = C.c_int(), C.c_double()
        check(lib().cfd_smoothing(self._h, dptr(x), dptr(y), iptr(fx),
                                  None if dx is None else dptr(dx),
                                  None if dy is None else dptr(dy), int(order), C.byref(n),
                                  C.byref(h)))
        return x, y, n.value, h.value

    def set_tol_move(self, dX, dY):
        """smoothing.f90:386-401 setTol_move: tol_move = |dX| + |dY|"""
        check(lib().cfd_smoothing_set_tol_move(self._h, dptr(_f64(dX)), dptr(_f64(dY))))

    def force_visc(self, iset, X, Y, P, T, VEL_X, VEL_Y, dNx, dNy, t_inf):
        """ns2DComp.ALE.f90:819-893 for one set of edges (n1, n2, elem):
        (F_VX, F_VY)"""
        ist = _i32(iset).reshape(-1)
        fvx, fvy = C.c_double(), C.c_double()
        check(lib().cfd_force_visc(self._h, ist.size // 3, iptr(ist) if ist.size else None,
                                   dptr(_f64(X)), dptr(_f64(Y)), dptr(_f64(P)), dptr(_f64(T)),
                                   dptr(_f64(VEL_X)), dptr(_f64(VEL_Y)), dptr(_f64(dNx)),
                                   dptr(_f64(dNy)), C.c_double(t_inf), C.byref(fvx),
                                   C.byref(fvy)))
        return fvx.value, fvy.value

    def RK(self, consts: Consts, nrk, dtmin, dtl, st: dict):
        """subrutinas.f90:645-849 on a state dict (same keys as the oracle);
        updates U1, RHS, T, VEL_X, VEL_Y, RHO, E, P, RMACH, SHOC, T_SUGN1..3"""
        for k in ("U1", "RHS", "T", "VEL_X", "VEL_Y", "RHO", "E", "P", "RMACH", "SHOC", "T_SUGN1",
                  "T_SUGN2", "T_SUGN3"):
            st[k] = _f64(st[k]).copy()
        keepin = {k: _f64(st[k]) for k in ("U", "W_X", "W_Y", "M", "GAMM", "DNX", "DNY", "AREA")}
        dl = _f64(dtl)
        nd = RkNodes(U1=dptr(st["U1"]), RHS=dptr(st["RHS"]), T=dptr(st["T"]),
                     VEL_X=dptr(st["VEL_X"]), VEL_Y=dptr(st["VEL_Y"]), RHO=dptr(st["RHO"]),
                     E=dptr(st["E"]), P=dptr(st["P"]), RMACH=dptr(st["RMACH"]),
                     U=dptr(keepin["U"]), W_X=dptr(keepin["W_X"]), W_Y=dptr(keepin["W_Y"]),
                     M=dptr(keepin["M"]), GAMM=dptr(keepin["GAMM"]))
        el = RkElems(dNx=dptr(keepin["DNX"]), dNy=dptr(keepin["DNY"]), area=dptr(keepin["AREA"]),
                     shoc=dptr(st["SHOC"]), t_sugn1=dptr(st["T_SUGN1"]),
                     t_sugn2=dptr(st["T_SUGN2"]), t_sugn3=dptr(st["T_SUGN3"]))
        check(lib().cfd_rk(self._h, C.byref(consts), None, int(nrk), C.c_double(dtmin), dptr(dl),
                           C.byref(nd), C.byref(el)))
        return st

    # ---- gcl_mod ----------------------------------------------------------------------
    def putW(self, W_x, W_y):
        """gcl.f90:47-55"""
        check(lib().cfd_gcl_put_w(self._h, dptr(_f64(W_x)), dptr(_f64(W_y))))

    def putArea(self, area):
        """gcl.f90:56-62"""
        check(lib().cfd_gcl_put_area(self._h, dptr(_f64(area))))

    def gcl_main(self, M, W_x, W_y, dNx, dNy, area, dt, mode=0):
        """gcl.f90:8-46; updates M in place (mode 0: as written, 1: corrected)"""
        assert M.dtype == np.float64 and M.flags.c_contiguous
        check(lib().cfd_gcl_main(self._h, dptr(M), dptr(_f64(W_x)), dptr(_f64(W_y)),
                                 dptr(_f64(dNx)), dptr(_f64(dNy)), dptr(_f64(area)),
                                 C.c_double(dt), int(mode)))
        return M


def biCG(spMtx, spIdx, spRowptr, diagMtx, x, b, x_fix, x_fixIdx):
    """biconjGrad.f90:8-62; x (float64, updated in place) -> (iters, |r.z|)"""
    assert x.dtype == np.float64 and x.flags.c_contiguous
    n = x.size
    it, rz = C.c_int(), C.c_double()
    xf = _f64(x_fix)
    fi = _i32(x_fixIdx)
    check(lib().cfd_bicg(n, dptr(_f64(spMtx)), iptr(_i32(spIdx)), iptr(_i32(spRowptr)),
                         dptr(_f64(diagMtx)), dptr(x), dptr(_f64(b)), dptr(xf) if xf.size else None,
                         iptr(fi) if fi.size else None, int(fi.size), C.byref(it), C.byref(rz)))
    return it.value, rz.value


def write_restart(path: str, iter_: int, time_: float, U, T, GAMM) -> None:
    """PRINTREST (ns2DComp.ALE.f90:898-917): the reference's .RST file."""
    U = _f64(U).reshape(-1, 4)
    check(lib().cfd_write_restart(str(path).encode(), int(iter_), C.c_double(time_), U.shape[0],
                                  dptr(U), dptr(_f64(T)), dptr(_f64(GAMM))))


def read_restart(path: str, npoin: int):
    """RESTART's restart branch (ns2DComp.ALE.f90:423-433) -> (ITER, TIME, U,
    T, GAMM); the reference reads ITER and TIME into locals and drops them."""
    U, T, G = np.empty((npoin, 4)), np.empty(npoin), np.empty(npoin)
    it, tm = C.c_int(), C.c_double()
    check(lib().cfd_read_restart(str(path).encode(), int(npoin), C.byref(it), C.byref(tm),
                                 dptr(U), dptr(T), dptr(G)))
    return it.value, tm.value, U, T, G


def biCG_ex(spMtx, spIdx, spRowptr, diagMtx, x, b, x_fix, x_fixIdx, tol=1e-10, maxit=1000, rel=0):
    """biCG with the stopping rule as arguments (rel=1: tests relative to the
    residual of the guess that is zero off the fixed nodes) -> (iters, |r.z|,
    converged)"""
    assert x.dtype == np.float64 and x.flags.c_contiguous
    n = x.size
    it, rz, cv = C.c_int(), C.c_double(), C.c_int()
    xf = _f64(x_fix)
    fi = _i32(x_fixIdx)
    check(lib().cfd_bicg_ex(n, dptr(_f64(spMtx)), iptr(_i32(spIdx)), iptr(_i32(spRowptr)),
                            dptr(_f64(diagMtx)), dptr(x), dptr(_f64(b)),
                            dptr(xf) if xf.size else None, iptr(fi) if fi.size else None,
                            int(fi.size), C.c_double(tol), int(maxit), int(rel), C.byref(it),
                            C.byref(rz), C.byref(cv)))
    return it.value, rz.value, bool(cv.value)


def _f32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u16p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_ushort))


def mlp_forward(params: np.ndarray, x_bits: np.ndarray) -> np.ndarray:
    """Warm-start MLP on the matrix cores: params (float32, layout of
    cfd_mlp_forward), x_bits (n, 16) BF16 bit patterns -> outputs (n, 2)."""
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x_bits, np.uint16)
    n = x.shape[0]
    out = np.zeros((n, 2), np.float32)
    check(lib().cfd_mlp_forward(_f32p(p), _u16p(x), n, _f32p(out)))
    return out


def mlp_gradient(params: np.ndarray, x_bits: np.ndarray, t: np.ndarray):
    """(loss, gradient) of sum_nodes sum_{o<2} (y - t)^2 / n."""
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x_bits, np.uint16)
    tt = np.ascontiguousarray(t, np.float32)
    n = x.shape[0]
    g = np.zeros(p.size, np.float32)
    loss = C.c_float()
    check(lib().cfd_mlp_gradient(_f32p(p), _u16p(x), _f32p(tt), n, _f32p(g), C.byref(loss)))
    return loss.value, g
