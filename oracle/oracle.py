"""Test infrastructure only: ctypes bindings for the C restatement (liboracle.so)
and a restatement of the main program's step sequence on top of it.

`OracleSolver` follows ns2DComp.ALE.f90 (:59-136 start-up without smoothing,
:138-281 time loop without printing) and keeps its state under the same names
as oracle/refharness.py, so states from the compiled reference, this oracle
and the GPU library compare field by field.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_d = C.c_double
_i = C.c_int
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)


class OrcConsts(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("gama", "fr", "fmu", "fk", "fcv", "t_inf", "rho_inf", "u_inf", "v_inf", "cte", "fsafe")]


class OrcBcs(C.Structure):
    _fields_ = [("nfixv", _i), ("ifixv_node", _pi), ("rfixv_valuex", _pd), ("rfixv_valuey", _pd),
                ("nnorm", _i), ("n_ipoin", _pi), ("n_x", _pd), ("n_y", _pd),
                ("nfixrho", _i), ("ifixrho_node", _pi), ("rfixrho_value", _pd),
                ("nfixt", _i), ("ifixt_node", _pi), ("rfixt_value", _pd)]


class OrcNodes(C.Structure):
    _fields_ = [(n, _pd) for n in ("U", "U1", "RHS", "T", "VEL_X", "VEL_Y", "W_X", "W_Y", "RHO",
                                    "E", "P", "RMACH", "M", "GAMM")]


class OrcElems(C.Structure):
    _fields_ = [("inpoel", _pi)] + [(n, _pd) for n in ("dNx", "dNy", "area", "shoc", "t1", "t2", "t3")]


_lib = None


def available() -> bool:
    return os.path.isfile(LIB_PATH)


def set_threads(n: int) -> None:
    """1 (the default) reproduces the sequential reference order."""
    lib().orc_set_threads(int(n))


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.orc_get_esup.argtypes = [_i, _i, _ip, _ip, _ip]
        L.orc_get_psup.argtypes = [_i, _i, _ip, _ip, _ip, C.c_void_p, _ip]
        L.orc_get_psup.restype = _i
        L.orc_lap_structure.argtypes = [_i, _ip, _ip, _ip, _ip]
        L.orc_deriv.argtypes = [_i, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _pd]
        L.orc_masas.argtypes = [_i, _i, _ip, _dp, _dp]
        L.orc_deltat.argtypes = [C.POINTER(OrcConsts), _i, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _pd]
        L.orc_estab.argtypes = [C.POINTER(OrcConsts), _i, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                _dp, _dp, _d, _dp, _dp, _dp, _dp]
        L.orc_calcrhs.argtypes = [C.POINTER(OrcConsts), _i, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                  _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_fuente.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_rk.argtypes = [C.POINTER(OrcConsts), _i, _i, _i, _d, _dp, C.POINTER(OrcNodes),
                             C.POINTER(OrcElems), C.POINTER(OrcBcs)]
        L.orc_normales.argtypes = [_i, _i, _ip, _dp, _dp, _ip, _dp, _dp]
        L.orc_normales.restype = _i
        L.orc_laplace.argtypes = [_i, _i, _ip, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _dp, _dp]
        L.orc_bicg.argtypes = [_i, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _ip, _i, _pd]
        L.orc_bicg.restype = _i
        L.orc_forces.argtypes = [_i, _ip, _dp, _dp, _dp, _d, _d, _pd, _pd, _pd]
        L.orc_transf.argtypes = [_i, _ip, _dp, _dp, _d, _d, _d, _d, _dp, _dp]
        L.orc_normalvel.argtypes = [_i, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_force_visc.argtypes = [_i, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d,
                                     _pd, _pd]
        L.orc_set_threads.argtypes = [_i]
        L.orc_get_threads.restype = _i
        L.orc_set_threads(1)
        L.orc_gcl.argtypes = [_i, _i, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d, _i]
        L.orc_smoothing.argtypes = [_i, _i, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _ip, C.c_void_p,
                                    _pd, _i]
        L.orc_smoothing.restype = _i
        _lib = L
    return _lib


def _pp(a):
    return a.ctypes.data_as(_pd)


def _pip(a):
    return a.ctypes.data_as(_pi)


def consts_of(cs) -> OrcConsts:
    return OrcConsts(gama=cs.gama, fr=cs.fr, fmu=cs.fmu, fk=cs.fk, fcv=cs.fcv, t_inf=cs.t_inf,
                     rho_inf=cs.rho_inf, u_inf=cs.u_inf, v_inf=cs.v_inf, cte=cs.cte, fsafe=cs.fsafe)


# ---------------------------------------------------------------------------
# tables
# ---------------------------------------------------------------------------
def get_esup(inpoel: np.ndarray, npoin: int):
    ne = inpoel.shape[0]
    esup1 = np.zeros(3 * ne, np.int32)
    esup2 = np.zeros(npoin + 1, np.int32)
    lib().orc_get_esup(ne, npoin, np.ascontiguousarray(inpoel, np.int32), esup1, esup2)
    return esup1, esup2


def get_psup(inpoel, npoin, esup1, esup2):
    ne = inpoel.shape[0]
    psup2 = np.zeros(npoin + 1, np.int32)
    ip = np.ascontiguousarray(inpoel, np.int32)
    n = lib().orc_get_psup(ne, npoin, ip, esup1, esup2, None, psup2)
    psup1 = np.zeros(n, np.int32)
    lib().orc_get_psup(ne, npoin, ip, esup1, esup2, psup1.ctypes.data_as(C.c_void_p), psup2)
    return psup1, psup2


def lap_structure(npoin, psup1, psup2):
    rowptr = np.zeros(npoin + 1, np.int32)
    idx = np.zeros(psup1.size + npoin, np.int32)
    lib().orc_lap_structure(npoin, psup1, psup2, rowptr, idx)
    return rowptr, idx


def smoothing(inpoel, npoin, X, Y, fixed, tol_move=None, tabs=None, max_sweeps=100):
    """smoothing.f90 smoothing in the reference's node order on copies of X,
    Y: (X, Y, sweeps, H_MIN_GLOBAL).  tol_move: |dX| + |dY| per node when
    setTol_move has been called (smoothing.f90:386-401), else None.
    max_sweeps: the reference's 100 (fewer to look at intermediate states)."""
    ip = np.ascontiguousarray(inpoel, np.int32)
    t = tabs or tables(ip, npoin)
    Xs, Ys = np.array(X, np.float64), np.array(Y, np.float64)
    fx = np.ascontiguousarray(np.asarray(fixed) != 0, np.int32)
    tm = None if tol_move is None else np.ascontiguousarray(tol_move, np.float64)
    h = C.c_double()
    n = lib().orc_smoothing(ip.shape[0], npoin, ip, t["esup1"], t["esup2"], t["psup1"],
                            t["psup2"], Xs, Ys, fx,
                            None if tm is None else tm.ctypes.data_as(C.c_void_p), C.byref(h),
                            int(max_sweeps))
    if n < 0:
        raise ValueError("a node has more than 20 elements (MAX_ELEM_PER_NODE)")
    return Xs, Ys, n, h.value


def tables(inpoel, npoin):
    esup1, esup2 = get_esup(inpoel, npoin)
    psup1, psup2 = get_psup(inpoel, npoin, esup1, esup2)
    rowptr, idx = lap_structure(npoin, psup1, psup2)
    return dict(esup1=esup1, esup2=esup2, psup1=psup1, psup2=psup2, lap_rowptr=rowptr, lap_idx=idx)


# ---------------------------------------------------------------------------
# solver restatement
# ---------------------------------------------------------------------------
NODE_FIELDS = ["T", "VEL_X", "VEL_Y", "W_X", "W_Y", "X", "Y", "X1", "Y1", "RHO", "E", "P", "RMACH",
               "M", "GAMM", "XPOS", "YPOS", "POS_AUX", "B", "W_X_OLD", "W_Y_OLD", "LAP_DIAG"]
ELEM_FIELDS = ["DTL", "DT", "SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3", "AREA", "HH", "HHX", "HHY",
               "AREA_OLD"]


class OracleSolver:
    """The reference's start-up and time step, on the C restatement."""

    def __init__(self, cs, threads: int | None = None):
        self.cs = cs
        self.c = consts_of(cs)
        m = cs.mesh
        self.np_, self.ne = m.npoin, m.nelem
        self.inpoel = np.ascontiguousarray(m.inpoel, np.int32)
        self.tab = tables(self.inpoel, self.np_)
        npn, ne = self.np_, self.ne
        L = cs.lists
        self.lists = L
        s = {}
        s["U"] = np.zeros((npn, 4))
        s["U1"] = np.zeros((npn, 4))
        s["RHS"] = np.zeros((npn, 4))
        for f in NODE_FIELDS:
            s[f] = np.zeros(npn)
        for f in ELEM_FIELDS:
            s[f] = np.zeros(ne)
        s["DNX"] = np.zeros((ne, 3))
        s["DNY"] = np.zeros((ne, 3))
        s["LAP_SPARSE"] = np.zeros(self.tab["lap_idx"].size)
        s["X"][:] = m.x
        s["Y"][:] = m.y
        self.s = s
        # RESTART free-stream branch (ns2DComp.ALE.f90:404-420)
        s["GAMM"][:] = cs.gama
        rhoamb, uamb, vamb = cs.rho_inf, cs.u_inf, cs.v_inf
        pamb = cs.rho_inf * cs.fr * cs.t_inf
        energia = pamb / ((cs.gama - 1.0) * rhoamb) + 0.5 * (uamb * uamb + vamb * vamb)
        s["U"][:, 0] = rhoamb
        s["U"][:, 1] = rhoamb * uamb
        s["U"][:, 2] = rhoamb * vamb
        s["U"][:, 3] = energia * rhoamb
        s["VEL_X"][:] = uamb
        s["VEL_Y"][:] = vamb
        s["T"][:] = cs.t_inf
        self.normales()
        self.hmin = self.deriv()
        self.masas()
        self.laplace()
        s["W_X"][:] = -0.0
        s["W_Y"][:] = 0.0
        s["U1"][:] = s["U"]
        s["RHO"][:] = s["U"][:, 0]
        s["E"][:] = s["U"][:, 3] / s["U"][:, 0]
        s["AREA_OLD"][:] = s["AREA"]
        self.time = 0.0
        self.iter = 0
        self.dtmin = 0.0
        self.dtmin1 = 0.0
        self.bandera = 1
        self.disn2 = 0.0          # MeshMove DISN(2), previous pitch angle
        self.fx = self.fy = self.rm = 0.0
        self.fvx = self.fvy = 0.0
        self.last_pcg_iters = (0, 0)

    # ---- geometry ---------------------------------------------------------
    def normales(self):
        L = self.lists
        wall = np.ascontiguousarray(L["wall"], np.int32).reshape(-1)
        n_ip = np.zeros(self.np_, np.int32)
        nx = np.zeros(self.np_)
        ny = np.zeros(self.np_)
        m = lib().orc_normales(self.np_, L["wall"].shape[0], wall if wall.size else np.zeros(2, np.int32),
                               self.s["X"], self.s["Y"], n_ip, nx, ny)
        self.norm = (n_ip[:m].copy(), nx[:m].copy(), ny[:m].copy())

    def deriv(self):
        s = self.s
        h = C.c_double()
        lib().orc_deriv(self.ne, self.inpoel, s["X"], s["Y"], s["AREA"], s["DNX"], s["DNY"], s["HH"],
                        s["HHX"], s["HHY"], C.byref(h))
        return h.value

    def masas(self):
        lib().orc_masas(self.ne, self.np_, self.inpoel, self.s["AREA"], self.s["M"])

    def laplace(self):
        s, t = self.s, self.tab
        lib().orc_laplace(self.ne, self.np_, self.inpoel, s["X"], s["Y"], s["DNX"], s["DNY"],
                          t["esup1"], t["esup2"], t["lap_rowptr"], t["lap_idx"], s["LAP_SPARSE"],
                          s["LAP_DIAG"])

    # ---- units --------------------------------------------------------------
    def deltat(self):
        s = self.s
        d = C.c_double()
        lib().orc_deltat(C.byref(self.c), self.ne, self.inpoel, s["AREA"], s["T"], s["VEL_X"],
                         s["VEL_Y"], s["W_X"], s["W_Y"], s["DT"], C.byref(d))
        return d.value

    def estab(self, dtmin):
        s = self.s
        lib().orc_estab(C.byref(self.c), self.ne, self.inpoel, s["U"], s["T"], s["GAMM"], s["VEL_X"],
                        s["VEL_Y"], s["W_X"], s["W_Y"], s["DNX"], s["DNY"], dtmin, s["SHOC"],
                        s["T_SUGN1"], s["T_SUGN2"], s["T_SUGN3"])

    def calcrhs(self, rhs=None, theta=None):
        s = self.s
        rhs = np.zeros((self.np_, 4)) if rhs is None else rhs
        theta = np.zeros((self.np_, 4)) if theta is None else theta
        lib().orc_calcrhs(C.byref(self.c), self.ne, self.np_, rhs, s["U"], theta, s["T"], s["DNX"],
                          s["DNY"], s["AREA"], s["SHOC"], s["DTL"], s["T_SUGN1"], s["T_SUGN2"],
                          s["T_SUGN3"], self.inpoel)
        return rhs

    def fuente(self, rhs=None):
        s = self.s
        rhs = np.zeros((self.np_, 4)) if rhs is None else rhs
        lib().orc_fuente(self.ne, rhs, s["U"], s["W_X"], s["W_Y"], s["DNX"], s["DNY"], s["AREA"],
                         s["DTL"], self.inpoel)
        return rhs

    def _bcs(self):
        L = self.lists
        self._keep = [np.ascontiguousarray(L[k]) for k in
                      ("ifixv_node", "rfixv_valuex", "rfixv_valuey", "ifixrho_node", "rfixrho_value",
                       "ifixt_node", "rfixt_value")] + list(self.norm)
        k = self._keep
        return OrcBcs(nfixv=k[0].size, ifixv_node=_pip(k[0]), rfixv_valuex=_pp(k[1]),
                      rfixv_valuey=_pp(k[2]), nnorm=k[7].size, n_ipoin=_pip(k[7]), n_x=_pp(k[8]),
                      n_y=_pp(k[9]), nfixrho=k[3].size, ifixrho_node=_pip(k[3]),
                      rfixrho_value=_pp(k[4]), nfixt=k[5].size, ifixt_node=_pip(k[5]),
                      rfixt_value=_pp(k[6]))

    def rk(self, dtmin, nrk=4):
        s = self.s
        nd = OrcNodes(**{k: _pp(s[k]) for k in ("U", "U1", "RHS", "T", "VEL_X", "VEL_Y", "W_X",
                                                 "W_Y", "RHO", "E", "P", "RMACH", "M", "GAMM")})
        el = OrcElems(inpoel=_pip(self.inpoel), dNx=_pp(s["DNX"]), dNy=_pp(s["DNY"]),
                      area=_pp(s["AREA"]), shoc=_pp(s["SHOC"]), t1=_pp(s["T_SUGN1"]),
                      t2=_pp(s["T_SUGN2"]), t3=_pp(s["T_SUGN3"]))
        bc = self._bcs()
        lib().orc_rk(C.byref(self.c), self.ne, self.np_, nrk, dtmin, s["DTL"], C.byref(nd),
                     C.byref(el), C.byref(bc))

    def bicg(self, x, x_fix, fidx, b=None):
        s, t = self.s, self.tab
        b = np.zeros(self.np_) if b is None else b
        rz = C.c_double()
        it = lib().orc_bicg(self.np_, s["LAP_SPARSE"], t["lap_idx"], t["lap_rowptr"], s["LAP_DIAG"],
                            x, b, np.ascontiguousarray(x_fix, np.float64),
                            np.ascontiguousarray(fidx, np.int32), int(np.size(fidx)), C.byref(rz))
        return it, rz.value

    def forces(self):
        s = self.s
        iset = np.ascontiguousarray(self.lists["iset"], np.int32)
        fx, fy, rm = C.c_double(), C.c_double(), C.c_double()
        lib().orc_forces(iset.shape[0], iset, s["X"], s["Y"], s["P"], self.cs.xref, self.cs.yref,
                         C.byref(fx), C.byref(fy), C.byref(rm))
        return fx.value, fy.value, rm.value

    def normalvel(self, vx, vy, wx, wy):
        """NORMALVEL (subrutinas.f90:66-85) with the normals of the last
        normales(): the projected copies of vx, vy"""
        n_ip, nx, ny = self.norm
        vx, vy = np.array(vx, np.float64), np.array(vy, np.float64)
        lib().orc_normalvel(n_ip.size, np.ascontiguousarray(n_ip, np.int32), nx, ny, vx, vy,
                            np.ascontiguousarray(wx, np.float64), np.ascontiguousarray(wy, np.float64))
        return vx, vy

    def force_visc(self):
        """FORCE_VISC (ns2DComp.ALE.f90:819-893) for force set 1: F_VX, F_VY."""
        s = self.s
        iset = np.ascontiguousarray(self.lists["iset"], np.int32)
        fvx, fvy = C.c_double(), C.c_double()
        lib().orc_force_visc(iset.shape[0], iset, self.inpoel, s["X"], s["Y"], s["P"], s["T"],
                             s["VEL_X"], s["VEL_Y"], s["DNX"], s["DNY"], self.cs.t_inf,
                             C.byref(fvx), C.byref(fvy))
        return fvx.value, fvy.value

    def pitch_angle(self, time):
        """The body's pitch angle at `time`: the reference's law
        pi/8 sin(10 t) (meshMove.f90:55-57,70), or with cs.motion == 1 the
        parameterised law motion_amp sin(motion_omega t)."""
        cs = self.cs
        if getattr(cs, "motion", 0) == 1:
            return cs.motion_amp * math.sin(cs.motion_omega * time)
        return math.acos(-1.0) / 8.0 * math.sin(10.0 * time)

    def fluid_structure(self, dtmin, time):
        """meshMove.f90:28-142 for one force set (set 1)."""
        s, L = self.s, self.lists
        self.fx, self.fy, self.rm = self.forces()
        dxpos = np.zeros(self.np_)
        dypos = np.zeros(self.np_)
        alphav = self.disn2
        self.disn2 = self.pitch_angle(time)
        alpha = self.disn2 - alphav
        yposr = 0.0
        iset = np.ascontiguousarray(L["iset"], np.int32)
        lib().orc_transf(iset.shape[0], iset, s["X"], s["Y"], alpha, yposr, self.cs.xref,
                         self.cs.yref, dxpos, dypos)
        nmove, ilaux = L["nmove"], L["ilaux"]
        pos = np.zeros(ilaux.size)
        pos[:nmove] = dxpos[ilaux[:nmove] - 1]
        itx, _ = self.bicg(s["XPOS"], pos, ilaux)
        s["X"] += s["XPOS"]
        s["X1"] += s["XPOS"]
        s["W_X"][:] = s["XPOS"] / dtmin
        pos[:nmove] = dypos[ilaux[:nmove] - 1]
        ity, _ = self.bicg(s["YPOS"], pos, ilaux)
        s["Y"] += s["YPOS"]
        s["Y1"] += s["YPOS"]
        s["W_Y"][:] = s["YPOS"] / dtmin
        self.last_pcg_iters = (itx, ity)

    def gcl(self, dt, mode=0):
        s = self.s
        lib().orc_gcl(self.ne, self.np_, self.inpoel, s["M"], s["W_X"], s["W_Y"], s["W_X_OLD"],
                      s["W_Y_OLD"], s["DNX"], s["DNY"], s["AREA"], s["AREA_OLD"], dt, mode)

    # ---- time step (ns2DComp.ALE.f90:140-281) ----------------------------------
    def step(self):
        s, cs = self.s, self.cs
        self.iter += 1
        dtmin = self.deltat()
        if self.bandera == 1:
            self.dtmin1 = dtmin
            self.bandera = 2
        porc = abs((dtmin - self.dtmin1) / dtmin)
        if 100.0 * porc <= 1.0:
            dtmin = self.dtmin1
        else:
            self.dtmin1 = dtmin
            self.bandera = 2
        if cs.itlocal != 0:
            f = 1.0 - math.exp(-self.iter * 4.6 / cs.itlocal)
            s["DTL"][:] = dtmin * f + s["DT"] * (1.0 - f)
        else:
            s["DTL"][:] = dtmin
        self.dtmin = dtmin
        self.time = self.time + dtmin
        s["U1"][:] = s["U"]
        self.rk(dtmin)
        self.fluid_structure(dtmin, self.time)
        # FORCE_VISC as the print block calls it (ns2DComp.ALE.f90:231-236):
        # after the mesh move, before the geometry update
        if cs.fmu != 0.0:
            self.fvx, self.fvy = self.force_visc()
        self.bandera += 1
        if cs.moving == 1:
            self.normales()
            self.hmin = self.deriv()
            self.masas()
            self.laplace()
        s["U"][:] = s["U1"]

    # ---- state exchange with the reference harness ------------------------------
    def state(self) -> dict:
        st = {k: v.copy() for k, v in self.s.items()}
        st.update(npoin=self.np_, nelem=self.ne, nnz=int(self.tab["lap_idx"].size),
                  DTMIN=self.dtmin, TIME=self.time, DTMIN1=self.dtmin1, BANDERA=float(self.bandera),
                  ITER=float(self.iter), HMIN=self.hmin, FX1=self.fx, FY1=self.fy, RM1=self.rm,
                  FVX1=self.fvx, FVY1=self.fvy)
        return st

    def set_state(self, st: dict) -> None:
        for k, v in self.s.items():
            if k in st:
                v[...] = np.asarray(st[k]).reshape(v.shape)
        self.dtmin = st.get("DTMIN", self.dtmin)
        self.time = st.get("TIME", self.time)
        self.dtmin1 = st.get("DTMIN1", self.dtmin1)
        self.bandera = int(round(st.get("BANDERA", self.bandera)))
        self.iter = int(round(st.get("ITER", self.iter)))
