"""O-grid cases in the reference solver's vocabulary.

A case is a mesh (nodes, P1 triangles), the boundary-condition lists the
reference loader builds (dataLoader.f90:95-275) and the free-stream constants
it derives (dataLoader.f90:17-65).  `write_reference_inputs` writes the three
text files that loader parses (EULER.DAT, <name>-1.dat, <name>.dat) so that the
compiled reference and this package start from the same bits.

Index conventions follow the reference: node and element ids stored in the
arrays are 1-based; offsets are 0-based.  `inpoel` has numpy shape (nelem, 3)
in C order, which is the memory of the Fortran array inpoel(3, nelem).

O-grid construction
-------------------
Node (i_r, i_t) has id 1 + i_t + Nt*i_r (theta fastest).  Ring radii (in the
body-to-farfield blend parameter s) are geometric with ratio
q = 1.05**(63/(Nr-1)), so Nr=64 gives exactly 1.05 and every Nr has the same
total stretch.  Each quad (i_r, i_t)-(i_r+1, i_t+1) is split along the same
diagonal into two counter-clockwise triangles with ids
1 + 2*(i_t + Nt*i_r) + {0, 1}.  Interior rings get seeded jitter of +/-10% of
the local spacing in s (radial) and theta (tangential), drawn from splitmix64.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

SEED = 20261202
R_BODY = 0.5
R_FAR = 20.0

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Draws k = offset+1 .. offset+n of splitmix64(seed) as doubles in [-1, 1)."""
    k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def _ellipse_arclength_angle(t: np.ndarray, a: float, b: float) -> np.ndarray:
    """Ellipse parameter phi whose arc length from phi=0 is t/(2 pi) of the
    perimeter (tabulated on a fine grid, linearly interpolated)."""
    n = 1 << 16
    ph = np.linspace(0.0, 2.0 * math.pi, n + 1)
    ds = np.hypot(a * np.sin(ph), b * np.cos(ph))
    s = np.concatenate([[0.0], np.cumsum(0.5 * (ds[1:] + ds[:-1]) * (ph[1] - ph[0]))])
    target = np.mod(t, 2.0 * math.pi) / (2.0 * math.pi) * s[-1]
    return np.interp(target, s, ph) + (t - np.mod(t, 2.0 * math.pi))


@dataclass
class Mesh:
    x: np.ndarray          # (npoin,) float64
    y: np.ndarray          # (npoin,) float64
    inpoel: np.ndarray     # (nelem, 3) int32, 1-based, CCW
    nr: int
    nt: int
    body: str

    @property
    def npoin(self) -> int:
        return int(self.x.shape[0])

    @property
    def nelem(self) -> int:
        return int(self.inpoel.shape[0])

    def node(self, ir, it):
        return 1 + (np.asarray(it) % self.nt) + self.nt * np.asarray(ir)


def ogrid(nr: int, nt: int, body: str = "circle", seed: int = SEED,
          jitter: float = 0.1) -> Mesh:
    """O-grid around a circle (r=0.5) or an ellipse (semi-axes 0.5, 0.06)."""
    if nr < 2 or nt < 3:
        raise ValueError("O-grid needs nr >= 2 and nt >= 3")
    q = 1.05 ** (63.0 / (nr - 1)) if nr > 1 else 1.05
    k = np.arange(nr, dtype=np.float64)
    s = (q ** k - 1.0) / (q ** (nr - 1) - 1.0)            # 0 .. 1 geometric
    s[0], s[-1] = 0.0, 1.0
    th = 2.0 * math.pi * np.arange(nt, dtype=np.float64) / nt
    S = np.repeat(s, nt).reshape(nr, nt)
    TH = np.tile(th, nr).reshape(nr, nt)
    if nr > 2 and jitter > 0.0:
        ids = np.arange(nr * nt, dtype=np.int64)
        u = splitmix64_uniform(seed, 2 * nr * nt)
        ur = u[0::2].reshape(nr, nt)
        ut = u[1::2].reshape(nr, nt)
        ds = np.empty(nr)
        ds[1:-1] = np.minimum(s[1:-1] - s[:-2], s[2:] - s[1:-1])
        interior = np.zeros((nr, nt), bool)
        interior[1:-1, :] = True
        dth = 2.0 * math.pi / nt
        S = np.where(interior, S + jitter * ur * ds[:, None], S)
        TH = np.where(interior, TH + jitter * ut * dth, TH)
        del ids
    if body == "circle":
        bx, by = R_BODY * np.cos(TH), R_BODY * np.sin(TH)
    elif body == "ellipse":
        # body nodes equally spaced in arc length (angle-uniform nodes give
        # tip elements of quality ~0.1 on a 0.12-thick ellipse)
        phi = _ellipse_arclength_angle(TH, 0.5, 0.06)
        bx, by = 0.5 * np.cos(phi), 0.06 * np.sin(phi)
    else:
        raise ValueError(f"unknown body {body!r}")
    fx, fy = R_FAR * np.cos(TH), R_FAR * np.sin(TH)
    x = ((1.0 - S) * bx + S * fx).reshape(-1)
    y = ((1.0 - S) * by + S * fy).reshape(-1)

    ir, it = np.meshgrid(np.arange(nr - 1), np.arange(nt), indexing="ij")
    ir = ir.reshape(-1)
    it = it.reshape(-1)
    a = 1 + it + nt * ir
    b = 1 + (it + 1) % nt + nt * ir
    c = 1 + (it + 1) % nt + nt * (ir + 1)
    d = 1 + it + nt * (ir + 1)
    inpoel = np.empty((2 * (nr - 1) * nt, 3), dtype=np.int32)
    inpoel[0::2, 0], inpoel[0::2, 1], inpoel[0::2, 2] = a, d, c
    inpoel[1::2, 0], inpoel[1::2, 1], inpoel[1::2, 2] = a, c, b
    return Mesh(x=np.ascontiguousarray(x), y=np.ascontiguousarray(y),
                inpoel=inpoel, nr=nr, nt=nt, body=body)


@dataclass
class Case:
    """Mesh + loader-equivalent constants and boundary lists."""
    name: str
    mesh: Mesh
    mach: float
    re: float
    alpha_deg: float = 0.0
    moving: int = 0
    xref: float = 0.0
    yref: float = 0.0
    maxiter: int = 1
    iprint: int = 1000000
    fsafe: float = 0.5
    t_inf: float = 288.15
    gama: float = 1.4
    fmu: float = 0.017
    prandtl: float = 0.72
    cte_in: float = 1.0
    itlocal: int = 0
    slip_body: bool = False      # body edges as slip walls (NORMALES) instead of no-slip
    # pitching law of a moving body: 0 = the reference's pi/8 sin(10 t)
    # (meshMove.f90:55-57,70); 1 = motion_amp sin(motion_omega t) (radians)
    motion: int = 0
    motion_amp: float = 0.0
    motion_omega: float = 0.0
    # derived (filled by __post_init__) -------------------------------------
    fr: float = field(init=False)
    u_inf: float = field(init=False)
    v_inf: float = field(init=False)
    rho_inf: float = field(init=False)
    p_inf: float = field(init=False)
    c_inf: float = field(init=False)
    fcv: float = field(init=False)
    fk: float = field(init=False)
    cte: float = field(init=False)
    twall: float = field(init=False)
    lists: dict = field(init=False)

    def __post_init__(self):
        # Nondimensional units in which the solver's hard-coded viscosity
        # 0.017 (subrutinas.f90:196,400) is also the kinematic viscosity that
        # its time-step bound assumes: rho_inf = 1, body length 1, FMU = 0.017,
        # |U_inf| = FMU*Re, and the gas constant chosen so c_inf = |U_inf|/M.
        g = self.gama
        speed = self.fmu * self.re
        c = speed / self.mach
        self.fr = c * c / (g * self.t_inf)
        self.fcv = self.fr / (g - 1.0)
        self.fk = self.fmu * g * self.fcv / self.prandtl
        a = math.radians(self.alpha_deg)
        self.u_inf = speed * math.cos(a)
        self.v_inf = speed * math.sin(a)
        self.rho_inf = 1.0
        # dataLoader.f90:59-64 (with T_inf, RHO_inf given, P_inf = 0 in file)
        self.cte = 1.0 / self.cte_in
        self.p_inf = self.rho_inf * self.fr * self.t_inf
        self.c_inf = math.sqrt(g * self.fr * self.t_inf)
        # dataLoader.f90:161: TWALL is implicitly typed (loadMeshData has no
        # IMPLICIT NONE), so it is a default REAL and is rounded to binary32
        self.twall = float(np.float32(self.t_inf * (1.0 + (g - 1) / 2.0 * self.mach * self.mach)))
        self.lists = _boundary_lists(self)

    # sizes -----------------------------------------------------------------
    @property
    def npoin(self):
        return self.mesh.npoin

    @property
    def nelem(self):
        return self.mesh.nelem


def _boundary_lists(cs: Case) -> dict:
    """Boundary lists as the loader leaves them (dataLoader.f90:130-268)."""
    m = cs.mesh
    body = m.node(0, np.arange(m.nt)).astype(np.int32)
    outer = m.node(m.nr - 1, np.arange(m.nt)).astype(np.int32)
    one = np.ones(m.nt)
    # fixed rho on the outer ring (value 1 -> 1*RHO_inf)
    ifixrho = outer.copy()
    rfixrho = one * cs.rho_inf
    # inflow velocity on the outer ring, then no-slip on the body
    # (dataLoader.f90:149-176: inflow first, no-slip appended, NFIXV += NFIXVI)
    noslip = np.zeros(0, np.int32) if cs.slip_body else body
    ifixv = np.concatenate([outer, noslip]).astype(np.int32)
    rvx = np.concatenate([one * cs.u_inf, np.zeros(noslip.size)])
    rvy = np.concatenate([one * cs.v_inf, np.zeros(noslip.size)])
    # fixed T: no-slip nodes first (TWALL), then the listed nodes (value*T_inf)
    ifixt = np.concatenate([noslip, outer]).astype(np.int32)
    rfixt = np.concatenate([np.full(noslip.size, cs.twall), one * cs.t_inf])
    # slip walls: body edges in the wall list (normal velocity removed by NORMALVEL)
    if cs.slip_body:
        it0 = np.arange(m.nt)
        wall = np.stack([m.node(0, it0), m.node(0, it0 + 1)], axis=1).astype(np.int32)
    else:
        wall = np.zeros((0, 2), np.int32)
    # force set 1: body edges, listed so that FORCES gives the force on the body
    it = np.arange(m.nt)
    n1 = m.node(0, it + 1).astype(np.int32)
    n2 = m.node(0, it).astype(np.int32)
    eset = (1 + 2 * it + 1).astype(np.int32)
    iset = np.stack([n1, n2, eset], axis=1)       # (nsets, 3): n1 n2 elem
    if cs.moving:
        i_m = body.copy()
        ifm = outer.copy()
    else:
        i_m = np.zeros(0, np.int32)
        ifm = np.concatenate([body, outer]).astype(np.int32)
    ilaux = np.concatenate([i_m, ifm]).astype(np.int32)
    return dict(
        ifixrho_node=ifixrho, rfixrho_value=rfixrho,
        nfixvi=m.nt, ifixv_node=ifixv, rfixv_valuex=rvx, rfixv_valuey=rvy,
        ifixt_node=ifixt, rfixt_value=rfixt,
        wall=wall,
        iset=iset, nset_numb=1,
        i_m=i_m, ifm=ifm, ilaux=ilaux,
        nmove=int(i_m.size), nfix_move=int(ifm.size),
    )


def cylinder_case(nr=64, nt=128, mach=0.3, re=100.0, name="cyl", **kw) -> Case:
    return Case(name=name, mesh=ogrid(nr, nt, "circle"), mach=mach, re=re, **kw)


def acoustic_fsafe(mach: float, cfl: float = 0.5) -> float:
    """FSAFE giving an acoustic CFL of `cfl` where deltat's advective bound
    (which uses |v| but not the sound speed, subrutinas.f90:194-206) governs."""
    return cfl * mach / (1.0 + mach)


def ellipse_case(nr=512, nt=1024, mach=0.5, re=5000.0, name="ell", moving=1,
                 xref=-0.25, yref=0.0, fsafe=None, **kw) -> Case:
    fs = acoustic_fsafe(mach) if fsafe is None else fsafe
    return Case(name=name, mesh=ogrid(nr, nt, "ellipse"), mach=mach, re=re,
                moving=moving, xref=xref, yref=yref, fsafe=fs, **kw)


def sweep_cases(n: int = 256, seed: int = SEED, nr=(32, 256), nt=(64, 512), mach=(0.2, 0.85),
                alpha_deg=(-5.0, 5.0), re=(50.0, 5000.0), cfl: float = 0.5) -> list:
    """The batched sweep: n fixed-mesh cylinder cases of uneven size.  Case k
    draws, from splitmix64(seed + k) far past the mesh jitter's draws, Nr
    uniform in nr, Ntheta uniform in nt (inclusive), M uniform in mach, the
    angle of attack uniform in alpha_deg (set through the free-stream
    direction) and Re log-uniform in re; its O-grid is jittered with seed +
    k.  FSAFE gives each case an acoustic CFL of cfl (deltat bounds only the
    advective and viscous rates)."""
    cases = []
    for k in range(n):
        u = 0.5 * (splitmix64_uniform(seed + k, 5, offset=1 << 40) + 1.0)  # [0, 1)
        nr_k = nr[0] + min(int(u[0] * (nr[1] - nr[0] + 1)), nr[1] - nr[0])
        nt_k = nt[0] + min(int(u[1] * (nt[1] - nt[0] + 1)), nt[1] - nt[0])
        m_k = mach[0] + u[2] * (mach[1] - mach[0])
        a_k = alpha_deg[0] + u[3] * (alpha_deg[1] - alpha_deg[0])
        re_k = math.exp(math.log(re[0]) + u[4] * (math.log(re[1]) - math.log(re[0])))
        cases.append(Case(name=f"sweep{k}", mesh=ogrid(nr_k, nt_k, "circle", seed=seed + k),
                          mach=m_k, re=re_k, alpha_deg=a_k, fsafe=acoustic_fsafe(m_k, cfl)))
    return cases


# ---------------------------------------------------------------------------
# Reference text format (dataLoader.f90:17-65 and 95-275)
# ---------------------------------------------------------------------------
def _f(v: float) -> str:
    return "%.17e" % v


def write_reference_inputs(cs: Case, directory: str) -> None:
    """Writes EULER.DAT, <name>-1.dat and <name>.dat for the reference loader."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "EULER.DAT"), "w") as fh:
        fh.write(cs.name + "\n")
    with open(os.path.join(directory, cs.name + "-1.dat"), "w") as fh:
        w = fh.write
        w("IRESTART MAXITER IPRINT MOVIE ITLOCAL\n")
        w(f"0 {cs.maxiter} {cs.iprint} 0 {cs.itlocal}\n")
        w("FSAFE U_inf V_inf MACH_inf T_inf RHO_inf P_inf\n")
        w(" ".join(_f(v) for v in (cs.fsafe, cs.u_inf, cs.v_inf, cs.mach,
                                    cs.t_inf, cs.rho_inf, 0.0)) + "\n")
        w("FMU FGX FGY QH\n")
        w(f"{_f(cs.fmu)} 0.0 0.0 0.0\n")
        w("FK FR FCv GAMA NGAS\n")
        w(f"{_f(cs.fk)} {_f(cs.fr)} {_f(cs.fcv)} {_f(cs.gama)} 0\n")
        w("CTE\n")
        w(f"{_f(cs.cte_in)}\n")
        w("ALE\n")
        w("MOVING XREF YREF\n")
        w(f"{cs.moving} {_f(cs.xref)} {_f(cs.yref)}\n")
        w("PRINT\n")
        w("flags\n")
        w("rho vel mach pres temp ener pos\n")
        w("REFINEMENT\n")
        w("h1\n")
        w("h2\n")
        w("h3\n")
        w("0.0 0.0 0.0\n")
    L = cs.lists
    m = cs.mesh
    nfixv_noslip = L["ifixv_node"].size - L["nfixvi"]
    nfixt_listed = L["ifixt_node"].size - nfixv_noslip
    with open(os.path.join(directory, cs.name + ".dat"), "w") as fh:
        w = fh.write
        w("NPOIN NELEM\n")
        w(f"{m.npoin} {m.nelem}\n")
        w("NFIXRHO NFIXVI NFIXV NWALL NFIXT NSETS NMASTER NSLAVE NFIX_MOVE NMOVE\n")
        w(f"{L['ifixrho_node'].size} {L['nfixvi']} {nfixv_noslip} "
          f"{L['wall'].shape[0]} {nfixt_listed} {L['iset'].shape[0]} 0 0 "
          f"{L['nfix_move']} {L['nmove']}\n")
        w("h\nh\nh\nCOORDINATES\n")
        ids = np.arange(1, m.npoin + 1)
        np.savetxt(fh, np.column_stack([ids, m.x, m.y]), fmt=["%d", "%.17e", "%.17e"])
        w("ELEMENTS\n")
        np.savetxt(fh, np.column_stack([np.arange(1, m.nelem + 1), m.inpoel]), fmt="%d")
        w("FIXRHO\n")
        for n in L["ifixrho_node"]:
            w(f"{n} 1.0\n")
        w("INFLOW\n")
        for n in L["ifixv_node"][: L["nfixvi"]]:
            w(f"{n} 1.0 1.0\n")
        w("NOSLIP\n")
        for n in L["ifixv_node"][L["nfixvi"]:]:
            w(f"{n} 0.0 0.0\n")
        w("WALL\n")
        for a, b in L["wall"]:
            w(f"{a} {b}\n")
        w("FIXT\n")
        for n in L["ifixt_node"][nfixv_noslip:]:
            w(f"{n} 1.0\n")
        w("SETS\n")
        for n1, n2, e in L["iset"]:
            w(f"{e} {n1} {n2} 1\n")
        w("MASTER\n")
        w("SLAVE\n")
        w("FIXMOVE\n")
        for n in L["ifm"]:
            w(f"{n} 0\n")
        w("MOVE\n")
        for n in L["i_m"]:
            w(f"{n} 0\n")
