"""ctypes loader for libcfd_amd.so (the gfx950 HIP library).

There is no CPU fallback: if the library is not built, or no GPU is usable,
every call raises CfdError with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcfd_amd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "cfd_abi.h")


class CfdError(RuntimeError):
    pass


class Consts(C.Structure):
    """cfd_consts (InputData constants)."""
    _fields_ = [(n, C.c_double) for n in
                ("gama", "fr", "fmu", "fk", "fcv", "t_inf", "rho_inf", "u_inf", "v_inf", "cte",
                 "fsafe")]


_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)


class BcLists(C.Structure):
    _fields_ = [("nfixv", C.c_int), ("ifixv_node", _pi), ("rfixv_valuex", _pd),
                ("rfixv_valuey", _pd), ("nfixrho", C.c_int), ("ifixrho_node", _pi),
                ("rfixrho_value", _pd), ("nfixt", C.c_int), ("ifixt_node", _pi),
                ("rfixt_value", _pd), ("nwall", C.c_int), ("wall", _pi)]


class RkNodes(C.Structure):
    _fields_ = [(n, _pd) for n in ("U1", "RHS", "T", "VEL_X", "VEL_Y", "RHO", "E", "P", "RMACH",
                                    "U", "W_X", "W_Y", "M", "GAMM")]


class RkElems(C.Structure):
    _fields_ = [(n, _pd) for n in ("dNx", "dNy", "area", "shoc", "t_sugn1", "t_sugn2", "t_sugn3")]


class CaseDesc(C.Structure):
    _fields_ = [("c", Consts), ("nrk", C.c_int), ("moving", C.c_int), ("itlocal", C.c_int),
                ("xref", C.c_double), ("yref", C.c_double), ("nsetedges", C.c_int),
                ("iset", _pi), ("nmove", C.c_int), ("nfix_move", C.c_int), ("ilaux", _pi),
                ("bc", BcLists), ("X", _pd), ("Y", _pd), ("motion", C.c_int),
                ("motion_amp", C.c_double), ("motion_omega", C.c_double), ("pcg_rel", C.c_int),
                ("pcg_tol", C.c_double), ("pcg_maxit", C.c_int), ("gcl", C.c_int),
                ("precision", C.c_int), ("warm_start", C.c_int), ("patch_elems", C.c_int),
                ("patch_block", C.c_int), ("renumber", C.c_int)]


def _enum(name_prefix: str) -> dict:
    """Reads an enum block's names from the header (ids are their positions)."""
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for block in re.findall(r"enum\s*\{([^}]*)\}", txt):
        names = [n.strip().split("=")[0].strip() for n in block.split(",")]
        names = [n for n in names if n]
        if names and names[0].startswith(name_prefix):
            for i, n in enumerate(names):
                out[n] = i
    return out


FIELDS = _enum("CFD_F_")
SCALARS = _enum("CFD_S_")
KTIMERS = _enum("CFD_K_")

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(LIB_PATH):
        raise CfdError(f"{LIB_PATH} is not built (run __graft_entry__.build() or python -m cfd_amd.build)")
    L = C.CDLL(LIB_PATH)
    L.cfd_last_error.restype = C.c_char_p
    for name in _exported_names():
        fn = getattr(L, name)
        if name != "cfd_last_error":
            fn.restype = C.c_int
    _lib = L
    return L


def _exported_names() -> list[str]:
    txt = open(HEADER).read()
    return re.findall(r"^\s*(?:int|const char \*)\s*(cfd_[a-z_0-9]+)\(", txt, re.M)


def declared_symbols() -> list[str]:
    return _exported_names()


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().cfd_last_error().decode(errors="replace")
        raise CfdError(f"libcfd_amd error {rc}: {msg}")


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "float64 C-contiguous array expected"
    return a.ctypes.data_as(_pd)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous, "int32 C-contiguous array expected"
    return a.ctypes.data_as(_pi)
