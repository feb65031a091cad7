"""Test infrastructure only: drives oracle/_ref/ref_harness (the compiled
reference) and reads its binary outputs.

State layout written by ref_harness.f90 (save_state), all little-endian:
  int64  npoin, nelem, nnz
  float64 scalars[10]: DTMIN TIME DTMIN1 BANDERA ITER HMIN FX1 FY1 RM1 0
  nodal arrays, element arrays and CSR values in the order of STATE_FIELDS.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref", "ref_harness")
# the same harness linked against the Fortran shims (libcfd_amd.so) instead of
# the reference's hot-path objects (oracle/Makefile.ref shimharness)
SHIM_BIN = os.path.join(HERE, "_ref", "shim_harness")

SCALARS = ["DTMIN", "TIME", "DTMIN1", "BANDERA", "ITER", "HMIN", "FX1", "FY1", "RM1", "_pad"]
# (name, kind, width): kind n = npoin, e = nelem, z = nnz
STATE_FIELDS = [
    ("U", "n", 4), ("U1", "n", 4), ("RHS", "n", 4), ("T", "n", 1), ("VEL_X", "n", 1),
    ("VEL_Y", "n", 1), ("W_X", "n", 1), ("W_Y", "n", 1), ("X", "n", 1), ("Y", "n", 1),
    ("X1", "n", 1), ("Y1", "n", 1), ("RHO", "n", 1), ("E", "n", 1), ("P", "n", 1),
    ("RMACH", "n", 1), ("M", "n", 1), ("GAMM", "n", 1),
    ("XPOS", "n", 1), ("YPOS", "n", 1), ("POS_AUX", "n", 1), ("B", "n", 1),
    ("W_X_OLD", "n", 1), ("W_Y_OLD", "n", 1),
    ("DTL", "e", 1), ("DT", "e", 1), ("SHOC", "e", 1), ("T_SUGN1", "e", 1),
    ("T_SUGN2", "e", 1), ("T_SUGN3", "e", 1), ("AREA", "e", 1), ("DNX", "e", 3),
    ("DNY", "e", 3), ("HH", "e", 1), ("HHX", "e", 1), ("HHY", "e", 1), ("AREA_OLD", "e", 1),
    ("LAP_SPARSE", "z", 1), ("LAP_DIAG", "n", 1),
]


def available() -> bool:
    return os.path.isfile(REF_BIN) and os.access(REF_BIN, os.X_OK)


def state_nbytes(npoin, nelem, nnz):
    n = 24 + 80
    for _, k, w in STATE_FIELDS:
        n += 8 * w * {"n": npoin, "e": nelem, "z": nnz}[k]
    return n


def read_states(path: str) -> list[dict]:
    raw = open(path, "rb").read()
    out = []
    off = 0
    while off < len(raw):
        npoin, nelem, nnz = np.frombuffer(raw, np.int64, 3, off)
        off += 24
        st = {"npoin": int(npoin), "nelem": int(nelem), "nnz": int(nnz)}
        sc = np.frombuffer(raw, np.float64, 10, off)
        off += 80
        for name, v in zip(SCALARS, sc):
            st[name] = float(v)
        for name, k, w in STATE_FIELDS:
            cnt = w * {"n": npoin, "e": nelem, "z": nnz}[k]
            a = np.frombuffer(raw, np.float64, cnt, off).copy()
            off += 8 * cnt
            st[name] = a.reshape(-1, w) if w > 1 else a
        out.append(st)
    return out


def write_state(path: str, st: dict) -> None:
    with open(path, "wb") as fh:
        np.array([st["npoin"], st["nelem"], st["nnz"]], np.int64).tofile(fh)
        np.array([st.get(s, 0.0) for s in SCALARS], np.float64).tofile(fh)
        for name, k, w in STATE_FIELDS:
            np.ascontiguousarray(st[name], dtype=np.float64).reshape(-1).tofile(fh)


def read_tables(path: str) -> dict:
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype, n, off).copy()
        off += a.nbytes
        return a

    npoin, nelem, nesup, npsup, nnz, nfixrho, nfixv, nfixt = take(np.int64, 8)
    t = {}
    t["esup2"] = take(np.int32, npoin + 1)
    t["esup1"] = take(np.int32, nesup)
    t["psup2"] = take(np.int32, npoin + 1)
    t["psup1"] = take(np.int32, npsup)
    t["lap_rowptr"] = take(np.int32, npoin + 1)
    t["lap_idx"] = take(np.int32, nnz)
    t["ifixrho_node"] = take(np.int32, nfixrho)
    t["rfixrho_value"] = take(np.float64, nfixrho)
    t["ifixv_node"] = take(np.int32, nfixv)
    t["rfixv_valuex"] = take(np.float64, nfixv)
    t["rfixv_valuey"] = take(np.float64, nfixv)
    t["ifixt_node"] = take(np.int32, nfixt)
    t["rfixt_value"] = take(np.float64, nfixt)
    nset_numb, nmove, nfix_move, nwall = take(np.int64, 4)
    t["nset_numb"], t["nmove"], t["nfix_move"], t["nwall"] = map(int, (nset_numb, nmove, nfix_move, nwall))
    nsetedges = int(take(np.int64, 1)[0])
    t["iset"] = take(np.int32, 3 * nsetedges).reshape(-1, 3)
    t["ilaux"] = take(np.int32, nmove + nfix_move)
    c = take(np.float64, 20)
    names = ["U_inf", "V_inf", "T_inf", "RHO_inf", "P_inf", "C_inf", "FR", "FMU", "FK", "FCv",
             "GAMA", "FSAFE", "CTE", "XREF", "YREF", "MACH_inf", "HMIN"]
    t["consts"] = {n: float(v) for n, v in zip(names, c)}
    return t


class RefCase:
    """A case directory the reference harness can run in."""

    def __init__(self, case, workdir: str | None = None, binary: str = REF_BIN):
        from cfd_amd.mesh import write_reference_inputs
        self.case = case
        self.binary = binary
        self._tmp = None
        if workdir is None:
            self._tmp = tempfile.TemporaryDirectory(prefix="cfdref_")
            workdir = self._tmp.name
        self.dir = workdir
        write_reference_inputs(case, self.dir)

    def _run(self, *args, threads=1):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([self.binary, *map(str, args)], cwd=self.dir, env=env,
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"{os.path.basename(self.binary)} {args} failed: "
                               f"{r.stdout[-2000:]} {r.stderr[-2000:]}")
        return r

    def path(self, name):
        return os.path.join(self.dir, name)

    def tables(self) -> dict:
        self._run("tables", "tables.bin")
        return read_tables(self.path("tables.bin"))

    def init_state(self) -> dict:
        self._run("init", "init.bin")
        return read_states(self.path("init.bin"))[0]

    def run(self, nsteps: int, state: dict | None = None) -> list[dict]:
        src = "-"
        if state is not None:
            write_state(self.path("in.bin"), state)
            src = "in.bin"
        self._run("run", nsteps, src, "run.bin")
        return read_states(self.path("run.bin"))

    def unit(self, name: str, state: dict) -> dict:
        write_state(self.path("in.bin"), state)
        self._run(name, "in.bin", "out.bin")
        return read_states(self.path("out.bin"))[0]
