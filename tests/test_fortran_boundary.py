"""The Fortran side of the drop-in boundary.

CPU: the shim modules (cfd_amd/fortran/cfd_shims.f90), compiled against the
reference's own module sources by oracle/Makefile.ref, define every procedure
symbol of the reference objects they replace, so the reference's driver links
against them unchanged.

GPU: a Fortran program (cfd_amd/fortran/abi_demo.f90, built by build()) calls
libcfd_amd.so through the iso_c_binding interface module; its tables, deriv,
MASAS and laplace outputs match the oracle bit for bit and its biCG solve
matches the same solve made through the Python binding.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "seq")
SHIM_OBJ = os.path.join(ROOT, "oracle", "_ref", "shims", "cfd_shims.o")

# reference object -> procedure symbols the shims replace
REPLACED = {
    "pointNeighbor.o": ["_QMpointneighborPgetesup", "_QMpointneighborPgetpsup"],
    "biconjGrad.o": ["_QMbiconjgradPbicg"],
    "mLaplace.o": ["_QMmlaplacePlaplace"],
    "calcRHS.o": ["_QMcalcrhs_modPcalcrhs"],
    "gcl.o": ["_QMgcl_modPmain", "_QMgcl_modPputw", "_QMgcl_modPputarea"],
    "subrutinas.o": ["_QMmnormalesPnormales", "_QMmnormalesPnormalvel", "deriv_", "masas_",
                     "deltat_", "estab_", "fuente_", "rk_"],
}


def _defined(obj):
    out = subprocess.run(["nm", "--defined-only", obj], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


@pytest.mark.ref
def test_shims_define_the_reference_symbols():
    if not os.path.isdir("/root/reference") or shutil.which("flang") is None:
        pytest.skip("needs the reference sources and flang")
    subprocess.run(["make", "-s", "-C", ROOT, "-f", "oracle/Makefile.ref", "all", "shims"],
                   check=True, capture_output=True)
    shim = _defined(SHIM_OBJ)
    for obj, syms in REPLACED.items():
        ref = _defined(os.path.join(REF_OBJ, obj))
        for s in syms:
            assert s in ref, f"{s} not in reference {obj}"
            assert s in shim, f"shim does not define {s} (from {obj})"


@pytest.mark.gpu
def test_fortran_driver_through_the_c_abi(tmp_path):
    from cfd_amd.api import biCG
    from cfd_amd.build import ABI_DEMO
    from cfd_amd.mesh import ellipse_case
    from oracle.oracle import OracleSolver

    if not os.path.isfile(ABI_DEMO):
        pytest.skip("Fortran demo not built (flang absent at build time)")
    cs = ellipse_case(32, 64)
    o = OracleSolver(cs)
    m = cs.mesh
    fidx = np.ascontiguousarray(cs.lists["ilaux"], np.int32)
    xfix = np.random.default_rng(21).standard_normal(fidx.size) * 1e-3
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([m.npoin, m.nelem], np.int32).tofile(f)
        np.ascontiguousarray(m.inpoel, np.int32).tofile(f)
        m.x.astype(np.float64).tofile(f)
        m.y.astype(np.float64).tofile(f)
        np.array([fidx.size], np.int32).tofile(f)
        fidx.tofile(f)
        xfix.tofile(f)
    print(f"running {ABI_DEMO} on {m.nelem} triangles", flush=True)
    r = subprocess.run([ABI_DEMO, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = open(fout, "rb").read()
    npsup, nnz = np.frombuffer(raw[:8], np.int32)
    ne, npn = m.nelem, m.npoin
    off = 8
    ints = {}
    for k, n in (("esup1", 3 * ne), ("esup2", npn + 1), ("psup1", npsup), ("psup2", npn + 1),
                 ("lap_rowptr", npn + 1), ("lap_idx", nnz)):
        ints[k] = np.frombuffer(raw[off:off + 4 * n], np.int32)
        off += 4 * n
    dbl = {}
    for k, n in (("AREA", ne), ("DNX", 3 * ne), ("DNY", 3 * ne), ("HH", ne), ("HHX", ne),
                 ("HHY", ne), ("hmin", 1), ("M", npn), ("LAP_SPARSE", nnz), ("LAP_DIAG", npn),
                 ("x", npn)):
        dbl[k] = np.frombuffer(raw[off:off + 8 * n], np.float64)
        off += 8 * n
    iters = int(np.frombuffer(raw[off:off + 4], np.int32)[0])
    for k in ints:
        assert np.array_equal(ints[k], o.tab[k]), k
    for k in ("AREA", "DNX", "DNY", "HH", "HHX", "HHY", "M", "LAP_SPARSE", "LAP_DIAG"):
        assert np.array_equal(dbl[k].view(np.uint64), o.s[k].reshape(-1).view(np.uint64)), k
    # the same solve through the Python binding: identical bits
    x = np.zeros(npn)
    it_py, _ = biCG(o.s["LAP_SPARSE"], o.tab["lap_idx"], o.tab["lap_rowptr"], o.s["LAP_DIAG"], x,
                    np.zeros(npn), xfix, fidx)
    assert iters == it_py and iters > 5
    assert np.array_equal(dbl["x"].view(np.uint64), x.view(np.uint64))


def test_fortran_ids_match_the_header():
    """The field ids and scalar slots the Fortran module declares are the
    header's enums."""
    import re
    from cfd_amd import _lib
    src = open(os.path.join(ROOT, "cfd_amd", "fortran", "cfd_abi.f90")).read()
    ids = {k.upper(): int(v) for k, v in re.findall(r"\b(cfd_[fs]_\w+|cfd_nscalars)\s*=\s*(\d+)", src)}
    fields = {k: v for k, v in _lib.FIELDS.items() if k != "CFD_NFIELDS"}
    scalars = {k: v for k, v in _lib.SCALARS.items() if k != "CFD_NSCALARS"}
    assert {k: ids[k] for k in fields} == fields
    assert {k: ids[k] for k in scalars} == scalars
    assert ids["CFD_NSCALARS"] == _lib.SCALARS["CFD_NSCALARS"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_session_driver(tmp_path, precision):
    """A Fortran program runs the device-resident session (create, free
    stream, steps, fields, scalars) through the interface module; its state
    equals the Python binding's session bit for bit."""
    from cfd_amd.api import consts_of
    from cfd_amd.build import SESSION_DEMO
    from cfd_amd.mesh import acoustic_fsafe, cylinder_case
    from cfd_amd.solver import Session

    if not os.path.isfile(SESSION_DEMO):
        pytest.skip("Fortran demo not built (flang absent at build time)")
    cs = cylinder_case(24, 48, mach=0.5, re=300.0, alpha_deg=2.0, fsafe=acoustic_fsafe(0.5))
    m, L = cs.mesh, cs.lists
    nsteps = 10
    c = consts_of(cs)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    i32 = lambda a: np.ascontiguousarray(a, np.int32).tofile(f)
    f64 = lambda a: np.ascontiguousarray(a, np.float64).tofile(f)
    with open(fin, "wb") as f:
        i32([m.npoin, m.nelem, nsteps, precision])
        f64([getattr(c, n) for n, _ in c._fields_])
        i32(m.inpoel)
        f64(m.x)
        f64(m.y)
        for ids, vals in (("ifixv_node", ("rfixv_valuex", "rfixv_valuey")),
                          ("ifixrho_node", ("rfixrho_value",)), ("ifixt_node", ("rfixt_value",))):
            i32([L[ids].size])
            i32(L[ids])
            for v in vals:
                f64(L[v])
        i32([L["iset"].shape[0]])
        i32(L["iset"])
        i32([L["nmove"], L["nfix_move"]])
        i32(L["ilaux"])
        f64([cs.xref, cs.yref])
    print(f"running {SESSION_DEMO}, precision {precision}", flush=True)
    r = subprocess.run([SESSION_DEMO, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(fout, np.float64)
    U = raw[:4 * m.npoin].reshape(-1, 4)
    T = raw[4 * m.npoin:5 * m.npoin]
    sc = raw[5 * m.npoin:]
    S = Session(cs, precision=precision)
    S.step(nsteps)
    assert np.isfinite(U).all()
    assert np.array_equal(U, S.get("U")) and np.array_equal(T, S.get("T"))
    ref = S.scalars()
    from cfd_amd._lib import SCALARS
    for k in ("DTMIN", "TIME", "ITER", "FX1", "FY1"):
        assert sc[SCALARS["CFD_S_" + {"FX1": "FX", "FY1": "FY"}.get(k, k)]] == ref[k], k
    S.close()


@pytest.mark.ref
@pytest.mark.ref
def test_reference_main_program_links_against_the_shims():
    """The reference's main program (ns2DComp.ALE.f90, unchanged) compiles
    against the shim modules and links with them, the reference's module
    sources, its loader and meshMove.f90, and libcfd_amd.so, with no symbol
    left unresolved: every hot-path routine it calls (getEsup/getPsup through
    initialize, deriv, MASAS, laplace, NORMALES, deltat, RK, biCG through
    fluidStructure) and smoothing_mod (start-up smoothing; the reference's
    smoothing.f90 does not compile with flang) come from the shims."""
    if not os.path.isdir("/root/reference") or shutil.which("flang") is None:
        pytest.skip("needs the reference sources and flang")
    r = subprocess.run(["make", "-s", "-C", ROOT, "-f", "oracle/Makefile.ref", "mainprog"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    exe = os.path.join(ROOT, "oracle", "_ref", "ns2d_shims")
    main_o = os.path.join(ROOT, "oracle", "_ref", "mainp", "ns2DComp.ALE.o")
    undefined = subprocess.run(["nm", "--undefined-only", main_o], capture_output=True, text=True,
                               check=True).stdout
    shim = _defined(SHIM_OBJ)
    # what the main program takes from the shims
    for sym in ("_QMsmoothing_modPsmoothing", "_QMmlaplacePlaplace", "_QMmnormalesPnormales",
                "deltat_", "deriv_", "masas_", "rk_"):
        assert sym in undefined, sym
        assert sym in shim, sym
    assert "_QMsmoothing_modPsettol_move" in shim
    # the executable resolves everything (the link has no unresolved-symbol allowance)
    assert os.path.isfile(exe)
    dyn = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True,
                         check=True).stdout
    assert "cfd_rk" in dyn and "cfd_smoothing" in dyn


def test_kept_subrutinas_compile_against_the_shims(tmp_path):
    """The routines of subrutinas.f90 that a drop-in keeps (CUARTO_ORDEN,
    fixvel, FIX, ADAMSB: everything the shims do not replace) compile
    against the shim modules: ADAMSB's `use Mnormales` and its NORMALVEL call
    resolve to the shim's Mnormales.  The kept lines are cut from the
    reference source into a temporary directory at test time (nothing of the
    reference is stored in this repository) and linked with the shim object
    into a relocatable object that defines adamsb_."""
    if not os.path.isdir("/root/reference") or shutil.which("flang") is None:
        pytest.skip("needs the reference sources and flang")
    subprocess.run(["make", "-s", "-C", ROOT, "-f", "oracle/Makefile.ref", "shims"], check=True,
                   capture_output=True)
    lines = open("/root/reference/subrutinas.f90").read().splitlines(keepends=True)

    import re

    def block(name):
        start = re.compile(r"\s*subroutine\s+" + name + r"\b", re.I)
        end = re.compile(r"\s*end\s+subroutine\s+" + name + r"\b", re.I)
        lo = next(i for i, l in enumerate(lines) if start.match(l))
        hi = next(i for i in range(lo, len(lines)) if end.match(lines[i]))
        return "".join(lines[lo:hi + 1])

    kept = lines[0] + "".join(block(n) for n in ("CUARTO_ORDEN", "fixvel", "FIX", "ADAMSB"))
    assert "call normalvel" in kept.lower() and "use mnormales" in kept.lower()
    src = tmp_path / "kept.f90"
    src.write_text(kept)
    shims = os.path.join(ROOT, "oracle", "_ref", "shims")
    r = subprocess.run(["flang", "-cpp", "-O0", "-I", shims, "-I", os.path.join(shims, "ref"), "-c",
                        str(src), "-o", str(tmp_path / "kept.o")], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    undefined = subprocess.run(["nm", "--undefined-only", str(tmp_path / "kept.o")],
                               capture_output=True, text=True, check=True).stdout
    assert "_QMmnormalesPnormalvel" in undefined            # resolved by the shim object
    assert "_QMmnormalesPnormalvel" in _defined(SHIM_OBJ)
    assert "adamsb_" in _defined(str(tmp_path / "kept.o"))


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["slip", "ell"])
def test_reference_calling_sequence_through_the_shims(which, tmp_path):
    """The drop-in run: the reference's loader, start-up and step sequence
    (the harness program, with fluidStructure compiled from meshMove.f90)
    linked against the shim modules, so every hot-path call (NORMALES, DERIV,
    MASAS, laplace, DELTAT, ESTAB, calcRHS, FUENTE, RK, biCG, NORMALVEL) goes
    through cfd_shims.f90 and the C ABI to the GPU.  Against the oracle:
    start-up geometry bit for bit; each unit on the oracle's developed state
    within 1e-12 (on the pitching ellipse with its nonzero mesh velocity, and
    fluidStructure's mesh move with the shim biCG to its absolute rule);
    whole steps from the start-up within 1e-12 on the fixed mesh.  (On a
    moving mesh a free run is no parity test: the absolute PCG rule lets two
    differently rounded solves stop 1e-6 apart, and the mesh velocity
    W = XPOS / DTMIN carries that into the next step's momentum, 36 % of the
    y-momentum norm after one step here, on the oracle itself.)"""
    from cfd_amd.mesh import acoustic_fsafe, cylinder_case, ellipse_case
    from oracle import refharness as rh
    from oracle.oracle import OracleSolver
    from tests.helpers import bitdiff, relnorm, relnorm_cols
    if not os.path.isfile(rh.SHIM_BIN):
        pytest.skip("shim harness not built (needs the reference sources and flang at build time)")
    cs = (cylinder_case(32, 64, mach=0.5, re=500.0, slip_body=True, fsafe=acoustic_fsafe(0.5))
          if which == "slip" else ellipse_case(24, 48))
    rc = rh.RefCase(cs, str(tmp_path), binary=rh.SHIM_BIN)
    o = OracleSolver(cs)
    init = rc.init_state()
    print(f"shim harness start-up on {cs.nelem} triangles done", flush=True)
    for k in ("AREA", "DNX", "DNY", "HH", "M", "LAP_SPARSE", "LAP_DIAG"):
        assert bitdiff(init[k], o.s[k]) == 0, k
    for _ in range(6):
        o.step()
    st = o.state()
    if which == "ell":
        assert np.abs(st["W_Y"]).max() > 0                  # FUENTE and ESTAB see W
    # single units on the oracle's developed state
    for unit, fields in (("deltat", ("DT",)), ("estab", ("SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3")),
                         ("rk", ("U1", "RHS", "T", "VEL_X", "VEL_Y", "P")),
                         ("normales", ("VEL_X", "VEL_Y")), ("deriv", ("AREA", "DNX", "DNY")),
                         ("masas", ("M",)), ("laplace", ("LAP_SPARSE", "LAP_DIAG")),
                         ("fluid", ("X", "Y", "W_X", "W_Y"))):
        got = rc.unit(unit, st)
        ref_o = OracleSolver(cs)
        ref_o.set_state(st)
        if unit == "deltat":
            ref_o.deltat()
        elif unit == "estab":
            ref_o.estab(st["DTMIN"])
        elif unit == "rk":
            ref_o.rk(st["DTMIN"])
        elif unit == "normales":
            ref_o.normales()
            vx, vy = ref_o.normalvel(st["VEL_X"], st["VEL_Y"], st["W_X"], st["W_Y"])
            ref_o.s["VEL_X"][:], ref_o.s["VEL_Y"][:] = vx, vy
        elif unit == "deriv":
            ref_o.deriv()
        elif unit == "masas":
            ref_o.masas()
        elif unit == "laplace":
            ref_o.laplace()
        else:   # the harness sets the previous pitch angle by a call at TIME - DTMIN
            ref_o.disn2 = ref_o.pitch_angle(st["TIME"] - st["DTMIN"])
            ref_o.fluid_structure(st["DTMIN"], st["TIME"])
        ref = ref_o.state()
        if unit in ("deriv", "masas", "laplace"):       # gathers in element order: bitwise
            for f in fields:
                assert bitdiff(got[f], ref[f]) == 0, (unit, f)
        elif unit == "fluid":
            # the mesh move: both solves stop under the absolute rule, 1e-6 apart at most
            for f in ("XPOS", "YPOS"):
                assert relnorm(got[f], ref[f]) < 1e-6 or np.abs(ref[f]).max() == 0, (unit, f)
            for f in ("X", "Y"):
                assert relnorm(got[f], ref[f]) < 1e-12, (unit, f)
            if which == "ell":
                assert relnorm(ref["X"], st["X"]) > 0        # the body moved
        else:
            for f in fields:
                err = (relnorm_cols(got[f], ref[f]) if np.ndim(ref[f]) == 2
                       else [relnorm(got[f], ref[f])])
                assert np.all(np.asarray(err) < 1e-12), (unit, f, err)
        print(f"shim unit {unit}: ok", flush=True)
    if which == "ell":
        return
    # three whole steps of the reference's sequence from the start-up (fixed mesh)
    runs = rc.run(3)
    o2 = OracleSolver(cs)
    for k, got in enumerate(runs):
        o2.step()
        ref = o2.state()
        assert got["DTMIN"] == ref["DTMIN"]
        for f in ("U", "RHS"):
            assert np.all(relnorm_cols(got[f], ref[f]) < 1e-12), (k, f)
        for f in ("T", "P", "X", "Y", "M"):
            assert relnorm(got[f], ref[f]) < 1e-12, (k, f)
    print(f"shim harness: three steps match the oracle ({which})", flush=True)
