"""Restart files (.RST): RESTART (ns2DComp.ALE.f90:394-435) and PRINTREST
(:898-917).

CPU, with the compiled reference (oracle/_ref/ref_harness, whose RESTART and
PRINTREST are the reference's own, compiled from its source): the library's
writer produces byte for byte the reference's file for the same state; the
reference's RESTART reads the library's file into the same U, T, GAMM; the
library reads the reference's file.  CPU, without it: misframed, short and
missing files are refused with an error.  GPU: a session checkpoints and
restores its state through the file."""
import os
import struct

import numpy as np
import pytest

from cfd_amd.api import read_restart, write_restart
from cfd_amd.mesh import cylinder_case


def _state(cs, seed=0):
    rng = np.random.default_rng(seed)
    n = cs.npoin
    U = np.column_stack([1 + rng.random(n), rng.standard_normal((n, 2)), 3 + rng.random(n)])
    return 17, 0.123456789012345, U, 280 + 10 * rng.random(n), np.full(n, 1.4)


@pytest.mark.ref
def test_restart_matches_the_reference(tmp_path):
    from oracle import refharness as rh
    if not rh.available():
        pytest.skip("compiled reference not built")
    cs = cylinder_case(12, 24)
    rc = rh.RefCase(cs, str(tmp_path))
    st = rc.init_state()
    it, tm, U, T, G = _state(cs)
    st.update(U=U, T=T, GAMM=G, TIME=tm)
    # the reference writes <name>.RST from the state (PRINTREST)
    rh.write_state(rc.path("in.bin"), st)
    rc._run("rstwrite", "in.bin", it)
    ref_file = rc.path(cs.name + ".RST")
    ref_bytes = open(ref_file, "rb").read()
    assert len(ref_bytes) == 20 + 56 * cs.npoin
    mine = str(tmp_path / "mine.RST")
    write_restart(mine, it, tm, U, T, G)
    assert open(mine, "rb").read() == ref_bytes
    # the library reads the reference's file
    it2, tm2, U2, T2, G2 = read_restart(ref_file, cs.npoin)
    assert (it2, tm2) == (it, tm)
    assert np.array_equal(U2, U) and np.array_equal(T2, T) and np.array_equal(G2, G)
    # the reference's RESTART reads the library's file (U, T, GAMM)
    os.replace(mine, ref_file)
    rc._run("rstread", "out.bin")
    out = rh.read_states(rc.path("out.bin"))[0]
    assert np.array_equal(out["U"], U) and np.array_equal(out["T"], T)
    assert np.array_equal(out["GAMM"], G)


def test_restart_rejects_bad_files(tmp_path):
    from cfd_amd._lib import CfdError
    cs = cylinder_case(6, 12)
    it, tm, U, T, G = _state(cs, 1)
    p = str(tmp_path / "a.RST")
    write_restart(p, it, tm, U, T, G)
    raw = open(p, "rb").read()
    assert struct.unpack("<i", raw[:4])[0] == 12 and struct.unpack("<i", raw[16:20])[0] == 12
    # a file for more nodes is read up to npoin (the rest is ignored, as RESTART does)
    it3, _, U3, _, _ = read_restart(p, cs.npoin - 5)
    assert it3 == it and np.array_equal(U3, U[:-5])
    with pytest.raises(CfdError, match="ends before"):
        read_restart(p, cs.npoin + 1)
    bad = bytearray(raw)
    bad[20:24] = struct.pack("<i", 40)            # a node record claiming 40 bytes
    open(p, "wb").write(bytes(bad))
    with pytest.raises(CfdError, match="record 1"):
        read_restart(p, cs.npoin)
    open(p, "wb").write(raw[:-3])                  # truncated last record
    with pytest.raises(CfdError, match="truncated"):
        read_restart(p, cs.npoin)
    with pytest.raises(CfdError, match="cannot open"):
        read_restart(str(tmp_path / "missing.RST"), cs.npoin)


@pytest.mark.gpu
def test_session_checkpoint_through_a_restart_file(tmp_path):
    """A session writes its state after 12 steps; a fresh session restores it:
    U and T bit for bit, U1 = U, the primitives of U, the clock kept or
    restarted as asked; its next step lands within 1e-12 of the oracle's step
    from the same restored state."""
    from cfd_amd.solver import Session
    from oracle.oracle import OracleSolver
    from tests.helpers import relnorm_cols
    cs = cylinder_case(32, 64, mach=0.8, re=1000.0)
    A = Session(cs, precision=0, renumber=1)
    A.step(12)
    path = str(tmp_path / "cyl.RST")
    A.write_restart(path)
    sa = A.scalars()
    it, tm, U, T, G = read_restart(path, cs.npoin)
    assert it == 12 and tm == sa["TIME"]
    assert np.array_equal(U, A.get("U")) and np.array_equal(T, A.get("T"))
    for keep in (False, True):
        B = Session(cs, precision=0)
        B.read_restart(path, keep_clock=keep)
        sb = B.scalars()
        assert np.array_equal(B.get("U"), U) and np.array_equal(B.get("U1"), U)
        assert np.array_equal(B.get("T"), T)
        assert np.array_equal(B.get("VEL_X"), U[:, 1] / U[:, 0])
        assert (sb["TIME"], sb["ITER"]) == ((tm, 12.0) if keep else (0.0, 0.0))
        assert sb["BANDERA"] == 1.0
        B.close()
    # a step from the restored state against the oracle's from the same state
    B = Session(cs, precision=0)
    B.read_restart(path, keep_clock=True)
    o = OracleSolver(cs)
    o.set_state(B.state())
    o.step()
    B.step(1)
    ref, got = o.state(), B.state()
    assert got["DTMIN"] == ref["DTMIN"]
    assert np.all(relnorm_cols(got["U"], ref["U"]) < 1e-12)
    assert np.all(relnorm_cols(got["RHS"], ref["RHS"]) < 1e-12)
    A.close()
    B.close()


_FORTRAN_RESTART = """\
program restart_io
    use iso_c_binding
    use cfd_abi
    implicit none
    integer, parameter :: n = 7
    real(c_double) :: u(4, n), t(n), g(n), u2(4, n), t2(n), g2(n), tm
    integer(c_int) :: i, k, it
    do i = 1, n
        do k = 1, 4
            u(k, i) = i + k / 8.0d0
        end do
        t(i) = 280.0d0 + i / 3.0d0
        g(i) = 1.4d0
    end do
    call cfd_check(cfd_write_restart('f.RST' // c_null_char, 42_c_int, 0.375d0, n, u, t, g), &
                   'write')
    call cfd_check(cfd_read_restart('f.RST' // c_null_char, n, it, tm, u2, t2, g2), 'read')
    if (it /= 42 .or. tm /= 0.375d0 .or. any(u2 /= u) .or. any(t2 /= t) .or. any(g2 /= g)) &
        error stop 'restart round trip differs'
    print '(a)', 'ok'
end program restart_io
"""


def test_fortran_interface_writes_the_restart_file(tmp_path):
    """A Fortran caller writes and reads a restart file through the
    iso_c_binding interfaces of cfd_amd/fortran/cfd_abi.f90 (host-only
    entries: no GPU); its file is byte for byte the Python binding's."""
    import shutil
    import subprocess
    from cfd_amd import build as B
    abi_o = os.path.join(B.OBJ, "cfd_abi_f.o")
    if shutil.which("flang") is None or not os.path.isfile(abi_o):
        pytest.skip("needs flang and the built Fortran interface module")
    src = tmp_path / "restart_io.f90"
    src.write_text(_FORTRAN_RESTART)
    exe = str(tmp_path / "restart_io")
    subprocess.run(["flang", "-O0", "-I", B.OBJ, "-c", str(src), "-o", exe + ".o"],
                   check=True, capture_output=True, cwd=tmp_path)
    subprocess.run(["flang", exe + ".o", abi_o, "-L" + B.PKG, "-lcfd_amd",
                    "-Wl,-rpath," + B.PKG, "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
    i = np.arange(1, 8, dtype=np.float64)
    U = i[:, None] + np.arange(1, 5) / 8.0
    write_restart(str(tmp_path / "py.RST"), 42, 0.375, U, 280.0 + i / 3.0, np.full(7, 1.4))
    assert (tmp_path / "f.RST").read_bytes() == (tmp_path / "py.RST").read_bytes()
