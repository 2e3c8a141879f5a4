"""Density restart files (csrc/den_io.cpp, MOPAC's .den): the library's
writer and reader against the Fortran runtime using den_in_out's own read and
write statements (tests/fortran/den_roundtrip.F90), both forms, plus the
header check and the fallback to the other form."""
import os
import subprocess

import numpy as np
import pytest

from mopac_gpu_amd import lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("MHIP_FC", "/opt/rocm/bin/amdflang")


def _prog(tmp_path_factory):
    exe = os.path.join(HERE, "fortran", "_build", "den_roundtrip")
    src = os.path.join(HERE, "fortran", "den_roundtrip.F90")
    if not os.path.exists(exe) or os.path.getmtime(exe) < os.path.getmtime(src):
        if not os.path.exists(FC):
            pytest.skip("no Fortran compiler")
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        mod = str(tmp_path_factory.mktemp("fmod"))
        subprocess.run([FC, "-O1", "-module-dir", mod, src, "-o", exe], check=True)
    return exe


def _err():
    return (L.load().mhip_last_error() or b"").decode()


@pytest.mark.parametrize("formatted", [0, 1])
def test_den_round_trip_through_fortran(tmp_path, tmp_path_factory, formatted):
    lib = L.load()
    exe = _prog(tmp_path_factory)
    norbs, numat = 37, 11
    mpack = norbs * (norbs + 1) // 2
    rng = np.random.default_rng(formatted)
    pa = rng.standard_normal(mpack)
    pa[::5] = 0.0  # runs of equal values (list-directed output may write r*v)
    pa[7:40] = 0.25
    f1, f2 = str(tmp_path / "a.den"), str(tmp_path / "b.den")
    assert lib.mhip_den_write(f1.encode(), norbs, numat, L.ptr(pa), formatted) == 0, _err()
    subprocess.run([exe, "f" if formatted else "u", str(norbs), str(numat), f1, f2], check=True)
    back = np.zeros(mpack)
    assert lib.mhip_den_read(f2.encode(), norbs, numat, L.ptr(back), formatted) == 0, _err()
    want = pa * 0.5 + 0.125
    if formatted:
        assert np.max(np.abs(back - want)) <= 1e-15 * np.max(np.abs(want))
    else:
        assert np.array_equal(back, want)  # bit for bit
    # the reader falls back to the other form, as den_in_out does
    other = np.zeros(mpack)
    assert lib.mhip_den_read(f2.encode(), norbs, numat, L.ptr(other), 1 - formatted) == 0, _err()
    assert np.array_equal(other, back)


def test_den_header_mismatch_and_missing(tmp_path):
    lib = L.load()
    norbs, numat = 6, 3
    pa = np.arange(21, dtype=np.float64)
    f = str(tmp_path / "w.den")
    assert lib.mhip_den_write(f.encode(), norbs, numat, L.ptr(pa), 0) == 0
    buf = np.zeros(21)
    assert lib.mhip_den_read(f.encode(), norbs, numat, L.ptr(buf), 0) == 0
    assert np.array_equal(buf, pa)
    big = np.zeros(28)
    assert lib.mhip_den_read(f.encode(), 7, numat, L.ptr(big), 0) != 0
    assert "does not match" in _err()
    assert lib.mhip_den_read(str(tmp_path / "none.den").encode(), norbs, numat, L.ptr(buf), 0) != 0
    assert "missing or corrupt" in _err()
    (tmp_path / "bad.den").write_bytes(b"\x05\x00")
    assert lib.mhip_den_read(str(tmp_path / "bad.den").encode(), norbs, numat, L.ptr(buf), 0) != 0
