"""PDEP database in the reference's on-disk format (IO_kernel/pdep_db.f90,
IO_kernel/pdep_io.f90, IO_kernel/mod_west_io.f90:22-35).

The reference holds no PDEP files among its fixtures, so the format is pinned
by a Fortran program that writes and reads one eigenvector file with the
reference's stream-I/O statements (tests/pdep_stream.f90, compiled with
amdflang): the Python writer must produce the same bytes, and each side must
read the other's file.  The summary is checked through the json paths the
reference's reader queries (pdep_db.f90:288-301)."""
from __future__ import annotations

import json
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from west_amd import pdep_db

ROOT = Path(__file__).resolve().parents[1]
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if Path("/opt/rocm/bin/amdflang").exists() else None)


def _stream_vector(n: int) -> np.ndarray:
    k = np.arange(1, n + 1, dtype=np.float64)
    return k + 1j * (-k / 2 + 1.0 / 3.0)


@pytest.fixture(scope="module")
def fortran_stream(tmp_path_factory):
    if FC is None:
        pytest.skip("amdflang not installed")
    exe = tmp_path_factory.mktemp("pdep_fc") / "pdep_stream"
    subprocess.run([FC, "-O1", "-o", str(exe), str(ROOT / "tests" / "pdep_stream.f90")], check=True,
                   capture_output=True, timeout=120)
    return str(exe)


def test_file_names():
    # generate_pdep_fname: 'Q'//i9.9(iq)//'E'//i9.9(j)//'.dat', 25 characters
    assert pdep_db.pdep_fname(1) == "Q000000001E000000001.dat"
    assert pdep_db.pdep_fname(37, iq=2) == "Q000000002E000000037.dat"
    assert len(pdep_db.pdep_fname(123456789, iq=987654321)) == 24
    with pytest.raises(ValueError):
        pdep_db.pdep_fname(10**9)


@pytest.mark.parametrize("n", [1, 7, 4099])
def test_vector_bytes_match_fortran_stream(fortran_stream, tmp_path, n):
    ffile, pfile = tmp_path / "f.dat", tmp_path / "p.dat"
    subprocess.run([fortran_stream, "w", str(ffile), str(n)], check=True, timeout=60)
    pdep_db.write_vector(str(pfile), _stream_vector(n))
    fb, pb = ffile.read_bytes(), pfile.read_bytes()
    assert len(fb) == 4 * 32 + 16 * n
    assert fb == pb
    # Fortran reads the Python file; Python reads the Fortran file
    out = subprocess.run([fortran_stream, "r", str(pfile), str(n)], check=True, capture_output=True, text=True,
                         timeout=60).stdout.split()
    v = _stream_vector(n)
    assert int(out[0]) == 1                     # little-endian flag on this host
    assert float(out[1]) == pytest.approx(v.real.sum(), rel=1e-15)
    assert float(out[2]) == pytest.approx(v.imag.sum(), rel=1e-13)
    assert float(out[3]) == v.real[-1]
    np.testing.assert_array_equal(pdep_db.read_vector(str(ffile), n), v)


def test_header_fields_and_errors(tmp_path):
    f = tmp_path / "v.dat"
    v = np.array([1 + 2j, -3.5 + 0.25j, 0j])
    pdep_db.write_vector(str(f), v)
    hdr = np.frombuffer(f.read_bytes()[:128], dtype=np.int32)
    assert hdr[0] == 210405 and hdr[1] == 1 and hdr[2] == 3 and not hdr[3:].any()
    with pytest.raises(ValueError, match="Dimension mismatch"):
        pdep_db.read_vector(str(f), 4)
    bad = bytearray(f.read_bytes())
    bad[0:4] = np.int32(210404).tobytes()
    (tmp_path / "bad.dat").write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="Unknown file format"):
        pdep_db.read_vector(str(tmp_path / "bad.dat"), 3)
    bad[0:4] = np.int32(210405).tobytes()
    bad[4:8] = np.int32(0).tobytes()
    (tmp_path / "be.dat").write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="Endianness mismatch"):
        pdep_db.read_vector(str(tmp_path / "be.dat"), 3)
    with pytest.raises(OSError, match="Cannot read file"):
        pdep_db.read_vector(str(tmp_path / "missing.dat"), 3)


def test_db_round_trip_and_summary(tmp_path):
    rng = np.random.default_rng(7)
    npwq, npwqx, n = 53, 60, 6
    at = np.diag([10.0, 11.0, 12.5])
    bg = 2 * np.pi * np.linalg.inv(at).T
    ev = -np.sort(rng.random(n))[::-1]
    dvg = rng.standard_normal((n, npwqx)) + 1j * rng.standard_normal((n, npwqx))
    save = str(tmp_path / "west.wstat.save")
    names = pdep_db.pdep_db_write(save, ev, dvg, npwq, at, bg)
    assert names == [f"Q000000001E{j:09d}.dat" for j in range(1, n + 1)]
    doc = json.loads((tmp_path / "west.wstat.save" / "summary.json").read_text())
    dm = doc["dielectric_matrix"]
    np.testing.assert_allclose(dm["domain"]["a2"], [0, 11.0, 0])
    np.testing.assert_allclose(np.array([dm["domain"][f"b{i}"] for i in (1, 2, 3)]) @ at.T, 2 * np.pi * np.eye(3),
                               atol=1e-13)
    assert [r["iq"] for r in dm["pdep"]] == [1] and dm["pdep"][0]["q"] == [0.0, 0.0, 0.0]
    assert dm["pdep"][0]["eigenval"] == list(ev)
    for name in names:
        assert os.path.getsize(os.path.join(save, name)) == 128 + 16 * npwq

    ev2, dvg2 = pdep_db.pdep_db_read(save, npwq, npwqx)
    np.testing.assert_array_equal(ev2, ev)
    np.testing.assert_array_equal(dvg2[:, :npwq], dvg[:, :npwq])
    assert not dvg2[:, npwq:].any()
    ev3, dvg3 = pdep_db.pdep_db_read(save, npwq, nglob_to_be_read=4)
    assert ev3.shape == (4,) and dvg3.shape == (4, npwq)

    # a second q point is appended; rewriting q point 1 replaces its record
    pdep_db.pdep_db_write(save, ev[:2], dvg, npwq, at, bg, iq=2, q=(0.5, 0, 0))
    pdep_db.pdep_db_write(save, ev[:3] * 2, dvg, npwq, at, bg, iq=1)
    dm = json.loads((tmp_path / "west.wstat.save" / "summary.json").read_text())["dielectric_matrix"]
    assert [r["iq"] for r in dm["pdep"]] == [1, 2]
    assert dm["pdep"][0]["eigenval"] == list(ev[:3] * 2) and dm["pdep"][1]["q"] == [0.5, 0.0, 0.0]
    ev4, _ = pdep_db.pdep_db_read(save, npwq, iq=2)
    np.testing.assert_array_equal(ev4, ev[:2])
    with pytest.raises(ValueError, match="no record"):
        pdep_db.pdep_db_read(save, npwq, iq=3)
