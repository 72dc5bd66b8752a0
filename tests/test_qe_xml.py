"""Ground-state description from pw.x output (west_amd/qe_xml.py), pinned by
the pw.x runs the reference's test-suite keeps (test-suite/testNNN/ref/pw.xml,
copied to tests/golden/qe by copy_fixtures.sh): the basis built here from a
run's cell and cutoffs reproduces that run's FFT grid, G-vector count and
plane-wave count."""
from pathlib import Path

import numpy as np
import pytest

from west_amd.gsphere import count_sphere, fft_dims, make_gamma_basis
from west_amd.qe_xml import read_pw_xml, system_from_pw_xml

GOLD = Path(__file__).resolve().parent / "golden" / "qe"


def _run(t):
    return read_pw_xml(str(GOLD / f"qe_{t}.xml"))


def test_reader_fields():
    r = _run("test001")   # SiH4 in a 20 bohr cube, Gamma, 25 / 100 Ry
    assert r.species == ["Si", "H"] and r.atoms == ["Si", "H", "H", "H", "H"]
    np.testing.assert_allclose(r.at, 20.0 * np.eye(3))
    np.testing.assert_allclose(r.tau[1], [11.614581, 11.614581, 11.614581])
    assert (r.ecutwfc, r.ecutrho) == (25.0, 100.0)          # Hartree in the file
    assert r.gamma_only and not r.lsda and r.nbnd == 10 and r.nelec == 8.0
    (k,) = r.kpoints
    assert k.weight == 2.0 and k.npw == 8440 and k.occ.tolist() == [1.0] * 4 + [0.0] * 6
    assert k.et[0] == 2 * -4.864490346183685e-01            # Ry
    assert r.pseudo_files == ["Si_ONCV_PBE-1.1.upf", "H_ONCV_PBE-1.0.upf"]


@pytest.mark.parametrize("t", ["test001", "test008", "test014"])
def test_gamma_runs_basis(t):
    """Gamma runs (a cube, a free cell of MgO, an fcc cell): FFT grid, half
    density sphere (ngm) and half wave-function sphere (npwx, npw)."""
    r = _run(t)
    b = make_gamma_basis(r.at, r.ecutwfc, r.ecutrho, alat=r.alat)
    assert tuple(b.nr) == r.fft_grid
    assert count_sphere(r.at, r.ecutrho, half=True) == r.ngm
    assert b.npw == r.npwx == r.kpoints[0].npw


def test_kpoint_run_basis():
    """A k-point run (fcc Si): FFT grid, full density sphere and the full
    wave-function sphere of its k = 0 point."""
    r = _run("test004")
    assert not r.gamma_only
    assert fft_dims(r.at, r.ecutrho) == r.fft_grid
    assert count_sphere(r.at, r.ecutrho, half=False) == r.ngm
    k0 = [k for k in r.kpoints if not np.any(k.xk)]
    assert len(k0) == 1 and count_sphere(r.at, r.ecutwfc, half=False) == k0[0].npw


def test_system_from_run():
    s = system_from_pw_xml(_run("test001"))
    assert s.name == "SiH4_25Ry" and s.nelec == 8 and s.nbnd_occ == 4
    assert [sp.name for sp in s.species] == ["Si", "H"] and s.ityp.tolist() == [0, 1, 1, 1, 1]
    assert s.basis.npw == 8440 and tuple(s.basis.nr) == (64, 64, 64)
    with pytest.raises(ValueError, match="k points"):
        system_from_pw_xml(_run("test004"))
    with pytest.raises(ValueError, match="spin polarised"):
        system_from_pw_xml(_run("test014"))
    with pytest.raises(KeyError):
        system_from_pw_xml(_run("test008"))   # Mg, O: no pseudopotentials here
