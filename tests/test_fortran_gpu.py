"""The boundary as the Fortran host uses it: tests/fortran_hotpath.f90 (Gamma) and
tests/fortran_qpoint.f90 (q points) are
built with amdflang against include/west_hip.f90 and libwesthip.so, reads a
ground state, and calls wh_set_*, wh_apply_operator, wh_dfpt (optional ierr
array absent) and wh_davidson_diago (per-step ev/conv log) on the GPU; the
results are compared with the oracle."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def _write_gs(path, gs, sq, dvg, nvec, tr2):
    b = gs.basis
    nkb = gs.vkb.shape[0]
    g2 = np.zeros(gs.npwx)
    g2[:b.npw] = b.g2kin
    sqx = np.zeros(gs.npwx)
    sqx[:len(sq)] = sq
    with open(path, "wb") as f:
        np.array([*b.nr, b.npw, gs.npwx, b.gstart, nkb, gs.nbndval, dvg.shape[0], nvec], dtype=np.int32).tofile(f)
        np.array([b.omega, b.tpiba2, gs.weight, gs.alpha_pv, tr2], dtype=np.float64).tofile(f)
        np.asarray(b.nl, dtype=np.int32).tofile(f)
        np.asarray(b.nlm, dtype=np.int32).tofile(f)
        g2.tofile(f)
        np.asarray(gs.vrs, dtype=np.float64).tofile(f)
        if nkb:
            np.ascontiguousarray(gs.vkb, dtype=np.complex128).tofile(f)
            np.asarray(gs.dvan, dtype=np.float64).tofile(f)
        np.ascontiguousarray(gs.evc[:gs.nbndval], dtype=np.complex128).tofile(f)
        np.asarray(gs.et[:gs.nbndval], dtype=np.float64).tofile(f)
        np.asarray(gs.occ[:gs.nbndval], dtype=np.float64).tofile(f)
        sqx.tofile(f)
        np.ascontiguousarray(dvg, dtype=np.complex128).tofile(f)


@pytest.mark.gpu
@pytest.mark.skipif(not Path(FC).exists(), reason="amdflang not installed")
@pytest.mark.parametrize("key", ["3h2", "si8"])
def test_fortran_host_calls_the_hot_path(tmp_path, key, request):
    from conftest import random_wfc
    from oracle.davidson import davidson_diago_gamma
    from oracle.dfpt import apply_operator, dfpt, sqvc_init
    from west_amd import build_lib
    gs = request.getfixturevalue(f"gs_{key}")
    sq = sqvc_init(gs.basis, gs.npwx)
    m, nvec, tr2 = 3, 4, 1e-22
    dvg = random_wfc(gs, m, seed=61)
    _write_gs(tmp_path / "gs.bin", gs, sq, dvg, nvec, tr2)
    lib = build_lib.LIB
    subprocess.run([FC, "-J", str(tmp_path), "-c", str(ROOT / "include" / "west_hip.f90"), "-o",
                    str(tmp_path / "west_hip.o")], check=True, capture_output=True)
    exe = tmp_path / "hot"
    subprocess.run([FC, "-J", str(tmp_path), str(ROOT / "tests" / "fortran_hotpath.f90"), str(tmp_path / "west_hip.o"),
                    str(lib), f"-Wl,-rpath,{lib.parent}", "-o", str(exe)], check=True, capture_output=True)
    print(f"case fortran {key}: running the driver", flush=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    npwx = gs.npwx
    with open(tmp_path / "out.bin", "rb") as f:
        dng = np.fromfile(f, dtype=np.complex128, count=npwx * m).reshape(m, npwx)
        dng2 = np.fromfile(f, dtype=np.complex128, count=npwx * m).reshape(m, npwx)
        nlog = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        ev = np.fromfile(f, dtype=np.float64, count=nvec)
        conv = np.fromfile(f, dtype=np.int32, count=nvec)
        ev_log = np.fromfile(f, dtype=np.float64, count=nvec * nlog).reshape(nlog, nvec)
        conv_log = np.fromfile(f, dtype=np.int32, count=nvec * nlog).reshape(nlog, nvec)
        dav_iter = np.fromfile(f, dtype=np.int32, count=nlog)
        dfpt_dim = np.fromfile(f, dtype=np.int32, count=nlog)
        ncalls = np.fromfile(f, dtype=np.int64, count=nlog)

    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    assert rel(dng, apply_operator(gs, sq, dvg, tr2)) < 1e-9
    assert rel(dng2, dfpt(gs, dvg, tr2)) < 1e-9
    log_ref = []
    ev_ref, conv_ref, _ = davidson_diago_gamma(gs, sq, nvec, tr2_dfpt=1e-20, trev_pdep=1e-6, trev_pdep_rel=1e-4,
                                               log=log_ref)
    assert np.max(np.abs(ev - ev_ref) / np.abs(ev_ref)) < 1e-8
    assert np.array_equal(conv.astype(bool), conv_ref)
    assert list(dav_iter) == [rec["dav_iter"] for rec in log_ref]
    assert list(dfpt_dim) == [rec["dfpt_dim"] for rec in log_ref]
    np.testing.assert_array_equal(ev_log[-1], ev)
    np.testing.assert_array_equal(conv_log[-1].astype(bool), conv_ref)
    for row, rec in zip(ev_log, log_ref):
        assert np.max(np.abs(row - np.asarray(rec["ev"])) / np.abs(np.asarray(rec["ev"]))) < 1e-7
    assert np.all(ncalls > 0)


@pytest.mark.gpu
@pytest.mark.skipif(not Path(FC).exists(), reason="amdflang not installed")
def test_fortran_host_calls_the_qpoint_path(tmp_path):
    """tests/fortran_qpoint.f90: the k-grid registration, apply_operator at q
    point 2 and davidson_diago_k at q point 2, from Fortran, against the oracle."""
    from oracle import kpoint as K
    from west_amd import build_lib
    from west_amd.coulomb import sqvc_init
    from west_amd.kpoint import make_h2_crystal, make_k_ground_state
    kgs = make_k_ground_state(make_h2_crystal(10.0), (2, 1, 1), nbnd=4)
    m, nvec, tr2 = 2, 4, 1e-20
    qb = kgs.qpts[1].basis
    rng = np.random.default_rng(71)
    dvg = np.zeros((m, kgs.npwqx), complex)
    dvg[:, :qb.npw] = (rng.normal(size=(m, qb.npw)) + 1j * rng.normal(size=(m, qb.npw))) / (1.0 + qb.g2kin)
    with open(tmp_path / "kgs.bin", "wb") as f:
        np.array([*kgs.nr, kgs.nks, kgs.npwx, len(kgs.qpts), kgs.npwqx, kgs.nbndval, kgs.nkb, m, nvec],
                 dtype=np.int32).tofile(f)
        np.array([kgs.omega, kgs.alpha_pv, tr2]).tofile(f)
        np.asarray(kgs.vrs, dtype=np.float64).tofile(f)
        np.array([k.weight for k in kgs.kpts]).tofile(f)
        if kgs.nkb:
            np.asarray(kgs.dvan, dtype=np.float64).tofile(f)
        for kp in kgs.kpts:
            np.array([kp.basis.npw], dtype=np.int32).tofile(f)
            np.asarray(kp.basis.nl, dtype=np.int32).tofile(f)
            np.asarray(kp.basis.g2kin, dtype=np.float64).tofile(f)
            np.ascontiguousarray(kp.evc[:kgs.nbndval]).tofile(f)
            np.asarray(kp.et[:kgs.nbndval], dtype=np.float64).tofile(f)
            if kgs.nkb:
                np.ascontiguousarray(kp.vkb).tofile(f)
        for qp in kgs.qpts:
            b = qp.basis
            np.array([b.npw], dtype=np.int32).tofile(f)
            np.asarray(b.nl, dtype=np.int32).tofile(f)
            np.asarray(b.g2kin, dtype=np.float64).tofile(f)
            np.asarray(qp.ig_global, dtype=np.int32).tofile(f)
            sqvc_init(b).tofile(f)
            np.asarray(qp.ikq + 1, dtype=np.int32).tofile(f)
            np.asarray(qp.g0, dtype=np.int32).tofile(f)
        np.ascontiguousarray(dvg).tofile(f)
    lib = build_lib.LIB
    subprocess.run([FC, "-J", str(tmp_path), "-c", str(ROOT / "include" / "west_hip.f90"), "-o",
                    str(tmp_path / "west_hip.o")], check=True, capture_output=True)
    exe = tmp_path / "qpt"
    subprocess.run([FC, "-J", str(tmp_path), str(ROOT / "tests" / "fortran_qpoint.f90"), str(tmp_path / "west_hip.o"),
                    str(lib), f"-Wl,-rpath,{lib.parent}", "-o", str(exe)], check=True, capture_output=True)
    print("case fortran qpoint: running the driver", flush=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(tmp_path / "kout.bin", "rb") as f:
        dng = np.fromfile(f, dtype=np.complex128, count=m * kgs.npwqx).reshape(m, kgs.npwqx)
        ev = np.fromfile(f, dtype=np.float64, count=nvec)
        conv = np.fromfile(f, dtype=np.int32, count=nvec)
    sq = sqvc_init(qb, kgs.npwqx)
    ref = K.apply_operator_q(kgs, 1, sq, dvg, tr2)
    assert float(np.abs(dng - ref).max() / np.abs(ref).max()) < 1e-9
    ev_ref, conv_ref, _ = K.davidson_diago_k(kgs, 1, sq, nvec, n_pdep_times=4, n_pdep_maxiter=40, trev_pdep=1e-8,
                                             trev_pdep_rel=1e-7, tr2_dfpt=1e-18)
    assert np.max(np.abs(ev - ev_ref) / np.abs(ev_ref)) < 1e-8
    assert np.array_equal(conv.astype(bool), conv_ref)
