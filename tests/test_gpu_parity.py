"""Parity of the HIP path with the FP64 oracle, through the C ABI.

Tolerances (FP64 throughout):
* operator applications (FFT + GEMM): relative max error <= 1e-12;
* Sternheimer solutions and chi0 at a tight threshold: relative <= 1e-9
  (two CG trajectories that differ only by rounding);
* PDEP eigenvalues: relative <= 1e-8 (the bar of the plain FP64 path).
"""
import numpy as np
import pytest

from conftest import random_wfc

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _register(gpu, gs, kinetic_only=False, maxiter=250):
    from oracle.dfpt import sqvc_init
    sq = sqvc_init(gs.basis, gs.npwx)
    gpu.set_ground_state(gs, sqvc=sq, n_dfpt_maxiter=maxiter, l_kinetic_only=kinetic_only)
    gs.l_kinetic_only = kinetic_only
    return sq


@pytest.mark.parametrize("key", ["h2", "3h2", "si8"])
@pytest.mark.parametrize("m", [1, 2, 5])
def test_sternheimer_operator(gpu, key, m, request):
    from oracle.hamiltonian import apply_sternheimerop
    gs = request.getfixturevalue(f"gs_{key}")
    _register(gpu, gs)
    print(f"case {key} m={m}", flush=True)
    psi = random_wfc(gs, m, seed=11 + m)
    e = np.linspace(-0.7, 0.3, m)
    ref = apply_sternheimerop(gs, psi, e, gs.alpha_pv)
    out = gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, psi, e, gs.alpha_pv)
    assert _rel(out, ref) < 1e-12


@pytest.mark.parametrize("plane_variant", [None, 1, 2, 4, 5, 6, 9])
@pytest.mark.parametrize("nrep,ecut", [(1, 25.0), (1, 70.0), (2, 25.0)])
def test_operator_workload_grids(gpu, nrep, ecut, plane_variant, monkeypatch):
    """The operator on the workload grids (36^3, 60^3 and 72^3; the plane kernel
    of the last needs more than 64 KB of LDS), with the Harris potential and
    random orthonormal occupied states.  plane_variant fixes the plane kernel
    (WH_PLANE_OCC: 1-3 the LDS-pass kernel sized for that many workgroups per
    CU, where compiled; above, the register-line kernels with 2, 3, 4 planes
    per workgroup where they are built (FP64: grids up to 64) and their LDS
    fits, and last the quad-lane kernels (72^3: 4 with rows bounded by
    |i| <= 18, 5 by 16 where the sphere allows; 60^3: after the
    register-line kernels); a number beyond the last runs the last, so 9 is
    the last variant of each grid); None leaves the choice to the library."""
    if plane_variant is not None:
        monkeypatch.setenv("WH_PLANE_OCC", str(plane_variant))
    from oracle.hamiltonian import apply_sternheimerop
    from west_amd.groundstate import GroundState, _orthonormalize, from_real, to_real
    from west_amd.systems import atomic_density_g, kb_projectors, ks_potential, make_si_system
    s = make_si_system(nrep=nrep, ecutwfc=ecut, seed=9)
    b = s.basis
    vkb, dvan = kb_projectors(s)
    nv = 8
    ev = from_real(_orthonormalize(to_real(random_wfc_basis(b, nv, 3), b.npw, 2)), b.npw, b.npw, 2)
    gs = GroundState(basis=b, npwx=b.npw, vrs=ks_potential(s, atomic_density_g(s)), vkb=np.ascontiguousarray(vkb.T),
                     dvan=dvan, evc=ev, et=np.linspace(-0.5, 0.2, nv), nbndval=nv, occ=np.ones(nv), alpha_pv=1.4)
    _register(gpu, gs)
    print(f"case nrep={nrep} nr={b.nr} npw={b.npw} plane variant {plane_variant}", flush=True)
    psi = random_wfc(gs, 3, seed=7)
    e = np.array([-0.4, -0.1, 0.05])
    ref = apply_sternheimerop(gs, psi, e, gs.alpha_pv)
    out = gpu.apply_sternheimerop_to_m_wfcs(nv, psi, e, gs.alpha_pv)
    monkeypatch.delenv("WH_PLANE_OCC", raising=False)
    assert _rel(out, ref) < 1e-12


def random_wfc_basis(b, m, seed):
    rng = np.random.default_rng(seed)
    a = (rng.normal(size=(m, b.npw)) + 1j * rng.normal(size=(m, b.npw))) / (1.0 + b.g2kin)[None, :]
    a[:, 0] = a[:, 0].real
    return a


def test_runtime_size_fft_path(gpu, gs_si8):
    """The runtime-size kernels (used for grids without a compiled
    specialisation) agree with the oracle too."""
    import os
    from oracle.dfpt import dfpt
    from oracle.hamiltonian import apply_sternheimerop
    gs = gs_si8
    os.environ["WH_FFT_GENERIC"] = "1"
    try:
        _register(gpu, gs)
        psi = random_wfc(gs, 3, seed=17)
        e = np.array([-0.3, 0.0, 0.2])
        assert _rel(gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, psi, e, gs.alpha_pv),
                    apply_sternheimerop(gs, psi, e, gs.alpha_pv)) < 1e-12
        dvg = random_wfc(gs, 2, seed=18)
        assert _rel(gpu.dfpt(2, dvg, -1.0), dfpt(gs, dvg, -1.0)) < 1e-12
    finally:
        del os.environ["WH_FFT_GENERIC"]
        _register(gpu, gs)


def test_kinetic_only_operator(gpu, gs_si8):
    from oracle.hamiltonian import apply_sternheimerop
    gs = gs_si8
    _register(gpu, gs, kinetic_only=True)
    try:
        psi = random_wfc(gs, 4, seed=3)
        e = np.array([-0.5, -0.2, 0.1, 0.3])
        ref = apply_sternheimerop(gs, psi, e, gs.alpha_pv)
        out = gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, psi, e, gs.alpha_pv)
        assert _rel(out, ref) < 1e-13
    finally:
        gs.l_kinetic_only = False


def test_h_psi_without_states(gpu, gs_si8):
    """H psi with no registered states (the ground-state eigensolver's use)."""
    from oracle.hamiltonian import h_psi
    gs = gs_si8
    h = gpu.gpu_h_factory(gs)
    psi = random_wfc(gs, 3, seed=5)
    assert _rel(h(psi), h_psi(gs, psi)) < 1e-12


@pytest.mark.parametrize("key", ["3h2", "si8"])
def test_linsolve(gpu, key, request):
    from oracle.hamiltonian import precondition_m_wfcts, set_eprec
    from oracle.linsolve import linsolve_sternheimer_m_wfcts
    gs = request.getfixturevalue(f"gs_{key}")
    _register(gpu, gs)
    nv = gs.nbndval
    b = random_wfc(gs, nv, seed=21)
    eprec = set_eprec(gs, gs.evc[:nv])
    x0 = precondition_m_wfcts(gs, b, eprec)
    xr = x0.copy()
    ierr_ref = linsolve_sternheimer_m_wfcts(gs, nv, b.copy(), xr, gs.et[:nv], eprec, 1e-24)
    xg = x0.copy()
    ierr = gpu.linsolve_sternheimer_m_wfcts(nv, nv, b.copy(), xg, gs.et[:nv], eprec, 1e-24)
    assert ierr == ierr_ref == 0
    assert _rel(xg, xr) < 1e-9


def test_linsolve_nonconvergence_reported(gpu, gs_si8):
    """n_dfpt_maxiter = 1 leaves systems unconverged: ierr > 0, not an error."""
    from oracle.hamiltonian import precondition_m_wfcts, set_eprec
    gs = gs_si8
    _register(gpu, gs, maxiter=1)
    nv = gs.nbndval
    b = random_wfc(gs, nv, seed=4)
    eprec = set_eprec(gs, gs.evc[:nv])
    x = precondition_m_wfcts(gs, b, eprec)
    ierr = gpu.linsolve_sternheimer_m_wfcts(nv, nv, b, x, gs.et[:nv], eprec, 1e-30)
    assert ierr == nv
    _register(gpu, gs)


@pytest.mark.parametrize("key", ["h2", "3h2", "si8"])
@pytest.mark.parametrize("tr2", [-1.0, 1e-22])
def test_dfpt(gpu, key, tr2, request):
    from oracle.dfpt import dfpt
    gs = request.getfixturevalue(f"gs_{key}")
    _register(gpu, gs)
    print(f"case {key} tr2={tr2}", flush=True)
    m = 3
    dvg = random_wfc(gs, m, seed=31)
    ref = dfpt(gs, dvg, tr2)
    out = gpu.dfpt(m, dvg, tr2)
    assert np.all(out[:, 0] == 0)
    assert _rel(out, ref) < (1e-12 if tr2 < 0 else 1e-9)


def test_apply_operator(gpu, gs_si8):
    from oracle.dfpt import apply_operator
    gs = gs_si8
    sq = _register(gpu, gs)
    dvg = random_wfc(gs, 4, seed=41)
    ref = apply_operator(gs, sq, dvg, 1e-22)
    out = gpu.apply_operator(4, dvg, 1e-22)
    assert _rel(out, ref) < 1e-9


def test_dfpt_device_arrays(gpu, gs_3h2):
    """Device-resident input/output through the same entry point."""
    import torch
    from oracle.dfpt import dfpt
    gs = gs_3h2
    _register(gpu, gs)
    dvg = random_wfc(gs, 2, seed=51)
    ref = dfpt(gs, dvg, 1e-22)
    d_in = torch.from_numpy(dvg).to("cuda")
    d_out = torch.zeros_like(d_in)
    gpu.dfpt(2, d_in, 1e-22, dng=d_out)
    assert _rel(d_out.cpu().numpy(), ref) < 1e-9


@pytest.mark.parametrize("key", ["3h2", "si8"])
def test_davidson(gpu, key, request):
    from oracle.davidson import davidson_diago_gamma
    from oracle.dfpt import sqvc_init
    gs = request.getfixturevalue(f"gs_{key}")
    sq = _register(gpu, gs)
    nvec = 4
    log_ref = []
    ev_ref, conv_ref, _ = davidson_diago_gamma(gs, sq, nvec, tr2_dfpt=1e-20, trev_pdep=1e-6, trev_pdep_rel=1e-4,
                                               log=log_ref)
    ev, conv, dvg, log = gpu.davidson_diago(nvec, tr2_dfpt=1e-20, trev_pdep=1e-6, trev_pdep_rel=1e-4)
    print("ev", ev, "ref", ev_ref, flush=True)
    assert np.all(conv == conv_ref)
    assert np.max(np.abs(ev - ev_ref) / np.abs(ev_ref)) < 1e-8
    assert [r["dfpt_dim"] for r in log] == [r["dfpt_dim"] for r in log_ref]
    assert [r["diago_dim"] for r in log] == [r["diago_dim"] for r in log_ref]
    # the exec.davitr record: per-step eigenvalues and flags, and the
    # reference's 'stern' call count (one per perturbation's solve per CG step)
    # (a perturbation whose last rho lands next to tr2 may take one CG step
    # more or less than in the oracle's rounding)
    for r, rr in zip(log, log_ref):
        assert abs(r["sternop_ncalls"] - rr["sternop_ncalls"]) <= max(1, rr["sternop_ncalls"] // 50)
    for r, rr in zip(log, log_ref):
        assert r["conv"] == list(rr["conv"])
        assert np.max(np.abs(np.array(r["ev"]) - rr["ev"]) / np.abs(rr["ev"])) < 1e-7


def test_dfpt_density_large_grid(gpu):
    """100^3 grid (the 216-atom workload's size): the response density is
    accumulated in two x halves because an FP64 100 x 100 accumulator does not
    fit in LDS beside the plane."""
    from oracle.dfpt import dfpt, sqvc_init
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_h2_system
    s = make_h2_system(nmol=2, L=30.78, ecutwfc=25.0, seed=8)
    assert tuple(s.basis.nr) == (100, 100, 100)
    gs = make_ground_state(s, gpu.gpu_h_factory, nbnd=s.nbnd_occ + 2, scf_steps=0, tol=1e-6)
    gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))
    print(f"case 2H2 L=30.78 nr={s.basis.nr} npw={s.basis.npw}", flush=True)
    dvg = random_wfc(gs, 3, seed=61)
    ref = dfpt(gs, dvg, -1.0)
    out = gpu.dfpt(3, dvg, -1.0)
    assert _rel(out, ref) < 1e-12


def test_systems_concurrent(gpu, gs_h2, gs_3h2, gs_si8):
    """Three systems registered side by side, chi0 applied from three host
    threads at once: each matches the oracle, and system 0 keeps its state."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.dfpt import dfpt, sqvc_init
    cases = [gs_h2, gs_3h2, gs_si8]
    handles = [gpu.system_create() for _ in cases]
    try:
        for h, gs in zip(handles, cases):
            gpu.system_select(h)
            gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))
        gpu.system_select(0)
        dvgs = [random_wfc(gs, 3, seed=71 + i) for i, gs in enumerate(cases)]

        def work(i):
            gpu.system_select(handles[i])
            return gpu.dfpt(3, dvgs[i], 1e-22)

        with ThreadPoolExecutor(max_workers=3) as ex:
            outs = list(ex.map(work, range(3)))
        for gs, dvg, out in zip(cases, dvgs, outs):
            assert _rel(out, dfpt(gs, dvg, 1e-22)) < 1e-9
        res = gpu.davidson_batch([(handles[i], {"n_pdep_eigen": 2, "tr2_dfpt": 1e-18, "want_vectors": False})
                                  for i in range(3)])
        assert all(len(r[0]) == 2 for r in res)
    finally:
        gpu.system_select(0)
        for h in handles:
            gpu.system_destroy(h)


@pytest.mark.parametrize("key", ["h2", "3h2"])
def test_cg_graphs_match_eager(gpu, key, request, monkeypatch):
    """Small batches replay their CG iterations from HIP graphs (captured the
    second time an active-column count comes up): the response is bit-identical
    to the all-eager run (WH_GRAPH=0), over many replays with the loop's host
    synchronisations between them, and matches the FP64 oracle."""
    from oracle.dfpt import dfpt
    gs = request.getfixturevalue(f"gs_{key}")
    _register(gpu, gs)
    dvg = random_wfc(gs, 4, seed=37)
    ref = dfpt(gs, dvg, 1e-22)
    monkeypatch.setenv("WH_GRAPH", "0")
    eager = gpu.dfpt(4, dvg, 1e-22)
    monkeypatch.delenv("WH_GRAPH")
    c0 = gpu.counters()
    replay = gpu.dfpt(4, dvg, 1e-22)
    c1 = gpu.counters()
    print(f"{key}: {c1[1] - c0[1]} CG iterations, rel err {_rel(replay, ref):.2e}", flush=True)
    assert c1[1] - c0[1] > 4
    assert np.array_equal(replay, eager)
    assert _rel(replay, ref) < 1e-9


def test_pdep_db_from_davidson(gpu, gs_3h2, tmp_path):
    """Davidson eigenpairs saved in the PDEP database format and read back:
    the vectors read from disk, put through apply_operator again, give the
    saved eigenvalues as Gamma Rayleigh quotients <v|A v>/<v|v>."""
    from west_amd import pdep_db
    gs = gs_3h2
    _register(gpu, gs)
    b = gs.basis
    ev, conv, dvg, _ = gpu.davidson_diago(4, tr2_dfpt=1e-20, trev_pdep=1e-6, trev_pdep_rel=1e-4)
    save = str(tmp_path / "wstat.save")
    pdep_db.pdep_db_write(save, ev, dvg, b.npw, b.at, b.bg)
    ev2, v = pdep_db.pdep_db_read(save, b.npw, gs.npwx)
    np.testing.assert_array_equal(ev2, ev)
    np.testing.assert_array_equal(v[:, :b.npw], dvg[:, :b.npw])
    av = gpu.apply_operator(v.shape[0], v, tr2=1e-20)

    def dot(x, y):   # <x|y> at Gamma: 2 Re sum x* y - x(0) y(0)
        return 2 * np.real(np.sum(np.conj(x) * y, axis=1)) - np.real(x[:, 0] * y[:, 0])

    rq = dot(v, av) / dot(v, v)
    print("ev", ev, "rayleigh", rq, flush=True)
    assert np.max(np.abs(rq - ev) / np.abs(ev)) < 1e-6


def test_davidson_restart(gpu, gs_3h2, tmp_path):
    """Davidson restart files (Wstat/davidson_restart.f90): a run stopped
    after 3 steps leaves summary.json, hr_vr.dat (two Fortran records) and the
    basis V/N files of the step-2 state; the files agree with each other (the
    V rows orthonormal, hr(i,j) = <V_i|N_j> for i <= j < nbase), and a run
    restarted from them continues the stopped one: the same dims trajectory
    and eigenvalues as an uninterrupted run, with the files removed once it
    converges."""
    import json
    import os
    import warnings
    from west_amd import pdep_db
    gs = gs_3h2
    _register(gpu, gs)
    b = gs.basis
    nvec, ntimes = 4, 4
    nvecx = nvec * ntimes
    kw = dict(tr2_dfpt=1e-20, trev_pdep=1e-6, trev_pdep_rel=1e-4, n_pdep_times=ntimes, want_vectors=False)
    ev_full, conv_full, _, log_full = gpu.davidson_diago(nvec, **kw)
    assert all(conv_full) and len(log_full) > 5
    rdir = str(tmp_path / "wstat.restart")
    try:
        gpu.set_davidson_restart(rdir, n_steps_write_restart=1)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            _, _, _, log_a = gpu.davidson_diago(nvec, n_pdep_maxiter=3, **kw)
        doc = json.load(open(os.path.join(rdir, "summary.json")))
        nbase = doc["nbase"]
        print("summary", {k: doc[k] for k in ("dav_iter", "notcnv", "nbase")}, flush=True)
        assert doc["dav_iter"] == 2 and nbase == 3 * nvec
        assert len(doc["conv"]) == nvec and len(doc["ev"]) == nvec and len(doc["ew"]) == nvecx
        raw = open(os.path.join(rdir, "hr_vr.dat"), "rb").read()
        rec = nvecx * nvecx * 8
        assert len(raw) == 2 * (rec + 8)
        marks = np.frombuffer(raw[:4] + raw[4 + rec:8 + rec] + raw[8 + rec:12 + rec] + raw[-4:], dtype=np.int32)
        assert np.all(marks == rec)
        hr = np.frombuffer(raw[4:4 + rec], dtype=np.float64).reshape(nvecx, nvecx).T   # hr(i, j), column major
        V = np.array([pdep_db.read_vector(os.path.join(rdir, f"V{j:06d}.dat"), b.npw) for j in range(1, nbase + 1)])
        N = np.array([pdep_db.read_vector(os.path.join(rdir, f"N{j:06d}.dat"), b.npw) for j in range(1, nbase + 1)])
        assert not os.path.exists(os.path.join(rdir, f"V{nbase + 1:06d}.dat"))

        def gram(x, y):   # <x_i|y_j> at Gamma
            return 2 * np.real(np.conj(x) @ y.T) - np.real(np.outer(np.conj(x[:, 0]), y[:, 0]))

        assert np.abs(gram(V, V) - np.eye(nbase)).max() < 1e-12
        hv = gram(V, N)
        iu = np.triu_indices(nbase)
        assert np.abs(hr[:nbase, :nbase][iu] - hv[iu]).max() < 1e-12 * np.abs(hv).max()

        gpu.set_davidson_restart(rdir, n_steps_write_restart=0, restart_from=True)
        ev_b, conv_b, _, log_b = gpu.davidson_diago(nvec, **kw)
        print("ev", ev_b, "uninterrupted", ev_full, flush=True)
        assert log_b[0]["dav_iter"] == 3
        dims = [(r["dav_iter"], r["dfpt_dim"], r["diago_dim"]) for r in log_full]
        assert [(r["dav_iter"], r["dfpt_dim"], r["diago_dim"]) for r in log_a] == dims[:4]
        assert [(r["dav_iter"], r["dfpt_dim"], r["diago_dim"]) for r in log_b] == dims[3:]
        assert all(conv_b) and np.max(np.abs(ev_b - ev_full) / np.abs(ev_full)) < 1e-9
        assert not os.path.exists(rdir)
        # the restart applied to that run only: the next one starts from scratch
        ev_c, _, _, log_c = gpu.davidson_diago(nvec, **kw)
        assert log_c[0]["dav_iter"] == 0 and np.max(np.abs(ev_c - ev_full) / np.abs(ev_full)) < 1e-9
    finally:
        gpu.set_davidson_restart(None, 0, False)


def test_qe_run_system(gpu):
    """The SiH4 run of the reference's test-suite (tests/golden/qe/qe_test001.xml:
    its geometry, cell, cutoffs and 64^3 grid) through the path: the ground
    state computed here on that basis, then the operator and the lite dfpt
    against the oracle."""
    from pathlib import Path

    from oracle.dfpt import dfpt
    from oracle.hamiltonian import apply_sternheimerop
    from west_amd.groundstate import make_ground_state
    from west_amd.qe_xml import read_pw_xml, system_from_pw_xml
    s = system_from_pw_xml(read_pw_xml(str(Path(__file__).resolve().parent / "golden" / "qe" / "qe_test001.xml")))
    gs = make_ground_state(s, gpu.gpu_h_factory, nbnd=s.nbnd_occ + 2, scf_steps=0, tol=1e-6)
    _register(gpu, gs)
    print(f"case {s.name} nr={s.basis.nr} npw={s.basis.npw} et={gs.et}", flush=True)
    psi = random_wfc(gs, 3, seed=5)
    e = np.array([-0.8, -0.3, 0.1])
    ref = apply_sternheimerop(gs, psi, e, gs.alpha_pv)
    out = gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, psi, e, gs.alpha_pv)
    assert _rel(out, ref) < 1e-12
    dvg = random_wfc(gs, 2, seed=6)
    assert _rel(gpu.dfpt(2, dvg, -1.0), dfpt(gs, dvg, -1.0)) < 1e-12


def test_edge_cases_and_errors(gpu, gs_si8):
    """Empty batches, a batch smaller than its arrays, and the errors the
    reference raises through errore (non-Gamma iq, inconsistent band counts,
    a Davidson basis too small for nvec: davidson_diago.f90:108) surface as
    exceptions carrying the library's message, not as crashes."""
    from oracle.hamiltonian import apply_sternheimerop
    from west_amd.lib import WestHipError, check, load
    gs = gs_si8
    _register(gpu, gs)
    empty = np.zeros((0, gs.npwx), dtype=np.complex128)
    assert gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, empty, np.zeros(0), 0.5).shape == (0, gs.npwx)
    assert gpu.dfpt(0, empty, 1e-12).shape == (0, gs.npwx)
    assert gpu.apply_operator(0, empty, 1e-12).shape == (0, gs.npwx)
    assert gpu.linsolve_sternheimer_m_wfcts(gs.nbndval, 0, empty.copy(), empty.copy(), np.zeros(0), np.zeros(0),
                                            1e-12) == 0
    # m below the rows of the arrays: only the first m columns are formed
    psi = random_wfc(gs, 5, seed=21)
    e = np.linspace(-0.5, 0.1, 5)
    out = gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, psi, e, gs.alpha_pv, m=3)
    assert _rel(out[:3], apply_sternheimerop(gs, psi[:3], e[:3], gs.alpha_pv)) < 1e-12
    assert not out[3:].any()
    with pytest.raises(WestHipError, match="iq must be 1"):
        gpu.dfpt(1, psi[:1], 1e-12, iq=2)
    with pytest.raises(WestHipError, match="nbndval"):
        gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval + 1, psi, e, gs.alpha_pv)
    with pytest.raises(WestHipError, match="nvecx"):
        gpu.davidson_diago(4, n_pdep_times=1)
    with pytest.raises(WestHipError, match="directory"):
        gpu.set_davidson_restart(None, 1)
    with pytest.raises(WestHipError, match="precision mode"):
        check(load().wh_set_precision(3, 0.0))
    # the library still works after the errors
    out = gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, psi, e, gs.alpha_pv)
    assert _rel(out, apply_sternheimerop(gs, psi, e, gs.alpha_pv)) < 1e-12
