"""The q-point path on the GPU against the oracle (``oracle/kpoint.py``),
through the C ABI: the Sternheimer operator and CG at each k point, chi0(q)
(lite and tight) and apply_operator at each q point, and the Davidson at each
q point.  System: the H2 molecular crystal with s projectors on a 2x1x1, a
2x2x1 and a 3x1x1 k grid (q = b1/2 folds k = 0 onto k - q = -b1/2 with
g0 = -b1; on the 3x1x1 grid k and q are not their own inverses, so the
states are genuinely complex).  The
runs use a system of their own, destroyed at the end.  Each case prints its
name before it runs."""
import os

import numpy as np
import pytest

from west_amd.kpoint import make_h2_crystal, make_k_ground_state


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module", params=[(2, 1, 1), (2, 2, 1), (3, 1, 1)], ids=["k211", "k221", "k311"])
def ksys(gpu, request):
    nk = request.param
    print(f"case kpoint {nk}: ground state", flush=True)
    kgs = make_k_ground_state(make_h2_crystal(10.0), nk, nbnd=4)
    h = gpu.system_create()
    gpu.system_select(h)
    gpu.set_k_ground_state(kgs)
    yield kgs
    gpu.system_select(0)
    gpu.system_destroy(h)


def _smooth(n, npwx, m, g2, seed):
    rng = np.random.default_rng(seed)
    a = np.zeros((m, npwx), complex)
    a[:, :n] = (rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))) / (1.0 + g2[None, :n])
    return a


@pytest.mark.gpu
def test_sternheimer_operator_at_k(gpu, ksys):
    from oracle import kpoint as K
    kgs = ksys
    for ik in range(kgs.nks):
        print(f"case kpoint operator ik={ik + 1}", flush=True)
        v = K.KView(kgs, ik)
        psi = _smooth(v.basis.npw, kgs.npwx, 5, v.basis.g2kin, 10 + ik)
        e = np.linspace(-0.6, 0.1, 5)
        ref = K.apply_sternheimerop_k(v, psi, e, kgs.alpha_pv)
        gpu.set_current_k(ik + 1)
        out = gpu.apply_sternheimerop_to_m_wfcs(kgs.nbndval, psi, e, kgs.alpha_pv)
        assert _rel(out, ref) < 1e-12


@pytest.mark.gpu
def test_linsolve_at_k(gpu, ksys):
    from oracle import kpoint as K
    from oracle.linsolve import linsolve_sternheimer_m_wfcts
    kgs = ksys
    ik = kgs.nks - 1
    print(f"case kpoint linsolve ik={ik + 1}", flush=True)
    v = K.KView(kgs, ik)
    nb = kgs.nbndval
    b = K.apply_alpha_pc_k(v, _smooth(v.basis.npw, kgs.npwx, nb, v.basis.g2kin, 3), -1.0)
    eprec = K.set_eprec_k(v, v.evc[:nb])
    e = np.asarray(v.et[:nb])
    x_ref = K.precondition_k(v, b, eprec)
    x_gpu = x_ref.copy()
    ierr_ref = linsolve_sternheimer_m_wfcts(v, nb, b.copy(), x_ref, e, eprec, 1e-20)
    gpu.set_current_k(ik + 1)
    ierr = gpu.linsolve_sternheimer_m_wfcts(nb, nb, b.copy(), x_gpu, e, eprec, 1e-20)
    assert ierr == ierr_ref == 0
    assert _rel(x_gpu, x_ref) < 1e-10
    # the iterations run out (3 CG steps): every band is counted unconverged,
    # and the iterate is the oracle's after the same steps
    from west_amd.lib import check, load
    x3_ref = K.precondition_k(v, b, eprec)
    x3_gpu = x3_ref.copy()
    ierr_ref = linsolve_sternheimer_m_wfcts(v, nb, b.copy(), x3_ref, e, eprec, 1e-20, n_dfpt_maxiter=3)
    check(load().wh_set_params(3, 0))
    try:
        ierr = gpu.linsolve_sternheimer_m_wfcts(nb, nb, b.copy(), x3_gpu, e, eprec, 1e-20)
    finally:
        check(load().wh_set_params(250, 0))
    assert ierr == ierr_ref == nb
    assert _rel(x3_gpu, x3_ref) < 1e-10


@pytest.mark.gpu
def test_dfpt_and_apply_operator_at_q(gpu, ksys):
    from oracle import kpoint as K
    from west_amd.coulomb import sqvc_init
    kgs = ksys
    for iq, qp in enumerate(kgs.qpts):
        qb = qp.basis
        dvg = _smooth(qb.npw, kgs.npwqx, 3, qb.g2kin, 20 + iq)
        print(f"case kpoint dfpt iq={iq + 1} lite", flush=True)
        assert _rel(gpu.dfpt(3, dvg, -1.0, iq=iq + 1), K.dfpt_q(kgs, iq, dvg, -1.0)) < 1e-12
        print(f"case kpoint dfpt iq={iq + 1} tight", flush=True)
        out, ierr = gpu.dfpt(3, dvg, 1e-20, iq=iq + 1, return_ierr=True)
        assert not ierr.any()
        ref = K.dfpt_q(kgs, iq, dvg, 1e-20)
        assert _rel(out, ref) < 1e-9
        if qp.is_gamma:
            assert np.all(out[:, 0] == 0)
        # the k points one after another (full-grid transforms around a
        # solve per k point) instead of all at once on LDS grids
        print(f"case kpoint dfpt iq={iq + 1} tight, k points in turn", flush=True)
        os.environ["WH_KBATCH"] = "0"
        try:
            out1, ierr1 = gpu.dfpt(3, dvg, 1e-20, iq=iq + 1, return_ierr=True)
        finally:
            del os.environ["WH_KBATCH"]
        assert not ierr1.any()
        assert _rel(out1, ref) < 1e-9
        assert _rel(out1, out) < 1e-12
        print(f"case kpoint apply_operator iq={iq + 1}", flush=True)
        sq = sqvc_init(qb, kgs.npwqx)
        assert _rel(gpu.apply_operator(3, dvg, 1e-20, iq=iq + 1), K.apply_operator_q(kgs, iq, sq, dvg, 1e-20)) < 1e-9


@pytest.mark.gpu
def test_davidson_at_q(gpu, ksys):
    from oracle import kpoint as K
    from west_amd.coulomb import sqvc_init
    kgs = ksys
    kw = dict(n_pdep_times=4, n_pdep_maxiter=40, trev_pdep=1e-8, trev_pdep_rel=1e-7, tr2_dfpt=1e-18)
    for iq in range(len(kgs.qpts)):
        print(f"case kpoint davidson iq={iq + 1}", flush=True)
        log_ref = []
        ev_ref, conv_ref, _ = K.davidson_diago_k(kgs, iq, sqvc_init(kgs.qpts[iq].basis, kgs.npwqx), 6, log=log_ref,
                                                 **kw)
        ev, conv, dvg, log = gpu.davidson_diago_k(iq + 1, 6, **kw)
        assert np.max(np.abs(ev - ev_ref) / np.abs(ev_ref)) < 1e-8
        assert np.array_equal(conv, conv_ref)
        assert [r["dfpt_dim"] for r in log] == [r["dfpt_dim"] for r in log_ref]
        assert [r["diago_dim"] for r in log] == [r["diago_dim"] for r in log_ref]
        # 'stern' calls: one per perturbation's solve per k point per CG step
        # (a solve whose last rho lands next to tr2 may take one step more or less)
        for r, rr in zip(log, log_ref):
            assert abs(r["sternop_ncalls"] - rr["sternop_ncalls"]) <= max(1, rr["sternop_ncalls"] // 50)
        # orthonormal eigenvectors on the q sphere
        n = kgs.qpts[iq].basis.npw
        gram = dvg[:, :n].conj() @ dvg[:, :n].T
        assert np.abs(gram - np.eye(6)).max() < 1e-8


@pytest.mark.gpu
def test_kinetic_only_and_device_arrays_at_k(gpu):
    """l_kinetic_only at a k point (k_psi plus alpha P_v) and device (torch)
    arrays through the k-point entries."""
    import torch
    from oracle import kpoint as K
    kgs = make_k_ground_state(make_h2_crystal(10.0), (2, 1, 1), nbnd=4)
    h = gpu.system_create()
    gpu.system_select(h)
    try:
        print("case kpoint kinetic-only operator", flush=True)
        kgs.l_kinetic_only = True
        gpu.set_k_ground_state(kgs, l_kinetic_only=True)
        v = K.KView(kgs, 1)
        psi = _smooth(v.basis.npw, kgs.npwx, 3, v.basis.g2kin, 31)
        e = np.array([-0.3, -0.1, 0.2])
        gpu.set_current_k(2)
        ref = K.apply_sternheimerop_k(v, psi, e, kgs.alpha_pv)
        assert _rel(gpu.apply_sternheimerop_to_m_wfcs(kgs.nbndval, psi, e, kgs.alpha_pv), ref) < 1e-12
        print("case kpoint device arrays", flush=True)
        kgs.l_kinetic_only = False
        gpu.set_k_ground_state(kgs)
        qb = kgs.qpts[1].basis
        dvg = _smooth(qb.npw, kgs.npwqx, 2, qb.g2kin, 41)
        dvg_d = torch.from_numpy(dvg).to("cuda:0")
        out = gpu.dfpt(2, dvg_d, 1e-18, iq=2).cpu().numpy()
        assert _rel(out, K.dfpt_q(kgs, 1, dvg, 1e-18)) < 1e-9
        # back to Gamma on the same system: the k grid is dropped
        from oracle.dfpt import sqvc_init
        from west_amd.groundstate import make_ground_state
        from oracle.hamiltonian import h_psi
        gs = make_ground_state(make_h2_crystal(10.0), lambda g: (lambda p: h_psi(g, p)), nbnd=3, tol=1e-10)
        gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))
        from west_amd.lib import WestHipError
        with pytest.raises(WestHipError, match="iq must be 1"):
            gpu.dfpt(1, np.zeros((1, gs.npwx), complex), -1.0, iq=2)
    finally:
        gpu.system_select(0)
        gpu.system_destroy(h)
