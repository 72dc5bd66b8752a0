"""Pins of the q-point oracle (``oracle/kpoint.py``) on the CPU.

The strongest check is supercell folding: a cell repeated twice along a1 has,
at Gamma, exactly the states of the primitive cell at k = 0 and k = b1/2, and
its chi0 is block diagonal in q (q = 0 and q = b1/2).  The Gamma path of the
oracle is pinned independently (sum over states, dense solves, brute-force
DFT), so agreement of the supercell's Gamma chi0 with the primitive cell's
chi0(q) pins the whole q machinery: the k - q folding with its phase
exp(-i g0.r) (k = 0, q = b1/2 folds with g0 = -b1), the full-sphere complex
products, the k-point projectors, the energies at k - q, the weights 2 wk /
Omega and the G = 0 rule; on a 3x1x1 grid the same holds at q = b1/3 and
2 b1/3, where the Bloch states are genuinely complex.  The lite path (tr2 < 0) is not compared: its
preconditioner uses eprec of the states at k (dfpt_module.f90:211), which at
Gamma is a half-sphere sum (Tools/set_eprec.f90), so only converged responses
agree.
"""
import numpy as np
import pytest

from oracle import kpoint as K
from oracle.davidson import davidson_diago_gamma
from oracle.dfpt import apply_operator, dfpt, sqvc_init
from oracle.hamiltonian import h_psi
from west_amd.coulomb import sqvc_init as sqvc_q
from west_amd.groundstate import make_ground_state
from west_amd.kpoint import dense_hamiltonian_k, find_point, make_h2_crystal, make_k_ground_state


@pytest.fixture(scope="module")
def pair():
    """Primitive cell on the 2x1x1 k grid and its doubled cell at Gamma."""
    prim = make_h2_crystal(10.0)
    sup = make_h2_crystal(10.0, nrep=2)
    kgs = make_k_ground_state(prim, (2, 1, 1), nbnd=4)
    gs = make_ground_state(sup, lambda g: (lambda psi: h_psi(g, psi)), nbnd=6, tol=1e-12)
    return kgs, gs


def _sector_maps(kgs, gs, iq):
    """Supercell half-sphere index of each q+G (or of -(q+G), conjugated)."""
    ms = gs.basis.mill
    lut = {tuple(x): i for i, x in enumerate(ms)}
    qb = kgs.qpts[iq].basis
    own, mirror = np.full(qb.npw, -1), np.full(qb.npw, -1)
    for ig, G in enumerate(qb.mill):
        Gs = (2 * G[0] + iq, G[1], G[2])
        if Gs in lut:
            own[ig] = lut[Gs]
        else:
            mirror[ig] = lut[tuple(-np.array(Gs))]
    return own, mirror


def test_states_fold(pair):
    kgs, gs = pair
    et = np.sort(np.concatenate([k.et for k in kgs.kpts]))[:len(gs.et)]
    assert np.max(np.abs(et - gs.et)) < 1e-9
    assert kgs.meta["gap_ry"] > 0.3
    # the same potential: the supercell's grid samples the primitive one twice
    n1 = kgs.nr[0]
    vs = gs.vrs.reshape(kgs.nr[2], kgs.nr[1], 2 * n1)
    vp = kgs.vrs.reshape(kgs.nr[2], kgs.nr[1], n1)
    assert np.abs(vs[:, :, :n1] - vp).max() < 1e-10 and np.abs(vs[:, :, n1:] - vp).max() < 1e-10


def test_k_minus_q_map(pair):
    kgs, _ = pair
    for qp in kgs.qpts:
        for ik, kc in enumerate(kgs.kcryst):
            d = kc - qp.cryst - kgs.kcryst[qp.ikq[ik]] - qp.g0[ik]
            assert np.allclose(d, 0.0)
    assert find_point(kgs.kcryst, np.array([-0.5, 0.0, 0.0]))[1].tolist() == [-1, 0, 0]


def test_h_psi_k_dense_and_hermitian(pair):
    kgs, _ = pair
    rng = np.random.default_rng(3)
    for ik in range(kgs.nks):
        v = K.KView(kgs, ik)
        n = v.basis.npw
        psi = np.zeros((3, v.npwx), complex)
        psi[:, :n] = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
        H = dense_hamiltonian_k(v.basis, kgs.vrs, v.vkb, kgs.dvan)
        hp = K.h_psi_k(v, psi)
        assert np.abs(hp[:, :n] - psi[:, :n] @ H.T).max() < 1e-11 * np.abs(hp).max()
        a = np.vdot(psi[0, :n], hp[1, :n])
        b = np.conj(np.vdot(psi[1, :n], hp[0, :n]))
        assert abs(a - b) < 1e-10 * abs(a)
        # the states are eigenvectors with their energies
        ev = K.h_psi_k(v, v.evc[:kgs.nbndval])
        assert np.abs(ev - v.et[:kgs.nbndval, None] * v.evc[:kgs.nbndval]).max() < 1e-10


@pytest.mark.parametrize("iq", [0, 1])
def test_chi0_q_equals_supercell_gamma(pair, iq):
    kgs, gs = pair
    rng = np.random.default_rng(11 + iq)
    own, mirror = _sector_maps(kgs, gs, iq)
    ms = gs.basis.mill
    sel = (ms[:, 0] % 2) == iq
    idx = np.flatnonzero(sel)
    m = 2
    dvs = np.zeros((m, gs.npwx), complex)
    dvs[:, idx] = (rng.normal(size=(m, idx.size)) + 1j * rng.normal(size=(m, idx.size))) / (1 + gs.basis.g2kin[idx])
    if iq == 0:
        dvs[:, 0] = dvs[:, 0].real
    qb = kgs.qpts[iq].basis
    fq = np.zeros((m, kgs.npwqx), complex)
    fq[:, :qb.npw] = np.where(own >= 0, dvs[:, np.maximum(own, 0)], np.conj(dvs[:, np.maximum(mirror, 0)]))
    dn_s = dfpt(gs, dvs, 1e-22)
    dn_q = K.dfpt_q(kgs, iq, fq, 1e-22)
    ref = np.zeros_like(dn_s)
    ref[:, own[own >= 0]] = dn_q[:, :qb.npw][:, own >= 0]
    scale = np.abs(dn_s).max()
    assert np.abs(dn_s - ref).max() < 1e-9 * scale
    assert np.abs(dn_s[:, ~sel]).max() < 1e-10 * scale      # block diagonal in q
    # with the Coulomb factor: sqvc(q+G) = sqvc of the supercell's G_s = q+G
    sq_s = sqvc_init(gs.basis, gs.npwx)
    sq_q = sqvc_q(qb, kgs.npwqx)
    assert np.allclose(sq_q[:qb.npw][own >= 0], sq_s[own[own >= 0]])
    da_s = apply_operator(gs, sq_s, dvs, 1e-22)
    da_q = K.apply_operator_q(kgs, iq, sq_q, fq, 1e-22)
    assert np.abs(da_s[:, own[own >= 0]] - da_q[:, :qb.npw][:, own >= 0]).max() < 1e-9 * np.abs(da_s).max()


def test_chi0_q_hermitian_negative(pair):
    """chi0~(q) on a few vectors: <a|chi b> = conj <b|chi a>, <a|chi a> < 0."""
    kgs, _ = pair
    iq = 1
    qb = kgs.qpts[iq].basis
    sq = sqvc_q(qb, kgs.npwqx)
    rng = np.random.default_rng(5)
    v = np.zeros((2, kgs.npwqx), complex)
    v[:, :qb.npw] = (rng.normal(size=(2, qb.npw)) + 1j * rng.normal(size=(2, qb.npw))) / (1 + qb.g2kin)
    cv = K.apply_operator_q(kgs, iq, sq, v, 1e-22)
    a = np.vdot(v[0], cv[1])
    b = np.conj(np.vdot(v[1], cv[0]))
    assert abs(a - b) < 1e-8 * abs(a)
    assert np.vdot(v[0], cv[0]).real < 0


def test_davidson_k_union_equals_supercell(pair):
    kgs, gs = pair
    kw = dict(n_pdep_times=4, n_pdep_maxiter=60, trev_pdep=1e-9, trev_pdep_rel=1e-8, tr2_dfpt=1e-20)
    ev_s, conv_s, _ = davidson_diago_gamma(gs, sqvc_init(gs.basis, gs.npwx), 8, **kw)
    evq = []
    for iq in range(2):
        log = []
        ev, conv, _ = K.davidson_diago_k(kgs, iq, sqvc_q(kgs.qpts[iq].basis, kgs.npwqx), 6, log=log, **kw)
        assert conv.all()
        assert log[0]["dfpt_dim"] == 6 and log[-1]["conv"] == [True] * 6
        evq += list(ev)
    assert conv_s.all()
    evq = np.sort(evq)[:8]
    assert np.max(np.abs(evq - ev_s) / np.abs(ev_s)) < 1e-6


def test_randomize_q():
    kgs = make_k_ground_state(make_h2_crystal(10.0), (2, 1, 1), nbnd=2)
    for qp in kgs.qpts:
        a = np.zeros((3, kgs.npwqx), complex)
        K.do_randomize_q(qp, a, 1, 3)
        n = qp.basis.npw
        if qp.is_gamma:
            assert np.all(a[:, 0] == 0)
        # amplitude over |q+G|^2 + 1 (Ry): bounded by 1 / (|q+G|^2 + 1)
        assert np.all(np.abs(a[:, :n]) <= 1.0 / (qp.basis.g2kin + 1.0) + 1e-15)
        assert np.all(a[:, n:] == 0)
        # distinct vectors
        assert abs(np.vdot(a[0], a[1])) < 0.9 * np.linalg.norm(a[0]) * np.linalg.norm(a[1])


@pytest.mark.parametrize("iq", [1, 2])
def test_chi0_q_complex_points_equal_supercell_gamma(iq):
    """k and q points that are not their own inverses (the 3x1x1 grid,
    q = b1/3 and 2 b1/3): Bloch states are genuinely complex and q != -q, so
    a misplaced conjugate or a k + q in place of k - q would show.  A complex
    perturbation f on the q sphere is the sector q of the real supercell
    perturbation f e^{i(q+G)r} + c.c.; its chi0(q) f is the sector q of the
    supercell response."""
    prim = make_h2_crystal(8.0)
    sup = make_h2_crystal(8.0, nrep=3)
    kgs = make_k_ground_state(prim, (3, 1, 1), nbnd=4)
    gs = make_ground_state(sup, lambda g: (lambda psi: h_psi(g, psi)), nbnd=8, tol=1e-12)
    qb = kgs.qpts[iq].basis
    lut = {tuple(x): i for i, x in enumerate(gs.basis.mill)}
    rng = np.random.default_rng(3 + iq)
    m = 2
    fq = np.zeros((m, kgs.npwqx), complex)
    fq[:, :qb.npw] = (rng.normal(size=(m, qb.npw)) + 1j * rng.normal(size=(m, qb.npw))) / (1 + qb.g2kin)
    dvs = np.zeros((m, gs.npwx), complex)
    own = {}
    for ig, G in enumerate(qb.mill):
        Gs = (3 * G[0] + iq, G[1], G[2])
        if Gs in lut:
            dvs[:, lut[Gs]] += fq[:, ig]
            own[ig] = lut[Gs]
        else:
            dvs[:, lut[tuple(-np.array(Gs))]] += np.conj(fq[:, ig])
    assert len(own) > 10
    dn_s = dfpt(gs, dvs, 1e-22)
    dn_q = K.dfpt_q(kgs, iq, fq, 1e-22)
    ig = np.array(list(own.keys()))
    js = np.array(list(own.values()))
    assert np.abs(dn_s[:, js] - dn_q[:, ig]).max() < 1e-9 * np.abs(dn_s).max()
