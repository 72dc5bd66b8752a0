"""Hamiltonian pieces of the oracle against brute force and closed forms
(QE h_psi_ at Gamma; Hamiltonian_kernel/glbrak.f90, apply_alpha_pc/pv,
DFPT_kernel/precondition_m_wfcts.f90, Tools/set_eprec.f90)."""
import numpy as np

from conftest import random_wfc
from oracle.fft_at_gamma import fwfft
from oracle.hamiltonian import (apply_alpha_pc, apply_alpha_pv, apply_sternheimerop, get_alpha_pv, glbrak_gamma,
                                h_psi, k_psi, precondition_m_wfcts, set_eprec)


def _full(b, c):
    """half-sphere coefficients -> full sphere [G_0, G_1.., -G_1..]"""
    return np.concatenate([c[..., :b.npw], np.conj(c[..., 1:b.npw])], axis=-1)


def _dense_h(gs):
    b = gs.basis
    mill = np.concatenate([b.mill, -b.mill[1:]])
    g2 = np.concatenate([b.g2kin, b.g2kin[1:]])
    vg = fwfft(b.nr, gs.vrs.astype(complex))
    d = mill[:, None, :] - mill[None, :, :]
    idx = (np.mod(d[..., 0], b.nr[0]) + b.nr[0] * (np.mod(d[..., 1], b.nr[1]) + b.nr[1] * np.mod(d[..., 2], b.nr[2])))
    H = vg[idx] + np.diag(g2)
    if gs.nkb:
        beta = _full(b, gs.vkb)                 # (nkb, nfull), beta(-G) = conj(beta(G))
        H = H + (beta.T * gs.dvan[None, :]) @ beta.conj()
    return H


def test_glbrak_is_full_sphere_product(gs_si8):
    gs = gs_si8
    a = random_wfc(gs, 3, 1)
    c = random_wfc(gs, 4, 2)
    ref = np.real(_full(gs.basis, a).conj() @ _full(gs.basis, c).T)
    assert np.allclose(glbrak_gamma(a, c, gs.basis.npw, 2), ref, rtol=1e-13, atol=1e-13)


def test_h_psi_matches_dense_hamiltonian(gs_si8, gs_3h2):
    for gs in (gs_si8, gs_3h2):
        H = _dense_h(gs)
        assert np.allclose(H, H.conj().T, atol=1e-12)
        psi = random_wfc(gs, 3, 5)
        ref = (H @ _full(gs.basis, psi).T).T[:, :gs.basis.npw]
        out = h_psi(gs, psi)
        assert np.abs(out - ref).max() < 1e-11 * np.abs(ref).max()


def test_ground_state_is_eigenstate(gs_si8):
    gs = gs_si8
    hv = h_psi(gs, gs.evc)
    res = hv - gs.et[:, None] * gs.evc
    assert np.abs(res).max() < 1e-7
    S = glbrak_gamma(gs.evc, gs.evc, gs.basis.npw, 2)
    assert np.allclose(S, np.eye(S.shape[0]), atol=1e-12)


def test_projectors(gs_si8):
    gs = gs_si8
    n = gs.basis.npw
    f = random_wfc(gs, 5, 7)
    pc = apply_alpha_pc(gs, f, 1.0)                     # P_c f
    assert np.allclose(apply_alpha_pc(gs, pc, 1.0), pc, atol=1e-12)          # idempotent
    assert np.abs(glbrak_gamma(gs.evc[:gs.nbndval], pc, n, 2)).max() < 1e-12  # P_c psi_v = 0 (orthogonal)
    pv = apply_alpha_pv(gs, f, np.zeros_like(f), 1.0)   # P_v f
    assert np.allclose(pc + pv, f, atol=1e-12)
    assert np.allclose(apply_alpha_pc(gs, f, -1.0), -pc, atol=1e-12)
    # only the nbndval occupied states span P_v: empty eigenstates are in P_c
    empty = gs.evc[gs.nbndval:]
    assert empty.shape[0] > 0
    assert np.abs(apply_alpha_pv(gs, empty, np.zeros_like(empty), 1.0)).max() < 1e-12
    assert np.allclose(apply_alpha_pc(gs, empty, 1.0), empty, atol=1e-12)


def test_kinetic_only_and_operator(gs_si8):
    gs = gs_si8
    psi = random_wfc(gs, 3, 9)
    kp = k_psi(gs, psi)
    assert np.allclose(kp, gs.basis.g2kin[None, :] * psi)
    e = np.array([-0.3, 0.1, 0.2])
    out = apply_sternheimerop(gs, psi, e, 0.7)
    ref = h_psi(gs, psi) - e[:, None] * psi + 0.7 * apply_alpha_pv(gs, psi, np.zeros_like(psi), 1.0)
    assert np.allclose(out, ref, atol=1e-12)


def test_precondition_eprec_alpha(gs_si8):
    gs = gs_si8
    n = gs.basis.npw
    g2 = gs.basis.g2kin
    ep = set_eprec(gs, gs.evc)
    for v in range(3):
        s = 0.0
        for ig in range(n):            # set_eprec.f90:46-51, one term at a time
            s += (gs.evc[v, ig].conjugate() * gs.evc[v, ig]).real * g2[ig]
        assert np.isclose(ep[v], 1.35 * s, rtol=1e-13)
    f = random_wfc(gs, 2, 3)
    pf = precondition_m_wfcts(gs, f, ep[:2])
    for v in range(2):
        for ig in (0, 5, n - 1):
            assert np.isclose(pf[v, ig], f[v, ig] / max(1.0, g2[ig] / ep[v]))
    # alpha_pv = max(2 (max occupied - min all), 0.01)
    assert np.isclose(get_alpha_pv(gs.et, gs.nbndval), max(2 * (gs.et[:gs.nbndval].max() - gs.et.min()), 0.01))
    assert get_alpha_pv(np.array([1.0, 1.0]), 1) == 0.01


def test_set_eprec_is_kinetic_energy_from_real_space_gradients(gs_si8):
    """set_eprec (Tools/set_eprec.f90:46-51) sums |psi(G)|^2 g2kin over the
    half sphere without the Gamma doubling, so eprec = 1.35 T / 2 with T the
    kinetic energy <psi|-lap|psi> (Ry).  T here comes from another route: the
    gradient of psi(r), built with numpy's own FFT from the coefficients
    placed at +G and -G by Miller index, T = (1/N) sum_r |grad psi(r)|^2."""
    gs = gs_si8
    b = gs.basis
    nr, nnr = b.nr, b.nnr
    G = b.mill @ b.bg                                   # Cartesian G, bohr^-1
    pos, neg = np.mod(b.mill, nr), np.mod(-b.mill[1:], nr)
    ep = set_eprec(gs, gs.evc)
    for v in range(4):
        c = gs.evc[v, :b.npw]
        T = 0.0
        for a in range(3):
            grid = np.zeros(nr, dtype=np.complex128)
            grid[tuple(pos.T)] = 1j * G[:, a] * c
            grid[tuple(neg.T)] = np.conj(1j * G[1:, a] * c[1:])
            d = np.fft.ifftn(grid) * nnr                # unnormalised inverse, real function
            assert np.abs(d.imag).max() < 1e-12 * max(np.abs(d.real).max(), 1.0)
            T += np.sum(d.real ** 2) / nnr
        assert np.isclose(ep[v], 1.35 * T / 2, rtol=1e-12)


def test_precondition_limits(gs_si8):
    """precondition_m_wfcts.f90:41-46, pf = f / max(1, g2kin / eprec), in its
    two limits: eprec above every g2kin leaves f unchanged; a tiny eprec
    gives f * eprec / g2kin for G != 0 and leaves G = 0 (g2kin = 0); entries
    beyond npw are zero."""
    gs = gs_si8
    n = gs.basis.npw
    g2 = gs.basis.g2kin
    f = random_wfc(gs, 2, 11)
    f_pad = np.concatenate([f, np.ones((2, 3), dtype=f.dtype)], axis=1)
    pf = precondition_m_wfcts(gs, f_pad, [10.0 * g2.max(), 1e-8])
    np.testing.assert_array_equal(pf[0, :n], f[0])
    assert not pf[:, n:].any()
    assert pf[1, 0] == f[1, 0] and g2[0] == 0.0
    np.testing.assert_allclose(pf[1, 1:n], f[1, 1:] * 1e-8 / g2[1:], rtol=1e-14)


def test_alpha_pv_by_hand():
    """get_alpha_pv.f90:35-57: 2 (highest occupied - lowest of all bands),
    at least 0.01; the lowest is taken over every band, occupied or not."""
    assert np.isclose(get_alpha_pv(np.array([-0.5, -0.2, 0.1, 0.4]), 2), 0.6)
    # an empty band below the occupied ones sets the minimum
    assert np.isclose(get_alpha_pv(np.array([-0.2, -0.5, 0.1, -0.7]), 2), 1.0)
    assert get_alpha_pv(np.array([0.3, 0.3, 0.9]), 2) == 0.01
