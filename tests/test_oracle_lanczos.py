"""The oracle's Wfreq Lanczos (oracle/lanczos.py) pinned by what the
mathematics fixes, on tiny cells with exact (densely diagonalised) states:
the basis is orthonormal, Q^T (P_c H P_c) Q is the tridiagonal (alpha, beta),
the lowest Ritz value converges to the lowest eigenvalue outside the
deflated bands, the rotated braket reproduces the first two moments
<a|b> and <a|H|b> exactly, and the dmati update converges to the
sum over the conduction states."""
import numpy as np
import pytest

from conftest import dense_states, random_wfc
from oracle.hamiltonian import glbrak_gamma, h_psi
from oracle.lanczos import diago_lanczos, dmati_update, get_brak, solve_deflated_lanczos_w_full_ortho


def _exact(gs):
    """gs with its occupied states replaced by exact eigenstates; all eigenpairs."""
    w, U = dense_states(gs)
    g = type("GS", (), {})()
    for k, v in vars(gs).items():
        setattr(g, k, v)
    g.evc, g.et = U[:gs.nbndval], w[:gs.nbndval]
    return g, w, U


def _pc(g, f):
    ev = g.evc[:g.nbndval]
    ps = glbrak_gamma(ev, f, g.basis.npw, g.basis.gstart)
    return f - (ps.T @ ev.real + 1j * (ps.T @ ev.imag))


@pytest.mark.parametrize("key", ["h2", "si8"])
def test_lanczos_basis_and_tridiagonal(key, request):
    gs0 = request.getfixturevalue(f"gs_{key}")
    g, w, U = _exact(gs0)
    n, gst, nd = g.basis.npw, g.basis.gstart, g.nbndval
    b = _pc(g, random_wfc(g, 3, seed=2))
    nls = 12
    a, bt, q, bn = solve_deflated_lanczos_w_full_ortho(g, nd, b, nls)
    np.testing.assert_allclose(bn, np.sqrt(np.diag(glbrak_gamma(b, b, n, gst))), rtol=1e-12)
    for ip in range(b.shape[0]):
        Q = q[:, ip]
        assert np.abs(glbrak_gamma(Q, Q, n, gst) - np.eye(nls)).max() < 1e-10
        HQ = _pc(g, h_psi(g, _pc(g, Q)))
        T = glbrak_gamma(Q, HQ, n, gst)
        Tref = np.diag(a[ip]) + np.diag(bt[ip], 1) + np.diag(bt[ip], -1)
        assert np.abs(T - Tref).max() < 1e-9 * np.abs(Tref).max()


def test_lanczos_ritz_value_and_moments(gs_h2):
    g, w, U = _exact(gs_h2)
    n, gst, nd = g.basis.npw, g.basis.gstart, g.nbndval
    b = _pc(g, random_wfc(g, 2, seed=7))
    nls = 40
    a, bt, q, bn = solve_deflated_lanczos_w_full_ortho(g, nd, b, nls)
    # the lowest eigenvalue of H outside the deflated (occupied) states
    theta = np.linalg.eigvalsh(np.diag(a[0]) + np.diag(bt[0], 1) + np.diag(bt[0], -1))
    assert abs(theta[0] - w[nd]) < 1e-8 * abs(w[nd])
    brak = get_brak(g, b, q)                              # (ip, il, jglob)
    for ip in range(b.shape[0]):
        th, bk = diago_lanczos(bn[ip], a[ip], bt[ip], brak[ip])
        m0 = glbrak_gamma(b, b[ip:ip + 1], n, gst)[:, 0]
        m1 = glbrak_gamma(b, h_psi(g, b[ip:ip + 1]), n, gst)[:, 0]
        np.testing.assert_allclose(bk.sum(axis=0), m0, rtol=1e-10, atol=1e-12 * np.abs(m0).max())
        np.testing.assert_allclose(th @ bk, m1, rtol=1e-9, atol=1e-11 * np.abs(m1).max())


def test_dmati_update_converges_to_sum_over_states(gs_h2):
    g, w, U = _exact(gs_h2)
    n, gst, nd = g.basis.npw, g.basis.gstart, g.nbndval
    b = _pc(g, random_wfc(g, 2, seed=9))
    nls = 60
    a, bt, q, bn = solve_deflated_lanczos_w_full_ortho(g, nd, b, nls)
    brak = get_brak(g, b, q)
    diag = np.zeros((b.shape[0], nls))
    bks = np.zeros_like(brak)
    for ip in range(b.shape[0]):
        diag[ip], bks[ip] = diago_lanczos(bn[ip], a[ip], bt[ip], brak[ip])
    freqs = np.array([0.0, 0.3, 2.0])
    ev = g.et[nd - 1]
    dm = dmati_update(np.zeros((2, 2, 3)), bks, diag, ev, 1.0, freqs, g.weight, g.basis.omega)
    # sum over all conduction states of the dense spectrum
    ov = glbrak_gamma(U[nd:], b, n, gst)                    # (nc, nrhs)
    e = w[nd:, None] - ev
    f = (-g.weight / g.basis.omega) * 2.0 * e / (e ** 2 + freqs[None, :] ** 2)   # (nc, nfreq)
    ref = np.einsum("ci,cj,cf->ijf", ov, ov, f)
    assert np.abs(dm - ref).max() < 1e-6 * np.abs(ref).max()
