"""Wfreq's chi(i omega) Lanczos at Gamma (test infrastructure: the checker,
never the product).  Restatements of
``Wfreq/solve_deflated_lanczos_w_full_ortho.f90:14-357`` (gamma_only
branches), ``Wfreq/get_brak_hyper_parallel.f90:14-143`` on one image,
``Wfreq/diago_lanczos.f90:14-57`` and the Lanczos update of dmati in
``Wfreq/solve_wfreq.f90:624-645``."""
from __future__ import annotations

import numpy as np
from scipy.linalg import eigh_tridiagonal

from .hamiltonian import glbrak_gamma, h_psi


def _gdot(a, b, n, gstart):
    """Per-column Gamma inner product 2 Re sum a* b - Re a(1) Re b(1)."""
    r = 2.0 * np.sum(a[:, :n].real * b[:, :n].real + a[:, :n].imag * b[:, :n].imag, axis=1)
    if gstart == 2:
        r -= a[:, 0].real * b[:, 0].real
    return r


def solve_deflated_lanczos_w_full_ortho(gs, nbnd_to_deflate, b, nlsteps):
    """b: (NRHS, npwx).  Returns alpha_s (NRHS, NLSTEPS), beta_s (NRHS,
    NLSTEPS-1), q_s (NLSTEPS, NRHS, npwx), bnorm (NRHS)."""
    n = gs.basis.npw
    gstart = gs.basis.gstart
    nrhs = b.shape[0]
    r = b.copy()                                                     # (65-67)
    beta = np.sqrt(_gdot(r, r, n, gstart))                           # (76-112)
    bnorm = beta.copy()
    alpha_s = np.zeros((nrhs, nlsteps))
    beta_s = np.zeros((nrhs, max(nlsteps - 1, 0)))
    q_s = np.zeros((nlsteps, nrhs, b.shape[1]), dtype=np.complex128)
    evc = gs.evc[:nbnd_to_deflate]
    for il in range(nlsteps):
        q_s[il] = r / beta[:, None]                                  # (118-124)
        r = h_psi(gs, q_s[il])                                       # h_psi_ (143)
        if nbnd_to_deflate:                                          # apply_alpha_pc, alpha = 1 (145)
            ps = glbrak_gamma(evc, r, n, gstart)
            r = r - (ps.T @ evc.real + 1j * (ps.T @ evc.imag))
        if il > 0:                                                   # (149-160)
            r[:, :n] -= q_s[il - 1][:, :n] * beta[:, None]
        alpha = _gdot(q_s[il], r, n, gstart)                         # (164-200)
        alpha_s[:, il] = alpha
        r[:, :n] -= q_s[il][:, :n] * alpha[:, None]                  # (211-221)
        for ia in range(il - 1):                                     # full ortho, ia = 1..il-2 (225-287)
            t = _gdot(q_s[ia], r, n, gstart)
            r[:, :n] -= q_s[ia][:, :n] * t[:, None]
        beta = np.sqrt(_gdot(r, r, n, gstart))                       # (290-339)
        if il < nlsteps - 1:
            beta_s[:, il] = beta
    return alpha_s, beta_s, q_s, bnorm


def get_brak(gs, dvpsi, q_s):
    """brak(ip, il, jglob) = <dvpsi_jglob | q_s(il, ip)> (Gamma), i.e. the
    reference's brak(nglob, NLSTEPS, NRHS) in C order."""
    nls, nrhs, _ = q_s.shape
    c = glbrak_gamma(dvpsi, q_s.reshape(nls * nrhs, -1), gs.basis.npw, gs.basis.gstart)   # (nglob, nls*nrhs)
    return c.reshape(dvpsi.shape[0], nls, nrhs).transpose(2, 1, 0).copy()


def diago_lanczos(bnorm, diago, subdiago, braket):
    """DSTEQR('I') of the tridiagonal (diago, subdiago); braket(il, jp) ->
    bnorm * (sum_k braket(k, jp) z(k, il)) * z(1, il).  braket: (NLSTEPS, nglob).
    Returns (eigenvalues, braket)."""
    w, z = eigh_tridiagonal(np.asarray(diago, float), np.asarray(subdiago, float))
    bt = bnorm * (z.T @ braket)                                      # (il, jp)
    return w, bt * z[0, :, None]


def dmati_update(dmati, braket, diago, et_v, occ_v, imfreq, weight, omega, frac_occ=False):
    """dmati(jp, ip, ifreq) += sum_il braket(jp, il, ip) mwo 2 e / (e^2 + w^2),
    e = diago(il, ip) - et_v, mwo = -weight / omega  (solve_wfreq.f90:324, 624-645).
    dmati: (NRHS, nglob, nfreq); braket: (NRHS, NLSTEPS, nglob); diago: (NRHS, NLSTEPS)."""
    mwo = -weight / omega
    for ip in range(braket.shape[0]):
        e = diago[ip][:, None] - et_v                                # (il, 1)
        f = mwo * 2.0 * e / (e ** 2 + np.asarray(imfreq)[None, :] ** 2)
        if frac_occ:
            f = f * occ_v
        dmati[ip] += braket[ip].T @ f                                # (nglob, nfreq)
    return dmati
