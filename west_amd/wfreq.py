"""Wfreq's chi(i omega) by Lanczos chains at Gamma, through the C ABI.

Names and argument meaning follow the reference:
``solve_deflated_lanczos_w_full_ortho`` (``Wfreq/solve_deflated_lanczos_w_full_ortho.f90:14``)
and ``get_brak_hyper_parallel`` (``Wfreq/get_brak_hyper_parallel.f90:14``, one
image) run on the device; ``diago_lanczos`` (``Wfreq/diago_lanczos.f90:14``, a
tridiagonal eigensolve of a few tens of steps) and the dmati update of
``Wfreq/solve_wfreq.f90:624-645`` are small host operations.
Arrays are C-ordered views of the reference's Fortran arrays:
b (NRHS, npwx), alpha_s (NRHS, NLSTEPS), beta_s (NRHS, NLSTEPS-1),
q_s (NLSTEPS, NRHS, npwx), brak (NRHS, NLSTEPS, nglob).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
from scipy.linalg import eigh_tridiagonal

from .lib import check, load, ptr


def solve_deflated_lanczos_w_full_ortho(nbnd_to_deflate: int, b, nlsteps: int, want_basis: bool = True):
    """NLSTEPS Lanczos steps of P_c H P_c from the NRHS rows of b (the first
    nbnd_to_deflate registered states deflated).  Returns (alpha_s, beta_s,
    q_s or None, bnorm)."""
    lib = load()
    b = np.ascontiguousarray(b, dtype=np.complex128)
    nrhs, npwx = b.shape
    alpha_s = np.zeros((nrhs, nlsteps))
    beta_s = np.zeros((nrhs, max(nlsteps - 1, 1)))
    q_s = np.zeros((nlsteps, nrhs, npwx), dtype=np.complex128) if want_basis else None
    bnorm = np.zeros(nrhs)
    check(lib.wh_solve_deflated_lanczos_w_full_ortho(int(nbnd_to_deflate), int(nrhs), int(nlsteps), ptr(b),
                                                      ptr(alpha_s), ptr(beta_s), ptr(q_s) if want_basis else None,
                                                      ptr(bnorm)))
    return alpha_s, beta_s[:, :nlsteps - 1], q_s, bnorm


def glbrak_gamma(a, b):
    """c(i, j) = 2 Re <a_i|b_j> - Re a_i(G=0) Re b_j(G=0) on the device; a: (na, npwx), b: (nb, npwx)."""
    lib = load()
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    c = np.zeros((b.shape[0], a.shape[0]))          # column major (na, nb)
    check(lib.wh_glbrak_gamma(int(a.shape[0]), ptr(a), int(b.shape[0]), ptr(b), ptr(c)))
    return c.T.copy()


def get_brak(dvpsi, q_s):
    """brak(ip, il, jglob) = <dvpsi_jglob | q_s(il, ip)>."""
    nls, nrhs, npwx = q_s.shape
    c = glbrak_gamma(dvpsi, q_s.reshape(nls * nrhs, npwx))          # (nglob, nls*nrhs)
    return c.reshape(dvpsi.shape[0], nls, nrhs).transpose(2, 1, 0).copy()


def diago_lanczos(bnorm: float, diago, subdiago, braket):
    """Eigenpairs of the Lanczos tridiagonal (DSTEQR 'I') and the braket
    rotated onto them, scaled by bnorm and the first eigenvector components.
    braket: (NLSTEPS, nglob).  Returns (eigenvalues, braket)."""
    w, z = eigh_tridiagonal(np.asarray(diago, float), np.asarray(subdiago, float))
    return w, (bnorm * (z.T @ np.asarray(braket))) * z[0, :, None]


def dmati_update(dmati, braket, diago, et_v, occ_v, imfreq, weight, omega, frac_occ=False):
    """dmati(ip, jglob, ifreq) += sum_il braket(ip, il, jglob) mwo 2 e / (e^2 + w^2),
    e = diago(ip, il) - et_v, mwo = -weight / omega."""
    mwo = -weight / omega
    imfreq = np.asarray(imfreq, float)
    for ip in range(braket.shape[0]):
        e = np.asarray(diago[ip])[:, None] - et_v
        f = mwo * 2.0 * e / (e ** 2 + imfreq[None, :] ** 2)
        if frac_occ:
            f = f * occ_v
        dmati[ip] += braket[ip].T @ f
    return dmati
