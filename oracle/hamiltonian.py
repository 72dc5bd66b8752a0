"""Gamma-point Hamiltonian pieces used by the Sternheimer solver.

Arrays of wave functions are ``(m, npwx)`` complex128 in C order, which is
the memory layout of the reference's Fortran ``(npwx, m)`` arrays.

``gs`` is any object with the ground-state attributes ``basis`` (with
``npw, nnr, nr, nl, nlm, g2kin, gstart``), ``npwx``, ``vrs`` (nnr,),
``vkb`` (nkb, npwx), ``dvan`` (nkb,), ``evc`` (nbndval, npwx), ``et``,
``occ``, ``weight``, ``alpha_pv`` and ``l_kinetic_only``.
"""
from __future__ import annotations

import numpy as np

from .fft_at_gamma import double_fwfft_gamma, double_invfft_gamma


def glbrak_gamma(a, b, ng, gstart):
    """``Hamiltonian_kernel/glbrak.f90:14-43``:
    c = 2 Re<a|b> over ng rows (DGEMM on the real view), minus
    Re a(1) Re b(1) when gstart == 2 (DGER).  a: (na, ngx), b: (nb, ngx)
    -> c (na, nb)."""
    a = a[:, :ng]
    b = b[:, :ng]
    c = 2.0 * (a.real @ b.real.T + a.imag @ b.imag.T)
    if gstart == 2:
        c -= np.outer(a[:, 0].real, b[:, 0].real)
    return c


def k_psi(gs, psi):
    """``Hamiltonian_kernel/k_psi.f90:14-85``: g2kin * psi, zero beyond npw,
    Im hpsi(G=0) = 0."""
    n = gs.basis.npw
    hpsi = np.zeros_like(psi)
    hpsi[:, :n] = gs.basis.g2kin[None, :n] * psi[:, :n]
    if gs.basis.gstart == 2:
        hpsi[:, 0] = hpsi[:, 0].real
    return hpsi


def vloc_psi_gamma(gs, psi, hpsi):
    """QE ``vloc_psi_gamma`` (called from ``h_psi_``, ``apply_sternheimerop_to_m_wfcs.f90:57-59``):
    bands taken two at a time, ``psic = invfft(psi_i + i psi_j)``,
    ``psic *= vrs``, ``fwfft`` and the fp/fm unpacking of
    ``fft_at_gamma.f90:106-109``; an odd last band is transformed alone."""
    b = gs.basis
    n = b.npw
    m = psi.shape[0]
    npair = m // 2
    if npair:
        psic = double_invfft_gamma(b, n, psi[0:2 * npair:2], psi[1:2 * npair:2])
        psic *= gs.vrs[None, :]
        h1, h2 = double_fwfft_gamma(b, n, psi.shape[1], psic)
        hpsi[0:2 * npair:2, :n] += h1[:, :n]
        hpsi[1:2 * npair:2, :n] += h2[:, :n]
    if m % 2:
        psic = double_invfft_gamma(b, n, psi[m - 1:m], np.zeros_like(psi[m - 1:m]))
        psic *= gs.vrs[None, :]
        h1, _ = double_fwfft_gamma(b, n, psi.shape[1], psic)
        hpsi[m - 1, :n] += h1[0, :n]
    return hpsi


def h_psi(gs, psi):
    """QE ``h_psi_`` at Gamma for a local + Kleinman-Bylander Hamiltonian:
    kinetic, ``vloc_psi``, ``calbec`` + ``add_vuspsi`` (becp = glbrak(vkb, psi),
    hpsi += vkb (D becp)), then Im hpsi(G=0) = 0."""
    b = gs.basis
    n = b.npw
    hpsi = np.zeros_like(psi)
    hpsi[:, :n] = b.g2kin[None, :n] * psi[:, :n]
    vloc_psi_gamma(gs, psi, hpsi)
    if gs.vkb.shape[0] > 0:
        becp = glbrak_gamma(gs.vkb, psi, n, b.gstart)          # (nkb, m)
        ps = gs.dvan[:, None] * becp
        hv = ps.T @ gs.vkb.real + 1j * (ps.T @ gs.vkb.imag)    # (m, npwx) real DGEMM
        hpsi += hv
    if b.gstart == 2:
        hpsi[:, 0] = hpsi[:, 0].real
    return hpsi


def apply_alpha_pc(gs, f, alpha):
    """``Hamiltonian_kernel/apply_alpha_pc_to_m_wfcs.f90:43-60`` (Gamma):
    ps = glbrak(evc, f); f <- alpha f - alpha evc ps."""
    evc = gs.evc[:gs.nbndval]                                     # evc(:, 1:nbndval)
    ps = glbrak_gamma(evc, f, gs.basis.npw, gs.basis.gstart)      # (nbndval, m)
    ar = float(np.real(alpha))
    upd = ps.T @ evc.real + 1j * (ps.T @ evc.imag)
    return ar * f - ar * upd


def apply_alpha_pv(gs, f, g, alpha):
    """``Hamiltonian_kernel/apply_alpha_pv_to_m_wfcs.f90:56-72`` (Gamma):
    g <- g + alpha evc glbrak(evc, f)."""
    evc = gs.evc[:gs.nbndval]
    ps = glbrak_gamma(evc, f, gs.basis.npw, gs.basis.gstart)
    ar = float(np.real(alpha))
    upd = ps.T @ evc.real + 1j * (ps.T @ evc.imag)
    return g + ar * upd


def apply_sternheimerop(gs, psi, e, alpha):
    """``DFPT_kernel/apply_sternheimerop_to_m_wfcs.f90:14-84``:
    hpsi = H psi - e psi + alpha P_v psi (H = k_psi when l_kinetic_only)."""
    n = gs.basis.npw
    hpsi = k_psi(gs, psi) if gs.l_kinetic_only else h_psi(gs, psi)
    hpsi[:, :n] -= np.asarray(e)[:, None] * psi[:, :n]
    return apply_alpha_pv(gs, psi, hpsi, alpha)


def precondition_m_wfcts(gs, f, eprec):
    """``DFPT_kernel/precondition_m_wfcts.f90:35-51``: pf = f / max(1, g2kin/eprec),
    zero beyond npw."""
    n = gs.basis.npw
    pf = np.zeros_like(f)
    pf[:, :n] = f[:, :n] / np.maximum(1.0, gs.basis.g2kin[None, :n] / np.asarray(eprec)[:, None])
    return pf


def set_eprec(gs, wfc):
    """``Tools/set_eprec.f90:40-56``: eprec = 1.35 sum_{ig<=npw} |wfc|^2 g2kin
    (half-sphere sum, no Gamma doubling)."""
    n = gs.basis.npw
    return 1.35 * np.sum((wfc[:, :n].real ** 2 + wfc[:, :n].imag ** 2) * gs.basis.g2kin[None, :n], axis=1)


def get_alpha_pv(et_all, nbnd_occ):
    """``Tools/get_alpha_pv.f90:35-57``: alpha = max(2 (max_occ e - min_all e), 0.01)."""
    et_all = np.asarray(et_all)
    emin = float(et_all.min())
    emax = float(et_all[:nbnd_occ].max())
    return max(2.0 * (emax - emin), 0.01)
