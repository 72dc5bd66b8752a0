"""chi0 application at Gamma (restatement of ``DFPT_kernel/dfpt_module.f90:22-533``,
Gamma branches, one band group, integer occupations), ``Wstat/apply_operator.f90:14-69``
and the Coulomb factor of ``Coulomb_kernel/class_coulomb.f90:49-188``."""
from __future__ import annotations

import numpy as np

from .fft_at_gamma import (double_fwfft_gamma, double_invfft_gamma, single_fwfft_gamma,
                           single_invfft_gamma)
from .hamiltonian import apply_alpha_pc, precondition_m_wfcts, set_eprec
from .linsolve import linsolve_sternheimer_m_wfcts

E2 = 2.0
FPI = 4.0 * np.pi
EPS8 = 1.0e-8


def sqvc_init(basis, npwqx=None):
    """``class_coulomb.f90:139-176`` ('default' mode, Wave driver, Gamma):
    sqvc(G) = sqrt(e2 fpi / (|G|^2 tpiba2)), 0 where |G|^2 tpiba2 < eps8."""
    npwq = basis.npw
    npwqx = npwqx or npwq
    sq = np.zeros(npwqx)
    qg2 = basis.g2kin[:npwq]
    ok = qg2 >= EPS8
    sq[:npwq][ok] = np.sqrt(E2 * FPI / qg2[ok])
    return sq


def dfpt(gs, dvg, tr2, n_dfpt_maxiter=250, stats=None, ierr_out=None, bands=None):
    """dng = chi0 dvg.  dvg: (m, npwqx).  tr2 < 0: no Sternheimer solve
    (the initial guess M^-1 (-P_c dv psi) is used, ``dfpt_module.f90:130,348-361``).
    ``bands`` (timing samples only) restricts the band loop to a subset of the
    occupied bands; the projectors still span all of them."""
    b = gs.basis
    npw = b.npw
    npwx = gs.npwx
    m = dvg.shape[0]
    dng = np.zeros_like(dvg)
    sel = np.arange(gs.nbndval) if bands is None else np.asarray(bands)
    evc = gs.evc[sel]
    nbndval = len(sel)
    eprec = set_eprec(gs, evc)                                   # (211)
    et = np.asarray(gs.et)[sel]
    occ = np.asarray(gs.occ)[sel]
    npair = nbndval // 2
    for ipert in range(m):
        aux_r = single_invfft_gamma(b, npw, dvg[ipert:ipert + 1])[0]      # (229-235)
        dvpsi = np.zeros((nbndval, npwx), dtype=np.complex128)
        if npair:                                                # (245-260)
            psic = double_invfft_gamma(b, npw, evc[0:2 * npair:2], evc[1:2 * npair:2])
            psic *= aux_r.real[None, :]
            d1, d2 = double_fwfft_gamma(b, npw, npwx, psic)
            dvpsi[0:2 * npair:2] = d1
            dvpsi[1:2 * npair:2] = d2
        if nbndval % 2:                                          # (263-278)
            psic = single_invfft_gamma(b, npw, evc[nbndval - 1:nbndval])
            psic = (psic.real * aux_r.real[None, :]).astype(np.complex128)
            dvpsi[nbndval - 1] = single_fwfft_gamma(b, npw, npwx, psic)[0]
        dvpsi = apply_alpha_pc(gs, dvpsi, -1.0)                  # (346)
        dpsi = precondition_m_wfcts(gs, dvpsi, eprec)            # (348)
        if tr2 >= 0.0:                                           # (350-361)
            ierr = linsolve_sternheimer_m_wfcts(gs, nbndval, dvpsi, dpsi, et, eprec, tr2,
                                                n_dfpt_maxiter, stats)
            if ierr_out is not None:
                ierr_out.append(ierr)
        # response density (403-423): aux_r += occ Re(psic) Im(psic)
        psic = double_invfft_gamma(b, npw, evc, dpsi)
        aux = np.sum(occ[:, None] * psic.real * psic.imag, axis=0)
        aux_g = single_fwfft_gamma(b, npw, dvg.shape[1], aux.astype(np.complex128)[None, :])[0]
        dng[ipert, :npw] += 2.0 * gs.weight * aux_g[:npw] / b.omega   # (477-487)
    if b.gstart == 2:
        dng[:, 0] = 0.0                                          # (502-503)
    return dng


def apply_operator(gs, sqvc, dvg, tr2, n_dfpt_maxiter=250, stats=None, bands=None):
    """``Wstat/apply_operator.f90:49-67``: dng = sqvc * chi0(sqvc * dvg)."""
    npwq = gs.basis.npw
    aux = np.zeros_like(dvg)
    aux[:, :npwq] = dvg[:, :npwq] * sqvc[None, :npwq]
    dng = dfpt(gs, aux, tr2, n_dfpt_maxiter, stats, bands=bands)
    dng[:, :npwq] *= sqvc[None, :npwq]
    return dng
