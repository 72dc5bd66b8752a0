"""chi0(q) and the Wstat Davidson at q points (restatement of the non-Gamma
branches of ``DFPT_kernel/dfpt_module.f90:22-533``, ``Wstat/davidson_diago.f90:441-905``
(``davidson_diago_k``), ``do_mgs`` 909-1117 (complex products),
``do_randomize_q`` 1186-1250, the complex ``Wstat/wstat_tools.f90`` variants,
``FFT_kernel/fft_at_k.f90`` and QE's ``h_psi_`` at a k point).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Inputs come from :class:`west_amd.kpoint.KGroundState`.  Wave functions are
``(m, npwx)`` complex128 rows on the ``|k+G|`` sphere of their k point; a
grid array of length ``nnr`` is QE's layout (first index fastest).
"""
from __future__ import annotations

import time

import numpy as np

from .davidson import uniform_pair
from .fft_at_gamma import fwfft, invfft
from .linsolve import linsolve_sternheimer_m_wfcts


# ---------------------------------------------------------------------------
# FFT at k (fft_at_k.f90:29-77): psic(nl(igk)) = c, invfft; fwfft, c = psic(nl(igk))
# ---------------------------------------------------------------------------

def single_invfft_k(kb, c):
    c = np.asarray(c)
    lead = c.shape[:-1]
    b = np.zeros(lead + (kb.nnr,), dtype=np.complex128)
    b[..., kb.nl - 1] = c[..., :kb.npw]
    return invfft(kb.nr, b)


def single_fwfft_k(kb, f, nx):
    g = fwfft(kb.nr, f)
    out = np.zeros(f.shape[:-1] + (nx,), dtype=np.complex128)
    out[..., :kb.npw] = g[..., kb.nl - 1]
    return out


def compute_phase(nr, g0):
    """``Tools/types_bz_grid.f90:71-113``: phase(r) = exp(-i G0.r) on the
    grid, for G0 given by its integer crystal components."""
    n1, n2, n3 = nr
    i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    arg = 2.0 * np.pi * (g0[0] * i1 / n1 + g0[1] * i2 / n2 + g0[2] * i3 / n3)
    return np.exp(-1j * arg).ravel()


# ---------------------------------------------------------------------------
# one k point: the Hamiltonian pieces with complex (ZGEMM) projections
# ---------------------------------------------------------------------------

class KView:
    """The ground-state interface of the Gamma oracle for one k point, with
    full-sphere complex products (``gamma = False``)."""

    gamma = False

    def __init__(self, kgs, ik):
        kp = kgs.kpts[ik]
        self.kgs = kgs
        self.ik = ik
        self.basis = kp.basis
        self.npwx = kp.npwx
        self.vrs = kgs.vrs
        self.vkb = kp.vkb
        self.dvan = kgs.dvan
        self.evc = kp.evc
        self.et = kp.et
        self.nbndval = kgs.nbndval
        self.alpha_pv = kgs.alpha_pv
        self.l_kinetic_only = kgs.l_kinetic_only
        self.weight = kp.weight

    def sternheimer(self, psi, e, alpha):
        return apply_sternheimerop_k(self, psi, e, alpha)


def zbrak(a, b, n):
    """<a_i|b_j> = sum_G conj(a_i) b_j over n rows (ZGEMM 'C','N')."""
    return a[:, :n].conj() @ b[:, :n].T


def h_psi_k(v, psi):
    """QE ``h_psi_`` at k: |k+G|^2 psi + vloc_psi_k (invfft, times vrs, fwfft)
    + vkb (D (vkb^H psi))."""
    kb = v.basis
    n = kb.npw
    hpsi = np.zeros_like(psi)
    hpsi[:, :n] = kb.g2kin[None, :n] * psi[:, :n]
    if v.l_kinetic_only:
        return hpsi
    psic = single_invfft_k(kb, psi)
    psic *= v.vrs[None, :]
    hpsi += single_fwfft_k(kb, psic, psi.shape[1])
    if v.vkb.shape[0] > 0:
        becp = zbrak(v.vkb, psi, n)                      # (nkb, m)
        hpsi[:, :n] += (v.dvan[:, None] * becp).T @ v.vkb[:, :n]
    return hpsi


def apply_alpha_pv_k(v, f, g, alpha):
    """``apply_alpha_pv_to_m_wfcs.f90:56-78`` (k): g += alpha evc (evc^H f)."""
    evc = v.evc[:v.nbndval]
    ps = zbrak(evc, f, v.basis.npw)
    return g + alpha * (ps.T @ evc)


def apply_alpha_pc_k(v, f, alpha):
    """``apply_alpha_pc_to_m_wfcs.f90:56-78`` (k): f <- alpha f - alpha evc (evc^H f)."""
    evc = v.evc[:v.nbndval]
    ps = zbrak(evc, f, v.basis.npw)
    return alpha * f - alpha * (ps.T @ evc)


def apply_sternheimerop_k(v, psi, e, alpha):
    """``apply_sternheimerop_to_m_wfcs.f90:14-84`` at k: H_k psi - e psi + alpha P_v psi."""
    n = v.basis.npw
    hpsi = h_psi_k(v, psi)
    hpsi[:, :n] -= np.asarray(e)[:, None] * psi[:, :n]
    return apply_alpha_pv_k(v, psi, hpsi, alpha)


def set_eprec_k(v, wfc):
    """``Tools/set_eprec.f90:40-56``: 1.35 sum |wfc|^2 |k+G|^2."""
    n = v.basis.npw
    return 1.35 * np.sum(np.abs(wfc[:, :n]) ** 2 * v.basis.g2kin[None, :n], axis=1)


def precondition_k(v, f, eprec):
    n = v.basis.npw
    pf = np.zeros_like(f)
    pf[:, :n] = f[:, :n] / np.maximum(1.0, v.basis.g2kin[None, :n] / np.asarray(eprec)[:, None])
    return pf


# ---------------------------------------------------------------------------
# dfpt and apply_operator at q
# ---------------------------------------------------------------------------

def dfpt_q(kgs, iq, dvg, tr2, n_dfpt_maxiter=250, stats=None, ierr_out=None):
    """dng = chi0(q) dvg (``dfpt_module.f90`` non-Gamma branches).  dvg: (m, npwqx)
    on the |q+G| sphere of q point ``iq`` (0-based).  For each k: the states
    at k - q (``ikq``) times exp(-i g0.r) times dv_q(r) go to the k sphere
    (288-302), -P_c (346), M^-1 (348), the Sternheimer solve with H_k and the
    energies at k - q (222, 356), and the density sum_v conj(psi_{k-q} phase)
    dpsi (433-441) goes to the q sphere with 2 wk / Omega (480-486); G = 0 is
    zeroed at q = 0 (505)."""
    qp = kgs.qpts[iq]
    qb = qp.basis
    npwq = qb.npw
    m = dvg.shape[0]
    nbndval = kgs.nbndval
    dng = np.zeros((m, dvg.shape[1]), dtype=np.complex128)
    for ik in range(kgs.nks):
        v = KView(kgs, ik)
        kb = v.basis
        ikq = int(qp.ikq[ik])
        vq = KView(kgs, ikq)
        phase = compute_phase(kgs.nr, qp.g0[ik])
        eprec = set_eprec_k(v, v.evc[:nbndval])               # (211): evc at k
        et = np.asarray(vq.et[:nbndval])                      # (222): energies at k - q
        psi_kmq = single_invfft_k(vq.basis, vq.evc[:nbndval]) * phase[None, :]   # (288, 295)
        for ipert in range(m):
            aux_r = single_invfft_k(qb, dvg[ipert:ipert + 1])[0]                  # (237)
            dvpsi = single_fwfft_k(kb, psi_kmq * aux_r[None, :], v.npwx)          # (295-302)
            dvpsi = apply_alpha_pc_k(v, dvpsi, -1.0)                               # (346)
            dpsi = precondition_k(v, dvpsi, eprec)                                 # (348)
            if tr2 >= 0.0:
                ierr = linsolve_sternheimer_m_wfcts(v, nbndval, dvpsi, dpsi, et, eprec, tr2,
                                                    n_dfpt_maxiter, stats)
                if ierr_out is not None:
                    ierr_out.append(ierr)
            dpsic = single_invfft_k(kb, dpsi)                                      # (437)
            aux = np.sum(np.conj(psi_kmq) * dpsic, axis=0)                         # (441)
            aux_g = single_fwfft_k(qb, aux[None, :], dvg.shape[1])[0]             # (480)
            dng[ipert, :npwq] += 2.0 * v.weight * aux_g[:npwq] / kgs.omega        # (485-487)
    if qp.is_gamma and qb.gstart == 2:
        dng[:, 0] = 0.0                                                            # (505)
    return dng


def apply_operator_q(kgs, iq, sqvc, dvg, tr2, n_dfpt_maxiter=250, stats=None):
    """``Wstat/apply_operator.f90:49-67`` at q: dng = sqvc chi0(q) (sqvc dvg)."""
    npwq = kgs.qpts[iq].basis.npw
    aux = np.zeros_like(dvg)
    aux[:, :npwq] = dvg[:, :npwq] * sqvc[None, :npwq]
    dng = dfpt_q(kgs, iq, aux, tr2, n_dfpt_maxiter, stats)
    dng[:, :npwq] *= sqvc[None, :npwq]
    return dng


# ---------------------------------------------------------------------------
# Davidson at q
# ---------------------------------------------------------------------------

def do_randomize_q(qp, amat, mstart, mend):
    """``davidson_diago.f90:1186-1250``: amplitude rr, phase 2 pi u per global
    G (the counter generator of :mod:`oracle.davidson`), divided by
    |q+G|^2 tpiba2 + 1 (in Ry, not in tpiba units as at Gamma), skipping q+G = 0."""
    qb = qp.basis
    n = qb.npw
    for ivec in range(mstart, mend + 1):
        rr, ph = uniform_pair(ivec, qp.ig_global)
        col = np.zeros(amat.shape[1], dtype=np.complex128)
        ok = qb.g2kin >= 1e-8
        arg = 2.0 * np.pi * ph
        col[:n] = np.where(ok, (rr * np.cos(arg) + 1j * rr * np.sin(arg)) / (qb.g2kin + 1.0), 0.0)
        amat[ivec - 1] = col


def do_mgs_k(npwq, amat, ms, me):
    """``do_mgs`` (909-1117), complex branch: norm sum |v|^2, projection
    coefficients <v|a_j> (ZDOTC), a_j -= <v|a_j> v (ZGERU)."""
    for k in range(1, me + 1):
        v = amat[k - 1]
        if k >= ms:
            anorm = float(np.sum(np.abs(v[:npwq]) ** 2))
            v[:npwq] *= 1.0 / np.sqrt(anorm)
        j0 = max(k + 1, ms)
        if j0 > me:
            continue
        blk = amat[j0 - 1:me]
        br = blk[:, :npwq] @ v[:npwq].conj()
        blk -= br[:, None] * v[None, :]


def build_hr_k(npwq, ag, bg, c, l2s, l2e, g_e):
    """``build_hr`` (complex): hr(1:g_e, l2s:l2e) = ag^H bg, rows beyond g_e zeroed."""
    c[:, l2s - 1:l2e] = 0.0
    c[:g_e, l2s - 1:l2e] = ag[:g_e, :npwq].conj() @ bg[l2s - 1:l2e, :npwq].T


def diagox_k(nbase, nvec, hr, ew, vr):
    """``serial_diagox_zhe``: ZHEEV('V','U') of hr(1:nbase,1:nbase)."""
    w, z = np.linalg.eigh(hr[:nbase, :nbase], UPLO="U")
    ew[:nvec] = w[:nvec]
    vr[:nbase, :nvec] = z[:, :nvec]


def update_with_vr_k(npwq, ag, bg, nselect, n, vr, ew):
    cols = slice(n, n + nselect)
    hg = vr[:n, cols].T @ ag[:n, :npwq]
    hg2 = vr[:n, cols].T @ bg[:n, :npwq]
    ag[cols, :npwq] = -ew[cols, None] * hg + hg2


def refresh_with_vr_k(npwq, ag, nselect, n, vr):
    hg = vr[:n, :nselect].T @ ag[:n, :npwq]
    ag[:nselect, :npwq] = hg
    ag[nselect:, :] = 0.0


def davidson_diago_k(kgs, iq, sqvc, n_pdep_eigen, n_pdep_times=4, n_pdep_maxiter=100,
                     trev_pdep=1.0e-3, trev_pdep_rel=1.0e-1, tr2_dfpt=1.0e-16,
                     n_dfpt_maxiter=250, log=None):
    """One q point of ``davidson_diago_k`` (the QPOINTS_LOOP body, 551-894).
    Returns (ev, conv, dvg)."""
    qp = kgs.qpts[iq]
    npwq = qp.basis.npw
    nvec = n_pdep_eigen
    nvecx = n_pdep_eigen * n_pdep_times
    if nvec > nvecx // 2:
        raise ValueError("nvecx is too small")
    npwqx = kgs.npwqx
    dvg = np.zeros((nvecx, npwqx), dtype=np.complex128)
    dng = np.zeros_like(dvg)
    hr = np.zeros((nvecx, nvecx), dtype=np.complex128)
    vr = np.zeros((nvecx, nvecx), dtype=np.complex128)
    ew = np.zeros(nvecx)
    ev = np.zeros(nvec)
    conv = np.zeros(nvec, dtype=bool)
    nbase = nvec
    notcnv = nvec
    stats = {"nstern": 0, "ncalls": 0}
    t0 = time.perf_counter()

    def apply(block, tr2):
        return apply_operator_q(kgs, iq, sqvc, block, tr2, n_dfpt_maxiter, stats)

    def record(dav_iter, tr2, dfpt_dim, diago_dim, ns0, tit):
        if log is not None:
            log.append({"dav_iter": dav_iter, "ev": ev.copy().tolist(), "conv": conv.copy().tolist(),
                        "notcnv": int(notcnv), "time_elap:sec": time.perf_counter() - t0,
                        "time_iter:sec": time.perf_counter() - tit,
                        "sternop_ncalls": stats["ncalls"] - ns0, "dfpt_tr2": tr2,
                        "dfpt_dim": int(dfpt_dim), "diago_dim": int(diago_dim)})

    do_randomize_q(qp, dvg, 1, nvec)                                  # (633)
    do_mgs_k(npwq, dvg, 1, nvec)                                      # (637)
    dng[:nvec] = apply(dvg[:nvec], -1.0)                              # (655-660)
    tit = time.perf_counter()
    ns0 = stats["ncalls"]
    dvg[:nvec] = dng[:nvec]                                           # (671-672)
    do_mgs_k(npwq, dvg, 1, nvec)
    tr2 = min(1.0e-4, 1.0e8 * tr2_dfpt)
    dng[:nvec] = apply(dvg[:nvec], tr2)
    build_hr_k(npwq, dvg, dng, hr, 1, nvec, nvec)
    diagox_k(nbase, nvec, hr, ew, vr)
    ev[:] = ew[:nvec]
    record(0, tr2, nbase, nbase, ns0, tit)
    dav_iter = 0
    for kter in range(1, n_pdep_maxiter + 1):                         # (723-892)
        tit = time.perf_counter()
        ns0 = stats["ncalls"]
        dav_iter += 1
        dfpt_dim = notcnv
        diago_dim = nbase + notcnv
        np_ = 0
        for n in range(1, nvec + 1):
            if not conv[n - 1]:
                np_ += 1
                vr[:, nbase + np_ - 1] = vr[:, n - 1]
                ew[nbase + np_ - 1] = ev[n - 1]
        update_with_vr_k(npwq, dvg, dng, notcnv, nbase, vr, ew)
        do_mgs_k(npwq, dvg, nbase + 1, nbase + notcnv)
        tr2 = tr2_dfpt
        dng[nbase:nbase + notcnv] = apply(dvg[nbase:nbase + notcnv], tr2)
        build_hr_k(npwq, dvg, dng, hr, nbase + 1, nbase + notcnv, nbase + notcnv)
        nbase += notcnv
        diagox_k(nbase, nvec, hr, ew, vr)
        conv[:] = (np.abs((ew[:nvec] - ev) / ev) < trev_pdep_rel) & (np.abs(ew[:nvec] - ev) < trev_pdep)
        notcnv = int(np.count_nonzero(~conv))
        ev[:] = ew[:nvec]
        record(dav_iter, tr2, dfpt_dim, diago_dim, ns0, tit)
        if notcnv == 0 or nbase + notcnv > nvecx or kter == n_pdep_maxiter:
            if notcnv == 0:
                refresh_with_vr_k(npwq, dvg, nvec, nbase, vr)
                break
            if kter == n_pdep_maxiter:
                break
            refresh_with_vr_k(npwq, dvg, nvec, nbase, vr)
            refresh_with_vr_k(npwq, dng, nvec, nbase, vr)
            nbase = nvec
            hr[:] = 0.0
            vr[:] = 0.0
            for i in range(nbase):
                hr[i, i] = ev[i]
                vr[i, i] = 1.0
    return ev.copy(), conv.copy(), dvg[:nvec].copy()
