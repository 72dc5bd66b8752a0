"""Wstat Davidson driver at Gamma (restatement of ``Wstat/davidson_diago.f90:31-437``,
``do_mgs`` 909-1117, ``do_randomize`` 1121-1182, and the real variants of
``Wstat/wstat_tools.f90``: ``build_hr_real`` 345-468, ``serial_diagox_dsy``
208-273, ``redistribute_vr_distr_real`` 591-765, ``update_with_vr_distr_real``
945-1092, ``refresh_with_vr_distr_real`` 1245-1395) on one process.

Start vectors use the counter-based generator below in place of QE's
``randy`` (which is not in the reference tree): for basis vector ``ivec``
(1-based) and plane wave ``ig`` (1-based, in the sphere order given by ``nl``)
the two uniforms are ``u_k = mix64(SEED ^ (ivec << 32) ^ (2*ig + k)) / 2^64``
with ``mix64`` the SplitMix64 finaliser; amplitude ``u_0``, phase
``2 pi u_1``, divided by ``|G|^2 + 1`` in (2pi/alat)^2 units (1166-1174).
"""
from __future__ import annotations

import time

import numpy as np

from .dfpt import apply_operator

RNG_SEED = 0x5EED_0F_1E57


def mix64(x):
    """SplitMix64 finaliser on uint64 numpy arrays."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def uniform_pair(ivec: int, ig: np.ndarray):
    ig = np.asarray(ig, dtype=np.uint64)
    base = np.uint64(RNG_SEED) ^ (np.uint64(ivec) << np.uint64(32))
    u = []
    for k in (0, 1):
        z = mix64(base ^ (np.uint64(2) * ig + np.uint64(k)))
        u.append((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0))
    return u[0], u[1]


def do_randomize(basis, amat, mstart, mend):
    """Columns mstart..mend (1-based) of amat (nvecx, npwqx) get random start vectors."""
    npwq = basis.npw
    g2 = basis.g2_tpiba
    ig = np.arange(1, npwq + 1)
    gs0 = basis.gstart - 1
    for ivec in range(mstart, mend + 1):
        rr, ph = uniform_pair(ivec, ig)
        arg = 2.0 * np.pi * ph
        col = np.zeros(amat.shape[1], dtype=np.complex128)
        col[gs0:npwq] = (rr[gs0:] * np.cos(arg[gs0:]) + 1j * rr[gs0:] * np.sin(arg[gs0:])) / (g2[gs0:] + 1.0)
        amat[ivec - 1] = col


def do_mgs(basis, amat, ms, me):
    """Modified Gram-Schmidt of vectors ms..me (1-based) against 1..me, Gamma norms."""
    npwq = basis.npw
    g2 = basis.gstart == 2
    for k in range(1, me + 1):
        v = amat[k - 1]
        if k >= ms:
            anorm = 2.0 * np.sum(v[:npwq].real ** 2 + v[:npwq].imag ** 2)
            if g2:
                anorm -= v[0].real ** 2
            v[:npwq] *= 1.0 / np.sqrt(anorm)
        j0 = max(k + 1, ms)
        if j0 > me:
            continue
        blk = amat[j0 - 1:me]
        br = 2.0 * (blk[:, :npwq].real @ v[:npwq].real + blk[:, :npwq].imag @ v[:npwq].imag)
        if g2:
            br -= v[0].real * blk[:, 0].real
        blk -= br[:, None] * v[None, :]        # ZGERU over npwqx rows


def build_hr(basis, ag, bg, c, l2s, l2e, g_e):
    """hr(1:g_e, l2s:l2e) = <ag_i|bg_j> (Gamma), rows beyond g_e zeroed (1-based)."""
    npwq = basis.npw
    c[:, l2s - 1:l2e] = 0.0
    A = ag[:g_e, :npwq]
    B = bg[l2s - 1:l2e, :npwq]
    blk = 2.0 * (A.real @ B.real.T + A.imag @ B.imag.T)
    if basis.gstart == 2:
        blk -= np.outer(A[:, 0].real, B[:, 0].real)
    c[:g_e, l2s - 1:l2e] = blk


def diagox(nbase, nvec, hr, ew, vr):
    """``serial_diagox_dsy``: DSYEV('V','U') of hr(1:nbase,1:nbase); the nvec
    lowest eigenpairs go to ew(1:nvec), vr(1:nbase,1:nvec)."""
    w, z = np.linalg.eigh(hr[:nbase, :nbase], UPLO="U")
    ew[:nvec] = w[:nvec]
    vr[:nbase, :nvec] = z[:, :nvec]


def update_with_vr(basis, ag, bg, nselect, n, vr, ew):
    npwq = basis.npw
    cols = slice(n, n + nselect)
    hg = vr[:n, cols].T @ ag[:n, :npwq]
    hg2 = vr[:n, cols].T @ bg[:n, :npwq]
    ag[cols, :npwq] = -ew[cols, None] * hg + hg2


def refresh_with_vr(basis, ag, nselect, n, vr):
    npwq = basis.npw
    hg = vr[:n, :nselect].T @ ag[:n, :npwq]
    ag[:nselect, :npwq] = hg
    ag[nselect:, :] = 0.0


def davidson_diago_gamma(gs, sqvc, n_pdep_eigen, n_pdep_times=4, n_pdep_maxiter=100,
                         trev_pdep=1.0e-3, trev_pdep_rel=1.0e-1, tr2_dfpt=1.0e-16,
                         n_dfpt_maxiter=250, log=None):
    """Returns (ev, conv, dvg) with dvg the nvec PDEP eigenvectors (nvec, npwqx).
    ``log`` (a list) receives one record per Davidson step, with the keys of
    ``output_ev_and_time`` (``davidson_diago.f90:1310-1329``)."""
    basis = gs.basis
    nvec = n_pdep_eigen
    nvecx = n_pdep_eigen * n_pdep_times
    if nvec > nvecx // 2:
        raise ValueError("nvecx is too small")
    npwqx = gs.npwx
    dvg = np.zeros((nvecx, npwqx), dtype=np.complex128)
    dng = np.zeros_like(dvg)
    hr = np.zeros((nvecx, nvecx))
    vr = np.zeros((nvecx, nvecx))
    ew = np.zeros(nvecx)
    ev = np.zeros(nvec)
    conv = np.zeros(nvec, dtype=bool)
    nbase = nvec
    notcnv = nvec
    stats = {"nstern": 0, "ncalls": 0}
    t0 = time.perf_counter()

    def record(dav_iter, tr2, dfpt_dim, diago_dim, ns0, tit):
        if log is not None:
            log.append({"dav_iter": dav_iter, "ev": ev.copy().tolist(), "conv": conv.copy().tolist(),
                        "notcnv": int(notcnv), "time_elap:sec": time.perf_counter() - t0,
                        "time_iter:sec": time.perf_counter() - tit,
                        "sternop_ncalls": stats["ncalls"] - ns0, "dfpt_tr2": tr2,
                        "dfpt_dim": int(dfpt_dim), "diago_dim": int(diago_dim)})

    # from scratch (167-196)
    do_randomize(basis, dvg, 1, nvec)
    do_mgs(basis, dvg, 1, nvec)
    dng[:nvec] = apply_operator(gs, sqvc, dvg[:nvec], -1.0, n_dfpt_maxiter, stats)
    # extra step (205-249)
    tit = time.perf_counter()
    ns0 = stats["ncalls"]
    dvg[:nvec] = dng[:nvec]
    do_mgs(basis, dvg, 1, nvec)
    tr2 = min(1.0e-4, 1.0e8 * tr2_dfpt)
    dng[:nvec] = apply_operator(gs, sqvc, dvg[:nvec], tr2, n_dfpt_maxiter, stats)
    build_hr(basis, dvg, dng, hr, 1, nvec, nvec)
    diagox(nbase, nvec, hr, ew, vr)
    ev[:] = ew[:nvec]
    record(0, tr2, nbase, nbase, ns0, tit)
    dav_iter = 0
    for kter in range(1, n_pdep_maxiter + 1):                    # (259-424)
        tit = time.perf_counter()
        ns0 = stats["ncalls"]
        dav_iter += 1
        dfpt_dim = notcnv
        diago_dim = nbase + notcnv
        np_ = 0
        for n in range(1, nvec + 1):                             # (278-296)
            if not conv[n - 1]:
                np_ += 1
                vr[:, nbase + np_ - 1] = vr[:, n - 1]           # redistribute_vr_distr
                ew[nbase + np_ - 1] = ev[n - 1]
        update_with_vr(basis, dvg, dng, notcnv, nbase, vr, ew)   # (301)
        do_mgs(basis, dvg, nbase + 1, nbase + notcnv)            # (306)
        tr2 = tr2_dfpt
        dng[nbase:nbase + notcnv] = apply_operator(gs, sqvc, dvg[nbase:nbase + notcnv], tr2,
                                                   n_dfpt_maxiter, stats)
        build_hr(basis, dvg, dng, hr, nbase + 1, nbase + notcnv, nbase + notcnv)
        nbase += notcnv
        diagox(nbase, nvec, hr, ew, vr)
        conv[:] = (np.abs((ew[:nvec] - ev) / ev) < trev_pdep_rel) & (np.abs(ew[:nvec] - ev) < trev_pdep)
        notcnv = int(np.count_nonzero(~conv))
        ev[:] = ew[:nvec]
        record(dav_iter, tr2, dfpt_dim, diago_dim, ns0, tit)
        if notcnv == 0 or nbase + notcnv > nvecx or kter == n_pdep_maxiter:
            if notcnv == 0:
                refresh_with_vr(basis, dvg, nvec, nbase, vr)
                break
            if kter == n_pdep_maxiter:
                break
            refresh_with_vr(basis, dvg, nvec, nbase, vr)         # (398-415)
            refresh_with_vr(basis, dng, nvec, nbase, vr)
            nbase = nvec
            hr[:] = 0.0
            vr[:] = 0.0
            for i in range(nbase):
                hr[i, i] = ev[i]
                vr[i, i] = 1.0
    return ev.copy(), conv.copy(), dvg[:nvec].copy()
