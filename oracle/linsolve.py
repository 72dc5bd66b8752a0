"""Preconditioned CG for the Sternheimer equations (restatement of
``DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:14-382``)."""
from __future__ import annotations

import numpy as np

from .hamiltonian import apply_sternheimerop


def _gamma_dot(a, b, n, gstart, gamma=True):
    """Per-column 2 Re<a|b> - Re a(1) Re b(1)  (``linsolve...f90:162-181, 253-276``);
    away from Gamma the plain Re<a|b> over the full sphere (183-205, 277-305)."""
    r = np.sum(a[:, :n].real * b[:, :n].real + a[:, :n].imag * b[:, :n].imag, axis=1)
    if not gamma:
        return r
    r = 2.0 * r
    if gstart == 2:
        r -= a[:, 0].real * b[:, 0].real
    return r


def linsolve_sternheimer_m_wfcts(gs, nbndval, b, x, e, eprec, tr2, n_dfpt_maxiter=250, stats=None):
    """Solve (H - e_i + alpha_pv P_v) x_i = b_i.  ``b`` is destroyed (it holds
    ``hold``), ``x`` is refined in place.  Returns ierr = number of
    unconverged columns.  ``stats['nstern']`` counts operator applications
    (columns), ``stats['ncalls']`` calls of ``apply_sternheimerop`` (the
    reference's 'stern' clock, lines 127 and 251), ``stats['iters']``
    iterations.  A ``gs`` with ``gamma = False`` (one k point) supplies its
    operator as ``gs.sternheimer(psi, e, alpha)`` and uses full-sphere
    products."""
    bas = gs.basis
    n = bas.npw
    gstart = bas.gstart
    gamma = getattr(gs, "gamma", True)
    op = getattr(gs, "sternheimer", None) if not gamma else None
    if op is None:
        def op(psi, ee, al):
            return apply_sternheimerop(gs, psi, ee, al)
    m = b.shape[0]
    e = np.asarray(e, dtype=np.float64)
    eprec = np.asarray(eprec, dtype=np.float64)
    alpha = gs.alpha_pv
    g = np.zeros_like(b)
    t = np.zeros_like(b)
    h = np.zeros_like(b)
    rho = np.zeros(m)
    rhoold = np.zeros(m)
    is_conv = np.zeros(m, dtype=bool)
    nstern = 0
    ncalls = 1
    # Step 1 (127-138): g = A x - b
    g[:] = op(x, e, alpha)
    nstern += m
    g[:, :n] -= b[:, :n]
    l2i = np.flatnonzero(~is_conv)                     # map_lbnd2ibnd (356-382)
    it_done = 0
    for it in range(1, n_dfpt_maxiter + 1):
        it_done = it
        nc = l2i.size
        # preconditioning (150-160), only ig <= npw
        h[l2i, :n] = g[l2i, :n] / np.maximum(1.0, bas.g2kin[None, :n] / eprec[l2i, None])
        rho_c = _gamma_dot(h[l2i], g[l2i], n, gstart, gamma)   # (162-181), compacted
        # un-compact rho and test convergence (208-218)
        rho[l2i] = rho_c
        is_conv[l2i] = rho[l2i] < tr2
        l2i = np.flatnonzero(~is_conv)                  # (220)
        nc = l2i.size
        if nc == 0:
            break
        # new direction; hold is kept in b (226-240)
        if it != 1:
            h[l2i] = -h[l2i] + (rho[l2i] / rhoold[l2i])[:, None] * b[l2i]
        else:
            h[l2i] = -h[l2i]
        b[:nc] = h[l2i]
        eu = e[l2i]
        # t = A h on the compacted columns (251)
        t[:nc] = op(b[:nc], eu, alpha)
        nstern += nc
        ncalls += 1
        a_c = _gamma_dot(h[l2i], g[l2i], n, gstart, gamma)     # (253-276)
        c_c = _gamma_dot(h[l2i], t[:nc], n, gstart, gamma)
        step = (a_c / c_c)[:, None]
        x[l2i] -= step * h[l2i]                         # (313-321)
        g[l2i] -= step * t[:nc]
        b[l2i] = h[l2i]
        rhoold[l2i] = rho[l2i]                          # (324-328)
    ierr = int(l2i.size)
    if stats is not None:
        stats["nstern"] = stats.get("nstern", 0) + nstern
        stats["ncalls"] = stats.get("ncalls", 0) + ncalls
        stats["iters"] = stats.get("iters", 0) + it_done
    return ierr
