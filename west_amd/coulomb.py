"""Coulomb factor of the Wstat operator (``Coulomb_kernel/class_coulomb.f90:139-176``,
'default' singularity mode, 'Wave' driver, Gamma point):
``sqvc(G) = sqrt(e2 * 4 pi / (|G|^2 tpiba2))`` with e2 = 2 (Ry), and 0 where
``|G|^2 tpiba2 < 1e-8``.  Computed on the host once per ground state and
registered with :func:`west_amd.wstat.set_ground_state`."""
from __future__ import annotations

import numpy as np

E2 = 2.0
EPS8 = 1.0e-8


def sqvc_init(basis, npwqx: int | None = None) -> np.ndarray:
    npwq = basis.npw
    out = np.zeros(npwqx or npwq)
    g2 = basis.g2kin[:npwq]
    ok = g2 >= EPS8
    out[:npwq][ok] = np.sqrt(E2 * 4.0 * np.pi / g2[ok])
    return out
