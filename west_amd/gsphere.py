"""Plane-wave basis at the Gamma point: lattice, FFT grid, G-sphere and index maps.

This is the input side of the Wstat hot path: the quantities that the
reference reads from Quantum ESPRESSO's ``gvect``/``dffts`` modules
(``DFPT_kernel/dfpt_module.f90:27-34``, ``FFT_kernel/fft_at_gamma.f90:59-63``)
and that the C ABI receives explicitly (``nl``, ``nlm``, ``g2kin``, ``omega``,
``tpiba2``).

Conventions (Rydberg atomic units, as in the reference):

* lattice vectors are the rows of ``at`` in bohr; reciprocal vectors the rows of
  ``bg`` with ``a_i . b_j = 2 pi delta_ij``;
* ``g2kin(G) = |G|^2`` in Ry; the wave-function sphere is ``|G|^2 <= ecutwfc``;
* the FFT grid dimension along ``a_i`` is ``2 * max|m_i| + 1`` over the density
  sphere ``|G|^2 <= ecutrho``, rounded up to the next integer whose only prime
  factors are 2, 3 and 5;
* Gamma half-space: Miller ``(i, j, k)`` with ``i > 0``, or ``i == 0, j > 0``,
  or ``i == j == 0, k >= 0`` (so ``G = 0`` is the first vector, ``gstart = 2``);
* order inside the sphere: ascending ``|G|^2`` (ties closer than 1e-8 in
  tpiba^2 units are equal), ties kept in Miller-loop generation order.  The
  reference's in-shell order comes from QE's heap sort, which is not in its tree;
  every consumer here takes the order from the ``nl`` table it is given;
* ``nl``/``nlm`` are 1-based indices into the grid with the first index fastest,
  ``ir = 1 + i + j*nr1 + k*nr1*nr2`` (negative Miller indices wrapped), exactly
  the layout of QE's ``dffts%nl``/``dffts%nlm``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EPS8 = 1.0e-8


def good_fft_order(n: int) -> int:
    """Smallest m >= n whose prime factors are only 2, 3 and 5."""
    m = max(int(n), 1)
    while True:
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


def reciprocal(at: np.ndarray) -> np.ndarray:
    """Rows b_i with a_i . b_j = 2 pi delta_ij."""
    return 2.0 * np.pi * np.linalg.inv(np.asarray(at, dtype=np.float64)).T


def _miller_bounds(bg: np.ndarray, gcut: float) -> np.ndarray:
    # |m_i| <= |G| |a_i| / 2pi, with a_i = 2pi * (inverse of bg rows)
    at = 2.0 * np.pi * np.linalg.inv(bg).T
    return np.floor(np.sqrt(gcut) * np.linalg.norm(at, axis=1) / (2.0 * np.pi) + 1e-9).astype(int) + 1


def _sphere_miller(bg: np.ndarray, gcut: float, half: bool):
    """Miller indices with |G|^2 <= gcut (bohr^-2), in generation order."""
    nb = _miller_bounds(bg, gcut)
    i = np.arange(-nb[0], nb[0] + 1)
    j = np.arange(-nb[1], nb[1] + 1)
    k = np.arange(-nb[2], nb[2] + 1)
    I, J, K = np.meshgrid(i, j, k, indexing="ij")
    mill = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    if half:
        keep = (mill[:, 0] > 0) | ((mill[:, 0] == 0) & (mill[:, 1] > 0)) | (
            (mill[:, 0] == 0) & (mill[:, 1] == 0) & (mill[:, 2] >= 0))
        mill = mill[keep]
    g = mill @ bg
    g2 = np.einsum("ij,ij->i", g, g)
    sel = g2 <= gcut * (1.0 + 1e-12)
    return mill[sel], g2[sel]


def fft_dims(at: np.ndarray, ecutrho: float) -> tuple[int, int, int]:
    """FFT grid of the density sphere |G|^2 <= ecutrho (Ry)."""
    bg = reciprocal(at)
    mill, _ = _sphere_miller(bg, ecutrho, half=False)
    nmax = np.abs(mill).max(axis=0)
    return tuple(good_fft_order(2 * int(n) + 1) for n in nmax)


def count_sphere(at: np.ndarray, ecut: float, half: bool) -> int:
    return int(_sphere_miller(reciprocal(at), ecut, half)[0].shape[0])


@dataclass
class GammaBasis:
    """Gamma-point wave-function basis on one FFT grid."""

    at: np.ndarray            # (3,3) lattice vectors (rows, bohr)
    alat: float
    ecutwfc: float
    ecutrho: float
    nr: tuple                 # (nr1, nr2, nr3)
    mill: np.ndarray          # (npw,3) Miller indices, sorted
    g2kin: np.ndarray         # (npw,) |G|^2 in Ry
    nl: np.ndarray            # (npw,) int32, 1-based
    nlm: np.ndarray           # (npw,) int32, 1-based
    omega: float
    tpiba2: float
    bg: np.ndarray = field(repr=False, default=None)

    @property
    def npw(self) -> int:
        return int(self.mill.shape[0])

    @property
    def nnr(self) -> int:
        return int(self.nr[0] * self.nr[1] * self.nr[2])

    @property
    def gstart(self) -> int:
        return 2 if np.all(self.mill[0] == 0) else 1

    @property
    def gvec(self) -> np.ndarray:
        """Cartesian G (bohr^-1), (npw, 3)."""
        return self.mill @ self.bg

    @property
    def g2_tpiba(self) -> np.ndarray:
        """|G|^2 in (2pi/alat)^2 units (QE's ``gg``/``g`` convention)."""
        return self.g2kin / self.tpiba2


def grid_index(mill: np.ndarray, nr) -> np.ndarray:
    """0-based grid index of Miller indices, first index fastest."""
    i = np.mod(mill[:, 0], nr[0])
    j = np.mod(mill[:, 1], nr[1])
    k = np.mod(mill[:, 2], nr[2])
    return (i + nr[0] * (j + nr[1] * k)).astype(np.int64)


def make_gamma_basis(at, ecutwfc: float, ecutrho: float | None = None, alat: float | None = None,
                     nr=None) -> GammaBasis:
    at = np.asarray(at, dtype=np.float64)
    if ecutrho is None:
        ecutrho = 4.0 * ecutwfc
    if alat is None:
        alat = float(np.linalg.norm(at[0]))
    bg = reciprocal(at)
    if nr is None:
        nr = fft_dims(at, ecutrho)
    mill, g2 = _sphere_miller(bg, ecutwfc, half=True)
    tpiba2 = (2.0 * np.pi / alat) ** 2
    key = np.round(g2 / tpiba2 / EPS8) * EPS8
    order = np.argsort(key, kind="stable")
    mill = mill[order]
    g2 = g2[order]
    nl = (grid_index(mill, nr) + 1).astype(np.int32)
    nlm = (grid_index(-mill, nr) + 1).astype(np.int32)
    return GammaBasis(at=at, alat=alat, ecutwfc=ecutwfc, ecutrho=ecutrho, nr=tuple(int(n) for n in nr),
                      mill=mill.astype(np.int32), g2kin=g2.astype(np.float64), nl=nl, nlm=nlm,
                      omega=float(abs(np.linalg.det(at))), tpiba2=tpiba2, bg=bg)


def cubic(L: float):
    return np.eye(3) * float(L)


def r_grid(basis: GammaBasis) -> np.ndarray:
    """Cartesian real-space grid points, shape (nnr, 3), first index fastest."""
    n1, n2, n3 = basis.nr
    i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    frac = np.stack([i1.ravel() / n1, i2.ravel() / n2, i3.ravel() / n3], axis=1)
    return frac @ basis.at
