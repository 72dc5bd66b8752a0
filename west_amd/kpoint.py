"""k-point ground states and q-point maps for the Wstat path at q != 0.

The reference's q-point Wstat (``Wstat/davidson_diago.f90:441-905``,
``DFPT_kernel/dfpt_module.f90`` non-Gamma branches) reads from QE, per k
point, the plane waves ``igk_k``/``ngk``, the states ``evc`` and ``et``, the
projectors ``vkb`` (``init_us_2``) and the weights ``wk``; per q point it
builds the ``|q+G|^2 <= ecutwfc`` sphere ``igq_q`` (``Tools/set_npwq.f90:50-85``)
and, per k, the index of ``k - q`` on the k grid with the reciprocal vector
``g0`` that folds it back (``Tools/class_bz_grid.f90:166-214``).  Here the
same quantities are built for the synthetic systems of
:mod:`west_amd.systems`:

* the k grid is the unshifted Monkhorst-Pack grid ``(i/nk1, j/nk2, l/nk3)``
  in crystal units, all points kept (no symmetry), weights ``2/nks`` (nspin =
  1, ``sum wk = 2``), the q grid the same points (``class_bz_grid.f90:71-131``);
* the sphere of ``k`` is ``|k+G|^2 <= ecutwfc`` (Ry), ordered by ascending
  ``|k+G|^2`` (ties closer than 1e-8 in tpiba^2 units in generation order);
  ``nl`` maps each ``k+G`` to the FFT grid of ``G`` (1-based, first index
  fastest), as ``dffts%nl(igk_k(ig, ik))``;
* the states at each k are the lowest eigenpairs of the dense plane-wave
  Hamiltonian ``|k+G|^2 delta + V(G-G') + sum beta D beta^*`` (small cells),
  with ``V`` the potential of the Gamma machinery (the same ``vrs`` on the
  same grid), so that a supercell at Gamma and its primitive cell on the
  matching k grid describe the same crystal.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .gsphere import EPS8, grid_index
from .systems import System, kb_projectors_at, ks_potential, atomic_density_g, local_ionic_potential_g


@dataclass
class KBasis:
    """Plane waves ``k + G`` with ``|k+G|^2 <= ecutwfc`` on a grid."""

    xk: np.ndarray            # (3,) Cartesian, bohr^-1
    mill: np.ndarray          # (npw, 3) Miller indices of G
    g2kin: np.ndarray         # (npw,) |k+G|^2, Ry
    nl: np.ndarray            # (npw,) int32, 1-based grid index of G
    gvec: np.ndarray          # (npw, 3) Cartesian k + G
    nr: tuple
    omega: float
    tpiba2: float

    @property
    def npw(self) -> int:
        return int(self.mill.shape[0])

    @property
    def nnr(self) -> int:
        return int(self.nr[0] * self.nr[1] * self.nr[2])

    @property
    def gstart(self) -> int:
        """2 when the first vector is k + G = 0 (only at k = 0)."""
        return 2 if self.g2kin[0] < EPS8 else 1


def make_k_basis(gbasis, xk, ecut: float | None = None) -> KBasis:
    """Sphere ``|k+G|^2 <= ecut`` (default ecutwfc) on the grid of ``gbasis``."""
    ecut = gbasis.ecutwfc if ecut is None else float(ecut)
    xk = np.asarray(xk, dtype=np.float64)
    bg = gbasis.bg
    at = 2.0 * np.pi * np.linalg.inv(bg).T
    kmax = np.sqrt(ecut) + np.linalg.norm(xk)
    nb = np.floor(kmax * np.linalg.norm(at, axis=1) / (2.0 * np.pi) + 1e-9).astype(int) + 1
    i = np.arange(-nb[0], nb[0] + 1)
    j = np.arange(-nb[1], nb[1] + 1)
    k = np.arange(-nb[2], nb[2] + 1)
    I, J, K = np.meshgrid(i, j, k, indexing="ij")
    mill = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    kg = mill @ bg + xk[None, :]
    g2 = np.einsum("ij,ij->i", kg, kg)
    sel = g2 <= ecut * (1.0 + 1e-12)
    mill, g2, kg = mill[sel], g2[sel], kg[sel]
    key = np.round(g2 / gbasis.tpiba2 / EPS8) * EPS8
    order = np.argsort(key, kind="stable")
    mill, g2, kg = mill[order], g2[order], kg[order]
    nr = tuple(int(n) for n in gbasis.nr)
    if np.any(np.abs(mill).max(axis=0) * 2 + 1 > np.array(nr)):
        raise ValueError("the k sphere does not fit the FFT grid")
    nl = (grid_index(mill, nr) + 1).astype(np.int32)
    return KBasis(xk=xk, mill=mill.astype(np.int32), g2kin=g2.astype(np.float64), nl=nl, gvec=kg, nr=nr,
                  omega=float(gbasis.omega), tpiba2=float(gbasis.tpiba2))


def mp_grid(gbasis, nk) -> tuple[np.ndarray, np.ndarray]:
    """Unshifted nk1 x nk2 x nk3 grid: crystal coordinates (nks, 3) in the
    loop order of ``class_bz_grid.f90:107-119`` and Cartesian k (bohr^-1)."""
    nk1, nk2, nk3 = nk
    cr = np.array([[i / nk1, j / nk2, l / nk3] for i in range(nk1) for j in range(nk2) for l in range(nk3)])
    return cr, cr @ gbasis.bg


def find_point(cryst_grid: np.ndarray, p_cryst) -> tuple[int, np.ndarray]:
    """``findp`` (``class_bz_grid.f90:166-214``): index (0-based) of the grid
    point equivalent to ``p`` and the integer crystal vector ``g0 = p - p_i``."""
    p = np.asarray(p_cryst, dtype=np.float64)
    for i, c in enumerate(cryst_grid):
        d = p - c - np.rint(p - c)
        if np.all(np.abs(d) < EPS8):
            return i, np.rint(p - c).astype(np.int64)
    raise ValueError("point not commensurate with the grid")


def dense_hamiltonian_k(kb: KBasis, vrs: np.ndarray, vkb: np.ndarray, dvan: np.ndarray) -> np.ndarray:
    """``|k+G|^2 delta + V(G-G') + sum_i beta_i D_i beta_i^*`` on the k sphere."""
    n1, n2, n3 = kb.nr
    vg = (np.fft.fftn(vrs.reshape(n3, n2, n1)) / vrs.size).ravel()
    d = kb.mill[:, None, :] - kb.mill[None, :, :]
    idx = grid_index(d.reshape(-1, 3), kb.nr).reshape(kb.npw, kb.npw)
    h = vg[idx].astype(np.complex128)
    h[np.diag_indices(kb.npw)] += kb.g2kin
    if vkb.shape[0]:
        b = vkb[:, :kb.npw]                          # (nkb, npw)
        h += (b.T * dvan[None, :]) @ b.conj()
    return 0.5 * (h + h.conj().T)


@dataclass
class KPoint:
    basis: KBasis
    npwx: int
    evc: np.ndarray           # (nbnd, npwx) complex
    et: np.ndarray            # (nbnd,) Ry
    vkb: np.ndarray           # (nkb, npwx) complex
    weight: float


@dataclass
class QPoint:
    xq: np.ndarray            # (3,) Cartesian, bohr^-1
    cryst: np.ndarray         # (3,) crystal coordinates
    basis: KBasis             # |q+G|^2 <= ecutwfc
    ikq: np.ndarray           # (nks,) 0-based index of k - q on the k grid
    g0: np.ndarray            # (nks, 3) integer crystal vectors, k - q = k_ikq + g0
    ig_global: np.ndarray     # (npwq,) 1-based global index of G (random-number counter)

    @property
    def is_gamma(self) -> bool:
        return bool(np.all(np.abs(self.cryst) < EPS8))


@dataclass
class KGroundState:
    """Ground state on a k grid (nspin = 1, insulating, integer occupations)."""

    gbasis: object            # GammaBasis of the grid (cell, grid, cutoffs)
    nk: tuple
    kcryst: np.ndarray        # (nks, 3)
    kpts: list                # list[KPoint]
    vrs: np.ndarray           # (nnr,) Ry
    dvan: np.ndarray          # (nkb,)
    nbndval: int
    alpha_pv: float
    qpts: list = field(default_factory=list)   # list[QPoint]
    l_kinetic_only: bool = False
    meta: dict = field(default_factory=dict)

    @property
    def nks(self) -> int:
        return len(self.kpts)

    @property
    def npwx(self) -> int:
        return max(k.npwx for k in self.kpts)

    @property
    def nkb(self) -> int:
        return int(self.dvan.shape[0])

    @property
    def omega(self) -> float:
        return float(self.gbasis.omega)

    @property
    def nr(self) -> tuple:
        return tuple(int(n) for n in self.gbasis.nr)

    @property
    def npwqx(self) -> int:
        return max(q.basis.npw for q in self.qpts)


def _global_g_index(gbasis, mill: np.ndarray) -> np.ndarray:
    """1-based index of each Miller triple in a fixed enumeration of the
    grid's frequencies (plays the role of QE's ``ig_l2g``: the random start
    vectors of ``do_randomize_q`` are drawn per global G)."""
    nr = gbasis.nr
    return (grid_index(mill, nr) + 1).astype(np.int64)


def make_qpoint(gbasis, kcryst: np.ndarray, q_cryst) -> QPoint:
    q_cryst = np.asarray(q_cryst, dtype=np.float64)
    xq = q_cryst @ gbasis.bg
    qb = make_k_basis(gbasis, xq)
    ikq, g0 = [], []
    for kc in kcryst:
        i, g = find_point(kcryst, kc - q_cryst)
        ikq.append(i)
        g0.append(g)
    return QPoint(xq=xq, cryst=q_cryst, basis=qb, ikq=np.array(ikq, dtype=np.int64),
                  g0=np.array(g0, dtype=np.int64), ig_global=_global_g_index(gbasis, qb.mill))


def make_k_ground_state(system: System, nk, nbnd: int | None = None, npwx: int | None = None,
                        vrs: np.ndarray | None = None) -> KGroundState:
    """States on the unshifted ``nk`` grid by dense diagonalisation of H_k,
    with the atomic-superposition LDA potential of ``system`` (or ``vrs``).
    All grid points are also the q points."""
    gb = system.basis
    nbndval = system.nbnd_occ
    nbnd = nbnd or nbndval
    if vrs is None:
        vrs = ks_potential(system, atomic_density_g(system), local_ionic_potential_g(system))
    kcryst, kcart = mp_grid(gb, nk)
    bases = [make_k_basis(gb, xk) for xk in kcart]
    npwx = npwx or max(b.npw for b in bases)
    nks = len(bases)
    kpts, dvan = [], None
    for kb in bases:
        vkb, dvan = kb_projectors_at(system, kb.gvec, kb.omega)
        vkbx = np.zeros((vkb.shape[1], npwx), dtype=np.complex128)
        vkbx[:, :kb.npw] = vkb.T
        h = dense_hamiltonian_k(kb, vrs, vkbx, dvan)
        w, z = np.linalg.eigh(h)
        evc = np.zeros((nbnd, npwx), dtype=np.complex128)
        evc[:, :kb.npw] = z[:, :nbnd].T
        kpts.append(KPoint(basis=kb, npwx=npwx, evc=evc, et=w[:nbnd].copy(), vkb=vkbx, weight=2.0 / nks))
    et_all = np.concatenate([k.et for k in kpts])
    emax = max(float(k.et[:nbndval].max()) for k in kpts)
    alpha_pv = max(2.0 * (emax - float(et_all.min())), 0.01)   # Tools/get_alpha_pv.f90:35-57 over all k
    kgs = KGroundState(gbasis=gb, nk=tuple(nk), kcryst=kcryst, kpts=kpts, vrs=np.asarray(vrs, dtype=np.float64),
                       dvan=np.asarray(dvan), nbndval=nbndval, alpha_pv=alpha_pv,
                       meta={"system": system.name})
    kgs.qpts = [make_qpoint(gb, kcryst, qc) for qc in kcryst]
    gaps = [float(k.et[nbndval] - k.et[nbndval - 1]) for k in kpts if nbnd > nbndval]
    if gaps:
        kgs.meta["min_direct_gap_ry"] = min(gaps)
    homo = max(float(k.et[nbndval - 1]) for k in kpts)
    if nbnd > nbndval:
        kgs.meta["gap_ry"] = min(float(k.et[nbndval]) for k in kpts) - homo
    return kgs



# s-channel projector on hydrogen (D = 0.4 Ha, r = 0.35 bohr): synthetic, so
# that the molecular crystal below exercises the Kleinman-Bylander term at k
H_KB = None


def _h_kb_species():
    from .systems import H_HGH, Species
    global H_KB
    if H_KB is None:
        H_KB = Species("H", 1.0, 0.2, H_HGH.cloc, ((0, 0.35, 0.4),), 0.9)
    return H_KB


def make_h2_crystal(ecutwfc: float = 10.0, nrep: int = 1, seed: int = 7, cell=None, nmol: int = 2) -> System:
    """A molecular crystal of ``nmol`` H2 molecules (bond 1.4 bohr, seeded
    orientations) in a small triclinic cell, with an s projector on every
    atom; ``nrep`` repeats the cell along a1 on a commensurate FFT grid (the
    supercell whose Gamma point holds the primitive cell's k = (j/nrep, 0, 0)
    points).  An insulator with a gap of about 0.6 Ry at 10 Ry."""
    from .gsphere import fft_dims, make_gamma_basis
    at = np.diag([6.5, 6.0, 5.5]) + np.array([[0, 0.4, 0.3], [0, 0, 0.5], [0, 0, 0]]) if cell is None else np.asarray(cell)
    rng = np.random.default_rng(seed)
    cent = np.array([[0.25, 0.3, 0.2], [0.7, 0.75, 0.7], [0.5, 0.1, 0.55], [0.05, 0.6, 0.9]])[:nmol] @ at
    tau = []
    for c in cent:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        tau += [c + 0.7 * v, c - 0.7 * v]
    tau = np.array(tau)
    nr = fft_dims(at, 4.0 * ecutwfc)
    if nrep > 1:
        ats = at.copy()
        ats[0] *= nrep
        basis = make_gamma_basis(ats, ecutwfc, nr=(nrep * nr[0], nr[1], nr[2]), alat=float(np.linalg.norm(at[0])))
        tau = np.vstack([tau + i * at[0] for i in range(nrep)])
    else:
        basis = make_gamma_basis(at, ecutwfc, nr=nr)
    sp = _h_kb_species()
    return System(f"{nmol * nrep}H2_crystal_{ecutwfc:g}Ry", basis, [sp], np.zeros(len(tau), dtype=int), tau,
                  2 * nmol * nrep, {"seed": seed, "nrep": nrep})
