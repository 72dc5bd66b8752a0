"""Synthetic ground-state systems: atoms, GTH/HGH pseudopotentials and the
local potential / Kleinman-Bylander projectors of the Gamma-point Hamiltonian.

The Wstat path needs, from the ground state, only ``evc, et, vrs, vkb, D``
(QE's ``wavefunctions%evc``, ``wvfct%et``, ``scf%vrs``, ``uspp%vkb`` and the
``deeq`` coefficients used by ``h_psi_``).  No QE run is available here, so
these are produced from analytic norm-conserving pseudopotentials of the
Goedecker-Teter-Hutter / Hartwigsen-Goedecker-Hutter form:

* local part (Hartree units, ``x = |G| r_loc``)::

      V_loc(G) = [ -4 pi Z exp(-x^2/2) / G^2
                   + sqrt(8 pi^3) r_loc^3 exp(-x^2/2)
                     * (C1 + C2 (3 - x^2) + C3 (15 - 10 x^2 + x^4)) ] / Omega

  (the divergent ``-4 pi Z / G^2`` part at G = 0 cancels against the Hartree
  term of a neutral cell and is dropped, as usual);
* one projector per angular channel (first projector of the HGH form)::

      p_s(q) = 4 sqrt(2 r_0^3) pi^(5/4) exp(-q^2 r_0^2 / 2) / sqrt(Omega)
      p_p(q) = 8 sqrt(r_1^5 / 3) pi^(5/4) q exp(-q^2 r_1^2 / 2) / sqrt(Omega)

  with ``beta_lm(G) = (-i)^l p_l(|G|) Y_lm(G^) exp(-i G.tau)`` (real spherical
  harmonics), so that every projector is a real function at Gamma
  (``beta(-G) = conj(beta(G))``) and ``D = 2 h_11`` in Ry.

Parameter values are the published LDA ones of HGH (1998) for H and Si as
used here; the hot path does not depend on them beyond producing a
physically shaped Hamiltonian.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .gsphere import GammaBasis, cubic, make_gamma_basis


@dataclass(frozen=True)
class Species:
    name: str
    zion: float
    rloc: float
    cloc: tuple                 # (C1, C2, C3, C4), Hartree
    proj: tuple = ()            # ((l, r_l, h_11 in Hartree), ...)
    rho_width: float = 1.0      # width of the Gaussian atomic valence density (bohr)


H_HGH = Species("H", 1.0, 0.2, (-4.18023680, 0.72507482, 0.0, 0.0), (), 0.9)
SI_HGH = Species("Si", 4.0, 0.44, (-7.33610297, 0.0, 0.0, 0.0),
                 ((0, 0.42273813, 5.90692831), (1, 0.48427842, 2.72701346)), 1.6)


@dataclass
class System:
    name: str
    basis: GammaBasis
    species: list               # list[Species]
    ityp: np.ndarray            # (nat,) species index per atom
    tau: np.ndarray             # (nat,3) positions in bohr
    nelec: int
    meta: dict = field(default_factory=dict)

    @property
    def nat(self) -> int:
        return int(self.tau.shape[0])

    @property
    def nbnd_occ(self) -> int:
        """Occupied bands for nspin = 1 (``Tools/set_nbndocc.f90:104``)."""
        return int(round(self.nelec)) // 2

    @property
    def nkb(self) -> int:
        return int(sum(sum(2 * l + 1 for (l, _, _) in self.species[t].proj) for t in self.ityp))


# ----------------------------------------------------------------------------
# geometry builders
# ----------------------------------------------------------------------------

def _min_image_dist(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return float(np.linalg.norm(d))


def h2_molecules(nmol: int, L: float, seed: int, bond: float = 1.40, min_sep: float = 5.0):
    """``nmol`` H2 molecules with seeded random orientations (uniform on the
    sphere) and centres placed by rejection with minimum-image separation."""
    rng = np.random.default_rng(seed)
    centres = []
    tries = 0
    while len(centres) < nmol:
        tries += 1
        if tries > 200000:
            raise RuntimeError("cannot place molecules with the requested separation")
        c = rng.uniform(0.0, L, 3)
        if all(_min_image_dist(c, o, L) >= min_sep for o in centres):
            centres.append(c)
    tau = []
    for c in centres:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        tau.append(c + 0.5 * bond * v)
        tau.append(c - 0.5 * bond * v)
    return np.mod(np.array(tau), L)


def diamond_supercell(a: float, nrep: int, sigma: float, seed: int):
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0],
                      [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]])
    pos = []
    for i in range(nrep):
        for j in range(nrep):
            for k in range(nrep):
                for b in basis:
                    pos.append((np.array([i, j, k]) + b) * a)
    pos = np.array(pos)
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(scale=sigma, size=pos.shape)
    return np.mod(pos, a * nrep)


def make_h2_system(nmol=6, L=18.0, ecutwfc=25.0, seed=1, min_sep=5.0, name=None) -> System:
    basis = make_gamma_basis(cubic(L), ecutwfc)
    tau = h2_molecules(nmol, L, seed, min_sep=min_sep)
    return System(name or f"{nmol}H2_L{L:g}_{ecutwfc:g}Ry", basis, [H_HGH],
                  np.zeros(len(tau), dtype=int), tau, 2 * nmol, {"seed": seed})


def make_si_system(nrep=2, a=10.26, ecutwfc=25.0, sigma=0.05, seed=1, name=None) -> System:
    L = a * nrep
    basis = make_gamma_basis(cubic(L), ecutwfc)
    tau = diamond_supercell(a, nrep, sigma, seed)
    return System(name or f"Si{len(tau)}_{ecutwfc:g}Ry", basis, [SI_HGH],
                  np.zeros(len(tau), dtype=int), tau, 4 * len(tau), {"seed": seed, "sigma": sigma})


# ----------------------------------------------------------------------------
# Hamiltonian ingredients
# ----------------------------------------------------------------------------

def _full_grid_g(basis: GammaBasis):
    """Cartesian G of every FFT-grid frequency (first index fastest) and a mask
    of the density sphere."""
    n1, n2, n3 = basis.nr
    f = [np.fft.fftfreq(n, 1.0 / n).astype(int) for n in (n1, n2, n3)]
    K, J, I = np.meshgrid(f[2], f[1], f[0], indexing="ij")
    mill = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    g = mill @ basis.bg
    g2 = np.einsum("ij,ij->i", g, g)
    return g, g2, g2 <= basis.ecutrho * (1.0 + 1e-12)


def _vloc_species_g(sp: Species, g2: np.ndarray, omega: float) -> np.ndarray:
    """Local pseudopotential form factor in Ry (G = 0: finite part)."""
    x2 = g2 * sp.rloc ** 2
    e = np.exp(-0.5 * x2)
    c1, c2, c3, c4 = sp.cloc
    poly = c1 + c2 * (3 - x2) + c3 * (15 - 10 * x2 + x2 ** 2) + c4 * (105 - 105 * x2 + 21 * x2 ** 2 - x2 ** 3)
    gauss = np.sqrt(8 * np.pi ** 3) * sp.rloc ** 3 * e * poly
    with np.errstate(divide="ignore", invalid="ignore"):
        coul = np.where(g2 > 1e-12, -4 * np.pi * sp.zion * e / np.where(g2 > 1e-12, g2, 1.0),
                        2 * np.pi * sp.zion * sp.rloc ** 2)
    return 2.0 * (coul + gauss) / omega


def _struct_fac(g: np.ndarray, tau: np.ndarray) -> np.ndarray:
    return np.exp(-1j * (g @ tau.T)).sum(axis=1)


def local_ionic_potential_g(system: System) -> np.ndarray:
    g, g2, mask = _full_grid_g(system.basis)
    v = np.zeros(g.shape[0], dtype=np.complex128)
    for t, sp in enumerate(system.species):
        sel = system.ityp == t
        if not np.any(sel):
            continue
        v += _vloc_species_g(sp, g2, system.basis.omega) * _struct_fac(g, system.tau[sel])
    v[~mask] = 0.0
    return v


def atomic_density_g(system: System) -> np.ndarray:
    g, g2, mask = _full_grid_g(system.basis)
    n = np.zeros(g.shape[0], dtype=np.complex128)
    for t, sp in enumerate(system.species):
        sel = system.ityp == t
        if not np.any(sel):
            continue
        n += sp.zion * np.exp(-0.25 * g2 * sp.rho_width ** 2) * _struct_fac(g, system.tau[sel])
    n[~mask] = 0.0
    return n / system.basis.omega


def grid_to_real(basis: GammaBasis, fg: np.ndarray) -> np.ndarray:
    """Sum_G f(G) e^{iGr} on the grid (first index fastest), real part."""
    n1, n2, n3 = basis.nr
    a = fg.reshape(n3, n2, n1)
    return np.real(np.fft.ifftn(a) * a.size).ravel()


def real_to_grid(basis: GammaBasis, fr: np.ndarray) -> np.ndarray:
    n1, n2, n3 = basis.nr
    a = fr.reshape(n3, n2, n1)
    return (np.fft.fftn(a) / a.size).ravel()


def lda_pz81(n: np.ndarray):
    """Perdew-Zunger 1981 LDA exchange-correlation potential (Ry), unpolarised."""
    n = np.maximum(n, 1e-14)
    rs = (3.0 / (4.0 * np.pi * n)) ** (1.0 / 3.0)
    ex = -0.4581652932831429 / rs            # Hartree
    vx = 4.0 / 3.0 * ex
    g, b1, b2 = -0.1423, 1.0529, 0.3334
    A, B, C, D = 0.0311, -0.048, 0.0020, -0.0116
    srs = np.sqrt(rs)
    den = 1 + b1 * srs + b2 * rs
    ec_hi = g / den
    vc_hi = ec_hi * (1 + 7.0 / 6.0 * b1 * srs + 4.0 / 3.0 * b2 * rs) / den
    lnrs = np.log(rs)
    vc_lo = A * lnrs + (B - A / 3.0) + 2.0 / 3.0 * C * rs * lnrs + (2 * D - C) / 3.0 * rs
    vc = np.where(rs >= 1.0, vc_hi, vc_lo)
    return 2.0 * (vx + vc)


def ks_potential(system: System, n_g: np.ndarray, vion_g: np.ndarray | None = None) -> np.ndarray:
    """V_ion + V_H + V_xc on the real-space grid (Ry), from density n(G)."""
    b = system.basis
    g, g2, mask = _full_grid_g(b)
    if vion_g is None:
        vion_g = local_ionic_potential_g(system)
    with np.errstate(divide="ignore", invalid="ignore"):
        vh = np.where(g2 > 1e-12, 8 * np.pi * n_g / np.where(g2 > 1e-12, g2, 1.0), 0.0)
    vh[~mask] = 0.0
    n_r = grid_to_real(b, n_g * mask)
    return grid_to_real(b, vion_g + vh) + lda_pz81(n_r)


def _real_ylm(l: int, g: np.ndarray):
    """Real spherical harmonics for l = 0, 1 at directions of g (rows)."""
    gn = np.linalg.norm(g, axis=1)
    if l == 0:
        return [np.full(g.shape[0], 1.0 / np.sqrt(4 * np.pi))]
    safe = np.where(gn > 1e-12, gn, 1.0)
    c = np.sqrt(3.0 / (4 * np.pi))
    return [c * g[:, 0] / safe, c * g[:, 1] / safe, c * g[:, 2] / safe]


def kb_projectors(system: System):
    """vkb (npw, nkb) complex128 on the wave-function half sphere and the
    diagonal D (nkb,) in Ry.  Order: atom-major, then channel, then m."""
    b = system.basis
    vkb, dvals = kb_projectors_at(system, b.gvec, b.omega)
    # G = 0 of a real function is real: make it exactly so
    if b.gstart == 2 and vkb.shape[1]:
        vkb[0, :] = vkb[0, :].real
    return vkb, dvals


def kb_projectors_at(system: System, g: np.ndarray, omega: float):
    """Projectors at the wave vectors ``g`` (rows, Cartesian bohr^-1; k + G at
    a k point, as QE's ``init_us_2``): beta(|g|) Y_lm(g) (-i)^l e^{-i g.tau}
    / sqrt(omega), shape (len(g), nkb), and D (nkb,) in Ry."""
    q = np.sqrt(np.einsum("ij,ij->i", g, g))
    cols, dvals = [], []
    for ia in range(system.nat):
        sp = system.species[system.ityp[ia]]
        phase = np.exp(-1j * (g @ system.tau[ia]))
        for (l, rl, h) in sp.proj:
            if l == 0:
                p = 4.0 * np.sqrt(2 * rl ** 3) * np.pi ** 1.25 * np.exp(-0.5 * (q * rl) ** 2)
            elif l == 1:
                p = 8.0 * np.sqrt(rl ** 5 / 3.0) * np.pi ** 1.25 * q * np.exp(-0.5 * (q * rl) ** 2)
            else:
                raise ValueError("only s and p projectors are supported")
            p = p / np.sqrt(omega)
            for y in _real_ylm(l, g):
                cols.append(((-1j) ** l) * p * y * phase)
                dvals.append(2.0 * h)
    if not cols:
        return np.zeros((g.shape[0], 0), dtype=np.complex128), np.zeros(0)
    return np.ascontiguousarray(np.stack(cols, axis=1)), np.array(dvals)


def make_h2_clusters(nsys: int = 24, seed: int = 2024, ecutwfc: float = 25.0, min_sep: float = 4.0,
                     nmol_range=(2, 24), cell_range=(12.0, 24.0), pdep_range=(16, 64)):
    """A batch of independent H2 clusters of uneven size: for each, a seeded
    molecule count (uniform in ``nmol_range``), cubic cell edge (uniform in
    ``cell_range`` bohr, enlarged by 5 % steps while the molecules cannot be
    placed ``min_sep`` apart) and PDEP count (uniform in ``pdep_range``).
    Returns a list of (System, n_pdep)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nsys):
        nmol = int(rng.integers(nmol_range[0], nmol_range[1] + 1))
        L = float(rng.uniform(*cell_range))
        npdep = int(rng.integers(pdep_range[0], pdep_range[1] + 1))
        s_seed = int(rng.integers(1 << 30))
        while True:
            try:
                s = make_h2_system(nmol=nmol, L=L, ecutwfc=ecutwfc, seed=s_seed, min_sep=min_sep,
                                   name=f"cluster{k}_{nmol}H2_L{L:.2f}")
                break
            except RuntimeError:
                L *= 1.05
        out.append((s, npdep))
    return out
