"""Ground state for the Wstat path: Kohn-Sham potential and occupied states.

The reference reads the ground state from a QE run (``Tools/init_pw_arrays.f90``);
no QE run is possible here, so the states are computed for the synthetic
systems of :mod:`west_amd.systems`:

* ``vrs`` is the LDA (PZ81) Kohn-Sham potential of a density: either the
  superposition of Gaussian atomic valence densities, or the self-consistent
  density when ``scf_steps > 0`` (Pulay mixing of the density);
* the lowest bands of ``H = -lap + vrs + sum |beta> D <beta|`` are found by a
  preconditioned block Davidson in the real representation of Gamma-point
  wave functions (or by dense diagonalisation for tiny bases).  ``apply_h`` is a caller-supplied
  callable ``psi (m, npwx) -> H psi``; the GPU library provides one.

A Gamma wave function with half-sphere coefficients ``c`` (``c(G=0)`` real) is
mapped to the real vector ``u = [c_0, sqrt2 Re c_{1:}, sqrt2 Im c_{1:}]``, so
that ``<a|b> = 2 Re sum a* b - a_0 b_0`` becomes the Euclidean product.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .gsphere import GammaBasis
from .systems import (System, atomic_density_g, grid_to_real, kb_projectors, ks_potential,
                      local_ionic_potential_g, real_to_grid)


@dataclass
class GroundState:
    basis: GammaBasis
    npwx: int
    vrs: np.ndarray            # (nnr,) real, Ry
    vkb: np.ndarray            # (nkb, npwx) complex
    dvan: np.ndarray           # (nkb,)
    evc: np.ndarray            # (nbnd, npwx) complex, orthonormal
    et: np.ndarray             # (nbnd,) Ry
    nbndval: int
    occ: np.ndarray            # (nbndval,) occupations (1 for insulators, nspin = 1)
    weight: float = 2.0        # k-point weight, sum = 2 for nspin = 1
    alpha_pv: float = 0.0
    l_kinetic_only: bool = False
    meta: dict = field(default_factory=dict)

    @property
    def nkb(self) -> int:
        return int(self.vkb.shape[0])


# ----------------------------------------------------------------------------
# real representation
# ----------------------------------------------------------------------------

def to_real(c: np.ndarray, npw: int, gstart: int) -> np.ndarray:
    c = c[..., :npw]
    s = np.sqrt(2.0)
    if gstart == 2:
        return np.concatenate([c[..., :1].real, s * c[..., 1:].real, s * c[..., 1:].imag], axis=-1)
    return np.concatenate([s * c.real, s * c.imag], axis=-1)


def from_real(u: np.ndarray, npw: int, npwx: int, gstart: int) -> np.ndarray:
    s = 1.0 / np.sqrt(2.0)
    out = np.zeros(u.shape[:-1] + (npwx,), dtype=np.complex128)
    if gstart == 2:
        k = npw - 1
        out[..., 0] = u[..., 0]
        out[..., 1:npw] = s * (u[..., 1:1 + k] + 1j * u[..., 1 + k:1 + 2 * k])
    else:
        out[..., :npw] = s * (u[..., :npw] + 1j * u[..., npw:2 * npw])
    return out


def _orthonormalize(X: np.ndarray, cutoff: float = 1e-14) -> np.ndarray:
    """Rows of X orthonormalised (eigen-decomposition of the Gram matrix of the
    row-normalised block; directions with relative weight below ``cutoff``
    are dropped), followed by a second pass for orthogonality to rounding."""
    if X.shape[0] == 0:
        return X
    n = np.linalg.norm(X, axis=1)
    X = X[n > 0] / n[n > 0, None]
    for _ in range(2):
        S = X @ X.T
        w, V = np.linalg.eigh(0.5 * (S + S.T))
        keep = w > w.max() * cutoff
        X = (V[:, keep] / np.sqrt(w[keep])).T @ X
    return X


def _expand(W: np.ndarray, V: np.ndarray) -> np.ndarray:
    """New orthonormal directions of W outside span(V) (V orthonormal rows)."""
    n = np.linalg.norm(W, axis=1)
    W = W[n > 0] / n[n > 0, None]
    for _ in range(2):
        W = W - (W @ V.T) @ V
    n = np.linalg.norm(W, axis=1)
    W = W[n > 1e-8] / n[n > 1e-8, None]
    W = _orthonormalize(W, cutoff=1e-10)
    W = W - (W @ V.T) @ V
    return W / np.linalg.norm(W, axis=1)[:, None] if W.shape[0] else W


def block_davidson(apply_real, precond, X0: np.ndarray, nev: int, tol: float = 1e-9, maxiter: int = 400,
                   max_blocks: int = 4, verbose: bool = False):
    """Block Davidson with preconditioned residuals and thick restart for the
    ``nev`` lowest eigenpairs of a symmetric operator.  The basis is kept
    orthonormal, so the restart combinations are well conditioned.
    X0: (nblock, dim) start block (nblock >= nev).  Returns (eigvals, X, iters)."""
    V = _orthonormalize(X0)
    HV = apply_real(V)
    nb = X0.shape[0]
    maxdim = max_blocks * nb
    it = 0
    for it in range(maxiter):
        T = V @ HV.T
        w, C = np.linalg.eigh(0.5 * (T + T.T))
        X = C[:, :nb].T @ V
        HX = C[:, :nb].T @ HV
        lam = w[:nb]
        R = HX - lam[:, None] * X
        rn = np.linalg.norm(R, axis=1)
        if verbose and it % 10 == 0:
            print(f"  davidson it {it}: dim {V.shape[0]} max res {rn[:nev].max():.3e}")
        if rn[:nev].max() < tol:
            break
        active = rn > tol * 0.1
        W = precond(R[active], lam[active])
        if V.shape[0] + W.shape[0] > maxdim:
            V, HV = X, HX
        W = _expand(W, V)
        if W.shape[0] == 0:
            break
        V = np.concatenate([V, W])
        HV = np.concatenate([HV, apply_real(W)])
    return lam[:nev], X[:nev], it


def solve_bands(basis: GammaBasis, apply_h, nbnd: int, npwx: int | None = None, tol: float = 1e-9,
                seed: int = 7, x0: np.ndarray | None = None, nextra: int | None = None,
                maxiter: int = 400, verbose: bool = False):
    """Lowest ``nbnd`` eigenpairs of H.  Returns (et, evc (nbnd, npwx))."""
    npw = basis.npw
    npwx = npwx or npw
    gstart = basis.gstart
    dim = 2 * npw - 1 if gstart == 2 else 2 * npw
    g2 = basis.g2kin

    def apply_real(U):
        out = np.empty_like(U)
        bs = 256
        for s in range(0, U.shape[0], bs):
            c = from_real(U[s:s + bs], npw, npwx, gstart)
            out[s:s + bs] = to_real(apply_h(c), npw, gstart)
        return out

    g2r = to_real(g2.astype(np.complex128) * (1 + 1j), npw, gstart) / np.sqrt(2.0)
    g2r[0] = g2[0] if gstart == 2 else g2r[0]
    g2r = np.abs(g2r)

    def precond(R, lam):
        ek = np.maximum(np.abs(lam), 1.0)[:, None]
        return R / (1.0 + g2r[None, :] / ek)

    if dim <= 1500:
        U = np.eye(dim)
        H = apply_real(U)
        H = 0.5 * (H + H.T)
        w, V = np.linalg.eigh(H)
        X = V[:, :nbnd].T
        et = w[:nbnd]
    else:
        nb = nbnd + (nextra if nextra is not None else max(4, nbnd // 8))
        if x0 is None:
            rng = np.random.default_rng(seed)
            X0 = rng.normal(size=(nb, dim)) / (1.0 + g2r[None, :])
        else:
            X0 = to_real(x0, npw, gstart)
            if X0.shape[0] < nb:
                rng = np.random.default_rng(seed)
                X0 = np.concatenate([X0, rng.normal(size=(nb - X0.shape[0], dim)) / (1.0 + g2r[None, :])])
        et, X, _ = block_davidson(apply_real, precond, X0, nbnd, tol=tol, maxiter=maxiter, verbose=verbose)
    X = _orthonormalize(X)
    # final Rayleigh-Ritz so that (et, X) are exact Ritz pairs of the returned basis
    HX = apply_real(X)
    T = X @ HX.T
    w, V = np.linalg.eigh(0.5 * (T + T.T))
    X = V.T @ X
    evc = from_real(X, npw, npwx, gstart)
    return w[:nbnd], evc[:nbnd]


def make_ground_state(system: System, apply_h_factory, nbnd: int | None = None, scf_steps: int = 0,
                      tol: float = 1e-9, mix_beta: float = 0.5, verbose: bool = False,
                      npwx: int | None = None) -> GroundState:
    """Build the ground state of ``system``.

    ``apply_h_factory(gs)`` returns a callable ``psi -> H psi`` for the partial
    ground state ``gs`` (its ``vrs``, ``vkb``, ``dvan`` are set; ``evc`` is a
    placeholder).  The potential starts from the superposition of atomic
    densities; ``scf_steps`` rounds of Pulay mixing refine it."""
    b = system.basis
    npwx = npwx or b.npw
    nbndval = system.nbnd_occ
    nbnd = nbnd or nbndval
    vkb, dvan = kb_projectors(system)
    vkbx = np.zeros((vkb.shape[1], npwx), dtype=np.complex128)
    vkbx[:, :b.npw] = vkb.T
    vion = local_ionic_potential_g(system)
    n_g = atomic_density_g(system)
    vrs = ks_potential(system, n_g, vion)
    gs = GroundState(basis=b, npwx=npwx, vrs=vrs, vkb=vkbx, dvan=dvan,
                     evc=np.zeros((nbnd, npwx), dtype=np.complex128), et=np.zeros(nbnd),
                     nbndval=nbndval, occ=np.ones(nbndval), meta={"system": system.name})
    x0 = None
    hist_in, hist_res = [], []
    res_norm = None
    for it in range(scf_steps + 1):
        gs.vrs = vrs
        et, evc = solve_bands(b, apply_h_factory(gs), nbnd, npwx, tol=tol, x0=x0, verbose=verbose)
        x0 = evc
        if it == scf_steps:
            break
        # output density (nspin = 1, occupation 2 per band)
        rho_r = np.zeros(b.nnr)
        for v in range(nbndval):
            psir = _psi_r(b, evc[v])
            rho_r += 2.0 * psir ** 2 / b.omega
        n_out = real_to_grid(b, rho_r)
        res = n_out - n_g
        res_norm = float(np.linalg.norm(res))
        hist_in.append(n_g.copy())
        hist_res.append(res)
        if len(hist_in) > 5:
            hist_in.pop(0)
            hist_res.pop(0)
        k = len(hist_res)
        if k > 1:
            Rm = np.array(hist_res)
            A = np.real(Rm.conj() @ Rm.T)
            A += 1e-12 * np.trace(A) * np.eye(k)
            ainv1 = np.linalg.solve(A, np.ones(k))
            c = ainv1 / ainv1.sum()
            n_opt = sum(c[i] * hist_in[i] for i in range(k))
            r_opt = sum(c[i] * hist_res[i] for i in range(k))
            n_g = n_opt + mix_beta * r_opt
        else:
            n_g = n_g + mix_beta * res
        if verbose:
            print(f" scf {it}: |res| = {np.linalg.norm(res):.3e}  e_homo = {et[nbndval - 1]:.6f}")
        vrs = ks_potential(system, n_g, vion)
    gs.evc = evc
    gs.et = et
    gs.alpha_pv = max(2.0 * (float(et[:nbndval].max()) - float(et.min())), 0.01)
    gs.meta["scf_steps"] = scf_steps
    # |n_out - n_in| of the last mixing step (density Fourier components, 1/bohr^3)
    gs.meta["scf_density_residual"] = res_norm
    return gs


def _psi_r(basis: GammaBasis, c: np.ndarray) -> np.ndarray:
    """Real-space values of one Gamma wave function (first index fastest)."""
    grid = np.zeros(basis.nnr, dtype=np.complex128)
    n = basis.npw
    grid[basis.nl[:n] - 1] = c[:n]
    grid[basis.nlm[:n] - 1] = np.conj(c[:n])
    return grid_to_real(basis, grid)


# ----------------------------------------------------------------------------
# the same eigensolver on the device (large systems: the dense algebra of the
# block Davidson on (nbnd x 2 npw) blocks moves to the GPU, FP64)
# ----------------------------------------------------------------------------

def _t_orthonormalize(X, cutoff: float = 1e-14):
    import torch
    if X.shape[0] == 0:
        return X
    n = torch.linalg.norm(X, dim=1)
    X = X[n > 0] / n[n > 0, None]
    for _ in range(2):
        S = X @ X.T
        w, V = torch.linalg.eigh(0.5 * (S + S.T))
        keep = w > w.max() * cutoff
        X = (V[:, keep] / torch.sqrt(w[keep])).T @ X
    return X


def _t_expand(W, V):
    import torch
    n = torch.linalg.norm(W, dim=1)
    W = W[n > 0] / n[n > 0, None]
    for _ in range(2):
        W = W - (W @ V.T) @ V
    n = torch.linalg.norm(W, dim=1)
    W = W[n > 1e-8] / n[n > 1e-8, None]
    W = _t_orthonormalize(W, cutoff=1e-10)
    W = W - (W @ V.T) @ V
    return W / torch.linalg.norm(W, dim=1)[:, None] if W.shape[0] else W


def solve_bands_device(basis: GammaBasis, apply_h_dev, nbnd: int, npwx: int | None = None, tol: float = 1e-9,
                       seed: int = 7, maxiter: int = 400, max_blocks: int = 3, verbose: bool = False,
                       device: str = "cuda:0"):
    """:func:`solve_bands` with the block Davidson on the device.

    ``apply_h_dev`` maps a device complex128 tensor (m, npwx) to H psi on the
    device.  Returns (et, evc) as numpy arrays."""
    import torch
    npw = basis.npw
    npwx = npwx or npw
    gstart = basis.gstart
    dim = 2 * npw - 1 if gstart == 2 else 2 * npw
    s2 = np.sqrt(2.0)
    k = npw - 1

    def to_real_t(c):
        c = c[:, :npw]
        if gstart == 2:
            return torch.cat([c[:, :1].real, s2 * c[:, 1:].real, s2 * c[:, 1:].imag], dim=1)
        return torch.cat([s2 * c.real, s2 * c.imag], dim=1)

    def from_real_t(u):
        out = torch.zeros((u.shape[0], npwx), dtype=torch.complex128, device=u.device)
        if gstart == 2:
            out[:, 0] = u[:, 0].to(torch.complex128)
            out[:, 1:npw] = (u[:, 1:1 + k] + 1j * u[:, 1 + k:1 + 2 * k]) / s2
        else:
            out[:, :npw] = (u[:, :npw] + 1j * u[:, npw:2 * npw]) / s2
        return out

    def apply_real(U):
        out = torch.empty_like(U)
        bs = 512
        for s in range(0, U.shape[0], bs):
            c = from_real_t(U[s:s + bs]).contiguous()
            if c.is_cuda:
                torch.cuda.synchronize()
            out[s:s + bs] = to_real_t(apply_h_dev(c))
        return out

    g2 = torch.from_numpy(np.asarray(basis.g2kin, dtype=np.float64)).to(device)
    g2r = torch.cat([g2[:1], g2[1:], g2[1:]]) if gstart == 2 else torch.cat([g2, g2])

    def precond(R, lam):
        ek = torch.clamp(torch.abs(lam), min=1.0)[:, None]
        return R / (1.0 + g2r[None, :] / ek)

    nb = nbnd + max(4, nbnd // 8)
    rng = np.random.default_rng(seed)
    X0 = torch.from_numpy(rng.normal(size=(nb, dim))).to(device) / (1.0 + g2r[None, :])
    V = _t_orthonormalize(X0)
    HV = apply_real(V)
    maxdim = max_blocks * nb
    for it in range(maxiter):
        T = V @ HV.T
        w, C = torch.linalg.eigh(0.5 * (T + T.T))
        X = C[:, :nb].T @ V
        HX = C[:, :nb].T @ HV
        lam = w[:nb]
        R = HX - lam[:, None] * X
        rn = torch.linalg.norm(R, dim=1)
        rmax = float(rn[:nbnd].max())
        if verbose and it % 5 == 0:
            print(f"  davidson (device) it {it}: dim {V.shape[0]} max res {rmax:.3e}", flush=True)
        if rmax < tol:
            break
        active = rn > tol * 0.1
        W = precond(R[active], lam[active])
        if V.shape[0] + W.shape[0] > maxdim:
            V, HV = X, HX
        W = _t_expand(W, V)
        if W.shape[0] == 0:
            break
        V = torch.cat([V, W])
        HV = torch.cat([HV, apply_real(W)])
    X = _t_orthonormalize(X[:nbnd])
    HX = apply_real(X)
    T = X @ HX.T
    w, Vv = torch.linalg.eigh(0.5 * (T + T.T))
    X = Vv.T @ X
    evc = from_real_t(X).cpu().numpy()
    return w[:nbnd].cpu().numpy(), evc[:nbnd]


def make_ground_state_device(system: System, apply_h_factory, nbnd: int | None = None, tol: float = 1e-6,
                             verbose: bool = False, npwx: int | None = None) -> GroundState:
    """Non-self-consistent ground state (LDA potential of the superposed atomic
    densities) with the bands found on the device: for the largest workloads,
    where the host eigensolver of :func:`make_ground_state` is too slow.
    ``apply_h_factory(gs)`` must return a callable accepting device tensors."""
    b = system.basis
    npwx = npwx or b.npw
    nbndval = system.nbnd_occ
    nbnd = nbnd or nbndval
    vkb, dvan = kb_projectors(system)
    vkbx = np.zeros((vkb.shape[1], npwx), dtype=np.complex128)
    vkbx[:, :b.npw] = vkb.T
    vrs = ks_potential(system, atomic_density_g(system), local_ionic_potential_g(system))
    gs = GroundState(basis=b, npwx=npwx, vrs=vrs, vkb=vkbx, dvan=dvan,
                     evc=np.zeros((nbnd, npwx), dtype=np.complex128), et=np.zeros(nbnd),
                     nbndval=nbndval, occ=np.ones(nbndval), meta={"system": system.name, "scf_steps": 0})
    et, evc = solve_bands_device(b, apply_h_factory(gs), nbnd, npwx, tol=tol, verbose=verbose)
    gs.evc = evc
    gs.et = et
    gs.alpha_pv = max(2.0 * (float(et[:nbndval].max()) - float(et.min())), 0.01)
    return gs
