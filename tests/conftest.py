import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


# ---------------------------------------------------------------------------
# small ground states, computed on the CPU with the oracle Hamiltonian
# ---------------------------------------------------------------------------
_CACHE = {}


def _gs(key):
    if key in _CACHE:
        return _CACHE[key]
    from oracle.hamiltonian import h_psi
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_h2_system, make_si_system
    factory = lambda g: (lambda psi: h_psi(g, psi))  # noqa: E731
    if key == "h2":        # 1 H2, odd band count (1)
        s = make_h2_system(nmol=1, L=8.0, ecutwfc=10.0, seed=3)
        gs = make_ground_state(s, factory, scf_steps=4)
    elif key == "3h2":     # 3 H2, odd band count (3), 12 bohr
        s = make_h2_system(nmol=3, L=10.0, ecutwfc=8.0, seed=5, min_sep=3.0)
        gs = make_ground_state(s, factory, scf_steps=3)
    elif key == "si8":     # 8 Si, KB projectors, 16 bands
        s = make_si_system(nrep=1, ecutwfc=8.0, seed=2)
        gs = make_ground_state(s, factory, scf_steps=3, nbnd=20)
    else:
        raise KeyError(key)
    _CACHE[key] = gs
    return gs


@pytest.fixture(scope="session")
def gs_h2():
    return _gs("h2")


@pytest.fixture(scope="session")
def gs_3h2():
    return _gs("3h2")


@pytest.fixture(scope="session")
def gs_si8():
    return _gs("si8")


def random_wfc(gs, m, seed):
    rng = np.random.default_rng(seed)
    n = gs.basis.npw
    a = np.zeros((m, gs.npwx), dtype=np.complex128)
    a[:, :n] = (rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))) / (1.0 + gs.basis.g2kin)[None, :]
    a[:, 0] = a[:, 0].real
    return a


def dense_states(gs, vrs=None, nb=None):
    """All (or the lowest nb) eigenpairs of H(vrs) on the sphere, by dense
    diagonalisation of H in the real representation of Gamma wave functions."""
    from oracle.hamiltonian import h_psi
    from west_amd.groundstate import from_real, to_real
    b = gs.basis
    n = b.npw
    dim = 2 * n - 1 if b.gstart == 2 else 2 * n

    class G:
        pass

    g = G()
    for k in ("basis", "npwx", "vkb", "dvan", "l_kinetic_only"):
        setattr(g, k, getattr(gs, k))
    g.vrs = gs.vrs if vrs is None else vrs
    cols = from_real(np.eye(dim), n, gs.npwx, b.gstart)
    H = to_real(h_psi(g, cols), n, b.gstart)
    w, U = np.linalg.eigh(0.5 * (H + H.T))
    nb = dim if nb is None else nb
    return w[:nb], from_real(U[:, :nb].T, n, gs.npwx, b.gstart)


@pytest.fixture(scope="session")
def gpu():
    """Initialises the library on device 0 (GPU tests only)."""
    from west_amd import wstat
    wstat.init(0)
    return wstat
