"""Writes tests/golden/c1_6h2_davidson.npz: the oracle's Wstat Davidson on the
6-H2 workload, computed on the CPU from a ground state also computed on the
CPU (the oracle's Hamiltonian), for the GPU parity test of that workload.

Workload: 6 H2 molecules (bond 1.40 bohr, seeded random orientations and
centres at least 5 bohr apart) in an 18 bohr cube, Gamma, 25 Ry, HGH H;
ground state by 3 Pulay-mixed SCF steps (bands to tol 1e-9); Davidson for 32
PDEP eigenpairs with n_pdep_times 4 and the reference's defaults tr2_dfpt
1e-12, trev_pdep 1e-3, trev_pdep_rel 0.1 (Pytools/west_fetch_input.py).

The fixture holds the eigenvalues, convergence flags and the dims trajectory
of every step, plus the ground state's band energies and a checksum of its
potential so that the test can confirm it rebuilt the same ground state.
Run: python tests/golden/make_c1_fixture.py  (minutes on a few cores).
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

OUT = Path(__file__).with_name("c1_6h2_davidson.npz")
DAVIDSON = {"n_pdep_eigen": 32, "n_pdep_times": 4, "tr2_dfpt": 1e-12, "trev_pdep": 1e-3, "trev_pdep_rel": 0.1}


def c1_ground_state():
    """The 6-H2 ground state, on the CPU with the oracle's Hamiltonian."""
    from oracle.hamiltonian import h_psi
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_h2_system
    s = make_h2_system(nmol=6, L=18.0, ecutwfc=25.0, seed=2024)
    return make_ground_state(s, lambda g: (lambda psi: h_psi(g, psi)), nbnd=s.nbnd_occ + 4, scf_steps=3, tol=1e-9)


def main():
    from oracle.davidson import davidson_diago_gamma
    from oracle.dfpt import sqvc_init
    t = time.perf_counter()
    gs = c1_ground_state()
    print(f"ground state: {time.perf_counter() - t:.1f}s, et = {gs.et[:gs.nbndval]}", flush=True)
    sq = sqvc_init(gs.basis, gs.npwx)
    log = []
    t = time.perf_counter()
    ev, conv, _ = davidson_diago_gamma(gs, sq, DAVIDSON["n_pdep_eigen"], n_pdep_times=DAVIDSON["n_pdep_times"],
                                       tr2_dfpt=DAVIDSON["tr2_dfpt"], trev_pdep=DAVIDSON["trev_pdep"],
                                       trev_pdep_rel=DAVIDSON["trev_pdep_rel"], log=log)
    print(f"davidson: {time.perf_counter() - t:.1f}s, {len(log)} steps, ev[:4] = {ev[:4]}", flush=True)
    np.savez(OUT, ev=ev, conv=conv, et=np.asarray(gs.et[:gs.nbndval]), vrs_sum=np.sum(gs.vrs),
             vrs_sq=np.sum(gs.vrs ** 2), dav_iter=[r["dav_iter"] for r in log],
             dfpt_dim=[r["dfpt_dim"] for r in log], diago_dim=[r["diago_dim"] for r in log],
             notcnv=[r["notcnv"] for r in log], ev_steps=np.array([r["ev"] for r in log]))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
