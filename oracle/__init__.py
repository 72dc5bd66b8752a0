"""FP64 CPU oracle for the Wstat dielectric-response hot path.

TEST INFRASTRUCTURE ONLY.  This package is a CPU restatement (NumPy/SciPy) of
the reference algorithm, following the reference files cited in each
function.  It is imported by ``tests/``, by ``__graft_entry__.smoke()`` and by
the ``cpu_baseline`` leg of ``bench.py`` -- as the checker and as the timed CPU
baseline -- and by nothing in the product path (``west_amd``), which runs on
the HIP library and fails when that library is missing.

Parity status per function (what pins it; see ``tests/test_oracle_*.py``):

* ``fft_at_gamma``: textbook DFT by brute force on tiny grids, Hermitian
  symmetry, pack/unpack round trip, Parseval.
* ``glbrak_gamma``: equals the full-sphere complex inner product.
* ``h_psi`` / ``k_psi``: dense plane-wave Hamiltonian assembled by brute force
  (``|G|^2 delta + V(G-G') + sum beta D beta^*``), Hermiticity.
* ``apply_alpha_pc/pv``: idempotence, ``P_c psi_v = 0``.
* ``linsolve_sternheimer_m_wfcts``: dense ``numpy.linalg.solve`` of the same
  operator on tiny cells.
* ``dfpt`` / ``apply_operator``: sum-over-states chi0 from full
  diagonalisation on tiny cells; symmetry and negative semidefiniteness.
* ``davidson_diago_gamma``: ``numpy.linalg.eigh`` of the explicitly assembled
  chi0~ matrix on tiny cells; iteration-control semantics restated line by line.
* G-sphere counts: the reference test-suite fixtures (npw/ngm/FFT grid).
* ``kpoint`` (chi0(q), the k-point operator and CG, ``davidson_diago_k``):
  supercell folding -- the Gamma chi0 of a doubled cell equals chi0(q) of the
  primitive cell on the matching k grid, sector by sector, and the union of
  the q-point Davidson spectra equals the supercell's; dense H_k; Hermiticity
  and negativity of chi0~(q).
"""
