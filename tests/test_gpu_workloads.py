"""Parity at the workload sizes: the GPU path against the oracle on the
systems the benchmark configurations name (not only on small test cells).

* 64-atom Si (2x2x2 diamond, 72^3 grid, 9,139 plane waves, 128 bands, 256 KB
  projectors): one perturbation through apply_operator at tight tr2 in FP64
  and in the mixed mode, and the FP32 operator alone (FP32 FFT kernels,
  split-BF16 projections; no FP64 refinement involved) on 128 bands.
* 216-atom Si (3x3x3, 100^3, 30,723 plane waves, 864 projectors) with a few
  of its lowest bands as the occupied manifold: lite and tight dfpt.
* 6 H2 (18 bohr, 60^3): the 32-PDEP Davidson against eigenvalues the oracle
  computed on the CPU (tests/golden/make_c1_fixture.py), FP64 and mixed.
* H2 clusters of the uneven batch: four of them concurrently, each against
  the oracle's Davidson.
Each case prints its name before it runs.
"""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SI64 = {"nrep": 2, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}
SI216 = {"nrep": 3, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _smooth_perts(basis, npwx, m, seed):
    rng = np.random.default_rng(seed)
    n = basis.npw
    c = np.zeros((m, npwx), dtype=np.complex128)
    c[:, 1:n] = (rng.uniform(size=(m, n - 1)) * np.exp(2j * np.pi * rng.uniform(size=(m, n - 1)))
                 / (basis.g2_tpiba[1:] + 1.0))
    return c / np.sqrt(np.sum(np.abs(c) ** 2, axis=1, keepdims=True))


@pytest.fixture(scope="module")
def si64(gpu):
    """The 64-atom Si ground state of the benchmark (3 SCF steps, bands on the device)."""
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_si_system
    s = make_si_system(**SI64)
    print(f"case si64: ground state npw={s.basis.npw} nr={s.basis.nr}", flush=True)
    gs = make_ground_state(s, gpu.gpu_h_factory, nbnd=s.nbnd_occ + 4, scf_steps=3, tol=1e-7)
    assert (s.basis.npw, s.basis.nr, gs.nbndval, gs.nkb) == (9139, (72, 72, 72), 128, 256)
    return gs


@pytest.mark.gpu
def test_si64_chi0_fp64_and_mixed_vs_oracle(gpu, si64):
    from oracle.dfpt import apply_operator, sqvc_init
    gs = si64
    sq = sqvc_init(gs.basis, gs.npwx)
    gpu.set_ground_state(gs, sqvc=sq)
    dvg = _smooth_perts(gs.basis, gs.npwx, 1, seed=5)
    tr2 = 1e-14
    print("case si64 chi0: oracle (1 perturbation x 128 bands)", flush=True)
    ref = apply_operator(gs, sq, dvg, tr2)
    print("case si64 chi0: fp64", flush=True)
    gpu.set_precision("fp64")
    out = gpu.apply_operator(1, dvg, tr2)
    print(f"case si64 chi0: fp64 rel err {_rel(out, ref):.2e}", flush=True)
    assert _rel(out, ref) < 1e-9
    print("case si64 chi0: mixed6", flush=True)
    try:
        gpu.set_precision("mixed6")
        out_m = gpu.apply_operator(1, dvg, tr2)
    finally:
        gpu.set_precision("fp64")
    print(f"case si64 chi0: mixed6 rel err {_rel(out_m, ref):.2e}", flush=True)
    assert _rel(out_m, ref) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol", [("mixed6", 1e-5), ("mixed", 1e-4)])
def test_si64_fp32_operator_vs_oracle(gpu, si64, mode, tol):
    """The FP32 operator of the mixed solve on all 128 bands at 72^3: FP32
    stick/plane FFT kernels and the split-BF16 projector products at nq = 384,
    K = 18,278, with no FP64 refinement to hide an error."""
    from oracle.dfpt import sqvc_init
    from oracle.hamiltonian import apply_sternheimerop
    gs = si64
    gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))
    nv = gs.nbndval
    rng = np.random.default_rng(9)
    # the occupied states plus a smooth admixture: the shape of the CG's vectors
    psi = gs.evc[:nv].copy()
    psi[:, :gs.basis.npw] += 0.3 * _smooth_perts(gs.basis, gs.npwx, nv, seed=3)[:, :gs.basis.npw]
    e = np.asarray(gs.et[:nv]) + 0.01 * rng.uniform(size=nv)
    print(f"case si64 fp32 operator {mode}", flush=True)
    ref = apply_sternheimerop(gs, psi, e, gs.alpha_pv)
    try:
        gpu.set_precision(mode)
        n0 = gpu.precision_counters()[0]
        out = gpu.apply_sternheimerop_to_m_wfcs(nv, psi, e, gs.alpha_pv)
        assert gpu.precision_counters()[0] - n0 == nv     # the FP32 operator ran
    finally:
        gpu.set_precision("fp64")
    print(f"case si64 fp32 operator {mode}: rel err {_rel(out, ref):.2e}", flush=True)
    assert _rel(out, ref) < tol


@pytest.mark.gpu
def test_si216_band_subset_vs_oracle(gpu):
    """216 atoms, 100^3, all 864 projectors; the occupied manifold is the
    lowest bands up to the widest gap among the first eight (the full 432
    would take the oracle far too long)."""
    from oracle.dfpt import dfpt, sqvc_init
    from west_amd.groundstate import make_ground_state_device
    from west_amd.systems import make_si_system
    s = make_si_system(**SI216)
    print(f"case si216: ground state npw={s.basis.npw} nr={s.basis.nr}", flush=True)
    gs = make_ground_state_device(s, gpu.gpu_h_factory, nbnd=12, tol=1e-9)
    assert (s.basis.npw, s.basis.nr, gs.nkb) == (30723, (100, 100, 100), 864)
    gaps = np.diff(gs.et[:9])
    nv = int(np.argmax(gaps)) + 1
    gs.nbndval = nv
    gs.occ = np.ones(nv)
    gs.alpha_pv = max(2.0 * (float(gs.et[:nv].max()) - float(gs.et.min())), 0.01)
    print(f"case si216: {nv} occupied bands, gap {gaps[nv - 1]:.4f} Ry", flush=True)
    gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))
    dvg = _smooth_perts(gs.basis, gs.npwx, 2, seed=11)
    print("case si216: lite dfpt", flush=True)
    err = _rel(gpu.dfpt(2, dvg, -1.0), dfpt(gs, dvg, -1.0))
    print(f"case si216: lite rel err {err:.2e}", flush=True)
    assert err < 1e-12
    print("case si216: tight dfpt", flush=True)
    err = _rel(gpu.dfpt(1, dvg[:1], 1e-18), dfpt(gs, dvg[:1], 1e-18))
    print(f"case si216: tight rel err {err:.2e}", flush=True)
    assert err < 1e-9


def _c1():
    spec = importlib.util.spec_from_file_location("make_c1_fixture", ROOT / "tests" / "golden" / "make_c1_fixture.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.gpu
def test_c1_davidson_vs_cpu_fixture(gpu):
    from oracle.dfpt import sqvc_init
    mod = _c1()
    fx = np.load(mod.OUT)
    print("case c1: ground state on the CPU", flush=True)
    gs = mod.c1_ground_state()
    assert (gs.basis.npw, gs.basis.nr) == (6147, (60, 60, 60))
    # the same ground state as the fixture's
    assert np.max(np.abs(gs.et[:gs.nbndval] - fx["et"])) < 1e-9
    assert abs(np.sum(gs.vrs) - fx["vrs_sum"]) < 1e-8 * abs(fx["vrs_sum"])
    gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))
    d = mod.DAVIDSON
    kw = dict(n_pdep_times=d["n_pdep_times"], tr2_dfpt=d["tr2_dfpt"], trev_pdep=d["trev_pdep"],
              trev_pdep_rel=d["trev_pdep_rel"], want_vectors=False)
    print("case c1: davidson fp64", flush=True)
    gpu.set_precision("fp64")
    ev, conv, _, log = gpu.davidson_diago(d["n_pdep_eigen"], **kw)
    print(f"case c1: fp64 ev rel err {np.max(np.abs(ev - fx['ev']) / np.abs(fx['ev'])):.2e}", flush=True)
    assert np.max(np.abs(ev - fx["ev"]) / np.abs(fx["ev"])) < 1e-8
    assert np.array_equal(conv, fx["conv"])
    assert [r["dfpt_dim"] for r in log] == list(fx["dfpt_dim"])
    assert [r["diago_dim"] for r in log] == list(fx["diago_dim"])
    print("case c1: davidson mixed6", flush=True)
    try:
        gpu.set_precision("mixed6")
        ev_m, _, _, _ = gpu.davidson_diago(d["n_pdep_eigen"], **kw)
    finally:
        gpu.set_precision("fp64")
    print(f"case c1: mixed6 ev rel err {np.max(np.abs(ev_m - fx['ev']) / np.abs(fx['ev'])):.2e}", flush=True)
    assert np.max(np.abs(ev_m - fx["ev"]) / np.abs(fx["ev"])) < 1e-6


@pytest.mark.gpu
def test_c4_clusters_concurrent_vs_oracle(gpu):
    """Four clusters of the uneven batch (the four smallest of its 24), with
    8 PDEP each, solved concurrently on their own streams."""
    from oracle.davidson import davidson_diago_gamma
    from oracle.dfpt import sqvc_init
    from oracle.hamiltonian import h_psi
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_h2_clusters
    batch = sorted(make_h2_clusters(), key=lambda t: t[0].basis.npw)[:4]
    kw = dict(n_pdep_times=4, tr2_dfpt=1e-16, trev_pdep=1e-6, trev_pdep_rel=1e-5)
    handles, refs = [], []
    try:
        for s, _ in batch:
            print(f"case c4: {s.name} npw={s.basis.npw} nr={s.basis.nr}", flush=True)
            gs = make_ground_state(s, lambda g: (lambda psi: h_psi(g, psi)), scf_steps=2, tol=1e-9)
            sq = sqvc_init(gs.basis, gs.npwx)
            refs.append(davidson_diago_gamma(gs, sq, 8, **kw)[0])
            h = gpu.system_create()
            gpu.system_select(h)
            gpu.set_ground_state(gs, sqvc=sq)
            handles.append(h)
        gpu.system_select(0)
        print("case c4: concurrent davidson", flush=True)
        res = gpu.davidson_batch([(h, dict(n_pdep_eigen=8, want_vectors=False, **kw)) for h in handles])
        for (ev, conv, _, _), ref in zip(res, refs):
            assert np.all(conv)
            print(f"case c4: ev rel err {np.max(np.abs(ev - ref) / np.abs(ref)):.2e}", flush=True)
            assert np.max(np.abs(ev - ref) / np.abs(ref)) < 1e-8
    finally:
        gpu.system_select(0)
        for h in handles:
            gpu.system_destroy(h)
