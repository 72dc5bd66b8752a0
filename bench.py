"""chi0 . dv throughput of the Wstat hot path on one MI355X.

A step is one chi0 application (apply_operator: dfpt with the Sternheimer CG
to tr2_dfpt, Delta-n accumulation and the Coulomb factor) to one batch of
perturbations of the 64-atom Si workload: diamond Si, a = 10.26 bohr, 2x2x2
cubic cell, seeded Gaussian displacements sigma = 0.05 bohr, GTH/HGH Si with
1 s + 1 p Kleinman-Bylander projector, ecutwfc 25 Ry, Gamma, 128 occupied
bands, 72^3 FFT grid, 64 perturbations per chi0 batch (8192 pert x band
pairs).  Inputs (ground state, perturbations) are resident on the device
before the timed region.  Prints one JSON line.

The CPU baseline is the FP64 oracle (NumPy/SciPy restatement of the same
algorithm) on one perturbation of the same batch, on the host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# Measured HBM bytes per transform item (band pair) of the plane kernels on the
# si64 workload: rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a streaming kernel
# of known size (profiles/plane_vloc_si64_hbm_traffic.md).
TRAFFIC_PER_ITEM = {"si64": {"plane_vloc": 1028065 + 973824, "plane_vloc32": 491706 + 487296}}
FP64_PEAK_TFLOPS = 78.6

WORKLOADS = {
    # name: (builder kwargs, perturbations per batch)
    "si64": ({"nrep": 2, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}, 64),
    "si8": ({"nrep": 1, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}, 16),
    # largest single-GPU workload: 3x3x3 cell, 216 atoms, 432 bands, 100^3, 128 perturbations per batch
    "si216": ({"nrep": 3, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}, 128),
    # batch of 24 uneven H2 clusters (2-24 molecules, 12-24 bohr cells, 16-64 PDEP each): Davidson wall time
    "c4": (None, None),
    # Wstat Davidson for 256 PDEP on the si64 ground state (nvecx 1024), FP64 and mixed
    "si64dav": ({"nrep": 2, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}, 64),
    # chi0 at a q point: H2 molecular crystal on a 3x3x3 k grid (the default line's "qpoint" field alone)
    "qpoint": (None, None),
    # one FP64 Davidson of the 6-H2 system at the converged settings, warm-up then timed (for traces)
    "c1dav": (None, None),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def start_vectors(basis, npwx, m, seed):
    """Smooth random perturbations (amplitude / (|G|^2 + 1), as the Davidson
    start vectors), orthonormal in the Gamma inner product."""
    from west_amd.groundstate import _orthonormalize, from_real, to_real
    rng = np.random.default_rng(seed)
    n = basis.npw
    c = np.zeros((m, npwx), dtype=np.complex128)
    amp = rng.uniform(size=(m, n - 1))
    ph = rng.uniform(size=(m, n - 1)) * 2 * np.pi
    c[:, 1:n] = amp * np.exp(1j * ph) / (basis.g2_tpiba[1:] + 1.0)
    u = _orthonormalize(to_real(c, n, 2))
    return from_real(u, n, npwx, 2)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="si64", choices=sorted(WORKLOADS))
    ap.add_argument("--tr2", type=float, default=1e-16)
    ap.add_argument("--scf", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-bands", type=int, default=48)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"],
                    help="fp64: the FP64 solve; mixed: FP32 CG with FP64 refinement; both: the FP64 line "
                         "with the mixed-precision run of the same batch beside it")
    ap.add_argument("--mixed-mode", default="mixed6", choices=["mixed", "mixed6"],
                    help="mixed-precision variant: split-BF16 projector GEMMs with 3 (mixed) or 6 (mixed6) products")
    ap.add_argument("--refine-delta", type=float, default=0.0,
                    help="mixed precision: FP64 residual recomputed when rho has dropped by this factor "
                         "(0: the library default)")
    ap.add_argument("--hbm-calibration", action="store_true",
                    help="run a plain streaming kernel of known size first (torch.neg over 1 GiB of float32: "
                         "1 GiB read, 1 GiB written) to scale HBM byte counters of a profiler run")
    ap.add_argument("--no-davidson", action="store_true",
                    help="skip the Wstat Davidson wall-time runs (6 H2, 32 PDEP; Si64, 256 PDEP)")
    ap.add_argument("--no-lanczos", action="store_true",
                    help="skip the Wfreq Lanczos measurement (128 chains of 30 steps on the si64 ground state)")
    ap.add_argument("--no-qpoint", action="store_true",
                    help="skip the q-point chi0 measurement (H2 molecular crystal on a 3x3x3 k grid)")
    ap.add_argument("--gs-cache", default=None,
                    help="npz file: load the ground state from it if present, else write it there "
                         "(lets a profiled run skip the ground-state kernels)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from west_amd import wstat
    from west_amd.coulomb import sqvc_init
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_si_system

    if args.hbm_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.neg(a, out=b)
        torch.cuda.synchronize()
        del a, b
        log("[bench] calibration kernel: neg over 2^28 float32 (1 GiB read + 1 GiB written)")

    if args.workload == "c4":
        _batch_c4(wstat, torch, args)
        return
    if args.workload == "c1dav":
        wstat.init(0)
        print(json.dumps(_davidson_c1_fp64(wstat, torch, args)), flush=True)
        return
    if args.workload == "qpoint":
        wstat.init(0)
        print(json.dumps({"metric": "chi0(q).dv applies/s (pert x band x k), Wstat dfpt at q",
                          "qpoint": _qpoint_measure(wstat, torch, args, with_cpu=not args.no_cpu_baseline)}),
              flush=True)
        return

    kw, npert = WORKLOADS[args.workload]
    t_setup = time.perf_counter()
    system = make_si_system(**kw)
    b = system.basis
    log(f"[bench] {system.name}: npw={b.npw} nr={b.nr} nbndval={system.nbnd_occ} nkb={system.nkb}")
    wstat.init(0)
    gs = None
    if args.gs_cache and Path(args.gs_cache).exists():
        gs = _load_gs(system, args.gs_cache)
    if gs is None:
        if b.npw > 20000:
            # largest workload: non-self-consistent potential, bands by the device eigensolver
            from west_amd.groundstate import make_ground_state_device
            gs = make_ground_state_device(system, wstat.gpu_h_factory, nbnd=system.nbnd_occ + 4, tol=1e-5,
                                          verbose=True)
        else:
            gs = make_ground_state(system, wstat.gpu_h_factory, nbnd=system.nbnd_occ + 4, scf_steps=args.scf,
                                   tol=1e-7)
        if args.gs_cache:
            _save_gs(gs, args.gs_cache)
    log(f"[bench] ground state in {time.perf_counter() - t_setup:.1f}s; "
        f"gap {gs.et[gs.nbndval] - gs.et[gs.nbndval - 1] if len(gs.et) > gs.nbndval else float('nan'):.4f} Ry")
    sq = sqvc_init(b, gs.npwx)
    wstat.set_ground_state(gs, sqvc=sq)
    if args.workload == "si64dav":
        _davidson_si64(wstat, torch, args, system, gs)
        return
    dvg_h = start_vectors(b, gs.npwx, npert, seed=7)
    dev = torch.device("cuda:0")
    dvg = torch.from_numpy(dvg_h).to(dev)
    dng = torch.zeros_like(dvg)
    nv = gs.nbndval
    pairs = npert * nv

    head = _run_mode(wstat, torch, args, args.mixed_mode if args.precision == "mixed" else "fp64", npert, dvg, dng)
    value, ms_step = head["value"], head["ms_per_step"]
    stern_per_step, iters_per_step, kernels = head["stern_per_step"], head["iters_per_step"], head["kernels"]

    # ---- roofline of the dominant kernel and the end-to-end figure
    n1, n2, n3 = b.nr
    model = _byte_model(b, gs, npert, head)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    roof = _roofline(dom, kernels[dom], model, args.steps, args.workload)
    e2e = _end_to_end(b, nv, value, head)
    kbytes = _kernel_bytes_end_to_end(model, ms_step)

    # ---- mixed precision beside the FP64 line: same batch, same tr2
    mixed = None
    if args.precision == "both":
        dng64 = dng.clone()
        dng32 = torch.zeros_like(dng)
        mx = _run_mode(wstat, torch, args, args.mixed_mode, npert, dvg, dng32)
        mmodel = _byte_model(b, gs, npert, mx)
        mdom = max(mx["kernels"], key=lambda k: mx["kernels"][k]["ms_per_step"])
        l2 = float(torch.linalg.norm(dng32 - dng64) / torch.linalg.norm(dng64))
        mixed = {"value": round(mx["value"], 3), "unit": "pairs/s", "ms_per_step": round(mx["ms_per_step"], 3),
                 "speedup_vs_fp64": round(mx["value"] / value, 3),
                 "dtype": "f32 FFT/CG + split-bf16 MFMA projections (f32 accumulate) + f64 refinement",
                 "mode": args.mixed_mode, "bf16_products": 3 if args.mixed_mode == "mixed" else 6,
                 "refine_delta": args.refine_delta or "library default", "dng_l2_rel_err_vs_fp64": l2,
                 "sternheimer_applications_per_step": mx["stern_per_step"],
                 "fp32_applications_per_step": mx["fp32_per_step"],
                 "fp64_refinements_per_step": mx["refine_per_step"],
                 "fp64_fallback_bands_per_step": mx["fallback_per_step"],
                 "cg_iterations_per_step": mx["iters_per_step"],
                 "kernels_ms_per_step": {k: round(v["ms_per_step"], 3) for k, v in mx["kernels"].items()},
                 "roofline": _roofline(mdom, mx["kernels"][mdom], mmodel, args.steps, args.workload),
                 "kernel_bytes_end_to_end": _kernel_bytes_end_to_end(mmodel, mx["ms_per_step"]),
                 "chi0_model_end_to_end": _end_to_end(b, nv, mx["value"], mx)}
        log(f"[bench] mixed: {mx['value']:.1f} pairs/s ({mixed['speedup_vs_fp64']}x), dng L2 rel err {l2:.2e}")

    # ---- CPU baseline: the FP64 oracle on a bounded sample of the same batch
    cpu = None
    if not args.no_cpu_baseline:
        import oracle.fft_at_gamma as ofa
        from oracle.dfpt import apply_operator as apply_operator_ref
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ofa.WORKERS = threads
        gs.l_kinetic_only = False
        nb = min(nv, max(2, args.cpu_bands))
        t = time.perf_counter()
        apply_operator_ref(gs, sq, dvg_h[:1], args.tr2, bands=np.arange(nb))
        tc = time.perf_counter() - t
        cpu = {"value": round(nb / tc, 4), "unit": "pairs/s", "cores": threads, "kind": "port",
               "cpu_model": _cpu_model(),
               "sample": f"perturbation 1 of the batch x occupied bands 1..{nb} (of {nv}), tr2={args.tr2:g}, "
                         f"FP64 oracle (SciPy FFT on {threads} threads, OpenBLAS GEMM), {tc:.1f}s"}

    # ---- Wstat Davidson wall time, 256 PDEP of this Si64 ground state, FP64 and mixed
    dav64 = None
    if not args.no_davidson and args.workload == "si64":
        wstat.set_ground_state(gs, sqvc=sq)
        dav64 = _davidson_si64_measure(wstat, torch, args)

    # ---- Wstat Davidson wall time (6 H2 molecules, 32 PDEP, tr2_dfpt 1e-12), FP64 and mixed,
    # and the same Davidson run by the FP64 oracle on the host cores
    dav = None
    if not args.no_davidson:
        dav, h2 = _davidson_h2(wstat, torch, args)
        if cpu is not None:
            cpu["davidson_6h2"] = _davidson_h2_cpu(h2, dav)

    # ---- Wfreq's Lanczos chains on the same ground state, FP64, with the oracle beside it
    lz = None
    if not args.no_lanczos and args.workload == "si64":
        wstat.set_ground_state(gs, sqvc=sq)
        lz = _lanczos_measure(wstat, torch, args, gs, with_cpu=not args.no_cpu_baseline)

    # ---- chi0 at a q point (k-point ground state), FP64, with the oracle beside it
    qpt = None
    if not args.no_qpoint:
        qpt = _qpoint_measure(wstat, torch, args, with_cpu=not args.no_cpu_baseline)

    rec = {
        "metric": "chi0.dv applies/s (pert x band), Wstat dfpt at tr2_dfpt",
        "value": round(value, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != "mixed" else
                 f"f32+bf16x{3 if args.mixed_mode == 'mixed' else 6}, f64 refinement",
        "data": "synthetic (seeded Si supercell ground state computed on the device; seeded smooth perturbations)",
        "config": {"workload": f"{args.workload}: {system.name}, npw={b.npw}, fft={n1}x{n2}x{n3}, "
                               f"nbndval={nv}, nkb={system.nkb}, {npert} perturbations/batch",
                   "pairs_per_step": pairs, "tr2_dfpt": args.tr2,
                   "ground_state": {"scf_steps": gs.meta.get("scf_steps"),
                                    "scf_density_residual": gs.meta.get("scf_density_residual"),
                                    "gap_ry": round(float(gs.et[nv] - gs.et[nv - 1]), 5) if len(gs.et) > nv else None},
                   "sternheimer_applications_per_step": stern_per_step, "cg_iterations_per_step": iters_per_step},
        "kernels_ms_per_step": {k: round(v["ms_per_step"], 3) for k, v in kernels.items()},
        "roofline": roof,
        "kernel_bytes_end_to_end": kbytes,
        "chi0_model_end_to_end": e2e,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "davidson": dav,
        "davidson_si64": dav64,
        "qpoint": qpt,
        "lanczos": lz,
    }
    print(json.dumps(rec), flush=True)


def _lanczos_measure(wstat, torch, args, gs, with_cpu=True, nrhs=128, nlsteps=30, ncpu=2):
    """Wfreq's chi(i omega) Lanczos chains (solve_deflated_lanczos_w_full_ortho,
    n_lanczos = 30 as WEST's default) for 128 right-hand sides on the
    registered ground state, all occupied bands deflated: smooth seeded
    vectors projected on the conduction space, the shape of the
    P_c [psi_v dv_i] of solve_wfreq.  A unit is one Lanczos step of one chain
    (an application of H, P_c, the recurrence and the re-orthogonalisation
    against every earlier vector of the chain).  The first ``ncpu`` chains
    are also run by the FP64 oracle on the host cores."""
    from oracle.hamiltonian import apply_alpha_pc
    from oracle.lanczos import solve_deflated_lanczos_w_full_ortho as lanczos_ref
    from west_amd import wfreq
    b = gs.basis
    rng = np.random.default_rng(19)
    v = np.zeros((nrhs, gs.npwx), complex)
    v[:, :b.npw] = (rng.normal(size=(nrhs, b.npw)) + 1j * rng.normal(size=(nrhs, b.npw))) / (1.0 + b.g2kin)[None, :]
    v[:, 0] = v[:, 0].real
    rhs = apply_alpha_pc(gs, v, 1.0)
    wfreq.solve_deflated_lanczos_w_full_ortho(gs.nbndval, rhs[:8], 4, want_basis=False)   # warm-up
    torch.cuda.synchronize()
    wstat.profile(True)
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        a, bt, _, bn = wfreq.solve_deflated_lanczos_w_full_ortho(gs.nbndval, rhs, nlsteps, want_basis=False)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    kern = {n: round(wstat.profile_query(n)[0] / reps, 3) for n in ("plane_vloc", "zinv", "zfwd", "gemm_bra",
                                                                     "gemm_ket")}
    wstat.profile(False)
    rec = {"workload": f"Wfreq Lanczos on the si64 ground state: {nrhs} chains x {nlsteps} steps, "
                       f"{gs.nbndval} bands deflated, npw={b.npw}",
           "value": round(nrhs * nlsteps / dt, 1), "unit": "chain steps/s", "wall_s": round(dt, 4),
           "kernels_ms": kern, "timed_runs": reps}
    if with_cpu:
        import oracle.fft_at_gamma as ofa
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ofa.WORKERS = threads
        tc = time.perf_counter()
        ar, btr, _, bnr = lanczos_ref(gs, gs.nbndval, rhs[:ncpu], nlsteps)
        tc = time.perf_counter() - tc
        rec["alpha_max_rel_err_vs_oracle"] = float(np.abs(a[:ncpu] - ar).max() / np.abs(ar).max())
        rec["beta_max_rel_err_vs_oracle"] = float(np.abs(bt[:ncpu] - btr).max() / np.abs(btr).max())
        rec["cpu_baseline"] = {"value": round(ncpu * nlsteps / tc, 3), "unit": "chain steps/s", "cores": threads,
                               "kind": "port", "sample": f"chains 1-{ncpu} of the batch, {nlsteps} steps, FP64 oracle, "
                                                         f"{tc:.1f}s"}
    log(f"[bench] lanczos: {rec['value']} chain steps/s ({dt:.3f}s for {nrhs}x{nlsteps})")
    return rec


def _qpoint_measure(wstat, torch, args, with_cpu=True, nk=(3, 3, 3), npert=16, ecut=25.0):
    """chi0(q) dv away from Gamma: the H2 molecular crystal (2 H2, s projectors)
    at ecutwfc 25 Ry on an unshifted 3x3x3 k grid (27 k points, every one
    explicit), one batch of 16 perturbations at the last q point
    (2/3, 2/3, 2/3), tr2 = args.tr2.  A pair is (perturbation, band, k point).
    Runs on a system of its own.  The grid-FFT figure counts 3 passes x
    (read + write) of every transformed grid column."""
    from oracle import kpoint as K
    from west_amd.kpoint import make_h2_crystal, make_k_ground_state
    t0 = time.perf_counter()
    kgs = make_k_ground_state(make_h2_crystal(ecut), nk, nbnd=4)
    iq = len(kgs.qpts) - 1
    qp = kgs.qpts[iq]
    qb = qp.basis
    nb, nks = kgs.nbndval, kgs.nks
    nnr = int(np.prod(kgs.nr))
    log(f"[bench] qpoint: {kgs.meta['system']} k grid {nk}, npwq={qb.npw}, fft={kgs.nr}, "
        f"ground state {time.perf_counter() - t0:.1f}s, gap {kgs.meta.get('gap_ry', float('nan')):.3f} Ry")
    h = wstat.system_create()
    wstat.system_select(h)
    try:
        wstat.set_k_ground_state(kgs)
        rng = np.random.default_rng(17)
        dvg_h = np.zeros((npert, kgs.npwqx), complex)
        dvg_h[:, :qb.npw] = (rng.normal(size=(npert, qb.npw)) + 1j * rng.normal(size=(npert, qb.npw))) / \
            (1.0 + qb.g2kin)[None, :]
        dvg = torch.from_numpy(dvg_h).to("cuda:0")
        dng = torch.zeros_like(dvg)
        wstat.dfpt(npert, dvg, args.tr2, iq=iq + 1, dng=dng)          # warm-up (kernel variants, workspaces)
        torch.cuda.synchronize()
        # wall time without the per-call timing events (its launches are a few
        # microseconds each), then the kernel times from a profiled pass
        cols0 = wstat.counters()[0]
        steps = 3
        t = time.perf_counter()
        for _ in range(steps):
            wstat.dfpt(npert, dvg, args.tr2, iq=iq + 1, dng=dng)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        op_cols = (wstat.counters()[0] - cols0) / steps
        wstat.profile(True)
        for _ in range(steps):
            wstat.dfpt(npert, dvg, args.tr2, iq=iq + 1, dng=dng)
        torch.cuda.synchronize()
        kern = {n: wstat.profile_query(n) for n in ("kfft", "kgrid_lds", "kvloc_lds", "cg_iter_lds", "gemm_bra",
                                                    "gemm_ket")}
        wstat.profile(False)
        pairs = npert * nb * nks
        # grid columns through the full-grid transforms in HBM (three passes,
        # read + write each): the perturbations, and unless every k point
        # runs at once on LDS grids (kgrid_lds: states at k - q, right-hand
        # sides, densities) those per k point; the operator's own columns take
        # the LDS-resident kernels when the grid allows it (the first
        # application as kvloc_lds, the CG iterations as cg_iter_lds) and are
        # then not among them either
        lds_op = kern["kvloc_lds"][1] > 0 or kern["cg_iter_lds"][1] > 0
        lds_grids = kern["kgrid_lds"][1] > 0
        fft_cols = npert + (0 if lds_grids else nks * (nb + 2 * npert * nb + npert)) + (0 if lds_op else 2 * op_cols)
        fft_bytes = fft_cols * 6 * 16 * nnr
        kfft_ms = kern["kfft"][0] / steps
        rec = {"workload": f"H2 molecular crystal (2 H2, 6.5x6.0x5.5 bohr triclinic, s projectors), {ecut:g} Ry, "
                           f"k grid {nk[0]}x{nk[1]}x{nk[2]} ({nks} k points), q = {qp.cryst.round(4).tolist()}, "
                           f"npwq={qb.npw}, fft={kgs.nr[0]}x{kgs.nr[1]}x{kgs.nr[2]}, {nb} bands, "
                           f"{npert} perturbations, tr2={args.tr2:g}",
               "value": round(pairs / dt, 2), "unit": "pairs/s (pert x band x k)", "ms_per_batch": round(dt * 1e3, 2),
               "operator_columns_per_batch": op_cols,
               "kernels_ms_per_batch": {n: round(v[0] / steps, 3) for n, v in kern.items()},
               "grid_fft": {"bound": "hbm", "achieved": round(fft_bytes / (kfft_ms * 1e-3) / 1e9, 1) if kfft_ms > 0 else None,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(fft_bytes / (kfft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if kfft_ms > 0 else None,
                            "bytes_per_batch": int(fft_bytes)}}
        out = dng.cpu().numpy()
        ref = K.dfpt_q(kgs, iq, dvg_h[:1], args.tr2)
        rec["rel_err_vs_oracle_pert1"] = float(np.abs(out[:1] - ref).max() / np.abs(ref).max())
        if with_cpu:
            import oracle.fft_at_gamma as ofa
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            ofa.WORKERS = threads
            tc = time.perf_counter()
            K.dfpt_q(kgs, iq, dvg_h[:4], args.tr2)
            tc = time.perf_counter() - tc
            rec["cpu_baseline"] = {"value": round(4 * nb * nks / tc, 3), "unit": "pairs/s", "cores": threads,
                                   "kind": "port", "sample": f"perturbations 1-4 of the batch, all k points and "
                                                             f"bands, FP64 oracle, {tc:.1f}s"}
        log(f"[bench] qpoint: {rec['value']} pairs/s, {rec['ms_per_batch']} ms per batch, "
            f"err vs oracle {rec['rel_err_vs_oracle_pert1']:.1e}")
        return rec
    finally:
        wstat.system_select(0)
        wstat.system_destroy(h)


# Davidson settings: the reference's defaults (trev_pdep 1e-3, trev_pdep_rel 0.1)
# with tr2_dfpt 1e-12, and a converged setting whose eigenvalues are the
# accuracy reference: the difference of two such runs is the precision error
# alone, free of the Davidson convergence error.
DAV_DEFAULT = {"tr2_dfpt": 1e-12}
DAV_CONVERGED = {"tr2_dfpt": 1e-20, "trev_pdep": 1e-9, "trev_pdep_rel": 1e-9, "n_pdep_maxiter": 200}
# the loosest of the settings tried whose Si64 PDEP eigenvalues stay within 1e-6 (relative) of the
# converged FP64 ones in both precisions (trev 1e-7 or tr2_dfpt 1e-16 with trev 1e-8: 4e-6 to 7e-6)
DAV_1E6 = {"tr2_dfpt": 1e-18, "trev_pdep": 1e-8, "trev_pdep_rel": 1e-8, "n_pdep_maxiter": 100}


def _max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def _batch_c4(wstat, torch, args):
    """24 independent H2 clusters of uneven size (seeded: 2-24 molecules in
    12-24 bohr cubic cells, 25 Ry, 16-64 PDEP each): Wstat Davidson on all of
    them, one after the other and concurrently (one host thread and HIP stream
    per system), in FP64 and in mixed precision, with the default and with the
    converged Davidson settings."""
    from west_amd.coulomb import sqvc_init
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_h2_clusters
    clusters = make_h2_clusters(nsys=24, seed=2024)
    wstat.init(0)
    handles = []
    t = time.perf_counter()
    for k, (s, npdep) in enumerate(clusters):
        h = 0 if k == 0 else wstat.system_create()
        wstat.system_select(h)
        gs = make_ground_state(s, wstat.gpu_h_factory, nbnd=s.nbnd_occ + 4, scf_steps=args.scf, tol=1e-7)
        wstat.set_ground_state(gs, sqvc=sqvc_init(s.basis, gs.npwx))
        handles.append(h)
        log(f"[bench] {s.name}: nr={s.basis.nr[0]} npw={s.basis.npw} bands={gs.nbndval} pdep={npdep}")
    wstat.system_select(0)
    log(f"[bench] 24 ground states in {time.perf_counter() - t:.1f}s")

    def jobs(setting):
        return [(h, dict(setting, n_pdep_eigen=npdep, n_pdep_times=4, want_vectors=False))
                for h, (s, npdep) in zip(handles, clusters)]

    res = {}
    evs = {}
    for mode in ("fp64", args.mixed_mode):
        for h in handles:
            wstat.system_select(h)
            wstat.set_precision(mode)
        wstat.system_select(0)
        key = "fp64" if mode == "fp64" else "mixed"
        wstat.davidson_batch(jobs(DAV_DEFAULT), concurrent=True)   # warm-up: kernel loads, occupancy and GEMM timing
        torch.cuda.synchronize()
        t = time.perf_counter()
        seq = wstat.davidson_batch(jobs(DAV_DEFAULT), concurrent=False)
        t_seq = time.perf_counter() - t
        t = time.perf_counter()
        con = wstat.davidson_batch(jobs(DAV_DEFAULT), concurrent=True)
        t_con = time.perf_counter() - t
        t = time.perf_counter()
        tight = wstat.davidson_batch(jobs(DAV_CONVERGED), concurrent=True)
        t_tight = time.perf_counter() - t
        dseq = max(_max_rel(a[0], b[0]) for a, b in zip(seq, con))
        evs[key] = ([r[0] for r in con], [r[0] for r in tight])
        res[key] = {"default": {"wall_s_sequential": round(t_seq, 3), "wall_s_concurrent": round(t_con, 3),
                                "speedup_concurrent": round(t_seq / t_con, 3),
                                "converged": int(sum(int(r[1].sum()) for r in con)),
                                "davidson_iterations": [len(r[3]) for r in con],
                                "ev_max_rel_diff_sequential_vs_concurrent": dseq},
                    "converged_setting": {"wall_s_concurrent": round(t_tight, 3),
                                          "converged": int(sum(int(r[1].sum()) for r in tight)),
                                          "davidson_iterations": [len(r[3]) for r in tight]}}
        log(f"[bench] c4 {mode}: sequential {t_seq:.2f}s, concurrent {t_con:.2f}s, converged setting {t_tight:.2f}s")
    for h in handles:
        wstat.system_select(h)
        wstat.set_precision("fp64")
    wstat.system_select(0)
    ref = evs["fp64"][1]
    for key in ("fp64", "mixed"):
        res[key]["default"]["ev_max_rel_err_vs_converged_fp64"] = max(
            _max_rel(a, b) for a, b in zip(evs[key][0], ref))
    err = max(_max_rel(a, b) for a, b in zip(evs["mixed"][1], ref))
    npdep_total = sum(n for _, n in clusters)
    rec = {"metric": "Wstat Davidson wall time, batch of 24 uneven H2 clusters",
           "value": res["fp64"]["default"]["wall_s_concurrent"],
           "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic (seeded clusters)",
           "config": {"workload": "c4: 24 H2 clusters, 2-24 molecules, 12-24 bohr cells, 25 Ry, 16-64 PDEP each, "
                                  "n_pdep_times 4",
                      "davidson_default": DAV_DEFAULT, "davidson_converged": DAV_CONVERGED,
                      "systems": [{"name": s.name, "nr": s.basis.nr[0], "npw": s.basis.npw, "bands": s.nbnd_occ,
                                   "pdep": n} for s, n in clusters],
                      "pdep_total": npdep_total},
           "fp64": res["fp64"],
           "mixed_precision": dict(res["mixed"], mode=args.mixed_mode,
                                   ev_max_rel_err_vs_fp64_converged_setting=err)}
    print(json.dumps(rec), flush=True)


def _davidson_si64(wstat, torch, args, system, gs, npdep=256):
    """The si64dav workload's own JSON line (see _davidson_si64_measure)."""
    b = system.basis
    res = _davidson_si64_measure(wstat, torch, args, npdep)
    rec = {"metric": "Wstat Davidson wall time, 256 PDEP of the 64-atom Si cell",
           "value": res["fp64"]["converged_setting"]["wall_s"], "unit": "s", "n_gpus": 1,
           "higher_is_better": False, "dtype": "f64", "data": "synthetic (seeded Si supercell ground state)",
           "config": {"workload": f"si64dav: {system.name}, npw={b.npw}, fft={'x'.join(map(str, b.nr))}, "
                                  f"nbndval={gs.nbndval}, nkb={system.nkb}, {npdep} PDEP, n_pdep_times 4",
                      "davidson_default": DAV_DEFAULT, "davidson_converged": dict(DAV_CONVERGED, n_pdep_maxiter=100)},
           **res}
    print(json.dumps(rec), flush=True)


def _davidson_si64_measure(wstat, torch, args, npdep=256):
    """Wstat Davidson for 256 PDEP of the 64-atom Si cell (n_pdep_times 4,
    nvecx 1024), FP64 and mixed, with the default, the converged and the
    1e-6 settings (the metric's eigenvalue accuracy); eigenvalue errors
    against the converged FP64 run."""
    res, evs = {}, {}
    for mode in ("fp64", args.mixed_mode):
        wstat.set_precision(mode)
        key = "fp64" if mode == "fp64" else "mixed"
        res[key] = {}
        for name, setting in (("default", DAV_DEFAULT), ("converged_setting", dict(DAV_CONVERGED, n_pdep_maxiter=100)),
                              ("setting_1e-6", DAV_1E6)):
            log(f"[bench] si64 davidson {mode} {name} ...")
            torch.cuda.synchronize()
            c0 = wstat.counters()
            t = time.perf_counter()
            ev, conv, _, recs = wstat.davidson_diago(npdep, n_pdep_times=4, want_vectors=False, **setting)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t
            c1 = wstat.counters()
            evs[key, name] = ev
            res[key][name] = {"wall_s": round(wall, 3), "iterations": len(recs), "converged": int(conv.sum()),
                              "sternheimer_applications": c1[0] - c0[0]}
            log(f"[bench] si64 davidson {mode} {name}: {wall:.1f}s, {len(recs)} iterations, "
                f"{int(conv.sum())}/{npdep} converged")
    wstat.set_precision("fp64")
    ref = evs["fp64", "converged_setting"]
    for key in ("fp64", "mixed"):
        res[key]["default"]["ev_max_rel_err_vs_converged_fp64"] = _max_rel(evs[key, "default"], ref)
        res[key]["setting_1e-6"]["ev_max_rel_err_vs_converged_fp64"] = _max_rel(evs[key, "setting_1e-6"], ref)
        res[key]["setting_1e-6"]["settings"] = DAV_1E6
    mixed_err = _max_rel(evs["mixed", "converged_setting"], ref)
    return {"fp64": res["fp64"],
            "mixed_precision": dict(res["mixed"], mode=args.mixed_mode,
                                    ev_max_rel_err_vs_fp64_converged_setting=mixed_err,
                                    speedup_vs_fp64_converged_setting=round(
                                        res["fp64"]["converged_setting"]["wall_s"]
                                        / res["mixed"]["converged_setting"]["wall_s"], 3)),
            "ev_fp64_lowest": [round(float(x), 8) for x in ref[:4]]}


def _davidson_h2(wstat, torch, args):
    """Wstat Davidson on 6 H2 molecules (bond 1.40 bohr, random orientations,
    >= 5 bohr apart, 18 bohr cell, 25 Ry, 6 occupied bands): 32 PDEP,
    n_pdep_times 4; wall time of the FP64 and of the mixed-precision solve
    with the default settings (tr2_dfpt 1e-12, default trev) and with the
    converged settings; the mixed PDEP eigenvalues against the FP64 ones."""
    from west_amd.coulomb import sqvc_init
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_h2_system
    s = make_h2_system(nmol=6, L=18.0, ecutwfc=25.0, seed=2024)
    t = time.perf_counter()
    gs = make_ground_state(s, wstat.gpu_h_factory, nbnd=s.nbnd_occ + 4, scf_steps=args.scf, tol=1e-7)
    log(f"[bench] {s.name}: npw={s.basis.npw} nr={s.basis.nr}; ground state in {time.perf_counter() - t:.1f}s")
    wstat.set_ground_state(gs, sqvc=sqvc_init(s.basis, gs.npwx))
    res = {}
    evs = {}
    for mode in ("fp64", args.mixed_mode):
        wstat.set_precision(mode)
        key = "fp64" if mode == "fp64" else "mixed"
        wstat.davidson_diago(32, n_pdep_times=4, want_vectors=False, **DAV_DEFAULT)   # warm-up
        res[key] = {}
        for name, setting in (("default", DAV_DEFAULT), ("converged_setting", DAV_CONVERGED)):
            torch.cuda.synchronize()
            c0 = wstat.counters()
            t = time.perf_counter()
            ev, conv, _, recs = wstat.davidson_diago(32, n_pdep_times=4, want_vectors=False, **setting)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t
            c1 = wstat.counters()
            evs[key, name] = ev
            res[key][name] = {"wall_s": round(wall, 3), "iterations": len(recs), "converged": int(conv.sum()),
                              "sternheimer_applications": c1[0] - c0[0]}
            log(f"[bench] davidson {mode} {name}: {wall:.2f}s, {len(recs)} iterations, {int(conv.sum())}/32 converged")
    wstat.set_precision("fp64")
    ref = evs["fp64", "converged_setting"]
    for key in ("fp64", "mixed"):
        res[key]["default"]["ev_max_rel_err_vs_converged_fp64"] = _max_rel(evs[key, "default"], ref)
    state = {"gs": gs, "ev_fp64_default": evs["fp64", "default"]}
    return {"workload": f"{s.name}: npw={s.basis.npw}, fft={'x'.join(map(str, s.basis.nr))}, nbndval={gs.nbndval}",
            "n_pdep_eigen": 32, "n_pdep_times": 4, "davidson_default": DAV_DEFAULT,
            "davidson_converged": DAV_CONVERGED, "fp64": res["fp64"], "mixed": res["mixed"],
            "ev_max_rel_err_mixed_vs_fp64_converged_setting": _max_rel(evs["mixed", "converged_setting"], ref),
            "ev_fp64_lowest": [round(float(x), 8) for x in ref[:4]]}, state


def _davidson_c1_fp64(wstat, torch, args):
    """The 6-H2 FP64 Davidson (32 PDEP) at the converged settings: one warm-up
    run, then one timed run between device synchronisations."""
    from west_amd.coulomb import sqvc_init
    from west_amd.groundstate import make_ground_state
    from west_amd.systems import make_h2_system
    s = make_h2_system(nmol=6, L=18.0, ecutwfc=25.0, seed=2024)
    gs = make_ground_state(s, wstat.gpu_h_factory, nbnd=s.nbnd_occ + 4, scf_steps=args.scf, tol=1e-7)
    wstat.set_ground_state(gs, sqvc=sqvc_init(s.basis, gs.npwx))
    wstat.davidson_diago(32, n_pdep_times=4, want_vectors=False, **DAV_CONVERGED)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ev, conv, _, recs = wstat.davidson_diago(32, n_pdep_times=4, want_vectors=False, **DAV_CONVERGED)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    log(f"[bench] 6H2 davidson fp64 converged_setting: {wall:.4f}s, {len(recs)} iterations")
    return {"metric": "Wstat Davidson wall time, 6 H2, 32 PDEP, converged settings", "value": round(wall, 4),
            "unit": "s", "higher_is_better": False, "iterations": len(recs), "converged": int(conv.sum()),
            "ev_lowest": [round(float(x), 10) for x in ev[:4]]}


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _davidson_h2_cpu(state, dav):
    """The 6-H2 Davidson (32 PDEP, the default settings) by the FP64 oracle on
    the host cores, from the same ground state: wall time against the GPU's,
    and the PDEP eigenvalues of the FP64 GPU run against it."""
    import oracle.fft_at_gamma as ofa
    from oracle.davidson import davidson_diago_gamma
    from oracle.dfpt import sqvc_init
    gs = state["gs"]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ofa.WORKERS = threads
    t = time.perf_counter()
    log_ref = []
    ev, conv, _ = davidson_diago_gamma(gs, sqvc_init(gs.basis, gs.npwx), 32, n_pdep_times=4, log=log_ref,
                                       **DAV_DEFAULT)
    wall = time.perf_counter() - t
    log(f"[bench] davidson 6H2 on the host oracle: {wall:.1f}s, {len(log_ref)} iterations")
    return {"wall_s": round(wall, 2), "iterations": len(log_ref), "converged": int(np.sum(conv)),
            "cores": threads, "gpu_fp64_wall_s": dav["fp64"]["default"]["wall_s"],
            "ev_max_rel_err_gpu_fp64_vs_cpu": _max_rel(state["ev_fp64_default"], ev),
            "settings": DAV_DEFAULT}


def _run_mode(wstat, torch, args, mode, npert, dvg, dng):
    """Warmup + timed chi0 batches in one precision mode; per-step counters and kernel times."""
    wstat.set_precision(mode, args.refine_delta if mode != "fp64" else 0.0)
    nv = wstat.current_ground_state().nbndval
    pairs = npert * nv
    for i in range(args.warmup):
        s0 = wstat.counters()
        t = time.perf_counter()
        wstat.apply_operator(npert, dvg, args.tr2, dng=dng)
        torch.cuda.synchronize()
        s1 = wstat.counters()
        log(f"[bench] {mode} warmup {i}: {time.perf_counter() - t:.3f}s, stern={s1[0] - s0[0]}, "
            f"cg_iters={s1[1] - s0[1]}")
    wstat.profile(True)
    s0, p0 = wstat.counters(), wstat.precision_counters()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        wstat.apply_operator(npert, dvg, args.tr2, dng=dng)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s1, p1 = wstat.counters(), wstat.precision_counters()
    elapsed = t1 - t0
    kernels = {}
    for name in ("zinv", "plane_vloc", "zfwd", "plane_density", "plane_to_real", "gemm_bra", "gemm_ket", "cg_prec",
                 "cg_dir", "cg_update", "precondition", "zinv32", "plane_vloc32", "zfwd32", "gemm_bra32",
                 "gemm_ket32", "cg_prec32", "cg_dir32", "cg_update32", "refine_fold", "refine_residual"):
        ms, cnt = wstat.profile_query(name)
        if cnt:
            kernels[name] = {"ms_per_step": ms / args.steps, "launches": cnt, "avg_ms": ms / cnt}
    wstat.profile(False)
    wstat.set_precision("fp64")
    log(f"[bench] {mode} kernel ms/step: " + ", ".join(f"{k}={v['ms_per_step']:.1f}" for k, v in kernels.items()))
    return {"value": pairs * args.steps / elapsed, "ms_per_step": 1000.0 * elapsed / args.steps,
            "stern_per_step": (s1[0] - s0[0]) / args.steps, "iters_per_step": (s1[1] - s0[1]) / args.steps,
            "fp32_per_step": (p1[0] - p0[0]) / args.steps, "refine_per_step": (p1[1] - p0[1]) / args.steps,
            "fallback_per_step": (p1[2] - p0[2]) / args.steps,
            "pairs": pairs, "kernels": kernels}


def _save_gs(gs, path):
    np.savez(path, evc=gs.evc, et=gs.et, vrs=gs.vrs, alpha_pv=gs.alpha_pv, nbndval=gs.nbndval)


def _load_gs(system, path):
    from west_amd.groundstate import GroundState
    from west_amd.systems import kb_projectors
    d = np.load(path)
    b = system.basis
    if d["evc"].shape[1] != b.npw or int(d["nbndval"]) != system.nbnd_occ:
        return None
    vkb, dvan = kb_projectors(system)
    nv = int(d["nbndval"])
    return GroundState(basis=b, npwx=b.npw, vrs=d["vrs"], vkb=np.ascontiguousarray(vkb.T), dvan=dvan,
                       evc=d["evc"], et=d["et"], nbndval=nv, occ=np.ones(nv), alpha_pv=float(d["alpha_pv"]),
                       meta={"system": system.name, "cached": True})


def _ncols(b):
    """stored stick columns (distinct (i, j) of the half sphere)"""
    return len({(int(m[0]), int(m[1])) for m in b.mill})


def _byte_model(b, gs, npert, run):
    """Compulsory HBM bytes (or MFMA flops) per step of each kernel class:
    (bound, per-step amount, amount added per launch, transform items per step).

    Units: a transform item is one band pair (two real bands packed in one
    complex transform); its stick block is n3 x (2*ncols - 1) complex numbers
    (16 bytes in FP64, 8 in FP32).
    plane_vloc : read + write one stick block per item, read the real-space
                 potential once per launch (nnr reals; the right-hand-side
                 launches read each perturbation's dv(r) once more)
    zinv       : read two bands (2 npw complex), write one stick block
    zfwd       : read one stick block, read two bands, write two bands
    gemm_bra / gemm_ket : 2 * (2 npw) * nq flops per Sternheimer application
    cg_update  : read h, g, t, x and write x, g per application (6 npw complex)
    FP64 kernels of the mixed run are the refinement residuals and the
    right-hand sides; the FP32 kernels carry the CG iterations.
    """
    npw = b.npw
    nv = gs.nbndval
    nq = gs.nbndval + gs.nkb
    ncol = 2 * _ncols(b) - 1
    st64 = b.nr[2] * ncol * 16
    st32 = st64 // 2
    n32 = run["fp32_per_step"]
    n64 = run["stern_per_step"] - n32
    it64 = n64 / 2.0
    it32 = n32 / 2.0
    items_rhs = npert * ((nv + 1) // 2)
    return {
        "plane_vloc": ("hbm", (it64 + items_rhs) * 2 * st64 + npert * 8 * b.nnr, 8 * b.nnr, it64 + items_rhs),
        "zinv": ("hbm", it64 * (2 * 16 * npw + st64), 0, it64),
        "zfwd": ("hbm", it64 * (st64 + 4 * 16 * npw), 0, it64),
        "gemm_bra": ("mfma", 2.0 * 2 * npw * nq * n64, 0, 0),
        "gemm_ket": ("mfma", 2.0 * 2 * npw * nq * n64, 0, 0),
        "cg_update": ("hbm", 6 * 16 * npw * n64, 0, 0),
        "plane_vloc32": ("hbm", it32 * 2 * st32, 4 * b.nnr, it32),
        "zinv32": ("hbm", it32 * (2 * 8 * npw + st32), 0, it32),
        "zfwd32": ("hbm", it32 * (st32 + 4 * 8 * npw), 0, it32),
        "cg_update32": ("hbm", 6 * 8 * npw * n32, 0, 0),
    }


def _roofline(name, k, model, steps, workload=None):
    """achieved = compulsory bytes (flops) per launch / average launch time;
    traffic = measured HBM bytes per item x items per launch, where measured."""
    if name not in model:
        return {"kernel": name, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}
    bound, per_step, per_launch_extra, items_per_step = model[name]
    launches_per_step = k["launches"] / steps
    per_launch = per_step / launches_per_step + per_launch_extra
    avg_s = k["avg_ms"] * 1e-3
    if bound == "mfma":
        ach = per_launch / avg_s / 1e12
        return {"kernel": name, "bound": "mfma", "achieved": round(ach, 2), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "algorithmic_flops_per_launch": per_launch, "avg_launch_ms": k["avg_ms"]}
    ach = per_launch / avg_s / 1e9
    traffic = None
    meas = TRAFFIC_PER_ITEM.get(workload, {}).get(name)
    if meas and items_per_step:
        traffic = int(meas * items_per_step / launches_per_step)
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "compulsory_bytes_per_launch": int(per_launch), "avg_launch_ms": k["avg_ms"],
            "bytes": "stick blocks read and written once per item, the potential read once per launch"}


def _kernel_bytes_end_to_end(model, ms_per_step):
    """Sum of the HBM-bound kernels' compulsory bytes per step over the step
    time: the HBM bandwidth the whole chi0 step sustains on bytes it must move."""
    total = sum(v[1] for v in model.values() if v[0] == "hbm")
    ach = total / (ms_per_step * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_step": int(total)}


def _end_to_end(b, nv, pairs_per_s, run):
    """The chi0 byte model with one full-grid HBM round trip per 3D transform
    (which the pruned stick transforms here never make, so this is a model
    equivalent, not a measured bandwidth).  Bytes per (pert, band) pair, b
    bytes per complex and
    s per real (FP64: b = 16, s = 8; FP32: 8, 4):
    B_iter = 2b N_r + (s/2) N_r + 16 b N_pw per operator application and CG step,
    B_pair = n_it B_iter + 4b N_r + 1.5 s N_r + 4b N_pw (FP64 right-hand side and
    density), with n_it the measured operator applications per pair, split into
    FP32 and FP64 ones for the mixed-precision run."""
    nr, npw = b.nnr, b.npw
    pairs = run["pairs"]
    n32 = run["fp32_per_step"] / pairs
    n64 = (run["stern_per_step"] - run["fp32_per_step"]) / pairs

    def biter(bc, s):
        return 2 * bc * nr + s / 2 * nr + 16 * bc * npw

    bpair = n64 * biter(16, 8) + n32 * biter(8, 4) + 64 * nr + 12 * nr + 64 * npw
    ach = pairs_per_s * bpair / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_pair": int(bpair), "n_it": round(n32 + n64, 2),
            "n_it_fp32": round(n32, 2)}


if __name__ == "__main__":
    main()
