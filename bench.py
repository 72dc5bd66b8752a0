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
FP64_PEAK_TFLOPS = 78.6

WORKLOADS = {
    # name: (builder kwargs, perturbations per batch)
    "si64": ({"nrep": 2, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}, 64),
    "si8": ({"nrep": 1, "a": 10.26, "ecutwfc": 25.0, "sigma": 0.05, "seed": 2024}, 16),
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
        gs = make_ground_state(system, wstat.gpu_h_factory, nbnd=system.nbnd_occ + 4, scf_steps=args.scf, tol=1e-7)
        if args.gs_cache:
            _save_gs(gs, args.gs_cache)
    log(f"[bench] ground state in {time.perf_counter() - t_setup:.1f}s; "
        f"gap {gs.et[gs.nbndval] - gs.et[gs.nbndval - 1] if len(gs.et) > gs.nbndval else float('nan'):.4f} Ry")
    sq = sqvc_init(b, gs.npwx)
    wstat.set_ground_state(gs, sqvc=sq)
    dvg_h = start_vectors(b, gs.npwx, npert, seed=7)
    dev = torch.device("cuda:0")
    dvg = torch.from_numpy(dvg_h).to(dev)
    dng = torch.zeros_like(dvg)
    nv = gs.nbndval
    pairs = npert * nv

    for i in range(args.warmup):
        s0 = wstat.counters()
        t = time.perf_counter()
        wstat.apply_operator(npert, dvg, args.tr2, dng=dng)
        torch.cuda.synchronize()
        s1 = wstat.counters()
        log(f"[bench] warmup {i}: {time.perf_counter() - t:.3f}s, stern={s1[0] - s0[0]}, cg_iters={s1[1] - s0[1]}")

    wstat.profile(True)
    s0 = wstat.counters()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        wstat.apply_operator(npert, dvg, args.tr2, dng=dng)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s1 = wstat.counters()
    elapsed = t1 - t0
    ms_step = 1000.0 * elapsed / args.steps
    value = pairs * args.steps / elapsed
    stern_per_step = (s1[0] - s0[0]) / args.steps
    iters_per_step = (s1[1] - s0[1]) / args.steps
    kernels = {}
    for name in ("zinv", "plane_vloc", "zfwd", "plane_density", "plane_to_real", "gemm_bra", "gemm_fix",
                 "gemm_ket", "cg_prec", "cg_dir", "cg_update", "precondition"):
        ms, cnt = wstat.profile_query(name)
        if cnt:
            kernels[name] = {"ms_per_step": ms / args.steps, "launches": cnt, "avg_ms": ms / cnt}
    wstat.profile(False)
    log("[bench] kernel ms/step: " + ", ".join(f"{k}={v['ms_per_step']:.1f}" for k, v in kernels.items()))

    # ---- roofline of the dominant kernel and the end-to-end figure
    n1, n2, n3 = b.nr
    model = _byte_model(b, gs, npert, stern_per_step)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    roof = _roofline(dom, kernels[dom], model, args.steps)
    e2e = _end_to_end(b, nv, value, stern_per_step / pairs)

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
               "sample": f"perturbation 1 of the batch x occupied bands 1..{nb} (of {nv}), tr2={args.tr2:g}, "
                         f"FP64 oracle (SciPy FFT on {threads} threads, OpenBLAS GEMM), {tc:.1f}s"}

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
        "dtype": "f64",
        "data": "synthetic (seeded Si supercell ground state computed on the device; seeded smooth perturbations)",
        "config": {"workload": f"{args.workload}: {system.name}, npw={b.npw}, fft={n1}x{n2}x{n3}, "
                               f"nbndval={nv}, nkb={system.nkb}, {npert} perturbations/batch",
                   "pairs_per_step": pairs, "tr2_dfpt": args.tr2,
                   "sternheimer_applications_per_step": stern_per_step, "cg_iterations_per_step": iters_per_step},
        "kernels_ms_per_step": {k: round(v["ms_per_step"], 3) for k, v in kernels.items()},
        "roofline": roof,
        "roofline_end_to_end": e2e,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


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


def _byte_model(b, gs, npert, stern_per_step):
    """Compulsory HBM bytes (or MFMA flops) per step of each kernel class.

    Units: a transform item is one band pair (two real bands packed in one
    complex transform); its stick block is n3 x (2*ncols - 1) complex128.
    plane_vloc : read + write one stick block, read V(r) (nnr f64) per item
    zinv       : read two bands (2 npw complex128), write one stick block
    zfwd       : read one stick block, read two bands, write two bands
    gemm_bra / gemm_ket : 2 * (2 npw) * nq flops per Sternheimer application
    cg_update  : read h, g, t, x and write x, g per application (6 npw complex128)
    """
    npw = b.npw
    nv = gs.nbndval
    nq = gs.nbndval + gs.nkb
    sticks = b.nr[2] * (2 * _ncols(b) - 1) * 16
    items_cg = stern_per_step / 2.0
    items_rhs = npert * ((nv + 1) // 2)
    return {
        "plane_vloc": ("hbm", (items_cg + items_rhs) * (2 * sticks + 8 * b.nnr)),
        "zinv": ("hbm", items_cg * (2 * 16 * npw + sticks)),
        "zfwd": ("hbm", items_cg * (sticks + 4 * 16 * npw)),
        "gemm_bra": ("mfma", 2.0 * 2 * npw * nq * stern_per_step),
        "gemm_ket": ("mfma", 2.0 * 2 * npw * nq * stern_per_step),
        "cg_update": ("hbm", 6 * 16 * npw * stern_per_step),
    }


def _roofline(name, k, model, steps):
    """achieved = algorithmic bytes (flops) per launch / average launch time."""
    if name not in model:
        return {"kernel": name, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}
    bound, per_step = model[name]
    launches_per_step = k["launches"] / steps
    per_launch = per_step / launches_per_step
    avg_s = k["avg_ms"] * 1e-3
    if bound == "mfma":
        ach = per_launch / avg_s / 1e12
        return {"kernel": name, "bound": "mfma", "achieved": round(ach, 2), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "algorithmic_flops_per_launch": per_launch, "avg_launch_ms": k["avg_ms"]}
    ach = per_launch / avg_s / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes_per_launch": int(per_launch),
            "avg_launch_ms": k["avg_ms"]}


def _end_to_end(b, nv, pairs_per_s, n_it):
    """Bytes per (pert, band) pair of the chi0 model (FP64, b = 16, s = 8):
    B_iter = 2b N_r + (s/2) N_r + 16 b N_pw per operator application and CG step,
    B_pair = n_it B_iter + 4b N_r + 1.5 s N_r + 4b N_pw, with n_it the measured
    number of operator applications per pair."""
    nr, npw = b.nnr, b.npw
    biter = 32 * nr + 4 * nr + 256 * npw
    bpair = n_it * biter + 64 * nr + 12 * nr + 64 * npw
    ach = pairs_per_s * bpair / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_pair": int(bpair), "n_it": round(n_it, 2)}


if __name__ == "__main__":
    main()
