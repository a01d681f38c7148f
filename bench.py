"""Benchmark of the all-pairs KSG MI matrix on one MI355X.

Workload: 247 atoms x 2000 frames, k = 100 (the reference README's typical
shape), float32 coordinates from the seeded generator.  One step = one pass
of the hot path over all 30381 atom pairs with the per-atom distance rows
already resident in HBM: pair kernel + FP64 recount of flagged rows +
finalise.  The one-off distance-row build is reported separately.

Prints ONE JSON line.  `roofline` refers to the pair kernel: algorithmic
bytes per launch = npairs * 2 * N^2 * 4 (both atoms' N x N FP32 distance
matrices read once per pair) divided by the kernel's mean duration from HIP
events on the plan's stream.  `cpu_baseline` times the reference's own CPU
KSG (compiled from its sources into oracle/_ref, or the oracle restatement
when that build is absent) on a bounded sample of the same pairs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KSG atom-pairs/s at 2000 frames k=100; hybrid speedup at max |Δr|≤1e-3"
HBM_PEAK_GBS = 8000.0
README_CPU_PAIRS_PER_S = 232.5    # README.md:162-166: 30381 pairs in 130.69 s (10 threads)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(coords: np.ndarray, k: int, natoms: int, budget_s: float) -> dict:
    """Time the reference CPU KSG on a bounded, evenly spaced pair sample."""
    from oracle import oracle as O
    nthreads = max(1, min(os.cpu_count() or 1, 16))
    a1, a2 = np.tril_indices(natoms, -1)
    kind = "reference" if O.ref_available() else "port"

    def run(idx):
        if kind == "reference":
            return O.ref_ksg_pairs(coords, k, a1[idx], a2[idx], nthreads)
        from gmx_correlation_amd import ksg
        rows, psi = ksg.kraskov_prepare(ksg.center(coords))
        return O.ksg_pairs(rows, psi, k, a1[idx], a2[idx], nthreads)

    npairs = len(a1)
    probe = np.linspace(0, npairs - 1, 2 * nthreads).astype(np.int64)
    t0 = time.perf_counter()
    run(probe)
    t_probe = time.perf_counter() - t0
    per_pair = t_probe / len(probe)
    n = int(min(npairs, max(len(probe), budget_s / max(per_pair, 1e-6))))
    idx = np.linspace(0, npairs - 1, n).astype(np.int64)
    t0 = time.perf_counter()
    vals = run(idx)
    dt = time.perf_counter() - t0
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": n / dt,
        "unit": "pairs/s",
        "cores": nthreads,
        "kind": kind,
        "sample": f"{n} of {npairs} pairs, evenly spaced, same workload; {dt:.1f} s wall"
                  f" with OpenMP over pairs; host cpu: {cpu_model}",
        "_values": vals,
        "_idx": idx,
    }


def bench_dense(args) -> int:
    """Gaussian and GCMI at 1000 atoms x 10000 frames (Student-t noise):
    split-FP16 MFMA GEMM time and its error against an FP64 GEMM."""
    from gmx_correlation_amd import dense, ksg, synth
    from oracle import oracle as O
    spec = synth.DENSE_1000
    log(f"bench: generating workload {spec}")
    traj = ksg.center(synth.generate(spec))
    n, N = spec.natoms, spec.nframes
    out = {"workload": "gauss+gcmi 1000 atoms x 10000 frames", "natoms": n, "nframes": N}
    for name, fn in (("gauss", dense.gauss_corrmatrix), ("gcmi", dense.gcmi_corrmatrix)):
        runs = []
        for s in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            res = fn(traj, return_cov=True, stats=True)
            runs.append((time.perf_counter() - t0, res))
        mat, cov, st = runs[-1][1]
        gemm_ms = float(np.mean([r[1][2]["ms_gemm"] for r in runs[args.warmup:]]))
        V = traj.transpose(1, 0, 2).reshape(N, -1) if name == "gauss" else None
        if name == "gcmi":
            V = O.copula(traj).T
        t0 = time.perf_counter()
        c64 = V.T @ V / N
        cpu_gemm_s = time.perf_counter() - t0
        d = np.sqrt(np.outer(np.diag(c64), np.diag(c64)))
        rel = float(np.max(np.abs(cov - c64) / d))
        mi64 = O.gauss_matrix_from_cov(c64, n) if name == "gauss" else None
        if name == "gcmi":
            mi64 = _gcmi_from_cov(c64, n)
        iu = np.triu_indices(n, 1)
        out[name] = {
            "ms_total_host": float(np.mean([r[0] for r in runs[args.warmup:]]) * 1e3),
            "ms_gemm": gemm_ms, "ms_copula": st["ms_copula"], "ms_pairs": st["ms_pairs"],
            "gemm_tflops": st["gemm_flops"] / (gemm_ms * 1e-3) / 1e12,
            "mfma_frac_fp16_dense_2.5PF": st["gemm_flops"] / (gemm_ms * 1e-3) / 2.5e15,
            "cov_max_rel_err_vs_fp64": rel,
            "mi_max_abs_err_vs_fp64_gemm": float(np.max(np.abs(mat[iu] - mi64[iu]))),
            "numpy_fp64_gemm_s": cpu_gemm_s,
        }
        log(f"bench: {name}: {out[name]}")
    print(json.dumps(out), flush=True)
    return 0


def _gcmi_from_cov(c, n):
    """GCMI pair MI from an FP64 second-moment matrix, batched determinants."""
    a1, a2 = np.triu_indices(n, 1)
    idx1 = 3 * a1[:, None] + np.arange(3)[None, :]
    idx2 = 3 * a2[:, None] + np.arange(3)[None, :]
    idx = np.concatenate([idx1, idx2], axis=1)                 # [npairs][6]
    cj = c[idx[:, :, None], idx[:, None, :]]                   # [npairs][6][6]
    d1 = np.linalg.det(cj[:, :3, :3])
    d2 = np.linalg.det(cj[:, 3:, 3:])
    dj = np.linalg.det(cj)
    ok = (d1 > 0) & (d2 > 0) & (dj > 0)
    v = np.where(ok, 0.5 * (np.log(np.where(ok, d1, 1)) + np.log(np.where(ok, d2, 1))
                            - np.log(np.where(ok, dj, 1))), 0.0)
    mat = np.full((n, n), 2000.0)
    mat[a1, a2] = v
    mat[a2, a1] = v
    return mat


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU-baseline work (0 disables)")
    ap.add_argument("--workload", default="ksg", choices=["ksg", "dense"],
                    help="ksg: the headline metric; dense: Gaussian/GCMI GEMM path")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU pair sharding is a design in DESIGN.md, not a measured path.",
              file=sys.stderr)
        return 2
    if args.workload == "dense":
        return bench_dense(args)

    from gmx_correlation_amd import _lib, ksg, synth

    spec, k = synth.README_247, 100
    log(f"bench: generating workload {spec}")
    coords = synth.generate(spec)
    rows, psi = ksg.kraskov_prepare(ksg.center(coords))
    natoms, N = spec.natoms, spec.nframes
    npairs = natoms * (natoms - 1) // 2
    if not _lib.probe():
        raise RuntimeError("no gfx950 device: the benchmark has no CPU fallback")

    plan = ksg.KsgPlan(natoms, N, k)
    plan.load(rows, psi)
    st_load = plan.stats()
    log(f"bench: rows built: load {st_load['ms_load']:.2f} ms, build {st_load['ms_build']:.2f} ms,"
        f" {st_load['row_bytes'] / 1e9:.2f} GB resident")
    for w in range(args.warmup):
        plan.run_all_device()
        log(f"bench: warmup {w + 1}/{args.warmup} pair kernel {plan.stats()['ms_pairs']:.2f} ms")
    pair_ms = []
    fix_ms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        plan.run_all_device()
        st = plan.stats()
        pair_ms.append(st["ms_pairs"])
        fix_ms.append(st["ms_fixup"])
    wall = time.perf_counter() - t0
    st = plan.stats()
    mi = plan.run_all()
    log(f"bench: {args.steps} steps in {wall:.3f} s; pair kernel mean {np.mean(pair_ms):.2f} ms")

    parity = None
    gpath = os.path.join(ROOT, "tests", "golden", "readme247_k100_mi_triu.npy")
    if os.path.exists(gpath):
        ref_triu = np.load(gpath, allow_pickle=False)
        mat = np.zeros((natoms, natoms))
        a1, a2 = np.tril_indices(natoms, -1)
        order = np.lexsort((a2, a1))          # the plan's (a1, a2 < a1) order
        mat[a1[order], a2[order]] = mi
        mat = mat + mat.T
        got = mat[np.triu_indices(natoms, 1)]
        parity = {
            "max_abs_dmi": float(np.max(np.abs(got - ref_triu))),
            "max_abs_dr": float(np.max(np.abs(ksg.pearsify(got) - ksg.pearsify(ref_triu)))),
            "vs": "reference float64 CPU output (tests/golden)",
        }
    plan.close()

    ms_step = wall / args.steps * 1e3
    value = npairs / (wall / args.steps)
    algo_bytes = npairs * 2.0 * N * N * 4.0
    kern_s = float(np.mean(pair_ms)) / 1e3
    achieved = algo_bytes / kern_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ksg_pair_247x2000_k100_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if args.cpu_budget > 0:
        log(f"bench: CPU baseline, ~{args.cpu_budget:.0f} s budget")
        cpu = cpu_baseline(coords, k, natoms, args.cpu_budget)
        idx = cpu.pop("_idx")
        vals = cpu.pop("_values")
        a1, a2 = np.tril_indices(natoms, -1)
        # the CPU sample uses the tril order (a1 > a2) as the plan does
        order = np.lexsort((a2, a1))
        inv = np.empty_like(order)
        inv[order] = np.arange(len(order))
        cpu["max_abs_dmi_vs_gpu"] = float(np.max(np.abs(vals - mi[inv[idx]])))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": value / README_CPU_PAIRS_PER_S,
        "dtype": "f32 select/count, f64 recount+digamma sums",
        "data": "synthetic (seeded generator, gmx_correlation_amd/synth.py)",
        "config": {"workload": "ksg_all_pairs 247 atoms x 2000 frames k=100", "natoms": natoms,
                   "nframes": N, "k": k, "npairs": npairs},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "ksg_pair_kernel", "kernel_ms": kern_s * 1e3,
                     "algorithmic_bytes_per_launch": algo_bytes},
        "cpu_baseline": cpu,
        "detail": {"fixup_ms": float(np.mean(fix_ms)), "fixup_rows": st["fixup_rows"],
                   "exact_ties": st["ties"], "row_build_ms": st_load["ms_build"],
                   "row_bytes": st_load["row_bytes"], "parity": parity},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
