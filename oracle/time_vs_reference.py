"""Speed of the oracle's grid restatement against the reference's own CPU KSG.

    python oracle/time_vs_reference.py [--threads N] [--pairs P]

The baseline restatement (oracle_ksg_pair_baseline in oracle/ksg_oracle.c)
is the timed CPU baseline of bench.py, because the reference-compiled library never leaves
this checkout.  For that to be fair it must run the reference's algorithm at
the reference's speed: both sides get the same pairs of the same workload,
OpenMP over pairs with the same thread count, and the same compiler flags
(oracle/Makefile).  Prints one JSON line per workload with both wall times
and their ratio; also checks that the MI values agree bit for bit.

Needs oracle/_ref/libref.so (the reference checkout).  Test infrastructure.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main() -> int:
    from gmx_correlation_amd import ksg, synth
    from oracle import oracle as O
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0, help="0: all cores (omp_get_max_threads)")
    ap.add_argument("--pairs", type=int, default=400, help="pairs of the 247-atom workload")
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    if not O.ref_available():
        print("oracle/_ref/libref.so is not built (needs the reference checkout)", file=sys.stderr)
        return 2
    nth = args.threads or O.max_threads()
    for spec, k, npairs in ((synth.SMALL_CPU, 10, None), (synth.README_247, 100, args.pairs)):
        coords = synth.generate(spec)
        rows, psi = ksg.kraskov_prepare(ksg.center(coords))
        a1, a2 = np.tril_indices(spec.natoms, -1)
        if npairs is not None:
            idx = np.linspace(0, len(a1) - 1, npairs).astype(np.int64)
            a1, a2 = a1[idx], a2[idx]
        t_ref, t_port = [], []
        for _ in range(args.repeats):
            t0 = time.perf_counter()
            v_ref = O.ref_ksg_pairs(coords, k, a1, a2, nth)
            t_ref.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            v_port = O.ksg_pairs_baseline(rows, psi, k, a1, a2, nth)
            t_port.append(time.perf_counter() - t0)
        print(json.dumps({
            "workload": f"{spec.natoms} atoms x {spec.nframes} frames, k={k}",
            "pairs": int(len(a1)), "threads": nth,
            "reference_s": min(t_ref), "restatement_s": min(t_port),
            "ratio_restatement_over_reference": min(t_port) / min(t_ref),
            "mi_bit_identical": bool(np.array_equal(v_ref, v_port)),
        }), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
