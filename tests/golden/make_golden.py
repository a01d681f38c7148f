"""Write the golden fixtures under tests/golden/ from the reference's own CPU
code, compiled from its sources by oracle/Makefile into oracle/_ref/libref.so.

Run in the build container (where the reference checkout exists):
    python tests/golden/make_golden.py [--skip-large]

Every fixture is data: inputs are rebuilt from the seeded generator in
gmx_correlation_amd/synth.py (their SHA-256 is stored so a generator change is
caught), outputs are what the reference computed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gmx_correlation_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def tie_spec() -> np.ndarray:
    """12 atoms x 80 frames with coordinates quantised to 0.02 nm, which makes
    exact distance ties common (the XTC-like case)."""
    c = synth.generate(synth.TrajSpec(natoms=12, nframes=80, nmodes=3, seed=11))
    return np.round(c / 0.02).astype(np.float32) * np.float32(0.02)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-large", action="store_true")
    args = ap.parse_args()
    O.build(with_ref=True)
    meta = {}

    def save(name, arr):
        np.save(os.path.join(OUT, name), arr, allow_pickle=False)

    # Small CPU workload: 64 atoms x 500 frames, k = 10.
    c0 = synth.generate(synth.SMALL_CPU)
    meta["small_cpu_sha256"] = synth.coords_sha256(c0)
    save("small_cpu_k10_mi.npy", O.ref_ksg_matrix(c0, 10))
    kx, psi = O.ref_prepare(c0)
    save("small_cpu_prepare_rows_a0a1.npy", kx[:6].copy())
    save("small_cpu_psi.npy", psi)
    save("small_cpu_gauss_mi.npy", O.ref_gauss_matrix(c0))
    save("small_cpu_gcmi_mi.npy", O.ref_gcmi_matrix(c0))
    for k in (1, 3, 25, 499):
        save(f"small_cpu_k{k}_pairs.npy", np.array(
            [O.ref().ref_ksg_pair(c0.ctypes.data_as(O._fp), 500, 64, a1, a2, k)
             for a1, a2 in ((1, 0), (17, 5), (63, 62))]))

    # Transfer entropy on a 16-atom, 300-frame slice of the small workload.
    cte = np.ascontiguousarray(c0[:300, :16])
    for lag in (1, 5, 10):
        save(f"te16_k5_lag{lag}.npy", O.ref_te_matrix(cte, 5, lag))

    # Transfer entropy at k above 128 (the GPU's radix-select pass), on a
    # 6-atom, 400-frame slice, and on the same slice quantised to 0.02 nm
    # (exact ties at the k-th distance).
    cte6 = np.ascontiguousarray(c0[:400, 16:22])
    save("te6_k200_lag3.npy", O.ref_te_matrix(cte6, 200, 3))
    cq = np.round(cte6 / 0.02).astype(np.float32) * np.float32(0.02)
    save("te6q_k150_lag1.npy", O.ref_te_matrix(cq, 150, 1))

    # Quantised coordinates: exact ties in the neighbour search.
    ct = tie_spec()
    meta["ties_sha256"] = synth.coords_sha256(ct)
    save("ties12_k4_mi.npy", O.ref_ksg_matrix(ct, 4))

    # correl.dat text written by the reference's write_matrix.
    m = O.pearsify(O.ref_ksg_matrix(c0[:, :8], 10))
    path = os.path.join(OUT, "small8_correl.dat")
    O.ref_write_matrix(m, path)
    # the same r(MI) matrix as XPM (write_xpm_matrix), and a 5-atom matrix
    # whose off-diagonal values are all equal (the writer's 0..1 fallback)
    O.ref_write_xpm(m, os.path.join(OUT, "small8_correl.xpm"), "C-alpha correlation matrix",
                    "r(MI)")
    flat = np.full((5, 5), 0.25)
    np.fill_diagonal(flat, 1.0)
    O.ref_write_xpm(flat, os.path.join(OUT, "flat5.xpm"), "flat correlation matrix",
                    "MI (bits)")

    if not args.skip_large:
        c1 = synth.generate(synth.README_247)
        meta["readme247_sha256"] = synth.coords_sha256(c1)
        t0 = time.time()
        m1 = O.ref_ksg_matrix(c1, 100)
        meta["readme247_ref_seconds_8threads"] = time.time() - t0
        iu = np.triu_indices(247, 1)
        save("readme247_k100_mi_triu.npy", m1[iu])
    else:
        old = os.path.join(OUT, "meta.json")
        if os.path.exists(old):
            prev = json.load(open(old))
            for key in ("readme247_sha256", "readme247_ref_seconds_8threads"):
                if key in prev:
                    meta[key] = prev[key]
    with open(os.path.join(OUT, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
