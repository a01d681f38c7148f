"""Seeded synthetic trajectories for the KSG, Gaussian/GCMI, surrogate and
batched workloads.

The reference reads coordinates from GROMACS trajectories
(src/gmx_correlation.cpp:158-182: one float rvec per selected atom per frame).
Here frames come from a documented generator instead, so that the same seeded
input can be rebuilt on any machine with the same numpy:

1. Mean structure: a random-walk chain r_0 = 0, r_{a+1} = r_a + 0.38 nm * u,
   u uniform on the unit sphere.
2. Per-frame displacement = sum_m a_m(t) * L_m + eps, with a_m(t) ~ N(0, s_m^2)
   i.i.d. per frame, s_m linearly spaced over [0.05, 0.3] nm, L_m a per-atom
   i.i.d. unit 3-vector ("random unit loadings"), and eps ~ N(0, 0.05^2) per
   coordinate (or 0.05 * Student-t(nu) when `student_nu` is given).
3. Optional non-Gaussian couplings on disjoint atom pairs:
   - switch pairs: s(t) ~ Bernoulli(0.3) moves both atoms by 0.4 nm * s(t)
     along one shared unit direction;
   - rotor pairs: the two atoms rotate rigidly about their midpoint around a
     random unit axis by theta(t) = 0.5 * N(+-0.6 rad, 0.1^2) with the sign
     +-1 drawn with p = 1/2 per frame.
4. Cast to float32 last.  The output is frame-major [nframes][natoms][3].

The draw order below is part of the definition: it fixes the numbers.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np


@dataclasses.dataclass(frozen=True)
class TrajSpec:
    natoms: int
    nframes: int
    nmodes: int
    seed: int
    n_switch: int = 0
    n_rotor: int = 0
    student_nu: float | None = None
    # None: one stream for everything.  Otherwise the per-frame draws (mode
    # amplitudes, noise, switch states, rotor angles) come from this seed and
    # the structure (chain, loadings, coupled pairs) from `seed`: replicas of
    # one system with independent frames.
    frame_seed: int | None = None


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _rotate(vecs: np.ndarray, axis: np.ndarray, theta: np.ndarray) -> np.ndarray:
    """Rodrigues rotation of vecs[t] about the unit axis by theta[t]."""
    c = np.cos(theta)[:, None]
    s = np.sin(theta)[:, None]
    kxv = np.cross(axis[None, :], vecs)
    kdv = (vecs @ axis)[:, None] * axis[None, :]
    return vecs * c + kxv * s + kdv * (1.0 - c)


def generate(spec: TrajSpec) -> np.ndarray:
    """float32 coordinates [nframes][natoms][3] in nm."""
    rng = np.random.Generator(np.random.PCG64(spec.seed))
    n, T, M = spec.natoms, spec.nframes, spec.nmodes
    steps = _unit_vectors(rng, n - 1) * 0.38
    mean = np.concatenate([np.zeros((1, 3)), np.cumsum(steps, axis=0)], axis=0)
    sig = np.linspace(0.05, 0.3, M) if M > 1 else np.array([0.05])
    loads = np.stack([_unit_vectors(rng, n) for _ in range(M)])       # [M][n][3]
    fr = rng if spec.frame_seed is None else np.random.Generator(np.random.PCG64(spec.frame_seed))
    amps = fr.standard_normal((T, M)) * sig[None, :]                   # [T][M]
    disp = np.einsum("tm,mnd->tnd", amps, loads)
    if spec.student_nu is None:
        disp += fr.standard_normal((T, n, 3)) * 0.05
    else:
        disp += fr.standard_t(spec.student_nu, size=(T, n, 3)) * 0.05
    pos = mean[None, :, :] + disp

    need = 2 * (spec.n_switch + spec.n_rotor)
    if need > n:
        raise ValueError(f"{need} coupled atoms requested but only {n} atoms")
    if need:
        perm = rng.permutation(n)[:need]
        sw = perm[: 2 * spec.n_switch].reshape(-1, 2)
        rot = perm[2 * spec.n_switch:].reshape(-1, 2)
        for a, b in sw:
            direction = _unit_vectors(rng, 1)[0]
            s = (fr.random(T) < 0.3).astype(np.float64)
            shift = 0.4 * s[:, None] * direction[None, :]
            pos[:, a, :] += shift
            pos[:, b, :] += shift
        for a, b in rot:
            axis = _unit_vectors(rng, 1)[0]
            sign = np.where(fr.random(T) < 0.5, -1.0, 1.0)
            theta = 0.5 * fr.normal(0.6 * sign, 0.1)
            mid = 0.5 * (pos[:, a, :] + pos[:, b, :])
            pos[:, a, :] = mid + _rotate(pos[:, a, :] - mid, axis, theta)
            pos[:, b, :] = mid + _rotate(pos[:, b, :] - mid, axis, theta)
    return np.ascontiguousarray(pos.astype(np.float32))


def coords_sha256(coords: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(coords, dtype=np.float32).tobytes()).hexdigest()


def coupled_counts(natoms: int) -> tuple[int, int]:
    """Switch and rotor pair counts of the 247-atom workload scaled to natoms."""
    return min(30, natoms // 8), min(20, natoms // 12)


# The named workloads.  Seeds: the workload index; the surrogate inference
# set uses 33; batched trajectory t uses 4000 + t.
SMALL_CPU = TrajSpec(natoms=64, nframes=500, nmodes=8, seed=0)           # k = 10
README_247 = TrajSpec(natoms=247, nframes=2000, nmodes=20, seed=1, n_switch=30, n_rotor=20)  # k = 100
DENSE_1000 = TrajSpec(natoms=1000, nframes=10000, nmodes=50, seed=2, n_switch=30, n_rotor=20,
                      student_nu=5.0)
SURR_TRAIN = TrajSpec(natoms=500, nframes=4000, nmodes=20, seed=3, n_switch=30, n_rotor=20)
SURR_INFER = TrajSpec(natoms=2000, nframes=5000, nmodes=20, seed=33, n_switch=30, n_rotor=20)


def batched_specs(count: int = 32) -> list[tuple[TrajSpec, int]]:
    """The uneven batch: atoms ~ U{50..600}, frames ~ U{1000..8000},
    k ~ U{10, 50, 100}, drawn from seed 4; trajectory t uses seed 4000+t."""
    rng = np.random.Generator(np.random.PCG64(4))
    atoms = rng.integers(50, 601, size=count)
    frames = rng.integers(1000, 8001, size=count)
    ks = rng.choice(np.array([10, 50, 100]), size=count)
    out = []
    for t in range(count):
        n = int(atoms[t])
        ns, nr = coupled_counts(n)
        out.append((TrajSpec(natoms=n, nframes=int(frames[t]), nmodes=20, seed=4000 + t,
                             n_switch=ns, n_rotor=nr), int(ks[t])))
    return out
