"""Seeded inputs of the N=2^16, L=24, dnum=3 batched and serving workloads,
shared by the digest generator (tests/golden/make_workload_digests.py, which
runs the oracle on them) and the GPU test (tests/test_gpu_workload.py, which
runs the engine on them and compares SHA-256 digests).

Everything is regenerated from seeds with the oracle's sampler (uniform
residues per prime, one xoshiro256** stream per (seed, prime)), so the GPU box
rebuilds the exact inputs without running any of the oracle's ops.  Keys are
uniform residues in the RawKeySwitchKey layout: the digests pin residue
arithmetic, which any key exercises alike.

Workloads (the batched and serving configurations of the benchmark):
  * 128 HMult+relin+rescale pairs at levels drawn uniformly from 4..24 (seed 0)
  * the first ciphertext of each pair rotated by 2^U{0..15} (seed 1), batched
  * one level-24 ciphertext rotated by the 27 distinct indices
    {1..16} u {2^k, k <= 15}, one hoisted set
  * a serving mix: 2 HMult+relin+rescale, 2 rotations, 2 logistic-regression
    scoring jobs at level 24 (8 rotate-and-sum steps, degree-3 polynomial)
"""
from __future__ import annotations

import hashlib

import numpy as np

from oracle import oracle as O

N_PAIRS = 128
HOIST_IDX = sorted(set(range(1, 17)) | {1 << k for k in range(16)})
LOGREG_ROT = 8
LOGREG_COEFFS = (0.5, 0.197, -0.004)


def params() -> O.Params:
    return O.params_production()


def sample(o: O.Oracle, seed: int, nlimbs: int, primeids=None) -> np.ndarray:
    pid = O.ids(nlimbs) if primeids is None else np.asarray(primeids, dtype=np.int32)
    out = np.zeros((nlimbs, o.N), dtype=np.uint64)
    O.lib().orc_sample_uniform(o.ctx, seed, nlimbs, pid, out)
    return out


def ct(o: O.Oracle, seed: int, level: int):
    return sample(o, seed, level + 1), sample(o, seed + 1, level + 1)


def key(o: O.Oracle, seed: int):
    P = o.P
    np_ = P.L + 1 + P.K
    a = np.stack([sample(o, seed + 2 * d, np_) for d in range(P.dnum)])
    b = np.stack([sample(o, seed + 2 * d + 1, np_) for d in range(P.dnum)])
    return a, b


def relin_key(o):
    return key(o, 500_000)


def rotation_key(o, r: int):
    return key(o, 700_000 + 16 * r)


def rotation_indices():
    """Every rotation any of the workloads uses."""
    return sorted(set(HOIST_IDX) | {1 << k for k in range(16)})


def pair_levels():
    return [int(x) for x in np.random.default_rng(0).integers(4, 25, N_PAIRS)]


def pair(o, i: int, level: int):
    return ct(o, 10_000 + 4 * i, level), ct(o, 10_000 + 4 * i + 2, level)


def batch_rotations():
    return [int(1) << int(x) for x in np.random.default_rng(1).integers(0, 16, N_PAIRS)]


def hoist_input(o):
    return ct(o, 900_000, o.P.L)


def serve_jobs(o):
    """Six jobs: (kind, inputs...) with kind in mult / rot / logreg."""
    P = o.P
    jobs = [("mult", ct(o, 910_000, P.L), ct(o, 910_002, P.L)),
            ("mult", ct(o, 910_004, 9), ct(o, 910_006, 9)),
            ("rot", ct(o, 910_010, P.L), 1 << 13),
            ("rot", ct(o, 910_012, 6), 3),
            ("logreg", ct(o, 910_020, P.L), sample(o, 910_022, P.L + 1)),
            ("logreg", ct(o, 910_030, P.L), sample(o, 910_032, P.L + 1))]
    return jobs


def digest(c) -> str:
    h = hashlib.sha256()
    for x in c:
        h.update(np.ascontiguousarray(x, dtype=np.uint64).tobytes())
    return h.hexdigest()
