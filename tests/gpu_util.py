"""Helpers shared by the GPU parity tests: oracle-generated inputs, an engine
context built from the same parameters, device buffers through the C ABI."""
from __future__ import annotations

import sys

import numpy as np

from oracle import oracle as O


def log(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


def make_env(P: O.Params, seed: int = 0, rotations=(), relin=True, scaling_technique=0, hamming=None):
    import pyfideslib_amd as F

    o = O.Oracle(P, seed=seed, hamming=hamming)
    # FIXEDMANUAL: one scaling factor 2^scale_bits at every level (OpenFHE's ScalingFactorReal)
    raw = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, scaling_technique,
                                 [2.0 ** P.scale_bits] * (P.L + 1))
    ctx = F.Context(raw)
    keys = {}
    if relin:
        keys["relin"] = o.relin_key()
        ctx.load_eval_key(keys["relin"])
    for r in rotations:
        keys[r] = o.rotation_key(r)
        ctx.load_rotation_key(r, keys[r])
    return o, ctx, keys


def to_dev(ctx, a: np.ndarray):
    return ctx.device_array(a.shape, a)


def empty_dev(ctx, shape):
    return ctx.device_array(shape, np.zeros(shape, dtype=np.uint64))


def to_host(d) -> np.ndarray:
    return d.numpy()


def sync(ctx):
    ctx.synchronize()
