"""Write the golden fixture of the small workload (N=2^13, q0 60-bit, 3 x 40-bit
scaling primes, dnum=2, K=2 special 60-bit primes, scale 2^40, seed 0).

The fixture pins the oracle's per-stage residues so that any later change of
the oracle (or of the parameter rules) is caught: for every stage it stores
the SHA-256 of each limb and the first 16 coefficients of each limb.  Inputs
are regenerated from the seed by the oracle itself, so the fixture is data
only (a few hundred KB of JSON).

Run:  python tests/golden/make_golden.py   (writes tests/golden/small_seed0.json)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "small_seed0.json")


def digest(a: np.ndarray) -> dict:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    a2 = a.reshape(-1, a.shape[-1])
    return {"shape": list(a.shape), "sha256": [hashlib.sha256(r.tobytes()).hexdigest() for r in a2],
            "head": [[int(x) for x in r[:16]] for r in a2]}


def stages(seed: int = 0) -> dict:
    P = O.params_small()
    o = O.Oracle(P, seed=seed)
    rng = np.random.default_rng(seed)
    z1, z2 = rng.uniform(-1, 1, P.N // 2), rng.uniform(-1, 1, P.N // 2)
    ct1 = o.encrypt_slots(z1, P.L, 1)
    ct2 = o.encrypt_slots(z2, P.L, 2)
    rk = o.relin_key()
    k1 = o.rotation_key(1)
    d0, d1, d2 = o.tensor(*ct1, *ct2)
    ext = o.modup(d2)
    oa, ob = o.ksk_inner(ext, rk)
    ks_b, ks_a = o.moddown(np.concatenate([ob[: P.L + 1], ob[P.L + 1:]])), o.moddown(oa)
    mult = o.mult(ct1, ct2, rk)
    resc = o.rescale(mult)
    rot = o.rotate(ct1, 1, k1)
    hoist = o.rotate_hoisted(ct1, [1], [k1])[0]
    ntt_in = o.random_poly(P.L + 1 + P.K, seed=99, primeids=list(range(P.L + 1 + P.K)))
    ntt_out = o.ntt(ntt_in, list(range(P.L + 1 + P.K)), inverse=False)
    return {
        "params": {"logN": P.logN, "L": P.L, "K": P.K, "dnum": P.dnum, "moduli": P.moduli, "roots": P.roots,
                   "scale_bits": P.scale_bits},
        "secret_coeff_sha256": hashlib.sha256(o.s_coef.tobytes()).hexdigest(),
        "relin_key_a": digest(rk[0]), "relin_key_b": digest(rk[1]),
        "ct1": digest(np.stack(ct1)), "ct2": digest(np.stack(ct2)),
        "tensor": digest(np.stack([d0, d1, d2])),
        "modup": digest(ext),
        "ksk_a": digest(oa), "ksk_b": digest(ob),
        "moddown_b": digest(ks_b), "moddown_a": digest(ks_a),
        "mult": digest(np.stack(mult)), "rescale": digest(np.stack(resc)),
        "rotate1": digest(np.stack(rot)), "hoisted1": digest(np.stack(hoist)),
        "ntt_in": digest(ntt_in), "ntt_out": digest(ntt_out),
    }


if __name__ == "__main__":
    data = stages(0)
    with open(OUT, "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print(OUT, os.path.getsize(OUT), "bytes")
