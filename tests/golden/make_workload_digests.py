"""Write SHA-256 digests of the oracle's outputs on the N=2^16, L=24, dnum=3
batched and serving workloads (tests/workload_n65536.py has the seeded inputs):

  * the 128 HMult+relin+rescale pairs at levels U{4..24}
  * the 128 batched rotations by 2^U{0..15}
  * the 27 outputs of the hoisted rotation set at level 24
  * the six jobs of the serving mix (HMult, rotation, logistic-regression
    scoring: y = x w rescaled, 8 rotate-and-sum steps, c0 + c1 y + c3 y^3 with
    the scalar encodings of ElemForEvalMult / ElemForEvalAddOrSub)

tests/test_gpu_workload.py regenerates the inputs on the GPU box, runs the
engine's batched, hoisted and serving entry points and compares digests, so
the oracle's heavy ops never run there.  Digest = SHA-256 over the c0 limbs
then the c1 limbs (uint64 little-endian).

Run:  python tests/golden/make_workload_digests.py
      (writes tests/golden/workload_n65536_digests.json; a few minutes on 8 cores)
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests import workload_n65536 as W  # noqa: E402

OUT = os.path.join(HERE, "workload_n65536_digests.json")


def oracle_rotate(o, c, r, key):
    """Rotation by r slots; r = 0 mod N/2 is the automorphism X -> X (index
    5^(N/2) = 1 mod 2N): the input itself, with no key switch (the reference
    asserts a nonzero normalized index, src/CKKS/Ciphertext.cpp:157-164, :762)."""
    if r % (o.N // 2) == 0:
        return c
    return o.rotate(c, r, key)


def oracle_logreg(o, P, keys, x, w):
    """The request loop's scoring job composed of the oracle's ops, in the
    engine's order (serve.hip): y = rescale(x w); y += rot(y, 2^i) for i < 8;
    y2 = rescale(y y); u = rescale(y c3); y3 = rescale(y2 u);
    v = rescale(y c1); out = y3 + v (at y3's level) + c0."""
    sf = 2.0 ** P.scale_bits
    c0, c1, c3 = W.LOGREG_COEFFS
    y = o.rescale(o.mult_plain(x, w))
    for i in range(W.LOGREG_ROT):
        t = o.rotate(y, 1 << i, keys[1 << i])
        y = (o.add(y[0], t[0]), o.add(y[1], t[1]))
    ly = y[0].shape[0] - 1
    y2 = o.rescale(o.mult(y, y, keys["relin"]))
    u = o.rescale(o.mult_scalar(y, O.elem_for_mult(P.moduli, ly, c3, sf)))
    y3 = o.rescale(o.mult(y2, u, keys["relin"]))
    v = o.rescale(o.mult_scalar(y, O.elem_for_mult(P.moduli, ly, c1, sf)))
    n3 = y3[0].shape[0]
    out = (o.add(y3[0], np.ascontiguousarray(v[0][:n3])), o.add(y3[1], np.ascontiguousarray(v[1][:n3])))
    return o.add_scalar(out, O.elem_for_add(P.moduli, n3 - 1, abs(c0), 1, sf))


def main():
    t0 = time.time()
    P = W.params()
    o = O.Oracle(P, seed=0)
    keys = {"relin": W.relin_key(o)}
    for r in W.rotation_indices():
        keys[r] = W.rotation_key(o, r)
    print(f"keys {time.time() - t0:.1f}s", flush=True)
    levels = W.pair_levels()
    rots = W.batch_rotations()
    mult, rot = [], []
    for i, lv in enumerate(levels):
        a, b = W.pair(o, i, lv)
        mult.append(W.digest(o.rescale(o.mult(a, b, keys["relin"]))))
        rot.append(W.digest(oracle_rotate(o, a, rots[i], keys[rots[i]])))
    print(f"batch {time.time() - t0:.1f}s", flush=True)
    x = W.hoist_input(o)
    hoist = [W.digest(x if r % (o.N // 2) == 0 else h)
             for r, h in zip(W.HOIST_IDX, o.rotate_hoisted(x, W.HOIST_IDX, [keys[r] for r in W.HOIST_IDX]))]
    print(f"hoisted {time.time() - t0:.1f}s", flush=True)
    serve, kinds = [], []
    for job in W.serve_jobs(o):
        kinds.append(job[0])
        if job[0] == "mult":
            serve.append(W.digest(o.rescale(o.mult(job[1], job[2], keys["relin"]))))
        elif job[0] == "rot":
            serve.append(W.digest(oracle_rotate(o, job[1], job[2], keys[job[2]])))
        else:
            serve.append(W.digest(oracle_logreg(o, P, keys, job[1], job[2])))
    print(f"serve {time.time() - t0:.1f}s", flush=True)
    doc = {"params": {"logN": P.logN, "L": P.L, "K": P.K, "dnum": P.dnum, "moduli": [str(q) for q in P.moduli]},
           "pair_levels": levels, "batch_rotations": rots, "hoist_indices": W.HOIST_IDX,
           "mult_rescale": mult, "rotate": rot, "hoisted": hoist,
           "serve_kinds": kinds, "serve": serve}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {OUT} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
