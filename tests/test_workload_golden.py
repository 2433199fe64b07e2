"""CPU checks of the N=2^16 workload fixture (tests/golden/workload_n65536_digests.json):
it lists the seeded levels, rotations and indices the shared input module
generates, and a sample of its digests recomputes from the oracle."""
import json
import os

import numpy as np

from oracle import oracle as O
from tests import workload_n65536 as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workload_n65536_digests.json")


def test_fixture_matches_seeds_and_oracle():
    with open(GOLDEN) as f:
        g = json.load(f)
    P = W.params()
    assert g["params"]["moduli"] == [str(q) for q in P.moduli]
    assert g["pair_levels"] == W.pair_levels() and len(g["mult_rescale"]) == W.N_PAIRS
    assert all(4 <= lv <= 24 for lv in g["pair_levels"])
    assert g["batch_rotations"] == W.batch_rotations() and len(g["rotate"]) == W.N_PAIRS
    assert g["hoist_indices"] == W.HOIST_IDX and len(W.HOIST_IDX) == 27 and len(g["hoisted"]) == 27
    assert len(set(g["mult_rescale"])) == W.N_PAIRS  # distinct inputs give distinct outputs
    # recompute the lowest-level pair (cheapest) with the oracle
    o = O.Oracle(P, seed=0)
    i = int(np.argmin(g["pair_levels"]))
    a, b = W.pair(o, i, g["pair_levels"][i])
    assert W.digest(o.rescale(o.mult(a, b, W.relin_key(o)))) == g["mult_rescale"][i]
    r = g["batch_rotations"][i]
    assert W.digest(o.rotate(a, r, W.rotation_key(o, r))) == g["rotate"][i]


def test_zero_rotations_are_the_input():
    """A rotation by a multiple of N/2 slots is the automorphism X -> X: its
    expected output is the input ciphertext itself (no key switch, no noise)."""
    with open(GOLDEN) as f:
        g = json.load(f)
    P = W.params()
    o = O.Oracle(P, seed=0)
    zero = [i for i, r in enumerate(g["batch_rotations"]) if r % (P.N // 2) == 0]
    assert zero, "the seeded rotations include 2^15 = N/2"
    i = min(zero, key=lambda k: g["pair_levels"][k])
    a, _ = W.pair(o, i, g["pair_levels"][i])
    assert g["rotate"][i] == W.digest(a)
    k = W.HOIST_IDX.index(P.N // 2)
    assert g["hoisted"][k] == W.digest(W.hoist_input(o))
