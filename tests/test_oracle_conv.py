"""The oracle's composition of the convolution transform decrypts to the
slotwise closed form (CPU only): pins the rotation directions, the block
split and the giant-step order the GPU parity test compares against."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_conv import CASES, HALF, closed_form, oracle_conv


@pytest.mark.parametrize("case", [1, 3, 4])
def test_oracle_convolution_matches_closed_form(case):
    b, g, n, s, idx, level = CASES[case]
    P = O.params_small()
    o = O.Oracle(P, seed=3)
    rots = {r for r in idx if r % HALF}
    nb = -(-g // 8)
    for k in range(nb):
        G = min(8, g - 8 * k)
        rots |= {s * (G - j) for j in range(G)} | ({8 * s * (nb - 1 - k)} if k < nb - 1 else set())
    keys = {r: o.rotation_key(r) for r in rots if r % HALF}
    rng = np.random.default_rng(case)
    z = rng.uniform(-1, 1, HALF)
    ct = o.encrypt_slots(z, level, 500 + case)
    pz = [rng.uniform(-1, 1, HALF) for _ in range(n)]
    pl = [o.encode(d, P.L) for d in pz]
    out = oracle_conv(o, ct, pl, keys, b, g, s, idx)
    dec = o.decrypt_slots(*out, 2.0 ** (2 * P.scale_bits))
    assert np.abs(dec - closed_form(z, pz, b, g, s, idx)).max() < 1e-4
    # a dropped giant-step rotation would not match
    wrong = closed_form(z, pz, b, g, s + 1, idx) if s + 1 < HALF else None
    if wrong is not None:
        assert np.abs(dec - wrong).max() > 1e-2
