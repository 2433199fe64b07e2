"""Chebyshev series by Paterson-Stockmeyer (the polynomial step of EvalMod):
the decrypted result against the FP64 value of the series.  The evaluation is
a composition of ops whose residues are checked bit-exactly elsewhere
(mult + rescale, scalar ops, add/sub, drop to level); the operation order is
the engine's own, so the end-to-end check is numeric."""
from __future__ import annotations

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npc

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deep():
    P = O.make_params(12, 60, 40, 12, 60, 3)  # 12 levels for the series' depth
    o, ctx, keys = make_env(P, seed=7)
    yield P, o, ctx
    ctx.close()


@pytest.mark.parametrize("deg, a, b", [(15, -1.0, 1.0), (20, -2.0, 3.0), (5, -1.0, 1.0), (1, -1.0, 1.0),
                                       (0, -1.0, 1.0), (31, -1.0, 1.0)])
def test_chebyshev_series(deep, deg, a, b):
    P, o, ctx = deep
    log(f"chebyshev deg {deg} on [{a}, {b}]")
    rng = np.random.default_rng(deg)
    coeffs = [float(v) for v in rng.normal(0, 1, deg + 1) / (np.arange(deg + 1) + 1.0) ** 1.5]
    z = rng.uniform(a, b, P.N // 2)
    ct = o.encrypt_slots(z, P.L, 500 + deg)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    out = g.chebyshev_series(coeffs, a, b)
    info = out.info()
    assert info["noise_level"] == 1 and info["level"] >= 0
    dec = o.decrypt_slots(*out.store(), info["noise"])
    want = npc.chebval((2 * z - (a + b)) / (b - a), coeffs)
    assert np.abs(dec - want).max() < 2e-3, (np.abs(dec - want).max(), info)


def test_mod_reduction_composes(deep):
    """EvalMod (sparse): series, double angle, integer factor -- equal to the
    same three entry points called one after another, and to the FP64 value."""
    P, o, ctx = deep
    log("mod reduction")
    rng = np.random.default_rng(11)
    coeffs = [float(v) for v in rng.normal(0, 1, 12) / (np.arange(12) + 1.0) ** 2]
    z = rng.uniform(-1, 1, P.N // 2)
    ct = o.encrypt_slots(z, P.L, 900)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    r, post = 2, 3
    out = g.mod_reduction(coeffs, r, post)
    step = g.chebyshev_series(coeffs).double_angle(r).mult_int(post)
    c0, c1 = out.store()
    d0, d1 = step.store()
    assert np.array_equal(c0, d0) and np.array_equal(c1, d1)
    y = npc.chebval(z, coeffs)
    for j in range(1, r + 1):
        y = 2 * y * y - 1.0 / (2 * np.pi) ** (2.0 ** (j - r))
    dec = o.decrypt_slots(c0, c1, out.info()["noise"])
    assert np.abs(dec - post * y).max() < 5e-3
