"""Chebyshev series by Paterson-Stockmeyer (the polynomial step of EvalMod):
the decrypted result against the FP64 value of the series.  The evaluation is
a composition of ops whose residues are checked bit-exactly elsewhere
(mult + rescale, scalar ops, add/sub, drop to level); the operation order is
the engine's own, so the end-to-end check is numeric.  Scales are kept exact
(every addition meets operands on one tracked scale), so the tolerances are
those of the noise, not of FIXEDMANUAL's scale drift."""
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
    assert np.abs(dec - want).max() < 1e-5, (np.abs(dec - want).max(), info)
    assert abs(info["noise"] / 2.0 ** P.scale_bits - 1) < 1e-12  # lands on ScalingFactorReal of its level


def test_chebyshev_evalmod_degree_119(deep):
    """The bootstrap's EvalMod polynomial (degree 119, K = 28, r = 3)."""
    from pyfideslib_amd import bootstrap as B

    P, o, ctx = deep
    log("chebyshev degree 119")
    coeffs = B.evalmod_coefficients(28, 3, 119)
    rng = np.random.default_rng(119)
    z = rng.uniform(-0.7, 0.7, P.N // 2)
    ct = o.encrypt_slots(z, P.L, 119)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    out = g.chebyshev_series(coeffs)
    dec = o.decrypt_slots(*out.store(), out.info()["noise"])
    want = npc.chebval(z, coeffs)
    err = np.abs(dec - want).max()
    log(f"degree 119: level {out.level}, max error {err:.3e}")
    assert err < 1e-5


def test_mod_reduction_composes(deep):
    """EvalMod: series, double angle (constants at the tracked scale), integer
    factor -- the series part equal to chebyshev_series bit for bit, the whole
    against the FP64 value."""
    P, o, ctx = deep
    log("mod reduction")
    rng = np.random.default_rng(11)
    coeffs = [float(v) for v in rng.normal(0, 1, 12) / (np.arange(12) + 1.0) ** 2]
    z = rng.uniform(-1, 1, P.N // 2)
    ct = o.encrypt_slots(z, P.L, 900)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    s0, s1 = g.mod_reduction(coeffs, 0, 1).store()
    t0, t1 = g.chebyshev_series(coeffs).store()
    assert np.array_equal(s0, t0) and np.array_equal(s1, t1)
    r, post = 2, 3
    out = g.mod_reduction(coeffs, r, post)
    y = npc.chebval(z, coeffs)
    for j in range(1, r + 1):
        y = 2 * y * y - 1.0 / (2 * np.pi) ** (2.0 ** (j - r))
    dec = o.decrypt_slots(*out.store(), out.info()["noise"])
    assert np.abs(dec - post * y).max() < 1e-5


@pytest.mark.parametrize("tech", [O.FLEXIBLEAUTO, O.FIXEDAUTO])
def test_chebyshev_series_automatic_modes(tech):
    """Under FLEXIBLEAUTO / FIXEDAUTO the series runs with explicit rescales and
    exact scales; a FLEXIBLEAUTO result sits on its level's ScalingFactorReal
    (FIDESlib runs evalChebyshevSeries under either technique,
    src/CKKS/ApproxModEval.cu:370-700).  Inputs at noise degree 1 and 2."""
    import pyfideslib_amd as F

    P = O.make_params(12, 60, 40, 12, 60, 3)
    if tech == O.FLEXIBLEAUTO:
        sf, sfb = O.flexible_scaling_factors(P.q)
    else:
        sf = [2.0 ** P.scale_bits] * (P.L + 1)
        sfb = [x * x for x in sf]
    o = O.Oracle(P, seed=8)
    ctx = F.Context(F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, tech, sf, sfb))
    try:
        ctx.load_eval_key(o.relin_key())
        for deg, lv in ((15, P.L), (31, P.L - 1)):
            log(f"chebyshev deg {deg} under technique {tech}, input level {lv}")
            rng = np.random.default_rng(deg + 100)
            coeffs = [float(v) for v in rng.normal(0, 1, deg + 1) / (np.arange(deg + 1) + 1.0) ** 1.5]
            z = rng.uniform(-1, 1, P.N // 2)
            g = ctx.load(*o.encrypt_slots(z, lv, 600 + deg, sf[lv]), noise=sf[lv])
            out = g.chebyshev_series(coeffs)
            info = out.info()
            assert info["noise_level"] == 1
            if tech == O.FLEXIBLEAUTO:
                assert info["noise"] == sf[info["level"]]
            dec = o.decrypt_slots(*out.store(), info["noise"])
            assert np.abs(dec - npc.chebval(z, coeffs)).max() < 1e-5
        # a product left at noise degree 2 is rescaled first
        z1, z2 = rng.uniform(-1, 1, P.N // 2), rng.uniform(-1, 1, P.N // 2)
        a = ctx.load(*o.encrypt_slots(z1, P.L, 701, sf[P.L]), noise=sf[P.L])
        b = ctx.load(*o.encrypt_slots(z2, P.L, 702, sf[P.L]), noise=sf[P.L])
        prod = a.mult(b)
        assert prod.info()["noise_level"] == 2
        out = prod.chebyshev_series(coeffs)
        dec = o.decrypt_slots(*out.store(), out.info()["noise"])
        assert np.abs(dec - npc.chebval(z1 * z2, coeffs)).max() < 1e-5
    finally:
        ctx.close()
