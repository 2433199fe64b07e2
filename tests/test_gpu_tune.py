"""Every launch configuration the tuner may pick computes the same transform
(needs a GPU): forward and inverse NTTs over 1..48 limbs of cycling primes
(fides_ntt_limbs) under forced configurations, bit-exact against the oracle's
NTT, and the round trip returns the input."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log

pytestmark = pytest.mark.gpu

CONFIGS = [(4, 2, 4, 0), (8, 3, 16, 256), (16, 2, 8, 512), (16, 3, 4, 0), (4, 3, 8, 512), (8, 2, 16, 0)]


@pytest.fixture(scope="module")
def env():
    import pyfideslib_amd as F

    P = O.params_small()
    o = O.Oracle(P, seed=0)
    raw = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots)
    ctx = F.Context(raw)
    yield F, P, o, ctx
    F.tune_force(-1)
    ctx.close()


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("limbs", [1, 6, 25, 48])
def test_forced_configs_bitexact(env, cfg, limbs):
    F, P, o, ctx = env
    np_ = P.L + 1 + P.K
    pids = [i % np_ for i in range(limbs)]
    x = o.random_poly(limbs, 300 + limbs, pids)
    want = np.stack([o.ntt(np.ascontiguousarray(x[i:i + 1]), [pids[i]], inverse=False)[0] for i in range(limbs)])
    log(f"ntt_limbs {limbs} limbs under {cfg}")
    d = ctx.device_array(x.shape, x)
    try:
        F.tune_force(F.TUNE_NTT_FWD, cfg)
        F.tune_force(F.TUNE_NTT_INV, cfg)
        ctx.ntt_limbs(d, limbs, inverse=False)
        ctx.synchronize()
        assert np.array_equal(d.numpy(), want)
        ctx.ntt_limbs(d, limbs, inverse=True)
        ctx.synchronize()
        assert np.array_equal(d.numpy(), x)
    finally:
        F.tune_force(-1)


@pytest.mark.parametrize("which", ["small", "production"])
def test_modup_matrix_cores_bitexact(which):
    """The int8 matrix-core ModUp conversion (k_modup_mfma, selected by the
    tuning configuration v1 = 1) writes the same digits as the VALU one at
    levels with one, two and three digits (a ragged last digit included)."""
    import pyfideslib_amd as F

    P = O.params_small() if which == "small" else O.params_production()
    raw = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots)
    ctx = F.Context(raw)
    alpha = -(-(P.L + 1) // P.dnum)
    rng = np.random.default_rng(77)
    mods = list(P.moduli)
    try:
        for level in sorted({0, alpha - 1, min(P.L, alpha + 1), P.L}):
            nd = -(-(level + 1) // alpha)
            log(f"matrix-core ModUp, {which}, level {level}, {nd} digits")
            x = np.stack([rng.integers(0, mods[i], P.N, dtype=np.uint64) for i in range(level + 1)])
            dx = ctx.device_array(x.shape, x)
            zero = np.zeros((nd, level + 1 + P.K, P.N), dtype=np.uint64)
            want = ctx.device_array(zero.shape, zero)
            got = ctx.device_array(zero.shape, zero)
            F.tune_force(F.TUNE_MODUP, [1, 0])
            ctx.stage_modup(level, dx, want)
            F.tune_force(F.TUNE_MODUP, [1, 1])
            ctx.stage_modup(level, dx, got)
            ctx.synchronize()
            assert np.array_equal(got.numpy(), want.numpy()), level
    finally:
        F.tune_force(-1)
        ctx.close()
