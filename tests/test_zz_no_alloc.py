"""fides_stage_modup_noalloc (ModUp with the library's static 16 MiB device
workspace, no allocation) against fides_stage_modup (pool workspace), through
the C ABI, bit-exact (needs a GPU):

  * two calls in a row on one context, the second at a lower level, so the
    static workspace still holds the first call's limbs when the second runs;
  * the size limit at N = 2^16: level 31 fills the 2^21-word workspace exactly
    and matches stage_modup, level 32 is refused with FIDES_E_ARG before any
    device work.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import empty_dev, log, sync, to_dev, to_host

pytestmark = pytest.mark.gpu


def _context(P):
    import pyfideslib_amd as F

    raw = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, 0, [2.0 ** P.scale_bits] * (P.L + 1))
    return F.Context(raw)


def _both(ctx, o, P, level, seed):
    nl = level + 1
    x = o.random_poly(nl, seed)
    shape = (o.num_digits(level), nl + P.K, P.N)
    dx = to_dev(ctx, x)
    # both outputs start from the same fill: ModUp leaves each digit's own positions unwritten
    dwant = to_dev(ctx, np.full(shape, 7, np.uint64))
    dgot = to_dev(ctx, np.full(shape, 7, np.uint64))
    sync(ctx)
    ctx.stage_modup(level, dx, dwant)
    ctx.stage_modup_noalloc(level, dx, dgot)
    sync(ctx)
    return to_host(dwant), to_host(dgot)


def test_noalloc_modup_two_calls_second_smaller():
    P = O.params_production()
    o = O.Oracle(P, seed=3)
    ctx = _context(P)
    try:
        for level, seed in ((24, 101), (7, 102)):
            log(f"noalloc modup N=2^16 level {level}")
            want, got = _both(ctx, o, P, level, seed)
            assert np.array_equal(got, want), f"level {level}: static-workspace ModUp differs"
    finally:
        ctx.close()


def test_noalloc_modup_size_limit():
    import pyfideslib_amd as F

    P = O.make_params(16, 60, 50, 32, 60, 3)  # 33 limbs: (32 + 1) * 2^16 > 2^21
    o = O.Oracle(P, seed=4)
    ctx = _context(P)
    try:
        log("noalloc modup N=2^16 level 31 (workspace full)")
        want, got = _both(ctx, o, P, 31, 201)
        assert np.array_equal(got, want)
        log("noalloc modup N=2^16 level 32 (refused)")
        nl = 33
        dx = empty_dev(ctx, (nl, P.N))
        dext = empty_dev(ctx, (o.num_digits(32), nl + P.K, P.N))
        sync(ctx)
        with pytest.raises(F.FidesError) as e:
            ctx.stage_modup_noalloc(32, dx, dext)
        assert e.value.code == 1  # FIDES_E_ARG
    finally:
        ctx.close()
