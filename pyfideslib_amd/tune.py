"""Autotuner of the engine's launch configurations, run on the GPU it tunes.

For each ring dimension asked for it measures, with HIP events on the
context's stream:

  * the forward and inverse NTT at limb counts 1 .. 8 (L+1+K) (the launch
    shapes of single ops and of batched stages; fides_ntt_limbs) over every
    column tile width (4, 8, 16), values per column thread (4, 8), row tile
    height (4, 8, 16) and grid cap (one block per tile, or 1-2 blocks per CU
    looping over the tiles);
  * the ModUp conversion's target groups at 1, 2 and 3 digit conversions per
    launch (fides_stage_modup at levels with that many digits), and the int8
    matrix-core form of the same conversion against them, and the ModDown
    conversion's at the two conversions of a key switch (fides_stage_keyswitch
    at the top level);

keeps the fastest configuration of each size (the fixed rule's unless the
fastest beats it by more than --margin, so timing noise picks nothing), and
writes the table the
library loads (``tune_gfx950.txt`` next to libfides_amd.so, or
FIDES_TUNE_FILE).  Each size's entry covers the limb counts up to halfway to
the next measured size (geometric midpoint); the largest covers every larger
launch.  The report compares every pick with the fixed rules the library
falls back to.

    python -m pyfideslib_amd.tune [--logn 16 13] [--out FILE] [--reps 20]

The reference has no tuner: its launch shapes are compile-time constants
(src/CKKS/Parameters.cuh:30, src/CKKS/forwardDefs.cuh:38).
"""
from __future__ import annotations

import argparse
import itertools
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tune_gfx950.txt")


def _params(logN: int):
    from . import params as PP

    if logN == 16:
        return PP.production()
    if logN == 13:
        return PP.small()
    return PP.ckks_params(logN, 60, 50, max(3, min(24, (1 << (logN - 10)) + 2)), 60, 3)


def _time(ctx, fn, reps: int) -> float:
    """Median microseconds of `reps` calls, each bracketed by its own events."""
    fn()
    ctx.synchronize()
    evs = [(ctx.event(), ctx.event()) for _ in range(reps)]
    for a, b in evs:
        a.record()
        fn()
        b.record()
    ctx.synchronize()
    return float(np.median([a.elapsed_ms(b) for a, b in evs])) * 1e3


def _ntt_sizes(np_: int) -> list[int]:
    cap = 8 * np_
    sizes = sorted({1, 2, 4, 8, 16, 25, np_, 50, 74, 128, 192, cap})
    return [s for s in sizes if 1 <= s <= cap]


def _buckets(sizes: list[int]) -> list[tuple[int, int]]:
    out = []
    for i, s in enumerate(sizes):
        lo = 1 if i == 0 else out[-1][1] + 1
        hi = (1 << 30) if i + 1 == len(sizes) else int(math.floor(math.sqrt(s * sizes[i + 1])))
        out.append((lo, max(lo, hi)))
    return out


def tune_ntt(F, ctx, raw, reps: int, margin: float, log):
    N = 1 << raw.logN
    np_ = raw.L + 1 + raw.K
    mods = list(raw.moduli) + list(raw.special_moduli)
    rng = np.random.default_rng(1)
    sizes = _ntt_sizes(np_)
    cap = max(sizes)
    host = np.stack([rng.integers(0, mods[i % np_], N, dtype=np.uint64) for i in range(cap)])
    buf = ctx.device_array((cap, N), host)
    n_cu = 256
    cands = [c for c in itertools.product((4, 8, 16), (2, 3), (4, 8, 16), (0, n_cu, 2 * n_cu))
             if not (raw.logN == 17 and c[1] == 2)]
    rows = []
    for kind, inverse in ((F.TUNE_NTT_FWD, False), (F.TUNE_NTT_INV, True)):
        for size in sizes:
            F.tune_force(kind, None)
            rule, _ = ctx.tune_config(kind, size)
            best, best_t, t_rule = None, float("inf"), None
            for c in cands:
                F.tune_force(kind, c)
                t = _time(ctx, lambda: ctx.ntt_limbs(buf, size, inverse), reps)
                if list(c) == rule:
                    t_rule = t
                if t < best_t:
                    best, best_t = c, t
            F.tune_force(kind, None)
            if t_rule is None:
                F.tune_force(kind, rule)
                t_rule = _time(ctx, lambda: ctx.ntt_limbs(buf, size, inverse), reps)
                F.tune_force(kind, None)
            if best_t * (1 + margin) >= t_rule:
                best, best_t = tuple(rule), t_rule
            log(f"  {'inv' if inverse else 'fwd'} {size:4d} limbs: best {best} {best_t:8.2f} us, "
                f"rules {rule} {t_rule:8.2f} us ({t_rule / best_t:.3f}x)")
            rows.append((kind, size, best, best_t, rule, t_rule))
    # the data went through equal numbers of forward and inverse transforms
    # per size only when every timing ran the same count; restore the input
    buf.upload(host)
    return rows, sizes


def tune_conv(F, ctx, raw, reps: int, margin: float, log):
    N = 1 << raw.logN
    L, K, dnum = raw.L, raw.K, raw.dnum
    alpha = -(-(L + 1) // dnum)
    mods = list(raw.moduli) + list(raw.special_moduli)
    rng = np.random.default_rng(2)
    key = tuple(np.ascontiguousarray(np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods])
                                               for _ in range(dnum)])) for _ in range(2))
    ctx.load_eval_key(key)
    rows = []
    # ModUp: one job, nd digit conversions at the highest level with nd digits
    for nd in range(1, dnum + 1):
        level = min(L, nd * alpha - 1)
        if -(-(level + 1) // alpha) != nd:
            continue
        x = ctx.device_array((level + 1, N), np.stack([rng.integers(0, mods[i], N, dtype=np.uint64)
                                                        for i in range(level + 1)]))
        ext = ctx.device_array((nd, level + 1 + K, N))
        targets = L + 1 + K - alpha
        F.tune_force(F.TUNE_MODUP, None)
        rule, _ = ctx.tune_config(F.TUNE_MODUP, nd)
        # (target groups, 0) on the VALU, and (1, 1) the int8 matrix-core conversion
        res = {}
        for cfg in [(tg, 0) for tg in sorted({1, 2, 3, 4, 6, 8, 12, 16} & set(range(1, targets + 1)))] + [(1, 1)]:
            try:
                F.tune_force(F.TUNE_MODUP, list(cfg))
            except F.FidesError:
                continue
            res[cfg] = _time(ctx, lambda: ctx.stage_modup(level, x, ext), reps)
        F.tune_force(F.TUNE_MODUP, None)
        best = min(res, key=res.get)
        rk = (rule[0], 0)
        t_rule = res.get(rk, float("nan"))
        if rk in res and res[best] * (1 + margin) >= t_rule:
            best = rk
        mf = res.get((1, 1), float("nan"))
        log(f"  modup  {nd} conversions (level {level}): best {best} {res[best]:8.2f} us, "
            f"rules tg {rule[0]} {t_rule:8.2f} us, matrix cores {mf:8.2f} us")
        rows.append((F.TUNE_MODUP, nd, list(best), res[best], rule, t_rule))
    # ModDown: the two conversions of a key switch at the top level
    x = ctx.device_array((L + 1, N), np.stack([rng.integers(0, mods[i], N, dtype=np.uint64) for i in range(L + 1)]))
    kb, ka = ctx.device_array((L + 1, N)), ctx.device_array((L + 1, N))
    F.tune_force(F.TUNE_MODDOWN, None)
    rule, _ = ctx.tune_config(F.TUNE_MODDOWN, 2)
    res = {}
    for tg in sorted({1, 2, 3, 4, 6, 8, 12, 16} & set(range(1, L + 2))):
        F.tune_force(F.TUNE_MODDOWN, [tg])
        res[tg] = _time(ctx, lambda: ctx.stage_keyswitch(L, x, kb, ka), reps)
    F.tune_force(F.TUNE_MODDOWN, None)
    best = min(res, key=res.get)
    if rule[0] in res and res[best] * (1 + margin) >= res[rule[0]]:
        best = rule[0]
    log(f"  moddown 2 conversions (level {L}): best tg {best} {res[best]:8.2f} us, "
        f"rules tg {rule[0]} {res.get(rule[0], float('nan')):8.2f} us")
    rows.append((F.TUNE_MODDOWN, 2, [best, 0], res[best], rule, res.get(rule[0], float("nan"))))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--logn", type=int, nargs="+", default=[16, 13])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--margin", type=float, default=0.015, help="relative gain a pick needs over the fixed rule")
    args = ap.parse_args(argv)
    os.environ["FIDES_TUNE_FILE"] = ""  # measure against the fixed rules, not an older table
    import pyfideslib_amd as F

    def log(m):
        print(m, flush=True)

    names = {F.TUNE_NTT_FWD: "ntt_fwd", F.TUNE_NTT_INV: "ntt_inv", F.TUNE_MODUP: "modup", F.TUNE_MODDOWN: "moddown"}
    lines = ["# Launch configurations measured on the GPU by `python -m pyfideslib_amd.tune "
             + " ".join(argv if argv is not None else sys.argv[1:]) + "` (" + time.strftime("%Y-%m-%d")
             + "); loaded by libfides_amd.so.",
             "# kind logN size_lo size_hi v0 v1 v2 v3   (us best / us fixed rules)"]
    for logN in args.logn:
        raw = _params(logN)
        ctx = F.Context(raw, 0)
        log(f"logN {logN}: L {raw.L}, K {raw.K}, dnum {raw.dnum}")
        ntt_rows, sizes = tune_ntt(F, ctx, raw, args.reps, args.margin, log)
        conv_rows = tune_conv(F, ctx, raw, args.reps, args.margin, log)
        bk = dict(zip(sizes, _buckets(sizes)))
        for kind, size, best, t, rule, tr in ntt_rows:
            lo, hi = bk[size]
            lines.append(f"{names[kind]} {logN} {lo} {hi} " + " ".join(str(v) for v in best)
                         + f"   # {size} limbs: {t:.2f} / {tr:.2f}")
        conv_sizes = {}
        for kind, size, best, t, rule, tr in conv_rows:
            conv_sizes.setdefault(kind, []).append((size, best, t, tr))
        for kind, lst in conv_sizes.items():
            lst.sort()
            for i, (size, best, t, tr) in enumerate(lst):
                lo = 1 if i == 0 else lst[i - 1][0] + 1
                hi = size if i + 1 < len(lst) else size
                lines.append(f"{names[kind]} {logN} {lo} {hi} {best[0]} {best[1]} 0 0   # {size} conversions: "
                             f"{t:.2f} / {tr:.2f}")
        ctx.close()
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    log(f"wrote {args.out}")


if __name__ == "__main__":
    sys.exit(main())
