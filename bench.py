#!/usr/bin/env python3
"""Benchmark of the MI355X CKKS engine at N=2^16, L=24, dnum=3 (K=8).

Workload (synthetic, seeded; residues uniform in [0, q_i), keys uniform):
  * headline step: one HMult+relin+rescale of two level-24 ciphertexts, plain
    per-op form -> `value` = ops/s, p50/p99 per-op latency from HIP events
  * EvalRotate(1) at level 24, plain per-op
  * NTT: forward + inverse of a 25-limb polynomial (GB/s of algorithmic bytes)
  * batched: 128 ciphertext pairs at levels U{4..24} (seed 0), one launch per
    stage, against the plain per-op loop over the same pairs (bit-exact check)
  * batched rotations: the same 128 ciphertexts, each rotated by 2^U{0..14},
    one launch per stage against the plain per-op loop
  * hoisted rotation set: one level-24 ciphertext, 27 distinct indices
    ({1..16} u {2^k, k <= 15}) sharing one ModUp, against 27 plain rotations
  * BSGS linear transform (64 diagonals, 8 baby x 8 giant steps) at level 24,
    fused against the same ops issued one by one, 8 ciphertexts through one
    batched transform, and a degree-119 Chebyshev series (EvalMod's
    polynomial) at level 24
  * EvalBootstrap of a fully packed ciphertext (2^15 slots, sparse-secret
    encapsulation, level budget 3/3, degree-119 EvalMod) from level 0; a batch
    of 4 bootstrapped together; the variant with batched giant steps
  * serving: Poisson arrivals at a fixed offered rate (default 1000 jobs/s)
    into the request loop, a third each of logistic-regression scoring,
    HMult+relin+rescale and power-of-two rotations; throughput and p50/p99,
    then the same jobs submitted at once (saturated throughput)
  * the autotuned launch configurations (pyfideslib_amd/tune_gfx950.txt)
    the workload uses, and HMult / NTT on the library's fixed rules beside them
  * roofline of the dominant kernel (forward NTT, two launches) from HIP events
  * cpu_baseline: the CPU oracle's HMult+relin+rescale on a bounded sample
Prints ONE JSON line.  Only one GPU is supported.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# HBM bytes moved by one forward NTT of 25 limbs at N=2^16 (both launches),
# from rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this benchmark scaled by
# what the same counters read for the device-to-device copy kernel of a known
# 12.5 MiB: 12.5 MiB in and out per pass plus the row pass's 16 B x N twiddle
# table per limb.  Re-measure when the NTT kernels change.
def ntt_traffic_bytes():
    """HBM bytes moved by one forward NTT of 25 limbs at N=2^16 (both launches):
    the figure profiles/hbm_traffic_N65536_L24_kernels.md records from its
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this benchmark, scaled by
    what the same counters read for a device copy of a known 12.5 MiB; None
    when the file is absent."""
    import re

    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_N65536_L24_kernels.md")) as f:
            m = re.search(r"total HBM traffic per forward NTT: ([\d,]+) B", f.read())
        return int(m.group(1).replace(",", "")) if m else None
    except OSError:
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-hoisted", action="store_true")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-lt", action="store_true")
    ap.add_argument("--skip-boot", action="store_true")
    ap.add_argument("--conv-mfma", action="store_true",
                    help="run every ModUp base conversion on the int8 matrix cores (k_modup_mfma)")
    ap.add_argument("--serve-qps", type=float, default=1000.0, help="offered load (Poisson arrivals, jobs/s)")
    ap.add_argument("--serve-jobs", type=int, default=600)
    ap.add_argument("--profile-lt", choices=("q", "ext"), default=None,
                    help="run only --steps 8x8 linear transforms at level 24 (Q-basis or P*Q diagonals), for a "
                         "kernel profile")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def rand_poly(rng, moduli, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in moduli])


def timed(ctx, fn, steps, warmup):
    """Run fn warmup+steps times; returns (wall seconds of the timed steps,
    per-step event milliseconds)."""
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    evs = [(ctx.event(), ctx.event()) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        fn()
        b.record()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    return wall, [a.elapsed_ms(b) for a, b in evs]


def timed_block(ctx, fn, n):
    """Average event milliseconds of n back-to-back calls bracketed by one
    event pair (launch-to-launch time, without per-call event records)."""
    fn()
    ctx.synchronize()
    a, b = ctx.event(), ctx.event()
    a.record()
    for _ in range(n):
        fn()
    b.record()
    ctx.synchronize()
    return a.elapsed_ms(b) / n


def cpu_baseline(raw, seconds):
    """Oracle (C port of OpenFHE's CPU algorithms) HMult+relin+rescale at the
    same parameters, on this host's cores, on a time-bounded sample."""
    from oracle import oracle as O

    P = O.Params(raw.logN, raw.L, raw.K, raw.dnum, raw.moduli + raw.special_moduli,
                 raw.root_of_unity + raw.special_root_of_unity, 50)
    o = O.Oracle(P, 0)
    rng = np.random.default_rng(7)
    np_ = raw.L + 1 + raw.K
    key = tuple(np.ascontiguousarray(np.stack([rand_poly(rng, P.moduli, P.N) for _ in range(raw.dnum)]))
                for _ in range(2))
    c1 = tuple(rand_poly(rng, P.q, P.N) for _ in range(2))
    c2 = tuple(rand_poly(rng, P.q, P.N) for _ in range(2))

    def sample(budget, at_least):
        times = []
        t_end = time.perf_counter() + budget
        while time.perf_counter() < t_end or len(times) < at_least:
            t = time.perf_counter()
            o.rescale(o.mult(c1, c2, key))
            times.append(time.perf_counter() - t)
        return times

    times = sample(seconds * 2 / 3, 3)
    cores = int(O.lib().orc_get_threads())
    O.lib().orc_set_threads(1)
    try:
        times1 = sample(seconds / 3, 1)
    finally:
        O.lib().orc_set_threads(0)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(len(times) / sum(times), 3), "unit": "HMult+relin+rescale ops/s", "cores": cores,
            "kind": "port", "cpu_model": model, "value_1_thread": round(len(times1) / sum(times1), 3),
            "sample": f"{len(times)} HMult+relin+rescale ops at N=2^16 L=24 dnum=3 (random residues) on {cores} "
                      f"threads (OpenMP over limbs), median {np.median(times) * 1e3:.1f} ms; then {len(times1)} on "
                      f"1 thread, median {np.median(times1) * 1e3:.1f} ms; oracle built -O3 -march=x86-64-v3"}


def linear_transform(ctx, srcs, rand_poly, rng, Q, N, args):
    """BSGS linear transform at level 24 with 64 diagonals (8 baby x 8 giant
    steps, stride 1): the fused form against the same ops issued one by one
    (hoisted baby steps, then per-diagonal ct x pt and adds, Horner giant
    steps) -- bit-exact check between the two."""
    nd, b = 64, 8
    pts = [ctx.plaintext(rand_poly(rng, Q, N), noise=2.0 ** 50) for _ in range(nd)]
    a = srcs[3]
    out_f = ctx.ciphertext()

    out_g = ctx.ciphertext()

    def fused():
        a.linear_transform(pts, b, 1, 0, out=out_f)

    def fused_batched_giant():
        a.linear_transform(pts, b, 1, 0, out=out_g, batched_giant=True)

    baby = [ctx.ciphertext() for _ in range(b - 1)]
    prod, acc_j, rot = ctx.ciphertext(), [ctx.ciphertext() for _ in range(nd // b)], ctx.ciphertext()

    def per_op():
        a.rotate_hoisted(list(range(1, b)), baby)
        srcs_b = [a] + baby
        for j in range(nd // b):
            for i in range(b):
                if i == 0:
                    srcs_b[i].mult_plain(pts[b * j], out=acc_j[j])
                else:
                    srcs_b[i].mult_plain(pts[b * j + i], out=prod)
                    acc_j[j].add(prod, out=acc_j[j])
        acc = acc_j[-1]
        for j in range(nd // b - 2, -1, -1):
            acc.rotate(b, out=rot)
            acc_j[j].add(rot, out=acc_j[j])
            acc = acc_j[j]
        return acc

    # EvalMod's polynomial: a degree-119 Chebyshev series at level 24
    cheb = [float(v) for v in np.random.default_rng(3).normal(0, 1, 120) / (np.arange(120) + 1.0) ** 2]
    out_c = ctx.ciphertext()
    _, msc = timed(ctx, lambda: a.chebyshev_series(cheb, out=out_c), 5, 1)
    wf, msf = timed(ctx, fused, 10, 2)
    wg, msg = timed(ctx, fused_batched_giant, 10, 2)
    # 8 ciphertexts through one batched transform (one dot launch, batched giant steps)
    many = [a.copy() for _ in range(8)]
    outs8 = [ctx.ciphertext() for _ in many]
    _, ms8 = timed(ctx, lambda: ctx.linear_transform_batch(many, pts, b, 1, 0, outs=outs8), 5, 1)
    wp, msp = timed(ctx, per_op, 10, 2)
    # the same transform with P*Q diagonals: the extended-basis path (baby
    # steps and inner sums in P*Q, only c1 brought down per giant step)
    SP = list(ctx.raw.special_moduli)
    pts_x = [ctx.plaintext_ext(rand_poly(rng, list(Q) + SP, N), noise=2.0 ** 50) for _ in range(nd)]
    out_x, out_xg = ctx.ciphertext(), ctx.ciphertext()
    _, msx = timed(ctx, lambda: a.linear_transform(pts_x, b, 1, 0, out=out_x), 10, 2)
    _, msxg = timed(ctx, lambda: a.linear_transform(pts_x, b, 1, 0, out=out_xg, batched_giant=True), 10, 2)
    _, msx8 = timed(ctx, lambda: ctx.linear_transform_batch(many, pts_x, b, 1, 0, outs=outs8), 5, 1)
    del pts_x
    ref = acc_j[0].store()
    got = out_f.store()
    same = bool(np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]))
    out = {"lt_diagonals": nd, "lt_bstep": b, "lt_ms": round(float(np.median(msf)), 3),
           "lt_per_op_ms": round(float(np.median(msp)), 3),
           "lt_speedup": round(float(np.median(msp)) / float(np.median(msf)), 3), "lt_bitexact_vs_per_op": same,
           "lt_batched_giant_ms": round(float(np.median(msg)), 3),
           "lt_batched_giant_speedup": round(float(np.median(msp)) / float(np.median(msg)), 3),
           "chebyshev119_ms": round(float(np.median(msc)), 3), "chebyshev119_out_level": out_c.level,
           "lt_batch8_ms_per_ct": round(float(np.median(ms8)) / 8, 3),
           "lt_batch8_speedup": round(float(np.median(msf)) * 8 / float(np.median(ms8)), 3),
           "lt_ext_ms": round(float(np.median(msx)), 3),
           "lt_ext_batched_giant_ms": round(float(np.median(msxg)), 3),
           "lt_ext_batch8_ms_per_ct": round(float(np.median(msx8)) / 8, 3),
           "lt_ext_gain_over_q_basis": round(float(np.median(msf)) / float(np.median(msx)), 3)}
    log(f"linear transform 64 diagonals (8x8) level 24: fused {out['lt_ms']} ms, batched giant steps "
        f"{out['lt_batched_giant_ms']} ms vs per-op {out['lt_per_op_ms']} ms (x{out['lt_speedup']}, "
        f"x{out['lt_batched_giant_speedup']}), bit-exact {same}; 8 ciphertexts batched "
        f"{out['lt_batch8_ms_per_ct']} ms each (x{out['lt_batch8_speedup']}); degree-119 Chebyshev series "
        f"{out['chebyshev119_ms']} ms (level 24 -> {out['chebyshev119_out_level']})")
    log(f"  P*Q diagonals (extended-basis path): {out['lt_ext_ms']} ms (x{out['lt_ext_gain_over_q_basis']} over "
        f"Q-basis), batched giant steps {out['lt_ext_batched_giant_ms']} ms, 8 batched "
        f"{out['lt_ext_batch8_ms_per_ct']} ms each")
    return out


def serve(ctx, F, srcs, rand_poly, rng, Q, N, args):
    """Poisson arrivals at a fixed offered rate into the request loop; a third
    each of logistic-regression scoring (level 24, 8 rotate-and-sum steps,
    degree-3 polynomial), HMult+relin+rescale at a level U{4..24} and a
    rotation by 2^U{0..15}.  Latency: scheduled arrival -> completion seen by
    the host.  A second run submits every job at once (saturated throughput)."""
    srng = np.random.default_rng(1)
    pool_a, pool_b = [], []
    for i in range(24):  # resident inputs at levels 4..24 (device-side copies of the sources)
        lv = 4 + i % 21
        for pool, s in ((pool_a, srcs[i % 4]), (pool_b, srcs[(i + 1) % 4])):
            c = ctx.ciphertext()
            p0, p1 = s.device_ptrs()
            F._check(F.lib().fides_ct_load_device(ctx.h, c.h, p0, p1, lv + 1, 2.0 ** 50, 1, N // 2))
            pool.append(c)
    w = ctx.plaintext(rand_poly(rng, Q, N), noise=2.0 ** 50)
    xs = srcs[:4]
    coeffs = (0.5, 0.197, -0.004)
    kinds = srng.integers(0, 3, args.serve_jobs)
    picks = srng.integers(0, len(pool_a), args.serve_jobs)
    rots = [int(1 << x) for x in srng.integers(0, 16, args.serve_jobs)]
    free = []

    def take():
        return free.pop() if free else ctx.ciphertext()

    def submit(srv, i):
        o = take()
        k = int(kinds[i])
        if k == 0:
            jid = srv.submit_logreg(xs[i % 4], w, 8, coeffs, o)
        elif k == 1:
            jid = srv.submit_mult(pool_a[picks[i]], pool_b[picks[i]], o)
        else:
            jid = srv.submit_rotate(pool_a[picks[i]], rots[i], o)
        return jid, o

    def run(arrivals):
        srv = F.Server(ctx, max_slots=256, budget_bytes=96 << 30)
        outs, t_arr, lat, kind = {}, {}, [], []
        t_kind = {}
        n = len(arrivals)
        nxt = done = 0
        t0 = time.perf_counter()
        t_end = t0
        while done < n:
            now = time.perf_counter() - t0
            while nxt < n and arrivals[nxt] <= now:
                jid, o = submit(srv, nxt)
                outs[jid] = o
                t_arr[jid] = arrivals[nxt]
                t_kind[jid] = int(kinds[nxt])
                nxt += 1
            pending = srv.step()
            for jid in srv.poll():
                t = time.perf_counter() - t0
                lat.append(t - t_arr.pop(jid))
                kind.append(t_kind.pop(jid))
                free.append(outs.pop(jid))
                done += 1
                t_end = t
            if pending == 0 and nxt < n and done < nxt:
                ctx.synchronize()
            elif pending == 0 and nxt < n:
                while time.perf_counter() - t0 < arrivals[nxt]:
                    pass
        srv.close()
        return n / (t_end - arrivals[0]), (np.array(lat), np.array(kind))

    n = args.serve_jobs
    arrivals = np.cumsum(srng.exponential(1.0 / args.serve_qps, n))
    arrivals -= arrivals[0]
    run(np.zeros(min(n, 64)))  # warm-up (plan build paths, pool)
    thr, (lat, lk) = run(arrivals)
    sat, (sat_lat, _) = run(np.zeros(n))
    names = ("logreg", "mult", "rotate")
    by_kind = {nm: [round(float(np.percentile(lat[lk == k], q)) * 1e3, 3) for q in (50, 99)]
               for k, nm in enumerate(names) if np.any(lk == k)}
    out = {"serve_offered_qps": args.serve_qps, "serve_jobs": n,
           "serve_mix": "1/3 logreg (level 24, 8 rotations, degree 3), 1/3 HMult+relin+rescale U{4..24}, "
                        "1/3 rotation by 2^U{0..15}",
           "serve_jobs_per_s": round(thr, 1), "serve_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "serve_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "serve_p50_p99_ms_by_kind": by_kind,
           "serve_saturated_jobs_per_s": round(sat, 1),
           "serve_saturated_p50_ms": round(float(np.percentile(sat_lat, 50)) * 1e3, 3),
           "serve_saturated_p99_ms": round(float(np.percentile(sat_lat, 99)) * 1e3, 3)}
    log(f"serve @ {args.serve_qps:.0f} qps offered: {out['serve_jobs_per_s']} jobs/s, p50 {out['serve_p50_ms']} ms "
        f"p99 {out['serve_p99_ms']} ms; saturated {out['serve_saturated_jobs_per_s']} jobs/s "
        f"(p50 {out['serve_saturated_p50_ms']} ms, p99 {out['serve_saturated_p99_ms']} ms); p50/p99 by kind {by_kind}")
    return out


def bootstrap_leg(ctx, key, loaded, rand_poly, rng, Q, N):
    """EvalBootstrap of a fully packed ciphertext (2^15 slots) from level 0 to
    level 24 and back down: sparse-secret encapsulation around ModRaise,
    CoeffsToSlots and SlotsToCoeffs in 3 levels each (32 / 63 / 63
    diagonals), EvalMod of degree 119 (K = 28, 3 double-angle steps) on the
    real and the imaginary half.  Keys are uniform like the rest
    of the bench, so the kernels and sizes are those of a real bootstrap and the
    result is not decryptable."""
    import pyfideslib_amd as F
    from pyfideslib_amd import bootstrap as B

    t0 = time.perf_counter()
    bs = B.Bootstrapper(ctx, 2.0 ** 50, level_budget=(3, 3), K=28, double_angle=3, degree=119, sparse_encaps=True)
    ctx.load_switch_key(F.SSE_DENSE_TO_SPARSE, key)  # sparse-secret encapsulation: at level 0
    ctx.load_switch_key(F.SSE_SPARSE_TO_DENSE, key)  # and back after ModRaise
    for r in bs.rotations:  # keys already loaded stay (captured plans may hold them)
        if r not in loaded:
            ctx.load_rotation_key(r, key)
            loaded.add(r)
    ctx.load_automorph_key(2 * N - 1, key)
    setup_s = time.perf_counter() - t0
    ct = ctx.load(rand_poly(rng, Q[:1], N), rand_poly(rng, Q[:1], N), noise=2.0 ** 50)
    holder = {}

    def run():
        holder["out"] = bs.bootstrap(ct)

    _, ms = timed(ctx, run, 3, 1)
    # batched bootstrapping: 4 ciphertexts together (8 EvalMod halves in lockstep)
    cts4 = [ctx.load(rand_poly(rng, Q[:1], N), rand_poly(rng, Q[:1], N), noise=2.0 ** 50) for _ in range(4)]
    _, ms4 = timed(ctx, lambda: bs.bootstrap_batch(cts4), 3, 1)
    bs.close()
    # the same with each transform's giant steps as one batched rotation (more keys)
    bsg = B.Bootstrapper(ctx, 2.0 ** 50, level_budget=(3, 3), K=28, double_angle=3, degree=119, sparse_encaps=True,
                         batched_giant=True)
    for r in bsg.rotations:
        if r not in loaded:
            ctx.load_rotation_key(r, key)
            loaded.add(r)
    _, msg = timed(ctx, lambda: bsg.bootstrap(ct), 3, 1)
    bsg.close()
    # Q-basis diagonals (each rotation of the transforms ModDowns at once)
    bsq = B.Bootstrapper(ctx, 2.0 ** 50, level_budget=(3, 3), K=28, double_angle=3, degree=119, sparse_encaps=True,
                         q_basis_diagonals=True)
    _, msq = timed(ctx, lambda: bsq.bootstrap(ct), 3, 1)
    bsq.close()
    out = {"bootstrap_ms": round(float(np.median(ms)), 2), "bootstrap_setup_s": round(setup_s, 1),
           "bootstrap_batch4_ms_per_ct": round(float(np.median(ms4)) / 4, 2),
           "bootstrap_batched_giant_ms": round(float(np.median(msg)), 2),
           "bootstrap_batched_giant_rotation_keys": len(bsg.rotations) + 1,
           "bootstrap_q_basis_diagonals_ms": round(float(np.median(msq)), 2),
           "bootstrap_batch4_speedup": round(float(np.median(ms)) * 4 / float(np.median(ms4)), 3),
           "bootstrap_out_level": holder["out"].level, "bootstrap_rotation_keys": len(bs.rotations) + 1,
           "bootstrap_config": "2^15 slots, sparse-secret encapsulation, level budget (3, 3), EvalMod degree 119, "
                               "K 28, r 3, level 0 -> 24 -> "
                               f"{holder['out'].level}"}
    log(f"bootstrap (2^15 slots, SSE, budget 3/3, degree 119): {out['bootstrap_ms']} ms, output level "
        f"{out['bootstrap_out_level']}, setup {out['bootstrap_setup_s']} s; batch of 4: "
        f"{out['bootstrap_batch4_ms_per_ct']} ms per ciphertext (x{out['bootstrap_batch4_speedup']}); batched giant "
        f"steps {out['bootstrap_batched_giant_ms']} ms ({out['bootstrap_batched_giant_rotation_keys']} keys); "
        f"Q-basis diagonals {out['bootstrap_q_basis_diagonals_ms']} ms")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); multi-GPU sharding is not built "
              "-- see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import pyfideslib_amd as F
    from pyfideslib_amd import params as PP

    t_setup = time.perf_counter()
    raw = PP.production()
    N, L, K, dnum = raw.N, raw.L, raw.K, raw.dnum
    Q = raw.moduli
    ALL = raw.moduli + raw.special_moduli
    ctx = F.Context(raw, 0)
    if args.conv_mfma:
        F.tune_force(F.TUNE_MODUP, [1, 1])
    rng = np.random.default_rng(0)
    key = tuple(np.ascontiguousarray(np.stack([rand_poly(rng, ALL, N) for _ in range(dnum)])) for _ in range(2))
    ctx.load_eval_key(key)
    rot_idx = sorted(set(range(1, 17)) | {1 << k for k in range(16)})
    loaded = set(rot_idx) | {8 * j for j in range(1, 8)}  # + the linear transform's giant steps
    for r in sorted(loaded):
        ctx.load_rotation_key(r, key)
    srcs = []
    for i in range(4):
        srcs.append(ctx.load(rand_poly(rng, Q, N), rand_poly(rng, Q, N), noise=2.0 ** 50))
    log(f"setup {time.perf_counter() - t_setup:.1f}s, pool {ctx.pool_bytes() / 2**30:.2f} GiB")
    if args.profile_lt:
        SP = list(raw.special_moduli)
        mods = list(Q) + SP if args.profile_lt == "ext" else list(Q)
        load = ctx.plaintext_ext if args.profile_lt == "ext" else ctx.plaintext
        pts = [load(rand_poly(rng, mods, N), noise=2.0 ** 50) for _ in range(64)]
        out = ctx.ciphertext()
        w, ms = timed(ctx, lambda: srcs[3].linear_transform(pts, 8, 1, 0, out=out), args.steps, args.warmup)
        print(json.dumps({"profile_lt": args.profile_lt, "steps": args.steps, "ms_per_lt": round(w / args.steps * 1e3, 3),
                          "p50_ms": round(pct(ms, 50), 3)}))
        ctx.close()
        return
    res = {}

    # ---------------------------------------------------------------- NTT ----
    nl = L + 1
    poly = ctx.device_array((nl, N), rand_poly(rng, Q, N))
    ref_poly = poly.numpy()
    fwd = lambda: ctx.ntt(poly, nl, 0, False)  # noqa: E731
    inv = lambda: ctx.ntt(poly, nl, 0, True)  # noqa: E731
    w_f, ms_f = timed(ctx, fwd, args.steps, args.warmup)
    w_i, ms_i = timed(ctx, inv, args.steps, args.warmup)
    # HBM reference on the same box: device-to-device copy of a 25-limb
    # ciphertext (two 12.5 MiB copies, read + write)
    cdst = srcs[0].copy()
    cp_ms = timed_block(ctx, lambda: F._check(F.lib().fides_ct_copy(ctx.h, srcs[1].h, cdst.h)), 100)
    res["d2d_copy_GBps"] = round(2 * 2 * 8 * N * nl / (cp_ms * 1e-3) / 1e9, 1)
    # average forward launch time over a back-to-back run (the roofline figure)
    reps = max(4 * args.steps, 100)
    fwd_avg_ms = timed_block(ctx, fwd, reps)
    timed_block(ctx, inv, reps)
    # forward/inverse counts match, so the data is back to the input
    assert np.array_equal(poly.numpy(), ref_poly), "NTT round trip mismatch"
    ntt_bytes = 2 * 8 * N * nl  # one transform of 25 limbs: read + write each limb once
    fwd_ms = float(np.median(ms_f))
    inv_ms = float(np.median(ms_i))
    res["ntt_fwd_GBps"] = round(ntt_bytes / (fwd_ms * 1e-3) / 1e9, 1)
    res["ntt_inv_GBps"] = round(ntt_bytes / (inv_ms * 1e-3) / 1e9, 1)
    res["ntt_fwd_inv_us"] = round((fwd_ms + inv_ms) * 1e3, 2)
    res["ntt_GBps"] = round(2 * ntt_bytes / ((fwd_ms + inv_ms) * 1e-3) / 1e9, 1)
    log(f"NTT fwd {fwd_ms * 1e3:.1f} us inv {inv_ms * 1e3:.1f} us ({res['ntt_GBps']} GB/s)")

    # ----------------------------------------------- HMult+relin+rescale ----
    a, b = srcs[0], srcs[1]
    out = ctx.ciphertext()
    mult = lambda: a.mult(b, rescale=True, out=out)  # noqa: E731
    wall, ms = timed(ctx, mult, args.steps, args.warmup)
    value = args.steps / wall
    res["hmult_p50_ms"] = round(pct(ms, 50), 4)
    res["hmult_p99_ms"] = round(pct(ms, 99), 4)
    log(f"HMult+relin+rescale plain: {value:.1f} ops/s, p50 {pct(ms, 50):.3f} ms p99 {pct(ms, 99):.3f} ms")

    # ------------------------------------------------------------ tuning ----
    # the launch configurations the measured table picks for this workload,
    # and the same NTT / HMult on the fixed rules the library falls back to
    tuned = {"table_entries": F.tune_load()}
    for name, kind, size in (("ntt_fwd_25", F.TUNE_NTT_FWD, nl), ("ntt_inv_25", F.TUNE_NTT_INV, nl),
                             ("ntt_fwd_74", F.TUNE_NTT_FWD, 74), ("modup_3", F.TUNE_MODUP, 3),
                             ("moddown_2", F.TUNE_MODDOWN, 2)):
        cfg, from_table = ctx.tune_config(kind, size)
        tuned[name] = cfg if kind in (F.TUNE_NTT_FWD, F.TUNE_NTT_INV) else cfg[0]
    # tuned table and fixed rules measured alternately, back to back, the
    # median of three rounds each (so clock drift over the run cancels)
    nb = max(2 * args.steps, 50)
    t_ntt = {True: [], False: []}
    t_mul = {True: [], False: []}
    for _ in range(3):
        for use_table in (True, False):
            F.tune_load() if use_table else F.tune_load(None)
            t_ntt[use_table].append(timed_block(ctx, fwd, nb))
            timed_block(ctx, inv, nb)  # the poly back to its input
            t_mul[use_table].append(timed_block(ctx, mult, nb))
    F.tune_load()
    med = lambda v: float(np.median(v))  # noqa: E731
    tuned["ntt_fwd_25_us_table"] = round(med(t_ntt[True]) * 1e3, 2)
    tuned["ntt_fwd_25_us_fixed_rules"] = round(med(t_ntt[False]) * 1e3, 2)
    tuned["hmult_ms_table"] = round(med(t_mul[True]), 4)
    tuned["hmult_ms_fixed_rules"] = round(med(t_mul[False]), 4)
    tuned["hmult_gain_over_fixed_rules"] = round(med(t_mul[False]) / med(t_mul[True]), 3)
    res["tuned"] = tuned
    log(f"tuned: {tuned}")

    rout = ctx.ciphertext()
    rot = lambda: a.rotate(1, out=rout)  # noqa: E731
    wall_r, ms_r = timed(ctx, rot, args.steps, args.warmup)
    res["rotate_ops_per_s"] = round(args.steps / wall_r, 1)
    res["rotate_p50_ms"] = round(pct(ms_r, 50), 4)
    res["rotate_p99_ms"] = round(pct(ms_r, 99), 4)
    log(f"EvalRotate(1) plain: {res['rotate_ops_per_s']} ops/s, p50 {pct(ms_r, 50):.3f} ms")

    # ------------------------------------------------------------ batched ----
    if not args.skip_batch:
        brng = np.random.default_rng(0)
        levels = [int(x) for x in brng.integers(4, 25, args.batch)]
        c0a, c1a = srcs[0].device_ptrs()
        A, B, Op, Ob = [], [], [], []
        for i, lv in enumerate(levels):
            sa, sb = srcs[i % 4], srcs[(i + 1) % 4]
            A.append(ctx.ciphertext())
            B.append(ctx.ciphertext())
            for dst, s in ((A[-1], sa), (B[-1], sb)):
                p0, p1 = s.device_ptrs()
                F._check(F.lib().fides_ct_load_device(ctx.h, dst.h, p0, p1, lv + 1, 2.0 ** 50, 1, N // 2))
            Op.append(ctx.ciphertext())
            Ob.append(ctx.ciphertext())

        def plain_batch():
            for x, y, o in zip(A, B, Op):
                x.mult(y, rescale=True, out=o)

        def batched():
            ctx.mult_batch(A, B, Ob, rescale=True)

        wp, msp = timed(ctx, plain_batch, args.batch_steps, 1)
        wb, msb = timed(ctx, batched, args.batch_steps, 1)
        same = all(np.array_equal(x, y) for i in range(0, args.batch, max(1, args.batch // 16))
                   for x, y in zip(Op[i].store(), Ob[i].store()))
        res["batch_pairs"] = args.batch
        res["batch_levels"] = "U{4..24} seed 0"
        res["batch_plain_ops_per_s"] = round(args.batch * args.batch_steps / wp, 1)
        res["batch_batched_ops_per_s"] = round(args.batch * args.batch_steps / wb, 1)
        res["batch_speedup"] = round(res["batch_batched_ops_per_s"] / res["batch_plain_ops_per_s"], 3)
        res["batch_bitexact_vs_plain"] = bool(same)
        log(f"batch {args.batch}: plain {res['batch_plain_ops_per_s']} ops/s, batched "
            f"{res['batch_batched_ops_per_s']} ops/s (x{res['batch_speedup']}), bit-exact {same}")

        # batched rotations: the same 128 ciphertexts, each by a random power of two
        ridx = [int(1 << x) for x in brng.integers(0, 15, args.batch)]

        def plain_rot_batch():
            for x, r, o in zip(A, ridx, Op):
                x.rotate(r, out=o)

        def batched_rot():
            ctx.rotate_batch(A, ridx, Ob)

        wp, _ = timed(ctx, plain_rot_batch, args.batch_steps, 1)
        wb, _ = timed(ctx, batched_rot, args.batch_steps, 1)
        same = all(np.array_equal(x, y) for i in range(0, args.batch, max(1, args.batch // 16))
                   for x, y in zip(Op[i].store(), Ob[i].store()))
        res["batch_rot_plain_ops_per_s"] = round(args.batch * args.batch_steps / wp, 1)
        res["batch_rot_batched_ops_per_s"] = round(args.batch * args.batch_steps / wb, 1)
        res["batch_rot_speedup"] = round(res["batch_rot_batched_ops_per_s"] / res["batch_rot_plain_ops_per_s"], 3)
        res["batch_rot_bitexact_vs_plain"] = bool(same)
        log(f"batch rotations {args.batch}: plain {res['batch_rot_plain_ops_per_s']} ops/s, batched "
            f"{res['batch_rot_batched_ops_per_s']} ops/s (x{res['batch_rot_speedup']}), bit-exact {same}")
        del A, B, Op, Ob

    # ------------------------------------------------------------ hoisted ----
    if not args.skip_hoisted:
        src = srcs[2]
        outs_h = [ctx.ciphertext() for _ in rot_idx]
        outs_p = [ctx.ciphertext() for _ in rot_idx]

        def hoisted():
            src.rotate_hoisted(rot_idx, outs_h)

        def plain_rots():
            for r, o in zip(rot_idx, outs_p):
                src.rotate(r, out=o)

        wh, _ = timed(ctx, hoisted, max(2, args.batch_steps), 1)
        wpr, _ = timed(ctx, plain_rots, max(2, args.batch_steps), 1)
        same = all(np.array_equal(x, y) for i in (0, len(rot_idx) // 2, len(rot_idx) - 1)
                   for x, y in zip(outs_h[i].store(), outs_p[i].store()))
        nrep = max(2, args.batch_steps)
        res["hoisted_indices"] = len(rot_idx)
        res["hoisted_rot_per_s"] = round(len(rot_idx) * nrep / wh, 1)
        res["plain_rot_set_per_s"] = round(len(rot_idx) * nrep / wpr, 1)
        res["hoisted_speedup"] = round(res["hoisted_rot_per_s"] / res["plain_rot_set_per_s"], 3)
        res["hoisted_bitexact_vs_plain"] = bool(same)
        log(f"hoisted {len(rot_idx)}: {res['hoisted_rot_per_s']} rot/s vs plain {res['plain_rot_set_per_s']} "
            f"(x{res['hoisted_speedup']}), bit-exact {same}")

    # ---------------------------------------------------- linear transform ----
    if not args.skip_lt:
        res.update(linear_transform(ctx, srcs, rand_poly, rng, Q, N, args))
    if not args.skip_boot:
        res.update(bootstrap_leg(ctx, key, loaded, rand_poly, rng, Q, N))

    # ------------------------------------------------------------ serving ----
    if not args.skip_serve:
        res.update(serve(ctx, F, srcs, rand_poly, rng, Q, N, args))

    # ------------------------------------------------------------ roofline ----
    # forward NTT of 25 limbs (both launches), averaged over a back-to-back run
    achieved = ntt_bytes / (fwd_avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": "ntt_fwd (k_ntt_fwd_cols + k_ntt_fwd_rows, 25 limbs)",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": ntt_traffic_bytes(),
                "launch_us": round(fwd_avg_ms * 1e3, 2), "algorithmic_bytes": ntt_bytes}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(raw, args.cpu_seconds)
        log(f"cpu: {cpu['value']} ops/s on {cpu['cores']} threads")

    line = {
        "metric": "HMult+relin & rotations/s at N=2^16 L=24, batched vs plain; NTT GB/s; p50/p99",
        "value": round(value, 2), "unit": "HMult+relin+rescale ops/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic (uniform residues, seeded)",
        "config": {"workload": "plain per-op HMult+relin+rescale at N=2^16, L=24 (25 limbs), dnum=3, K=8; "
                               "plus EvalRotate(1), NTT fwd+inv 25 limbs, 128-pair batch, 27-index hoisted set",
                   "N": N, "L": L, "dnum": dnum, "K": K},
        **res,
        "modup_conversion": "int8 matrix cores (forced)" if args.conv_mfma else "tuned (VALU unless the table picks the matrix cores)",
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    ctx.synchronize()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
