"""The C++ binding (include/fides_amd.hpp): argument checks on the CPU, and on
the GPU a mult+rescale / rotate / hoisted rotate driven through it from
RawParams / RawKeySwitchKey / RawCipherText records, bit-exact against the
oracle.  The check program is tests/cpp/facade_check.cpp; __graft_entry__.build()
builds it next to the library as tests/cpp/_build/facade_check."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "facade_check")
LIBDIR = os.path.join(ROOT, "pyfideslib_amd")


def compile_check(out: str) -> str:
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, "-L", LIBDIR,
           "-lfides_amd", f"-Wl,-rpath,{LIBDIR}", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_facade_compiles_and_checks_arguments(tmp_path):
    exe = compile_check(str(tmp_path / "facade_check"))
    r = subprocess.run([exe, "errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors ok" in r.stdout


@pytest.mark.gpu
def test_facade_ops_match_oracle(tmp_path):
    from oracle import oracle as O
    from tests.gpu_util import log

    assert os.path.exists(BIN), f"{BIN} not built (run __graft_entry__.build())"
    log("cpp facade")
    P = O.params_small()
    o = O.Oracle(P, seed=5)
    rk, k1, k2, k3 = o.relin_key(), o.rotation_key(1), o.rotation_key(2), o.rotation_key(3)
    rng = np.random.default_rng(5)
    c1 = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), P.L, 51)
    c2 = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), P.L, 52)
    head = np.array([P.logN, P.L, P.K, P.dnum] + list(P.moduli) + list(P.roots), dtype=np.uint64)
    head.tofile(tmp_path / "params.bin")
    for name, k in (("relin.bin", rk), ("rot1.bin", k1), ("rot2.bin", k2), ("rot3.bin", k3)):
        np.concatenate([np.ascontiguousarray(k[0]).ravel(), np.ascontiguousarray(k[1]).ravel()]).tofile(tmp_path / name)
    for name, c in (("ct1.bin", c1), ("ct2.bin", c2)):
        np.concatenate([c[0].ravel(), c[1].ravel()]).tofile(tmp_path / name)
    pt = o.encode(rng.uniform(-1, 1, P.N // 2), P.L)
    np.ascontiguousarray(pt).tofile(tmp_path / "pt.bin")
    r = subprocess.run([BIN, "run", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    def load(name, nl):
        v = np.fromfile(tmp_path / name, dtype=np.uint64).reshape(2, nl, P.N)
        return v[0], v[1]

    ref_m = o.rescale(o.mult(c1, c2, rk))
    ref_r = o.rotate(c1, 1, k1)
    ref_p = o.mult_plain(c1, pt)
    r1 = o.rotate(c1, 1, k1)
    lt0, lt1 = o.mult_plain(c1, pt), o.mult_plain(r1, pt)  # two diagonals, one giant step
    ref_lt = (o.add(lt0[0], lt1[0]), o.add(lt0[1], lt1[1]))
    def add(x, y):
        return o.add(x[0], y[0]), o.add(x[1], y[1])

    ref_sum = add(add(c1, c2), c1)
    q = np.array(P.moduli[: P.L + 1], dtype=np.uint64)[:, None]
    ref_neg = tuple(np.where(x == 0, x, q - x) for x in c1)
    # AccumulateSum(4, 1): one radix-4 step, summed in P*Q and brought down once
    E = O.ExtBasis(o)
    tot = (E.extend(c1[0]), E.extend(c1[1]))
    for r, k in ((1, k1), (2, k2), (3, k3)):
        h = E.rotate(c1[0], c1[1], O.rot_index(P.logN, r), k)
        tot = (E.add(tot[0], h[0]), E.add(tot[1], h[1]))
    ref_acc = E.moddown(tot)
    ref_hx = E.moddown(E.rotate(c1[0], c1[1], O.rot_index(P.logN, 1), k1))
    ref_mx = o.mult(c1, c2, rk)
    t1, t2 = o.tensor(c1[0], c1[1], c2[0], c2[1]), o.tensor(c2[0], c2[1], c1[0], c1[1])
    dsum = [o.add(x, y) for x, y in zip(t1, t2)]
    kb, ka = o.keyswitch(dsum[2], rk)
    ref_dot = (o.add(dsum[0], kb), o.add(dsum[1], ka))
    # ConvolutionTransform(bStep 2, idx (0, 1), gStep 2, three plaintexts): u0 = pt a + pt rot1(a),
    # u1 = pt a; one block of 2: rot2(u0) + rot1(u1)
    u0 = add(o.mult_plain(c1, pt), o.mult_plain(r1, pt))
    u1 = o.mult_plain(c1, pt)
    ref_cv = add(o.rotate(u0, 2, k2), o.rotate(u1, 1, k1))
    q = np.array(P.moduli[: P.L + 1], dtype=np.uint64)[:, None]
    ref_sp = (o.add(c1[0], np.where(pt == 0, pt, q - pt)), c1[1])
    for name, got, ref in (("hoistext", load("hoistext.bin", P.L + 1), ref_hx),
                           ("multext", load("multext.bin", P.L + 1), ref_mx),
                           ("dot", load("dot.bin", P.L + 1), ref_dot),
                           ("conv", load("conv.bin", P.L + 1), ref_cv),
                           ("subpt", load("subpt.bin", P.L + 1), ref_sp),
                           ("mult", load("mult.bin", P.L), ref_m), ("rot", load("rot.bin", P.L + 1), ref_r),
                           ("hoist", load("hoist.bin", P.L + 1), ref_r), ("mpt", load("mpt.bin", P.L + 1), ref_p),
                           ("lt", load("lt.bin", P.L + 1), ref_lt), ("srv", load("srv.bin", P.L), ref_m),
                           ("addmany", load("addmany.bin", P.L + 1), ref_sum),
                           ("addalias", load("addalias.bin", P.L + 1), add(c1, c2)),
                           ("neg", load("neg.bin", P.L + 1), ref_neg), ("acc", load("acc.bin", P.L + 1), ref_acc)):
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), name
