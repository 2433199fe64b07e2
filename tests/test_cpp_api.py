"""The api-level C++ binding (include/fides_amd_api.hpp, FIDESlib's
api/CryptoContext.hpp surface: shared_ptr ciphertexts, Eval* returning new
results or working in place, OpenFHE levels): it compiles against the library
on the CPU, and on the GPU tests/cpp/api_check.cpp drives it from the raw
interop records; every result is compared bit-exact with the oracle's
composition of the same ops (the program itself checks EvalSub(double, ct),
the Chebyshev coefficient formula on x^2 and the series' c0 convention).
__graft_entry__.build() builds it as tests/cpp/_build/api_check."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "api_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "api_check")
LIBDIR = os.path.join(ROOT, "pyfideslib_amd")


def test_api_binding_compiles(tmp_path):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, "-L", LIBDIR,
           "-lfides_amd", f"-Wl,-rpath,{LIBDIR}", "-o", str(tmp_path / "api_check")]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([str(tmp_path / "api_check")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stdout  # no GPU call without arguments


@pytest.mark.gpu
def test_api_ops_match_oracle(tmp_path):
    from oracle import oracle as O
    from tests.gpu_util import log

    assert os.path.exists(BIN), f"{BIN} not built (run __graft_entry__.build())"
    log("cpp api binding")
    P = O.params_small()
    assert P.scale_bits == 40
    o = O.Oracle(P, seed=7)
    rk, k1, k2, k3 = o.relin_key(), o.rotation_key(1), o.rotation_key(2), o.rotation_key(3)
    rng = np.random.default_rng(7)
    c1 = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), P.L, 71)
    c2 = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), P.L, 72)
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
    assert "api ok" in r.stdout

    def load(name, nl):
        v = np.fromfile(tmp_path / f"api_{name}.bin", dtype=np.uint64).reshape(2, nl, P.N)
        return v[0], v[1]

    def add(x, y):
        return o.add(x[0], y[0]), o.add(x[1], y[1])

    q = np.array(P.moduli[: P.L + 1], dtype=np.uint64)[:, None]

    def neg(x):
        return tuple(np.where(v == 0, v, q[: v.shape[0]] - v) for v in x)

    n = P.N // 2
    i = np.arange(n)
    z = ((i * 37) % 201 - 100) / 128.0 + 1j * (((i * 11) % 97) - 48) / 64.0
    E = O.ExtBasis(o)
    tot = (E.extend(c1[0]), E.extend(c1[1]))
    for rot, k in ((1, k1), (2, k2), (3, k3)):
        hh = E.rotate(c1[0], c1[1], O.rot_index(P.logN, rot), k)
        tot = (E.add(tot[0], hh[0]), E.add(tot[1], hh[1]))
    m = o.mult(c1, c2, rk)
    refs = {
        "mult": (P.L + 1, m),
        "resc": (P.L, o.rescale(m)),
        "rot": (P.L + 1, o.rotate(c1, 1, k1)),
        "fr2": (P.L + 1, o.rotate(c1, 2, k2)),
        "fr3": (P.L + 1, o.rotate(c1, 3, k3)),
        "add": (P.L + 1, add(c1, c2)),
        "sub": (P.L + 1, add(c1, neg(c2))),
        "neg": (P.L + 1, neg(c1)),
        "sq": (P.L + 1, o.mult(c1, c1, rk)),
        "mpt": (P.L + 1, o.mult_plain(c1, pt)),
        "many": (P.L + 1, add(add(c1, c2), c1)),
        "acc": (P.L + 1, E.moddown(tot)),
        "setl": (P.L - 1, tuple(x[: P.L - 1] for x in add(c1, c2))),
        "enc": (P.L + 1, o.mult_plain(c1, o.encode(z, P.L))),
    }
    for name, (nl, ref) in refs.items():
        got = load(name, nl)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), name
