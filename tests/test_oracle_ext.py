"""The oracle's extended-basis (P*Q) restatement pinned by what the mathematics
fixes, on the CPU:

  * ModDown(extend(x)) == x exactly (the special limbs of P x are 0, so the
    conversion adds nothing and the P^-1 scale undoes P);
  * ModDown of an extended-basis rotation equals the Q-basis rotation up to
    the conversion's sign-flip carries: per coefficient the two differ by an
    integer of magnitude <= K (sigma permutes coefficients with signs, and
    the conversion of -v is -conv(v) + c P with c <= K nonzero terms);
  * the extended-basis linear transform decrypts to the Q-basis one's value
    up to key-switching noise (a dropped P scale, a wrong automorphism or a
    transposed diagonal would differ by a value of the size of Q).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_linear import oracle_lt

B, STRIDE, OFFSET, NDIAG = 4, 2, 3, 10


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    o = O.Oracle(P, seed=5)
    rots = sorted({i * STRIDE for i in range(1, B)} | {j * B * STRIDE for j in (1, 2)} | {OFFSET})
    keys = {r: o.rotation_key(r) for r in rots}
    return P, o, O.ExtBasis(o), keys


def coef_centered(o, x, primeids):
    """Coefficient form of x, centered per limb (int64)."""
    c = o.ntt(x, primeids, inverse=True).astype(np.int64)
    q = np.array([o.P.moduli[i] for i in primeids], dtype=np.int64)[:, None]
    return np.where(c > q // 2, c - q, c)


def test_extend_then_moddown_is_identity(env):
    P, o, E, _ = env
    for level in (0, 2, P.L):
        x = o.random_poly(level + 1, 11 + level)
        y = E.extend(x)
        assert y.shape == (level + 1 + P.K, P.N)
        assert not y[level + 1:].any()
        assert np.array_equal(o.moddown(y), x)


@pytest.mark.parametrize("r", [1, 3, -2])
def test_ext_rotation_matches_q_rotation(env, r):
    P, o, E, _ = env
    level = 2
    key = o.rotation_key(r)
    ct = o.encrypt_slots(np.linspace(-1, 1, P.N // 2), level, 40)
    k = O.rot_index(P.logN, r)
    e0, e1 = E.moddown(E.rotate(ct[0], ct[1], k, key))
    q0, q1 = o.rotate(ct, r, key)
    qids = list(range(level + 1))
    for a, b in ((e0, q0), (e1, q1)):
        d = o.add(a, np.where(b == 0, b, np.array(P.q[: level + 1], dtype=np.uint64)[:, None] - b))
        dc = coef_centered(o, d, qids)
        assert np.abs(dc).max() <= P.K
        assert np.array_equal(dc, np.broadcast_to(dc[0], dc.shape))  # one integer per coefficient


@pytest.mark.parametrize("batched", [False, True])
def test_ext_linear_transform_decrypts_like_q_basis(env, batched):
    P, o, E, keys = env
    level = 3
    rng = np.random.default_rng(3)
    ct = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), level, 77)
    pq = E.ids(level)
    diags_pq, diags_q = [], []
    for t in range(NDIAG):
        m = rng.integers(-(1 << 10), 1 << 10, P.N, dtype=np.int64)
        d = o.from_int(m, len(pq), pq)  # evaluation form over P*Q
        diags_pq.append(d)
        diags_q.append(np.ascontiguousarray(d[: level + 1]))
    got = E.linear_transform(ct, diags_pq, keys, B, STRIDE, OFFSET, batched)
    want = oracle_lt(o, ct, diags_q, keys, B, STRIDE, OFFSET, batched)
    assert got[0].shape == want[0].shape
    # decryptions of both, subtracted on the residues, then centered
    s = np.ascontiguousarray(o.s[: level + 1])
    qv = np.array(P.q[: level + 1], dtype=np.uint64)[:, None]

    def dec(c):
        return o.add(c[0], o.mul(c[1], s))

    a, b = dec(got), dec(want)
    diff = o.add(a, np.where(b == 0, b, qv - b))
    dc = coef_centered(o, diff[:1], [0])
    assert np.abs(dc).max() < 2 ** 40, np.abs(dc).max()
    # and the two results are not the same residues (the rounding differs)
    assert not np.array_equal(got[0], want[0])
