"""The autotuned launch table (CPU: the library's loader and validation, no
device work): the committed table loads entry for entry, covers every limb
count of both NTT directions at N=2^16 and 2^13 without gaps, and invalid
forced configurations are refused."""
import os

import pytest

import pyfideslib_amd as F

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyfideslib_amd", "tune_gfx950.txt")


def _entries():
    out = []
    for line in open(TABLE):
        line = line.split("#")[0].split()
        if line:
            out.append((line[0], int(line[1]), int(line[2]), int(line[3]), [int(v) for v in line[4:8]]))
    return out


def test_table_loads_every_entry():
    n = F.tune_load(TABLE)
    try:
        assert n == len(_entries()) > 0
    finally:
        F.tune_load(TABLE)


def test_ntt_entries_cover_all_limb_counts():
    e = _entries()
    for logn in (16, 13):
        for kind in ("ntt_fwd", "ntt_inv"):
            rng = sorted((lo, hi) for k, lg, lo, hi, _ in e if k == kind and lg == logn)
            assert rng and rng[0][0] == 1 and rng[-1][1] >= 1 << 20
            assert all(b[0] == a[1] + 1 for a, b in zip(rng, rng[1:])), (kind, logn, rng)
    for k, lg, lo, hi, v in e:
        if k.startswith("ntt"):
            assert v[0] in (4, 8, 16) and v[1] in (2, 3) and v[2] in (4, 8, 16) and v[3] >= 0
        else:
            assert 1 <= v[0] <= 64


def test_invalid_forced_configurations_are_refused():
    for kind, cfg in ((F.TUNE_NTT_FWD, [5, 2, 4, 0]), (F.TUNE_NTT_INV, [16, 4, 4, 0]), (F.TUNE_MODUP, [0])):
        with pytest.raises(F.FidesError) as e:
            F.tune_force(kind, cfg)
        assert e.value.code == 1
    with pytest.raises(F.FidesError):
        F.tune_force(7, [4, 2, 4, 0])
    F.tune_force(-1)  # releases everything


def test_unreadable_table_is_an_error():
    with pytest.raises(F.FidesError):
        F.tune_load("/nonexistent/tune.txt")
    F.tune_load(TABLE)
