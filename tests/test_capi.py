"""CPU-side checks of the C ABI library: it loads, exports every function
include/fides_amd.h declares, and rejects bad parameters before touching the
GPU (no compute calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyfideslib_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fides_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fides_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    names = declared_functions()
    assert len(names) >= 30
    L = C.CDLL(F.LIB_PATH)
    for n in names:
        assert hasattr(L, n), f"{n} missing from {F.LIB_PATH}"
    assert sorted(F.SYMBOLS) == names


def test_abi_version_and_errors_without_gpu():
    L = F.lib()
    assert L.fides_abi_version() == 4
    raw = F.RawParams(logN=9, L=0, K=1, dnum=1, moduli=[12289], root_of_unity=[3], special_moduli=[40961],
                      special_root_of_unity=[3])
    with pytest.raises(F.FidesError) as e:  # logN below the supported range: rejected before any HIP call
        F.Context(raw)
    assert e.value.code == 1
    assert "logN" in str(e.value)


def test_rejects_bad_roots_before_hip():
    q = 0x3ffffffffc001  # not checked for primality here; root check fails first
    raw = F.RawParams(logN=13, L=0, K=1, dnum=1, moduli=[q], root_of_unity=[2], special_moduli=[q],
                      special_root_of_unity=[2])
    with pytest.raises(F.FidesError) as e:
        F.Context(raw)
    assert e.value.code == 1


def test_oracle_not_imported_by_product():
    """The product package never loads the oracle."""
    import sys

    src = open(os.path.join(ROOT, "pyfideslib_amd", "__init__.py")).read()
    assert "oracle" not in src.replace("oracle.", "")
    for f in os.listdir(os.path.join(ROOT, "pyfideslib_amd", "csrc")):
        assert "oracle" not in open(os.path.join(ROOT, "pyfideslib_amd", "csrc", f)).read()


def test_param_generator_matches_oracle():
    """The product's parameter generator (primes, minimal roots, K) agrees with
    the oracle's independent implementation."""
    from oracle import oracle as O
    from pyfideslib_amd import params as PP

    for mine, ref in ((PP.small(), O.params_small()), (PP.production(), O.params_production())):
        assert mine.moduli + mine.special_moduli == ref.moduli
        assert mine.root_of_unity + mine.special_root_of_unity == ref.roots
        assert (mine.L, mine.K, mine.dnum) == (ref.L, ref.K, ref.dnum)


@pytest.mark.parametrize("size, stride", [(8, 1), (32, 2), (64, 3), (2, 5), (1, 1)])
def test_accumulate_rotations_follow_reference_rule(size, stride):
    """GetAccumulateRotationIndices(bStep 4, stride, size) (AccumulateBroadcast.cu:14-23):
    for s = stride, 4 stride, .. < stride size: s, 2s, 3s below 4 s and stride size."""
    want = []
    s = stride
    while s < stride * size:
        idx = s
        while idx < s * 4 and idx < stride * size:
            want.append(idx)
            idx += s
        s <<= 2
    assert F.accumulate_rotations(size, stride) == want
    with pytest.raises(F.FidesError):
        F.accumulate_rotations(6, 1)  # not a power of two


@pytest.mark.parametrize("stride, gstep", [(1, 1), (2, 3), (1, 8), (3, 10), (1, 17), (4, 24)])
def test_convolution_rotations_cover_every_rotation(stride, gstep):
    """GetConvolutionTransformRotationIndices (LinearTransform.cu:655-678) lists
    s k for k < min(g, 8) and 8 s k for k < ceil(g / 8); the transform's loops
    (LinearTransform.cu:365-369, :403-410) also rotate the first giant step of a
    block of G by s G, so the list here is every rotation those loops make."""
    got = F.convolution_rotations(stride, gstep)
    used = set()
    nb = -(-gstep // 8)
    for k in range(nb):
        G = min(8, gstep - 8 * k)
        used |= {stride * (G - j) for j in range(G)}
        if k < nb - 1:
            used.add(8 * stride * (nb - 1 - k))
    assert sorted(got) == sorted(used) and len(got) == len(set(got))
    ref = {stride * k for k in range(1, min(gstep, 8))} | {8 * stride * k for k in range(1, nb)}
    assert ref <= set(got) and set(got) - ref <= {stride * min(gstep, 8)}
    with pytest.raises(F.FidesError):
        F.convolution_rotations(1, 0)
