"""Serialization without a GPU: FIDESlib's `.dev` sidecar written byte for
byte in the reference's text form (api/Serialize.cpp:35-65) and parsed as its
reader does (:117-190), and the engine's parameter file round trip."""
from __future__ import annotations

import numpy as np
import pytest

import pyfideslib_amd as F
from oracle import oracle as O


def test_sidecar_text_is_the_reference_form(tmp_path):
    p = tmp_path / "cc.bin.dev"
    F.write_dev_sidecar(str(p), devices=[0, 3], auto_load_ciphertexts=True, auto_load_plaintexts=False,
                        rotation_indexes=[1, -2, 16], key_dist=1)
    assert p.read_bytes() == (b"2 { 0 3 }\nAutoLoadCiphertexts: 1\nAutoLoadPlaintexts: 0\n"
                              b"RotationIndexes: { 1 -2 16 }\nKeyDist: 1\n")
    assert F.read_dev_sidecar(str(p)) == {"devices": [0, 3], "auto_load_ciphertexts": True,
                                          "auto_load_plaintexts": False, "rotation_indexes": [1, -2, 16],
                                          "key_dist": 1}
    F.write_dev_sidecar(str(p), devices=[], rotation_indexes=[])
    assert p.read_bytes() == (b"0 { }\nAutoLoadCiphertexts: 1\nAutoLoadPlaintexts: 0\nRotationIndexes: { }\n"
                              b"KeyDist: 0\n")
    assert F.read_dev_sidecar(str(p))["rotation_indexes"] == []


def test_sidecar_reader_defaults_and_spacing(tmp_path):
    p = tmp_path / "x.dev"
    # a file cut after two lines keeps the defaults for the rest
    p.write_bytes(b"1 {  5 }\nAutoLoadCiphertexts: 0\n")
    assert F.read_dev_sidecar(str(p)) == {"devices": [5], "auto_load_ciphertexts": False,
                                          "auto_load_plaintexts": False, "rotation_indexes": [], "key_dist": 0}
    with pytest.raises(F.FidesError):
        F.read_dev_sidecar(str(tmp_path / "missing.dev"))
    p.write_bytes(b"3 { 0 1 }\n")  # fewer devices than announced
    with pytest.raises(F.FidesError):
        F.read_dev_sidecar(str(p))


def test_params_file_round_trip(tmp_path):
    P = O.params_small()
    raw = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, 1,
                                 [2.0 ** P.scale_bits] * (P.L + 1), [2.0 ** 80] * (P.L + 1), [0.5] * (P.L + 1))
    path = str(tmp_path / "params.bin")
    F.save_params(raw, path)
    back = F.load_params(path)
    assert back == raw
    raw2 = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots)
    F.save_params(raw2, path)
    assert F.load_params(path) == raw2
    b = bytearray(open(path, "rb").read())
    b[0] ^= 1
    open(path, "wb").write(bytes(b))
    with pytest.raises(ValueError):
        F.load_params(path)
