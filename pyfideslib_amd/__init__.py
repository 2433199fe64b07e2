"""pyfideslib_amd -- MI355X (gfx950) CKKS evaluation engine, Python host side.

A thin ctypes layer over the C ABI of ``libfides_amd.so`` (declared in
``include/fides_amd.h``).  It mirrors FIDESlib's CKKS layer: a
:class:`Context` is built from an OpenFHE-style ``RawParams``
(src/CKKS/openfhe-interface/RawCiphertext.cuh:51-92), ciphertexts are loaded
from and stored to ``RawCipherText`` limb arrays, keys from
``RawKeySwitchKey`` arrays, and :class:`Ciphertext` exposes the evaluation
methods of ``FIDESlib::CKKS::Ciphertext`` (mult, square, rescale, rotate,
rotate_hoisted, add, sub, multPt, addPt, multScalar, addScalar, dropToLevel);
:class:`Plaintext` mirrors ``FIDESlib::CKKS::Plaintext``.

The library must be built (``python -m pyfideslib_amd.build`` or
``__graft_entry__.build()``); importing the package does not fall back to
anything when it is missing -- every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

__all__ = ["RawParams", "Context", "Ciphertext", "Plaintext", "Server", "DeviceArray", "Event", "FidesError",
           "lib", "LIB_PATH"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfides_amd.so")

FIXEDMANUAL, FIXEDAUTO, FLEXIBLEAUTO, FLEXIBLEAUTOEXT = 0, 1, 2, 3
SSE_DENSE_TO_SPARSE, SSE_SPARSE_TO_DENSE = 1, 2  # switching-key ids of sparse-secret encapsulation

ERRORS = {1: "FIDES_E_ARG", 2: "FIDES_E_HIP", 3: "FIDES_E_STATE", 4: "FIDES_E_NOKEY", 5: "FIDES_E_LEVEL",
          6: "FIDES_E_OOM"}

# Exported symbols of the C ABI (tests check the library against include/fides_amd.h).
SYMBOLS = [
    "fides_abi_version", "fides_last_error", "fides_context_create", "fides_context_destroy",
    "fides_synchronize", "fides_context_stream", "fides_context_pool_bytes", "fides_context_trim_pool", "fides_device_alloc",
    "fides_device_free", "fides_memcpy_htod", "fides_memcpy_dtoh", "fides_event_create",
    "fides_event_destroy", "fides_event_record", "fides_event_elapsed_ms", "fides_ct_create",
    "fides_ct_destroy", "fides_ct_load", "fides_ct_load_device", "fides_ct_store", "fides_ct_info",
    "fides_ct_copy", "fides_ct_device_ptrs", "fides_key_load_eval", "fides_key_load_rotation",
    "fides_key_load_automorph", "fides_rotation_automorph_index", "fides_eval_mult", "fides_eval_square",
    "fides_rescale", "fides_eval_add", "fides_eval_sub", "fides_eval_rotate", "fides_eval_rotate_hoisted",
    "fides_eval_mult_batch", "fides_poly_ntt", "fides_stage_modup", "fides_stage_modup_noalloc", "fides_tune_load", "fides_tune_force", "fides_tune_config", "fides_ntt_limbs", "fides_stage_moddown",
    "fides_stage_keyswitch", "fides_pt_create", "fides_pt_destroy", "fides_pt_load", "fides_eval_mult_plain",
    "fides_eval_add_plain", "fides_eval_mult_scalar", "fides_eval_add_scalar", "fides_drop_to_level",
    "fides_scalar_encoding", "fides_eval_rotate_batch", "fides_server_create", "fides_server_destroy",
    "fides_server_submit_mult", "fides_server_submit_rotate", "fides_server_submit_logreg", "fides_server_step",
    "fides_server_poll", "fides_server_counts", "fides_eval_linear_transform", "fides_eval_automorph",
    "fides_eval_conjugate", "fides_mod_raise", "fides_eval_mult_monomial", "fides_eval_mult_int",
    "fides_eval_horner_series", "fides_eval_double_angle", "fides_eval_chebyshev_series",
    "fides_eval_mod_reduction", "fides_encode", "fides_bootstrap_setup", "fides_bootstrap_rotations",
    "fides_eval_bootstrap", "fides_bootstrap_destroy", "fides_key_load_switch", "fides_eval_key_switch",
    "fides_eval_bootstrap_batch", "fides_eval_linear_transform_batch", "fides_accumulate_rotations",
    "fides_convolution_rotations", "fides_eval_convolution_transform",
    "fides_eval_accumulate_sum", "fides_eval_negate", "fides_eval_add_many", "fides_pt_load_ext",
    "fides_encode_ext", "fides_ct_extend", "fides_ct_mod_down", "fides_eval_rotate_hoisted_ext",
    "fides_eval_add_ext", "fides_eval_dot_product_plain", "fides_eval_dot_product", "fides_ct_store_ext", "fides_ct_is_ext",
    "fides_params_save", "fides_context_save_params", "fides_context_create_from_file", "fides_ct_save",
    "fides_ct_load_file", "fides_pt_save", "fides_pt_load_file", "fides_key_save", "fides_key_load_file",
    "fides_dev_sidecar_write", "fides_dev_sidecar_read", "fides_eval_sub_plain", "fides_eval_add_mult_plain",
    "fides_eval_add_mult_scalar", "fides_eval_linear_wsum", "fides_eval_mult_ext", "fides_context_set_op_graphs",
]


class BootstrapParams(C.Structure):
    """fides_bootstrap_params (include/fides_amd.h)."""

    _fields_ = [("level_budget_cts", C.c_int32), ("level_budget_stc", C.c_int32), ("K", C.c_double),
                ("double_angle", C.c_int32), ("degree", C.c_int32), ("scale", C.c_double),
                ("top_level", C.c_int32), ("sparse_encaps", C.c_int32), ("slots", C.c_int32),
                ("batched_giant", C.c_int32), ("q_basis_diagonals", C.c_int32)]


class FidesError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _RawParamsC(C.Structure):
    _fields_ = [
        ("logN", C.c_int32), ("L", C.c_int32), ("K", C.c_int32), ("dnum", C.c_int32),
        ("moduli", C.POINTER(C.c_uint64)), ("root_of_unity", C.POINTER(C.c_uint64)),
        ("special_moduli", C.POINTER(C.c_uint64)), ("special_root_of_unity", C.POINTER(C.c_uint64)),
        ("scaling_technique", C.c_int32), ("mod_reduce_factor", C.POINTER(C.c_double)),
        ("scaling_factor_real", C.POINTER(C.c_double)), ("scaling_factor_real_big", C.POINTER(C.c_double)),
    ]


_lib = None


def lib():
    """Load libfides_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `python -m pyfideslib_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i32, u32, u64p = C.c_void_p, C.c_int32, C.c_uint32, C.POINTER(C.c_uint64)
        for name in SYMBOLS:
            getattr(L, name).restype = C.c_int
        L.fides_last_error.restype = C.c_char_p
        L.fides_context_stream.restype = vp
        L.fides_context_create.argtypes = [C.POINTER(_RawParamsC), i32, C.POINTER(vp)]
        L.fides_context_destroy.argtypes = [vp]
        L.fides_synchronize.argtypes = [vp]
        L.fides_context_stream.argtypes = [vp]
        L.fides_context_pool_bytes.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.fides_context_trim_pool.argtypes = [vp, C.c_uint64]
        L.fides_device_alloc.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
        L.fides_device_free.argtypes = [vp, vp]
        L.fides_memcpy_htod.argtypes = [vp, vp, vp, C.c_uint64]
        L.fides_memcpy_dtoh.argtypes = [vp, vp, vp, C.c_uint64]
        L.fides_event_create.argtypes = [vp, C.POINTER(vp)]
        L.fides_event_destroy.argtypes = [vp]
        L.fides_event_record.argtypes = [vp, vp]
        L.fides_event_elapsed_ms.argtypes = [vp, vp, C.POINTER(C.c_float)]
        L.fides_ct_create.argtypes = [vp, C.POINTER(vp)]
        L.fides_ct_destroy.argtypes = [vp]
        L.fides_ct_load.argtypes = [vp, vp, vp, vp, i32, C.c_double, i32, i32]
        L.fides_ct_load_device.argtypes = [vp, vp, vp, vp, i32, C.c_double, i32, i32]
        L.fides_ct_store.argtypes = [vp, vp, vp, vp]
        L.fides_ct_info.argtypes = [vp, C.POINTER(i32), C.POINTER(C.c_double), C.POINTER(i32), C.POINTER(i32)]
        L.fides_ct_copy.argtypes = [vp, vp, vp]
        L.fides_ct_device_ptrs.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
        L.fides_key_load_eval.argtypes = [vp, vp, vp]
        L.fides_key_load_rotation.argtypes = [vp, i32, vp, vp]
        L.fides_key_load_automorph.argtypes = [vp, u32, vp, vp]
        L.fides_rotation_automorph_index.argtypes = [vp, i32, C.POINTER(u32)]
        L.fides_eval_mult.argtypes = [vp, vp, vp, vp, i32]
        L.fides_eval_square.argtypes = [vp, vp, vp, i32]
        L.fides_rescale.argtypes = [vp, vp]
        L.fides_eval_add.argtypes = [vp, vp, vp, vp]
        L.fides_eval_sub.argtypes = [vp, vp, vp, vp]
        L.fides_eval_rotate.argtypes = [vp, vp, i32, vp]
        L.fides_eval_rotate_hoisted.argtypes = [vp, vp, C.POINTER(i32), i32, C.POINTER(vp)]
        L.fides_eval_mult_batch.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), i32, i32]
        L.fides_poly_ntt.argtypes = [vp, vp, i32, i32, i32]
        L.fides_stage_modup.argtypes = [vp, i32, vp, vp]
        L.fides_stage_modup_noalloc.argtypes = [vp, i32, vp, vp]
        L.fides_tune_load.argtypes = [C.c_char_p, C.POINTER(i32)]
        L.fides_ntt_limbs.argtypes = [vp, vp, i32, i32]
        L.fides_tune_force.argtypes = [i32, C.POINTER(i32)]
        L.fides_tune_config.argtypes = [vp, i32, i32, C.POINTER(i32), C.POINTER(i32)]
        L.fides_stage_moddown.argtypes = [vp, i32, vp, vp]
        L.fides_stage_keyswitch.argtypes = [vp, i32, vp, u32, vp, vp]
        L.fides_pt_create.argtypes = [vp, C.POINTER(vp)]
        L.fides_pt_destroy.argtypes = [vp]
        L.fides_pt_load.argtypes = [vp, vp, vp, i32, C.c_double, i32, i32]
        L.fides_pt_load_ext.argtypes = [vp, vp, vp, i32, C.c_double, i32, i32]
        L.fides_eval_mult_plain.argtypes = [vp, vp, vp, vp, i32]
        L.fides_eval_add_plain.argtypes = [vp, vp, vp, vp]
        L.fides_eval_mult_scalar.argtypes = [vp, vp, C.c_double, vp, i32]
        L.fides_eval_add_scalar.argtypes = [vp, vp, C.c_double, vp]
        L.fides_drop_to_level.argtypes = [vp, vp, i32]
        L.fides_scalar_encoding.argtypes = [vp, i32, C.c_double, i32, i32, vp]
        L.fides_eval_rotate_batch.argtypes = [vp, C.POINTER(vp), C.POINTER(i32), C.POINTER(vp), i32]
        i64p = C.POINTER(C.c_int64)
        L.fides_server_create.argtypes = [vp, i32, C.c_uint64, C.POINTER(vp)]
        L.fides_server_destroy.argtypes = [vp]
        L.fides_server_submit_mult.argtypes = [vp, vp, vp, vp, i64p]
        L.fides_server_submit_rotate.argtypes = [vp, vp, i32, vp, i64p]
        L.fides_server_submit_logreg.argtypes = [vp, vp, vp, i32, C.POINTER(C.c_double), vp, i64p]
        L.fides_server_step.argtypes = [vp, C.POINTER(i32)]
        L.fides_server_poll.argtypes = [vp, i64p, i32, C.POINTER(i32)]
        L.fides_server_counts.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(C.c_uint64)]
        L.fides_eval_linear_transform.argtypes = [vp, vp, C.POINTER(vp), i32, i32, i32, i32, i32, vp]
        L.fides_eval_automorph.argtypes = [vp, vp, u32, vp]
        L.fides_eval_conjugate.argtypes = [vp, vp, vp]
        L.fides_mod_raise.argtypes = [vp, vp, i32, vp]
        L.fides_eval_mult_monomial.argtypes = [vp, vp, C.c_int64, vp]
        L.fides_eval_mult_int.argtypes = [vp, vp, C.c_uint64, vp]
        L.fides_eval_horner_series.argtypes = [vp, vp, C.POINTER(C.c_double), i32, vp]
        L.fides_eval_double_angle.argtypes = [vp, vp, i32, vp]
        L.fides_eval_chebyshev_series.argtypes = [vp, vp, C.POINTER(C.c_double), i32, C.c_double, C.c_double, vp]
        L.fides_eval_mod_reduction.argtypes = [vp, vp, C.POINTER(C.c_double), i32, i32, C.c_uint64, vp]
        dp = C.POINTER(C.c_double)
        L.fides_encode.argtypes = [vp, dp, dp, i32, C.c_double, vp]
        L.fides_encode_ext.argtypes = [vp, dp, dp, i32, C.c_double, vp]
        L.fides_ct_extend.argtypes = [vp, vp, vp]
        L.fides_ct_mod_down.argtypes = [vp, vp, vp]
        L.fides_eval_rotate_hoisted_ext.argtypes = [vp, vp, C.POINTER(i32), i32, C.POINTER(vp)]
        L.fides_eval_add_ext.argtypes = [vp, vp, vp, vp]
        L.fides_eval_dot_product_plain.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), i32, vp]
        L.fides_eval_dot_product.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), i32, i32, vp]
        L.fides_ct_store_ext.argtypes = [vp, vp, vp, vp]
        L.fides_ct_is_ext.argtypes = [vp, C.POINTER(i32)]
        cs = C.c_char_p
        L.fides_eval_sub_plain.argtypes = [vp, vp, vp, vp]
        L.fides_eval_add_mult_plain.argtypes = [vp, vp, vp, vp, i32]
        L.fides_eval_add_mult_scalar.argtypes = [vp, vp, vp, C.c_double]
        L.fides_eval_linear_wsum.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_double), i32, vp]
        L.fides_eval_mult_ext.argtypes = [vp, vp, vp, vp]
        L.fides_context_set_op_graphs.argtypes = [vp, i32]
        L.fides_params_save.argtypes = [C.POINTER(_RawParamsC), cs]
        L.fides_context_save_params.argtypes = [vp, cs]
        L.fides_context_create_from_file.argtypes = [cs, i32, C.POINTER(vp)]
        L.fides_ct_save.argtypes = [vp, vp, cs]
        L.fides_ct_load_file.argtypes = [vp, vp, cs]
        L.fides_pt_save.argtypes = [vp, vp, cs]
        L.fides_pt_load_file.argtypes = [vp, vp, cs]
        L.fides_key_save.argtypes = [vp, i32, i32, cs]
        L.fides_key_load_file.argtypes = [vp, cs, C.POINTER(i32), C.POINTER(i32)]
        L.fides_dev_sidecar_write.argtypes = [cs, i32, C.POINTER(i32), i32, i32, i32, C.POINTER(i32), i32]
        L.fides_dev_sidecar_read.argtypes = [cs, C.POINTER(i32), i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                             C.POINTER(i32), i32, C.POINTER(i32), C.POINTER(i32)]
        L.fides_bootstrap_setup.argtypes = [vp, C.POINTER(BootstrapParams), C.POINTER(vp)]
        L.fides_bootstrap_rotations.argtypes = [vp, C.POINTER(i32), i32, C.POINTER(i32)]
        L.fides_eval_bootstrap.argtypes = [vp, vp, vp, vp]
        L.fides_bootstrap_destroy.argtypes = [vp]
        L.fides_key_load_switch.argtypes = [vp, i32, vp, vp]
        L.fides_eval_bootstrap_batch.argtypes = [vp, vp, C.POINTER(vp), C.POINTER(vp), i32]
        L.fides_accumulate_rotations.argtypes = [i32, i32, C.POINTER(i32), i32, C.POINTER(i32)]
        L.fides_convolution_rotations.argtypes = [i32, i32, C.POINTER(i32), i32, C.POINTER(i32)]
        L.fides_eval_convolution_transform.argtypes = [vp, vp, C.POINTER(vp), i32, i32, i32, C.POINTER(i32), i32,
                                                       vp]
        L.fides_eval_accumulate_sum.argtypes = [vp, vp, i32, i32, vp]
        L.fides_eval_negate.argtypes = [vp, vp, vp]
        L.fides_eval_add_many.argtypes = [vp, C.POINTER(vp), i32, vp]
        L.fides_eval_linear_transform_batch.argtypes = [vp, C.POINTER(vp), i32, C.POINTER(vp), i32, i32, i32, i32,
                                                        i32, C.POINTER(vp)]
        L.fides_eval_key_switch.argtypes = [vp, vp, i32, vp]
        L.fides_abi_version.argtypes = []
        _lib = L
    return _lib


def accumulate_rotations(size: int, stride: int = 1) -> list:
    """Rotation indices AccumulateSum(size, stride) needs keys for."""
    n = C.c_int32()
    _check(lib().fides_accumulate_rotations(size, stride, None, 0, C.byref(n)))
    idx = (C.c_int32 * max(1, n.value))()
    _check(lib().fides_accumulate_rotations(size, stride, idx, n.value, C.byref(n)))
    return [int(v) for v in idx[: n.value]]


def convolution_rotations(stride: int, gstep: int) -> list:
    """Rotation indices ConvolutionTransform(gstep, stride) needs keys for, besides its baby-step indexes."""
    n = C.c_int32()
    _check(lib().fides_convolution_rotations(stride, gstep, None, 0, C.byref(n)))
    idx = (C.c_int32 * max(1, n.value))()
    _check(lib().fides_convolution_rotations(stride, gstep, idx, n.value, C.byref(n)))
    return [int(v) for v in idx[: n.value]]


TUNE_NTT_FWD, TUNE_NTT_INV, TUNE_MODUP, TUNE_MODDOWN = 0, 1, 2, 3


TUNE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_gfx950.txt")


def tune_load(path: str | None = TUNE_TABLE) -> int:
    """Load a tuning table (pyfideslib_amd/tune.py writes one; None: fixed rules only); returns its entry count."""
    n = C.c_int32()
    _check(lib().fides_tune_load(None if path is None else os.fsencode(path), C.byref(n)))
    return n.value


def tune_force(kind: int, cfg=None):
    """Pin the launch configuration of `kind` for every later launch (None releases it; kind < 0 releases all)."""
    arr = None if cfg is None else (C.c_int32 * 4)(*[int(x) for x in list(cfg) + [0] * (4 - len(cfg))])
    _check(lib().fides_tune_force(kind, arr))


def _check(rc: int):
    if rc != 0:
        raise FidesError(rc, lib().fides_last_error().decode())


def _ptr(a) -> int:
    """Device address of a DeviceArray (or a raw integer address)."""
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        raise TypeError("stage entry points take device memory: wrap the array in DeviceArray")
    return int(a)


@dataclass
class RawParams:
    """OpenFHE-derived parameters as FIDESlib's RawParams carries them."""

    logN: int
    L: int
    K: int
    dnum: int
    moduli: list            # Q primes [L+1]
    root_of_unity: list     # [L+1]
    special_moduli: list    # [K]
    special_root_of_unity: list
    scaling_technique: int = FIXEDMANUAL
    mod_reduce_factor: list | None = None
    scaling_factor_real: list | None = None  # [L+1]; needed by the scalar ops
    scaling_factor_real_big: list | None = None  # [L+1]; FLEXIBLEAUTO level alignment

    @property
    def N(self) -> int:
        return 1 << self.logN

    @classmethod
    def from_lists(cls, logN, L, K, dnum, moduli_all, roots_all, scaling_technique=FIXEDMANUAL,
                   scaling_factor_real=None, scaling_factor_real_big=None, mod_reduce_factor=None):
        return cls(logN, L, K, dnum, list(moduli_all[: L + 1]), list(roots_all[: L + 1]),
                   list(moduli_all[L + 1:]), list(roots_all[L + 1:]), scaling_technique,
                   None if mod_reduce_factor is None else list(mod_reduce_factor),
                   None if scaling_factor_real is None else list(scaling_factor_real),
                   None if scaling_factor_real_big is None else list(scaling_factor_real_big))


def _raw_c(raw: RawParams):
    """fides_raw_params of `raw`, and the arrays it points into (keep them alive)."""
    keep = [np.array(x, dtype=np.uint64) for x in
            (raw.moduli, raw.root_of_unity, raw.special_moduli, raw.special_root_of_unity)]
    u = C.POINTER(C.c_uint64)
    p = _RawParamsC(raw.logN, raw.L, raw.K, raw.dnum, *[a.ctypes.data_as(u) for a in keep],
                    raw.scaling_technique, None, None)
    dp = C.POINTER(C.c_double)
    for name in ("mod_reduce_factor", "scaling_factor_real", "scaling_factor_real_big"):
        v = getattr(raw, name)
        if v is not None:
            a = np.array(v, dtype=np.float64)
            keep.append(a)
            setattr(p, name, a.ctypes.data_as(dp))
    return p, keep


KEY_KIND_EVAL, KEY_KIND_AUTOMORPH, KEY_KIND_SWITCH = 0, 1, 2
_MAGIC = b"FIDESAMD"


def save_params(raw: RawParams, path: str):
    """The engine's parameter file of `raw` (fides_params_save)."""
    p, keep = _raw_c(raw)
    _check(lib().fides_params_save(C.byref(p), os.fsencode(path)))


def load_params(path: str) -> RawParams:
    """RawParams from a parameter file: "FIDESAMD", u32 version 1, u32 kind 1, i32 logN L K dnum
    technique, u64 moduli[L+1+K], roots[L+1+K], u32 flags, then the f64 per-level tables present."""
    with open(path, "rb") as f:
        b = f.read()
    if b[:8] != _MAGIC or np.frombuffer(b, np.uint32, 2, 8).tolist() != [1, 1]:
        raise ValueError("not an engine parameter file")
    logN, L, K, dnum, tech = np.frombuffer(b, np.int32, 5, 16).tolist()
    np_ = L + 1 + K
    off = 36
    mod = np.frombuffer(b, np.uint64, np_, off).tolist()
    off += 8 * np_
    roots = np.frombuffer(b, np.uint64, np_, off).tolist()
    off += 8 * np_
    flags = int(np.frombuffer(b, np.uint32, 1, off)[0])
    off += 4
    tabs = []
    for bit in (1, 2, 4):
        if flags & bit:
            tabs.append(np.frombuffer(b, np.float64, L + 1, off).tolist())
            off += 8 * (L + 1)
        else:
            tabs.append(None)
    return RawParams.from_lists(logN, L, K, dnum, mod, roots, tech, scaling_factor_real=tabs[0],
                                scaling_factor_real_big=tabs[1], mod_reduce_factor=tabs[2])


def write_dev_sidecar(path: str, devices=(0,), auto_load_ciphertexts: bool = True, auto_load_plaintexts: bool = False,
                      rotation_indexes=(), key_dist: int = 0):
    """FIDESlib's `<file>.dev` text sidecar (api/Serialize.cpp:35-65)."""
    d = (C.c_int32 * max(1, len(devices)))(*devices)
    r = (C.c_int32 * max(1, len(rotation_indexes)))(*rotation_indexes)
    _check(lib().fides_dev_sidecar_write(os.fsencode(path), len(devices), d, int(auto_load_ciphertexts),
                                         int(auto_load_plaintexts), len(rotation_indexes), r, key_dist))


def read_dev_sidecar(path: str) -> dict:
    """Parse a `.dev` sidecar as DeserializeFromFile does (api/Serialize.cpp:117-190)."""
    nd, nr, a, b, kd = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    _check(lib().fides_dev_sidecar_read(os.fsencode(path), None, 0, C.byref(nd), C.byref(a), C.byref(b), None, 0,
                                        C.byref(nr), C.byref(kd)))
    d = (C.c_int32 * max(1, nd.value))()
    r = (C.c_int32 * max(1, nr.value))()
    _check(lib().fides_dev_sidecar_read(os.fsencode(path), d, nd.value, C.byref(nd), C.byref(a), C.byref(b), r,
                                        nr.value, C.byref(nr), C.byref(kd)))
    return {"devices": list(d[: nd.value]), "auto_load_ciphertexts": bool(a.value),
            "auto_load_plaintexts": bool(b.value), "rotation_indexes": list(r[: nr.value]), "key_dist": kd.value}


class DeviceArray:
    """uint64 device buffer owned by a context (allocation through the C ABI)."""

    def __init__(self, ctx: "Context", shape, host: np.ndarray | None = None):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        _check(lib().fides_device_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=np.uint64)
        assert h.nbytes == self.nbytes
        _check(lib().fides_memcpy_htod(self.ctx.h, self.ptr, h.ctypes.data, self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.uint64)
        _check(lib().fides_memcpy_dtoh(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            _check(lib().fides_device_free(self.ctx.h, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    """hipEvent recorded on a context's stream."""

    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        e = C.c_void_p()
        _check(lib().fides_event_create(ctx.h, C.byref(e)))
        self.e = e.value

    def record(self):
        _check(lib().fides_event_record(self.ctx.h, self.e))
        return self

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        _check(lib().fides_event_elapsed_ms(self.e, stop.e, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.e:
                lib().fides_event_destroy(self.e)
        except Exception:
            pass


class Context:
    """FIDESlib::CKKS::Context on one GPU (GenCryptoContextGPU)."""

    def __init__(self, raw: RawParams, device: int = 0):
        self.raw = raw
        self.N = raw.N
        self.L, self.K, self.dnum = raw.L, raw.K, raw.dnum
        p, self._keep = _raw_c(raw)
        h = C.c_void_p()
        _check(lib().fides_context_create(C.byref(p), device, C.byref(h)))
        self.h = h

    @classmethod
    def from_file(cls, path: str, device: int = 0) -> "Context":
        """A context from a parameter file (:func:`save_params` / RawParams.save)."""
        return cls(load_params(path), device)

    def set_op_graphs(self, on: bool):
        """Capture single ops' launch sequences into HIP graphs and replay them (off by default)."""
        _check(lib().fides_context_set_op_graphs(self.h, int(bool(on))))

    def save_params(self, path: str):
        _check(lib().fides_context_save_params(self.h, os.fsencode(path)))

    def load_ciphertext_file(self, path: str) -> "Ciphertext":
        ct = Ciphertext(self)
        _check(lib().fides_ct_load_file(self.h, ct.h, os.fsencode(path)))
        return ct

    def load_plaintext_file(self, path: str) -> "Plaintext":
        pt = Plaintext(self)
        _check(lib().fides_pt_load_file(self.h, pt.h, os.fsencode(path)))
        return pt

    def save_key(self, kind: int, index: int, path: str):
        """kind KEY_KIND_EVAL (index ignored), KEY_KIND_AUTOMORPH (index = automorphism index k)
        or KEY_KIND_SWITCH (index = the switching key's id)."""
        _check(lib().fides_key_save(self.h, kind, index, os.fsencode(path)))

    def load_key_file(self, path: str) -> tuple:
        """Loads the key under the kind and index the file records; returns (kind, index)."""
        k, i = C.c_int32(), C.c_int32()
        _check(lib().fides_key_load_file(self.h, os.fsencode(path), C.byref(k), C.byref(i)))
        return k.value, i.value

    def close(self):
        if getattr(self, "h", None):
            lib().fides_synchronize(self.h)
            _check(lib().fides_context_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def alpha(self) -> int:
        return -(-(self.L + 1) // self.dnum)

    def num_digits(self, level: int) -> int:
        return min(self.dnum, -(-(level + 1) // self.alpha))

    def synchronize(self):
        _check(lib().fides_synchronize(self.h))

    @property
    def stream(self) -> int:
        return lib().fides_context_stream(self.h)

    def pool_bytes(self) -> int:
        v = C.c_uint64()
        _check(lib().fides_context_pool_bytes(self.h, C.byref(v)))
        return v.value

    def trim_pool(self, keep_bytes: int = 0):
        """Give cached, unused device blocks back to the device (at most keep_bytes stay cached)."""
        _check(lib().fides_context_trim_pool(self.h, keep_bytes))

    # ---- keys (RawKeySwitchKey) ----
    def _key_arrays(self, key):
        a, b = (np.ascontiguousarray(k, dtype=np.uint64) for k in key)
        shape = (self.dnum, self.L + 1 + self.K, self.N)
        if a.shape != shape or b.shape != shape:
            raise ValueError(f"key arrays must have shape {shape}")
        return a, b

    def load_eval_key(self, key):
        a, b = self._key_arrays(key)
        _check(lib().fides_key_load_eval(self.h, a.ctypes.data, b.ctypes.data))

    def load_rotation_key(self, index: int, key):
        a, b = self._key_arrays(key)
        _check(lib().fides_key_load_rotation(self.h, index, a.ctypes.data, b.ctypes.data))

    def load_switch_key(self, key_id: int, key):
        """Secret-switching key (GetSecretSwitchingKey) under a caller id; bootstrapping's sparse-secret
        encapsulation uses SSE_DENSE_TO_SPARSE and SSE_SPARSE_TO_DENSE."""
        A, B = self._key_arrays(key)
        _check(lib().fides_key_load_switch(self.h, key_id, A.ctypes.data, B.ctypes.data))

    def load_automorph_key(self, k: int, key):
        """Key of automorphism index k (e.g. 2N - 1: conjugation)."""
        a, b = self._key_arrays(key)
        _check(lib().fides_key_load_automorph(self.h, k, a.ctypes.data, b.ctypes.data))

    def automorph_index(self, index: int) -> int:
        k = C.c_uint32()
        _check(lib().fides_rotation_automorph_index(self.h, index, C.byref(k)))
        return k.value

    # ---- ciphertexts (RawCipherText) ----
    def ciphertext(self) -> "Ciphertext":
        return Ciphertext(self)

    def load(self, c0, c1, noise: float = 1.0, noise_level: int = 1, slots: int | None = None) -> "Ciphertext":
        ct = Ciphertext(self)
        ct.load(c0, c1, noise, noise_level, slots)
        return ct

    def plaintext(self, limbs, noise: float = 1.0, noise_level: int = 1, slots: int | None = None) -> "Plaintext":
        return Plaintext(self).load(limbs, noise, noise_level, slots)

    def linear_wsum(self, cts: list, weights: list, out=None) -> "Ciphertext":
        """Ciphertext::evalLinearWSumMutable: sum_i weights[i] * cts[i] (one launch)."""
        n = len(cts)
        o = Ciphertext(self) if out is None else out
        hc = (C.c_void_p * n)(*[x.h for x in cts])
        w = (C.c_double * n)(*[float(v) for v in weights])
        _check(lib().fides_eval_linear_wsum(self.h, hc, w, n, o.h))
        return o

    def dot_product_plain(self, cts: list, pts: list, out=None) -> "Ciphertext":
        """Ciphertext::dotProductPt: sum_i cts[i] * pts[i] (P*Q ciphertexts need P*Q plaintexts)."""
        n = len(cts)
        o = Ciphertext(self) if out is None else out
        hc = (C.c_void_p * n)(*[x.h for x in cts])
        hp = (C.c_void_p * n)(*[p.h for p in pts])
        _check(lib().fides_eval_dot_product_plain(self.h, hc, hp, n, o.h))
        return o

    def dot_product(self, a: list, b: list, ext: bool = False, out=None) -> "Ciphertext":
        """Ciphertext::dotProduct: sum_i a[i] * b[i] with one key switch of the summed d2, not
        rescaled; ext leaves it in P*Q."""
        n = len(a)
        o = Ciphertext(self) if out is None else out
        ha = (C.c_void_p * n)(*[x.h for x in a])
        hb = (C.c_void_p * len(b))(*[x.h for x in b])
        if len(b) != n:
            raise ValueError("as many ciphertexts on both sides")
        _check(lib().fides_eval_dot_product(self.h, ha, hb, n, int(bool(ext)), o.h))
        return o

    def plaintext_ext(self, limbs, noise: float = 1.0, noise_level: int = 1, slots: int | None = None) -> "Plaintext":
        """A P*Q plaintext ([num_res + K, N]: Q limbs, then the special limbs)."""
        return Plaintext(self).load_ext(limbs, noise, noise_level, slots)

    def encode(self, z, level: int, scale: float, ext: bool = False) -> "Plaintext":
        """MakeCKKSPackedPlaintext: N/2 complex (or real) slot values at `level` and scaling factor `scale`;
        with ext the plaintext also holds the special-prime limbs (P*Q form)."""
        z = np.asarray(z)
        if z.shape != (self.N // 2,):
            raise ValueError(f"need {self.N // 2} slot values")
        re = np.ascontiguousarray(z.real, dtype=np.float64)
        im = np.ascontiguousarray(z.imag, dtype=np.float64) if np.iscomplexobj(z) else None
        pt = Plaintext(self)
        dp = C.POINTER(C.c_double)
        fn = lib().fides_encode_ext if ext else lib().fides_encode
        _check(fn(self.h, re.ctypes.data_as(dp), None if im is None else im.ctypes.data_as(dp), level, scale, pt.h))
        return pt

    def scalar_encoding(self, level: int, v: float, kind: str = "mult", noise_deg: int = 1) -> np.ndarray:
        """Residues of ElemForEvalMult(v) (kind "mult") or ElemForEvalAddOrSub(|v|) ("add") at `level`."""
        out = np.zeros(level + 1, dtype=np.uint64)
        _check(lib().fides_scalar_encoding(self.h, level, v, 0 if kind == "mult" else 1, noise_deg,
                                           out.ctypes.data))
        return out

    # ---- stage entry points on device memory (torch tensors / device addresses) ----
    def ntt(self, data, num_limbs: int, first_prime: int = 0, inverse: bool = False):
        _check(lib().fides_poly_ntt(self.h, _ptr(data), num_limbs, first_prime, int(inverse)))

    def event(self) -> Event:
        return Event(self)

    def device_array(self, shape, host: np.ndarray | None = None) -> DeviceArray:
        return DeviceArray(self, shape, host)

    def stage_modup(self, level: int, x, ext):
        _check(lib().fides_stage_modup(self.h, level, _ptr(x), _ptr(ext)))

    def ntt_limbs(self, data, limbs: int, inverse: bool = False):
        """NTT of data[limbs][N] in place, limb i over prime id i mod (L+1+K)."""
        _check(lib().fides_ntt_limbs(self.h, _ptr(data), limbs, int(inverse)))

    def tune_config(self, kind: int, size: int):
        """(configuration, from_table) a launch of `kind` over `size` limbs / conversions would use."""
        cfg, tab = (C.c_int32 * 4)(), C.c_int32()
        _check(lib().fides_tune_config(self.h, kind, size, cfg, C.byref(tab)))
        return [int(x) for x in cfg], bool(tab.value)

    def stage_modup_noalloc(self, level: int, x, ext):
        """ModUp with the library's static device workspace (no allocation)."""
        _check(lib().fides_stage_modup_noalloc(self.h, level, _ptr(x), _ptr(ext)))

    def stage_moddown(self, level: int, x_ext, y):
        _check(lib().fides_stage_moddown(self.h, level, _ptr(x_ext), _ptr(y)))

    def stage_keyswitch(self, level: int, x, ks_b, ks_a, rotation_k: int = 0):
        _check(lib().fides_stage_keyswitch(self.h, level, _ptr(x), rotation_k, _ptr(ks_b), _ptr(ks_a)))

    def mult_batch(self, a: list, b: list, out: list, rescale: bool = True):
        n = len(a)
        arr = lambda xs: (C.c_void_p * n)(*[x.h for x in xs])  # noqa: E731
        _check(lib().fides_eval_mult_batch(self.h, arr(a), arr(b), arr(out), n, int(rescale)))

    def add_many(self, cts: list, out=None) -> "Ciphertext":
        """EvalAddMany: the sum of the ciphertexts in one launch."""
        o = Ciphertext(self) if out is None else out
        h = (C.c_void_p * len(cts))(*[c.h for c in cts])
        _check(lib().fides_eval_add_many(self.h, h, len(cts), o.h))
        return o

    def linear_transform_batch(self, cts: list, diag: list, bstep: int, stride: int = 1, offset: int = 0,
                               outs: list | None = None, batched_giant: bool = False) -> list:
        """The BSGS linear transform of several ciphertexts at one level with the same diagonals (one dot
        launch, every giant step one batched rotation); each output equals the single transform."""
        m, n = len(cts), len(diag)
        outs = [Ciphertext(self) for _ in cts] if outs is None else outs
        ha = (C.c_void_p * m)(*[c.h for c in cts])
        ho = (C.c_void_p * m)(*[o.h for o in outs])
        hd = (C.c_void_p * n)(*[p.h for p in diag])
        _check(lib().fides_eval_linear_transform_batch(self.h, ha, m, hd, n, bstep, stride, offset,
                                                       1 if batched_giant else 0, ho))
        return outs

    def rotate_batch(self, a: list, indices: list, out: list):
        """out[i] = rotate(a[i], indices[i]), one launch per stage for the whole batch."""
        n = len(a)
        arr = lambda xs: (C.c_void_p * n)(*[x.h for x in xs])  # noqa: E731
        idx = (C.c_int32 * n)(*indices)
        _check(lib().fides_eval_rotate_batch(self.h, arr(a), idx, arr(out), n))


class Ciphertext:
    """FIDESlib::CKKS::Ciphertext: device-resident (c0, c1) with metadata."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().fides_ct_create(ctx.h, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib().fides_ct_destroy(self.h)
            self.h = None
        except Exception:
            pass

    def load(self, c0, c1, noise: float = 1.0, noise_level: int = 1, slots: int | None = None):
        slots = self.ctx.N // 2 if slots is None else slots
        if isinstance(c0, np.ndarray):
            c0 = np.ascontiguousarray(c0, dtype=np.uint64)
            c1 = np.ascontiguousarray(c1, dtype=np.uint64)
            if c0.shape != c1.shape or c0.ndim != 2 or c0.shape[1] != self.ctx.N:
                raise ValueError("c0/c1 must be [num_res, N] uint64")
            _check(lib().fides_ct_load(self.ctx.h, self.h, c0.ctypes.data, c1.ctypes.data, c0.shape[0], noise,
                                       noise_level, slots))
        else:  # DeviceArray limbs
            _check(lib().fides_ct_load_device(self.ctx.h, self.h, _ptr(c0), _ptr(c1), c0.shape[0], noise,
                                              noise_level, slots))
        return self

    def info(self):
        lv, nl, sl, nz = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        _check(lib().fides_ct_info(self.h, C.byref(lv), C.byref(nz), C.byref(nl), C.byref(sl)))
        return {"level": lv.value, "noise": nz.value, "noise_level": nl.value, "slots": sl.value}

    @property
    def level(self) -> int:
        return self.info()["level"]

    def store(self):
        n = self.level + 1
        c0 = np.empty((n, self.ctx.N), dtype=np.uint64)
        c1 = np.empty((n, self.ctx.N), dtype=np.uint64)
        _check(lib().fides_ct_store(self.ctx.h, self.h, c0.ctypes.data, c1.ctypes.data))
        return c0, c1

    def device_ptrs(self):
        a, b = C.c_void_p(), C.c_void_p()
        _check(lib().fides_ct_device_ptrs(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def copy(self) -> "Ciphertext":
        out = Ciphertext(self.ctx)
        _check(lib().fides_ct_copy(self.ctx.h, self.h, out.h))
        return out

    # ---- evaluation (out-of-place by default; *_ in place) ----
    def _out(self, out):
        return Ciphertext(self.ctx) if out is None else out

    def mult(self, other: "Ciphertext", rescale: bool = False, out: "Ciphertext | None" = None):
        o = self._out(out)
        _check(lib().fides_eval_mult(self.ctx.h, self.h, other.h, o.h, int(rescale)))
        return o

    def square(self, rescale: bool = False, out=None):
        o = self._out(out)
        _check(lib().fides_eval_square(self.ctx.h, self.h, o.h, int(rescale)))
        return o

    def rescale(self):
        _check(lib().fides_rescale(self.ctx.h, self.h))
        return self

    def add(self, other, out=None):
        o = self._out(out)
        _check(lib().fides_eval_add(self.ctx.h, self.h, other.h, o.h))
        return o

    def sub(self, other, out=None):
        o = self._out(out)
        _check(lib().fides_eval_sub(self.ctx.h, self.h, other.h, o.h))
        return o

    def rotate(self, index: int, out=None):
        o = self._out(out)
        _check(lib().fides_eval_rotate(self.ctx.h, self.h, index, o.h))
        return o

    def mult_plain(self, pt: "Plaintext", rescale: bool = False, out=None):
        o = self._out(out)
        _check(lib().fides_eval_mult_plain(self.ctx.h, self.h, pt.h, o.h, int(rescale)))
        return o

    def add_plain(self, pt: "Plaintext", out=None):
        o = self._out(out)
        _check(lib().fides_eval_add_plain(self.ctx.h, self.h, pt.h, o.h))
        return o

    def mult_scalar(self, v: float, rescale: bool = False, out=None):
        o = self._out(out)
        _check(lib().fides_eval_mult_scalar(self.ctx.h, self.h, float(v), o.h, int(rescale)))
        return o

    def add_scalar(self, v: float, out=None):
        o = self._out(out)
        _check(lib().fides_eval_add_scalar(self.ctx.h, self.h, float(v), o.h))
        return o

    def linear_transform(self, diag: list, bstep: int, stride: int = 1, offset: int = 0, out=None,
                         batched_giant: bool = False):
        """sum_j rot_{j b s}(sum_i diag[b j + i] * rot_{i s}(self)), then rot_offset (BSGS).  Giant steps
        Horner style (the reference's order), or all in one batched key switch with batched_giant."""
        o = self._out(out)
        n = len(diag)
        hs = (C.c_void_p * n)(*[p.h for p in diag])
        _check(lib().fides_eval_linear_transform(self.ctx.h, self.h, hs, n, bstep, stride, offset,
                                                 1 if batched_giant else 0, o.h))
        return o

    def convolution_transform(self, pts: list, bstep: int, indexes: list, gstep: int, stride: int = 1, out=None):
        """ConvolutionTransform: u_j = sum_i pts[b j + i] * rot_{indexes[i]}(self), giant steps in blocks of 8,
        u_j rotated by stride (G - j mod 8) within its block of G, block k of nb by 8 stride (nb - 1 - k), summed."""
        o = self._out(out)
        n = len(pts)
        hs = (C.c_void_p * n)(*[p.h for p in pts])
        ix = (C.c_int32 * max(1, len(indexes)))(*[int(v) for v in indexes])
        if len(indexes) != bstep:
            raise ValueError("need one baby-step index per baby step")
        _check(lib().fides_eval_convolution_transform(self.ctx.h, self.h, hs, n, bstep, stride, ix, gstep, o.h))
        return o

    def key_switch(self, key_id: int, out=None):
        """Ciphertext::keySwitch with the switching key loaded under key_id."""
        o = self._out(out)
        _check(lib().fides_eval_key_switch(self.ctx.h, self.h, key_id, o.h))
        return o

    def negate(self, out=None):
        """EvalNegate: -self exactly, metadata unchanged."""
        o = self._out(out)
        _check(lib().fides_eval_negate(self.ctx.h, self.h, o.h))
        return o

    def accumulate_sum(self, size: int, stride: int = 1, out=None):
        """AccumulateSum: sum over k < size of the rotation by k stride (radix-4 hoisted steps)."""
        o = self._out(out)
        _check(lib().fides_eval_accumulate_sum(self.ctx.h, self.h, size, stride, o.h))
        return o

    def automorph(self, k: int, out=None):
        o = self._out(out)
        _check(lib().fides_eval_automorph(self.ctx.h, self.h, k, o.h))
        return o

    def conjugate(self, out=None):
        o = self._out(out)
        _check(lib().fides_eval_conjugate(self.ctx.h, self.h, o.h))
        return o

    def mod_raise(self, level: int, out=None):
        o = self._out(out)
        _check(lib().fides_mod_raise(self.ctx.h, self.h, level, o.h))
        return o

    def mult_monomial(self, power: int, out=None):
        o = self._out(out)
        _check(lib().fides_eval_mult_monomial(self.ctx.h, self.h, power, o.h))
        return o

    def mult_int(self, s: int, out=None):
        o = self._out(out)
        _check(lib().fides_eval_mult_int(self.ctx.h, self.h, s, o.h))
        return o

    def horner_series(self, coeffs, out=None):
        """sum_i coeffs[i] * self^i (Horner, FIDESlib evalHornerSeries)."""
        o = self._out(out)
        cf = (C.c_double * len(coeffs))(*[float(v) for v in coeffs])
        _check(lib().fides_eval_horner_series(self.ctx.h, self.h, cf, len(coeffs), o.h))
        return o

    def chebyshev_series(self, coeffs, a: float = -1.0, b: float = 1.0, out=None):
        """sum_k coeffs[k] T_k((2 self - (a + b)) / (b - a)) (Paterson-Stockmeyer)."""
        o = self._out(out)
        cf = (C.c_double * len(coeffs))(*[float(v) for v in coeffs])
        _check(lib().fides_eval_chebyshev_series(self.ctx.h, self.h, cf, len(coeffs), a, b, o.h))
        return o

    def mod_reduction(self, coeffs, double_angle_its: int, post: int = 1, out=None):
        """EvalMod (sparse): Chebyshev series on [-1, 1], double-angle iterations, times post."""
        o = self._out(out)
        cf = (C.c_double * len(coeffs))(*[float(v) for v in coeffs])
        _check(lib().fides_eval_mod_reduction(self.ctx.h, self.h, cf, len(coeffs), double_angle_its, post, o.h))
        return o

    def double_angle(self, iterations: int, out=None):
        o = self._out(out)
        _check(lib().fides_eval_double_angle(self.ctx.h, self.h, iterations, o.h))
        return o

    def drop_to_level(self, level: int):
        _check(lib().fides_drop_to_level(self.ctx.h, self.h, level))
        return self

    def rotate_hoisted(self, indices: list, outs: list | None = None, ext: bool = False):
        """One ModUp shared by every rotation; with ext the outputs stay in P*Q form
        (rotate_hoisted(.., ext = true)) until :meth:`mod_down`."""
        n = len(indices)
        outs = [Ciphertext(self.ctx) for _ in indices] if outs is None else outs
        idx = (C.c_int32 * n)(*indices)
        hs = (C.c_void_p * n)(*[o.h for o in outs])
        fn = lib().fides_eval_rotate_hoisted_ext if ext else lib().fides_eval_rotate_hoisted
        _check(fn(self.ctx.h, self.h, idx, n, hs))
        return outs

    # ---- P*Q (extended-basis) form ----
    @property
    def is_ext(self) -> bool:
        v = C.c_int32()
        _check(lib().fides_ct_is_ext(self.h, C.byref(v)))
        return bool(v.value)

    def extend(self, out=None):
        """Ciphertext::extend: (P c0, P c1) in P*Q form."""
        o = self._out(out)
        _check(lib().fides_ct_extend(self.ctx.h, self.h, o.h))
        return o

    def mod_down(self, out=None):
        """Ciphertext::modDown: P*Q form back to the Q basis (out may be self)."""
        o = self._out(out)
        _check(lib().fides_ct_mod_down(self.ctx.h, self.h, o.h))
        return o

    def mult_ext(self, other: "Ciphertext", out=None):
        """Ciphertext::mult(b, false, moddown = false): the relinearized product left in P*Q."""
        o = self._out(out)
        _check(lib().fides_eval_mult_ext(self.ctx.h, self.h, other.h, o.h))
        return o

    def sub_plain(self, pt: "Plaintext", out=None):
        """Ciphertext::subPt: c0 - pt."""
        o = self._out(out)
        _check(lib().fides_eval_sub_plain(self.ctx.h, self.h, pt.h, o.h))
        return o

    def add_mult_plain_(self, a: "Ciphertext", pt: "Plaintext", rescale: bool = False):
        """Ciphertext::addMultPt in place: self += a * pt."""
        _check(lib().fides_eval_add_mult_plain(self.ctx.h, self.h, a.h, pt.h, int(rescale)))
        return self

    def add_mult_scalar_(self, b: "Ciphertext", d: float):
        """Ciphertext::addMultScalar in place: self += b * d."""
        _check(lib().fides_eval_add_mult_scalar(self.ctx.h, self.h, b.h, d))
        return self

    def add_ext(self, other: "Ciphertext", out=None):
        """Ciphertext::add with P*Q operands (a Q operand is extended first)."""
        o = self._out(out)
        _check(lib().fides_eval_add_ext(self.ctx.h, self.h, other.h, o.h))
        return o

    def save(self, path: str):
        """The engine's ciphertext file (either form)."""
        _check(lib().fides_ct_save(self.ctx.h, self.h, os.fsencode(path)))

    def store_ext(self):
        """Both parts of a P*Q ciphertext: [level + 1 + K, N] each."""
        n = self.level + 1 + self.ctx.raw.K
        c0 = np.empty((n, self.ctx.N), dtype=np.uint64)
        c1 = np.empty((n, self.ctx.N), dtype=np.uint64)
        _check(lib().fides_ct_store_ext(self.ctx.h, self.h, c0.ctypes.data, c1.ctypes.data))
        return c0, c1


class Plaintext:
    """FIDESlib::CKKS::Plaintext: one device-resident polynomial in evaluation form."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().fides_pt_create(ctx.h, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib().fides_pt_destroy(self.h)
            self.h = None
        except Exception:
            pass

    def load(self, limbs, noise: float = 1.0, noise_level: int = 1, slots: int | None = None):
        slots = self.ctx.N // 2 if slots is None else slots
        a = np.ascontiguousarray(limbs, dtype=np.uint64)
        if a.ndim != 2 or a.shape[1] != self.ctx.N:
            raise ValueError("plaintext limbs must be [num_res, N] uint64")
        _check(lib().fides_pt_load(self.ctx.h, self.h, a.ctypes.data, a.shape[0], noise, noise_level, slots))
        return self

    def save(self, path: str):
        _check(lib().fides_pt_save(self.ctx.h, self.h, os.fsencode(path)))

    def load_ext(self, limbs, noise: float = 1.0, noise_level: int = 1, slots: int | None = None):
        """A P*Q plaintext: [num_res + K, N], the Q limbs then the K special limbs."""
        slots = self.ctx.N // 2 if slots is None else slots
        a = np.ascontiguousarray(limbs, dtype=np.uint64)
        K = self.ctx.raw.K
        if a.ndim != 2 or a.shape[1] != self.ctx.N or a.shape[0] <= K:
            raise ValueError("P*Q plaintext limbs must be [num_res + K, N] uint64")
        _check(lib().fides_pt_load_ext(self.ctx.h, self.h, a.ctypes.data, a.shape[0] - K, noise, noise_level, slots))
        return self


class Server:
    """Request loop with continuous batching (fides_server_*): submit jobs on
    device-resident operands, call :meth:`step` to advance every active job by
    one stage, :meth:`poll` for the ids of completed jobs."""

    def __init__(self, ctx: Context, max_slots: int = 64, budget_bytes: int = 64 << 30):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().fides_server_create(ctx.h, max_slots, budget_bytes, C.byref(h)))
        self.h = h
        self._keep = {}  # job id -> operands kept alive until the job is done

    def close(self):
        if getattr(self, "h", None):
            _check(lib().fides_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _job(self, rc, jid, keep):
        _check(rc)
        self._keep[jid.value] = keep
        return jid.value

    def submit_mult(self, a: Ciphertext, b: Ciphertext, out: Ciphertext) -> int:
        jid = C.c_int64()
        return self._job(lib().fides_server_submit_mult(self.h, a.h, b.h, out.h, C.byref(jid)), jid, (a, b, out))

    def submit_rotate(self, a: Ciphertext, index: int, out: Ciphertext) -> int:
        jid = C.c_int64()
        return self._job(lib().fides_server_submit_rotate(self.h, a.h, index, out.h, C.byref(jid)), jid, (a, out))

    def submit_logreg(self, x: Ciphertext, w: "Plaintext", n_rot: int, coeffs, out: Ciphertext) -> int:
        jid = C.c_int64()
        cf = (C.c_double * 3)(*[float(v) for v in coeffs])
        return self._job(lib().fides_server_submit_logreg(self.h, x.h, w.h, n_rot, cf, out.h, C.byref(jid)), jid,
                         (x, w, out))

    def step(self) -> int:
        p = C.c_int32()
        _check(lib().fides_server_step(self.h, C.byref(p)))
        return p.value

    def poll(self) -> list:
        buf = (C.c_int64 * 256)()
        n = C.c_int32()
        out = []
        while True:
            _check(lib().fides_server_poll(self.h, buf, 256, C.byref(n)))
            out.extend(buf[: n.value])
            if n.value < 256:
                break
        for j in out:
            self._keep.pop(j, None)
        return out

    def counts(self):
        a, q, u = C.c_int32(), C.c_int32(), C.c_uint64()
        _check(lib().fides_server_counts(self.h, C.byref(a), C.byref(q), C.byref(u)))
        return {"active": a.value, "queued": q.value, "used_bytes": u.value}

    def drain(self) -> list:
        """Step until no job is active or queued, then wait for the device; all completed ids."""
        while self.step() > 0:
            pass
        self.ctx.synchronize()
        return self.poll()
