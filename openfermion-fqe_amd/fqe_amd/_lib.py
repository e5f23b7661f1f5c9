"""ctypes binding of libfqehip.so (the HIP kernels behind the C ABI of
``include/fqehip.h``).

torch is imported first so that the HIP runtime torch ships
(``libamdhip64.so.7``) is the one the library binds to: device pointers and
streams from torch are then valid in the library.  There is no CPU
fallback: a missing library or a missing GPU raises.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_void_p

import torch  # noqa: F401  (must precede the CDLL load, see module doc)

LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "libfqehip.so")

C128 = 0
C64 = 1

TILE_SPLIT_F64_DTILE = 32   # FQEHIP_TILE_SPLIT_F64_DTILE
TILE_OZAKI = 64   # FQEHIP_TILE_OZAKI

OK = 0
ERR_INVALID = 1
ERR_OOM = 2
ERR_UNSUPPORTED = 3
ERR_HIP = 4


class FqeHipError(RuntimeError):
    """A fqehip C-ABI call returned a non-zero status."""


class Sector(ctypes.Structure):
    _fields_ = [("norb", c_int), ("nalpha", c_int), ("nbeta", c_int),
                ("npair", c_int), ("lena", c_int64), ("lenb", c_int64),
                ("ptab_a", c_void_p), ("ptab_b", c_void_p),
                ("fwd_a", c_void_p), ("fwd_b", c_void_p),
                ("slot_a", c_void_p), ("nslot", c_int),
                ("hbb_hi", c_void_p), ("hbb_lo", c_void_p), ("ldh", c_int64),
                ("tile", c_int), ("hbb_f64", c_void_p), ("hbb_oz", c_void_p)]


class Operator(ctypes.Structure):
    _fields_ = [("complex_op", c_int), ("a", c_void_p), ("h1", c_void_p)]


class BatchItem(ctypes.Structure):
    _fields_ = [("sector", Sector), ("op", Operator), ("d_coeff0", c_void_p),
                ("d_apply_out", c_void_p), ("d_term", c_void_p),
                ("d_result", c_void_p), ("d_workspace", c_void_p)]


_SIGNATURES = {
    "fqehip_last_error": (ctypes.c_char_p, []),
    "fqehip_version": (c_int, []),
    "fqehip_string_tables": (c_int, [c_int, c_int, c_int64, c_void_p,
                                     c_void_p, c_void_p]),
    "fqehip_string_maps": (c_int, [c_int, c_int, c_int64, c_void_p, c_void_p,
                                   c_void_p, c_void_p]),
    "fqehip_pair_table": (c_int, [c_int, c_int, c_int64, c_void_p, c_int,
                                  c_void_p, c_void_p]),
    "fqehip_dvec_gather": (c_int, [c_int, c_int, c_int64, c_int64, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p]),
    "fqehip_pair_gemm": (c_int, [c_int, c_int, c_int, c_void_p, c_int64,
                                 c_void_p, c_void_p, c_void_p]),
    "fqehip_scatter": (c_int, [c_int, c_int, c_int64, c_int64, c_void_p,
                               c_void_p, c_void_p, c_double, c_double,
                               c_double, c_double, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p]),
    "fqehip_norm_partials_count": (c_int64, [c_int64]),
    "fqehip_apply1": (c_int, [c_int, c_int, c_int64, c_int64, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_double,
                              c_double, c_double, c_double, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p]),
    "fqehip_norm_finalize": (c_int, [c_void_p, c_int64, c_void_p, c_double,
                                     c_int, c_void_p, c_void_p, c_void_p]),
    "fqehip_convert_from_c128": (c_int, [c_int, c_int64, c_void_p, c_void_p,
                                         c_void_p]),
    "fqehip_convert_to_c128": (c_int, [c_int, c_int64, c_void_p, c_double,
                                       c_double, c_void_p, c_void_p]),
    "fqehip_fused_tables": (c_int, [c_int, c_int, c_int64, c_void_p, c_int,
                                    c_void_p, c_void_p, c_void_p, c_void_p]),
    "fqehip_fused_supported": (c_int, [c_int, POINTER(Sector),
                                       POINTER(Operator)]),
    "fqehip_apply_fused_rows": (c_int, [c_int, POINTER(Sector),
                                        POINTER(Operator), c_void_p, c_void_p,
                                        c_void_p, c_void_p]),
    "fqehip_apply_slot_reduce": (c_int, [c_int, POINTER(Sector), c_void_p,
                                         c_double, c_double, c_double,
                                         c_double, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p,
                                         c_void_p]),
    "fqehip_apply_workspace_bytes": (c_int64, [c_int, POINTER(Sector)]),
    "fqehip_beta_block_ld": (c_int64, [c_int64]),
    "fqehip_batch_create": (c_int, [c_int, c_int, POINTER(BatchItem), c_double,
                                    c_int, c_int, c_int, c_void_p,
                                    POINTER(c_void_p)]),
    "fqehip_batch_launch": (c_int, [c_void_p, c_void_p]),
    "fqehip_batch_nodes": (c_int, [c_void_p]),
    "fqehip_batch_grouped": (c_int, [c_void_p]),
    "fqehip_batch_destroy": (None, [c_void_p]),
    "fqehip_diagonal": (c_int, [c_int, c_int64, c_int64, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p]),
    "fqehip_block_dot_workspace_bytes": (c_int64, [c_int]),
    "fqehip_gram_workspace_bytes": (c_int64, [c_int64, c_int]),
    "fqehip_gram": (c_int, [c_int64, c_int, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p]),
    "fqehip_block_dot": (c_int, [c_int64, c_int, c_void_p, c_int64, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "fqehip_block_combine": (c_int, [c_int64, c_int, c_void_p, c_int64,
                                     c_void_p, c_double, c_double, c_void_p,
                                     c_void_p]),
    "fqehip_block_combine_multi": (c_int, [c_int64, c_int, c_void_p, c_int64,
                                           c_void_p, c_int, c_void_p, c_int64,
                                           c_void_p]),
    "fqehip_set_length": (c_int64, [c_int, c_int]),
    "fqehip_set_workspace_bytes": (c_int64, [c_int, c_int]),
    "fqehip_set_dvec": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "fqehip_set_contract": (c_int, [c_int, c_int, c_void_p, c_void_p,
                                    c_void_p]),
    "fqehip_set_apply": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "fqehip_mlp_features": (c_int, [c_int, c_int64, c_int64, c_void_p,
                                    c_void_p, c_void_p, c_double, c_double,
                                    c_int64, c_int64, c_int64, c_void_p,
                                    c_void_p, c_void_p]),
    "fqehip_mlp_correction": (c_int, [c_void_p, c_int, c_int64, c_int64,
                                      c_void_p, c_void_p, c_void_p, c_double,
                                      c_double, c_int64, c_int64, c_void_p,
                                      c_void_p, c_void_p]),
    "fqehip_mlp_forward": (c_int, [c_void_p, c_int64, c_void_p, c_void_p,
                                   c_void_p]),
    "fqehip_mlp_train_workspace_floats": (c_int64, [c_int64]),
    "fqehip_mlp_loss_grad": (c_int, [c_void_p, c_int64, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p]),
    "fqehip_mlp_adam": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, c_void_p]),
    "fqehip_davidson_precondition": (c_int, [c_int64, c_void_p, c_void_p,
                                             c_double, c_double, c_void_p,
                                             c_void_p, c_void_p]),
    "fqehip_beta_block": (c_int, [c_int, c_int64, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_int64,
                                  c_void_p]),
    "fqehip_beta_gemm": (c_int, [c_int64, c_int64, c_void_p, c_void_p,
                                 c_int64, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "fqehip_beta_gemm_f64": (c_int, [c_int64, c_int64, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]),
    "fqehip_beta_block_digits_bytes": (c_int64, [c_int64]),
    "fqehip_beta_block_digits": (c_int, [c_int64, c_void_p, c_void_p,
                                         c_void_p]),
    "fqehip_beta_gemm_digits": (c_int, [c_int64, c_int64, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    "fqehip_cheb_update": (c_int, [c_int, c_int64, c_void_p, c_void_p,
                                   c_double, c_double, c_double, c_double,
                                   c_void_p, c_void_p, c_void_p, c_void_p]),
    "fqehip_apply": (c_int, [c_int, POINTER(Sector), POINTER(Operator),
                             c_void_p, c_double, c_double, c_double, c_double,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p]),
    "fqehip_apply_panels": (c_int, [c_int, POINTER(Sector),
                                    POINTER(Operator), c_void_p, c_double,
                                    c_double, c_double, c_double, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_int64, c_void_p]),
    "fqehip_apply_noalloc": (c_int, [c_int, POINTER(Sector),
                                     POINTER(Operator), c_void_p, c_double,
                                     c_double, c_double, c_double, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "fqehip_taylor_graph_create": (c_int, [c_int, POINTER(Sector),
                                           POINTER(Operator), c_double,
                                           c_double, c_int, c_void_p,
                                           c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p,
                                           POINTER(c_void_p)]),
    "fqehip_taylor_graph_launch": (c_int, [c_void_p, c_void_p]),
    "fqehip_taylor_graph_destroy": (None, [c_void_p]),
    "fqehip_taylor": (c_int, [c_int, POINTER(Sector), POINTER(Operator),
                              c_double, c_double, c_int, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p]),
}

_lib = None


def lib():
    """The loaded library; raises if it is not built or no GPU is present."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise FqeHipError(
                f"{LIBPATH} is missing: build it (make -C openfermion-fqe_amd)")
        handle = ctypes.CDLL(LIBPATH)
        for name, (restype, argtypes) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


def device() -> torch.device:
    """The GPU the engine runs on; raises without one (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise FqeHipError("fqe_amd needs an AMD GPU (torch.cuda is not "
                          "available); there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def call(name: str, *args) -> int:
    rc = getattr(lib(), name)(*args)
    if name.endswith(("_count", "_bytes", "_supported", "_ld", "_floats",
                      "_nodes", "_grouped", "_destroy", "_length")):
        return rc
    if rc != OK:
        msg = lib().fqehip_last_error().decode(errors="replace")
        if rc == ERR_INVALID:
            raise ValueError(f"{name}: {msg}")
        raise FqeHipError(f"{name} failed with status {rc}: {msg}")
    return rc
