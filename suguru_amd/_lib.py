"""ctypes binding of libibx.so (the C ABI in include/ibx.h).

The library is the product: if it is missing, loading fails loudly -- there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libibx.so")
HEADER = os.path.join(ROOT, "include", "ibx.h")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)

IBX_OK = 0
IBX_E_ARG = -1
IBX_E_HIP = -2
IBX_E_DOMAIN = -3
IBX_E_CAPACITY = -4
IBX_E_STATE = -5


class IbxGrid(ctypes.Structure):
    _fields_ = [
        ("itot", ctypes.c_int32),
        ("jtot", ctypes.c_int32),
        ("ktot", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("lx", ctypes.c_double),
        ("ly", ctypes.c_double),
        ("lz", ctypes.c_double),
        ("dxi", ctypes.c_double),
        ("dyi", ctypes.c_double),
        ("dzi", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("dz", ctypes.c_double),
        ("boundleft", ctypes.c_double),
        ("boundfront", ctypes.c_double),
    ]


class IbxBatchBufs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "u", "v", "w", "x", "y", "z", "fx", "fy", "fz", "area",
        "ul", "vl", "wl", "forcex", "forcey", "forcez")]


def grid_struct(g) -> IbxGrid:
    return IbxGrid(g.itot, g.jtot, g.ktot, 0, g.lx, g.ly, g.lz, g.dxi, g.dyi, g.dzi, g.dx,
                   g.dy, g.dz, g.boundleft, g.boundfront)


_lib = None


def declared_symbols(header: str = HEADER):
    """Names of the functions include/ibx.h declares."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ibx_[a-z0-9_]+)\s*\(", text)))


def torch_runtime_first() -> None:
    """PyTorch ships its own HIP runtime beside the system one these libraries link.  In one
    process both work only when PyTorch's is initialised first, so initialise it before the
    first call into libibx / libmemb (a no-op without PyTorch or without a HIP device)."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m suguru_amd.build` "
            "(there is no fallback implementation)")
    torch_runtime_first()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    D = ctypes.c_double
    sig = {
        "ibx_create": (ctypes.c_int, [ctypes.POINTER(P), ctypes.POINTER(IbxGrid), ctypes.c_int]),
        "ibx_destroy": (None, [P]),
        "ibx_last_error": (ctypes.c_char_p, [P]),
        "ibx_set_stream": (ctypes.c_int, [P, P]),
        "ibx_sync": (ctypes.c_int, [P]),
        "ibx_version": (ctypes.c_char_p, []),
        "ibx_kernel": (D, [D]),
        "ibx_kernel_dev": (ctypes.c_int, [P, I64, P, P]),
        "ibx_kerneltest_dev": (ctypes.c_int, [P, ctypes.POINTER(D)]),
        "ibx_collisions_dev": (ctypes.c_int, [P, I64, P, P, P, D, D, D]),
        "ibx_shift_z_dev": (ctypes.c_int, [P, I64, P, D]),
        "ibx_eulr2lagr": (ctypes.c_int, [P, P, P, P, I64, P, P, P, P, P, P, P]),
        "ibx_lagr2eulr": (ctypes.c_int, [P, I64, P, P, P, P, P, P, P, P, P, P, P]),
        "ibx_lagr2eulrind": (ctypes.c_int, [P, I64, P, P, P, P, P, P, P, P, P, P, P]),
        "ibx_interp_f64_dev": (ctypes.c_int, [P, P, P, P, I64, P, I64, P, P, P, P, P, P]),
        "ibx_spread_f64_dev": (ctypes.c_int, [P, I64, P, I64, P, P, P, P, P, P, P, P, P, P]),
        "ibx_spread_cc_f64_dev": (ctypes.c_int, [P, I64, P, I64, P, P, P, P, P, P, P, P, P, P]),
        "ibx_spread_f64_dev_noalloc": (ctypes.c_int, [P, I64, P, I64, P, P, P, P, P, P, P, P, P, P]),
        "ibx_update_pos": (ctypes.c_int, [P, I64, P, D, P, P, P, P, P, P, P, P, P, P, P]),
        "ibx_update_pos_dev": (ctypes.c_int, [P, I64, P, I64, D, P, P, P, P, P, P, P, P, P, P]),
        "ibx_batch_create": (ctypes.c_int, [P, ctypes.POINTER(P), I64, P]),
        "ibx_batch_bind": (ctypes.c_int, [P, ctypes.POINTER(IbxBatchBufs)]),
        "ibx_batch_autotune": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
        "ibx_batch_set_config": (ctypes.c_int, [P, ctypes.c_int]),
        "ibx_set_config": (ctypes.c_int, [P, ctypes.c_int]),
        "ibx_batch_set_overlap": (ctypes.c_int, [P, ctypes.c_int]),
        "ibx_batch_num_configs": (ctypes.c_int, []),
        "ibx_batch_config_name": (ctypes.c_char_p, [ctypes.c_int]),
        "ibx_batch_capture": (ctypes.c_int, [P, ctypes.c_int]),
        "ibx_batch_capture_timed": (ctypes.c_int, [P, ctypes.c_int]),
        "ibx_batch_run": (ctypes.c_int, [P]),
        "ibx_batch_reset_slots": (ctypes.c_int, [P]),
        "ibx_batch_slot_times": (ctypes.c_int, [P, ctypes.POINTER(D)]),
        "ibx_batch_profile": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(D)]),
        "ibx_batch_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


# ---- libmemb.so: the membrane law and its learned surrogate (include/memb.h) ----------------
MEMB_PATH = os.path.join(HERE, "libmemb.so")
MEMB_HEADER = os.path.join(ROOT, "include", "memb.h")
_memb = None


def declared_memb_symbols(header: str = MEMB_HEADER):
    """Names of the functions include/memb.h declares."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(memb_[a-z0-9_]+)\s*\(", text)))


def load_memb() -> ctypes.CDLL:
    global _memb
    if _memb is not None:
        return _memb
    if not os.path.exists(MEMB_PATH):
        raise ImportError(
            f"{MEMB_PATH} is missing: build it with `python -m suguru_amd.build` "
            "(there is no fallback implementation)")
    torch_runtime_first()
    lib = ctypes.CDLL(MEMB_PATH)
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    D = ctypes.c_double
    F = ctypes.c_float
    I = ctypes.c_int
    sig = {
        "memb_plan_create": (I, [ctypes.POINTER(P), I, I]),
        "memb_plan_destroy": (None, [P]),
        "memb_last_error": (ctypes.c_char_p, [P]),
        "memb_set_stream": (I, [P, P]),
        "memb_sync": (I, [P]),
        "memb_npts": (I, [P]),
        "memb_ncoef": (I, [P]),
        "memb_traction_f64_dev": (I, [P, I64, D, D] + [P] * 10),
        "memb_coeffs_f64_dev": (I, [P, I64, P, P, P, P]),
        "memb_geometry_f64_dev": (I, [P, I64, P, P, P, P, P, P, P, P]),
        "memb_traction_select_f64_dev": (I, [P, I64, P, D, D] + [P] * 10),
        "memb_corot_f64_dev": (I, [P, I64] + [P] * 10),
        "memb_corot_apply_f64_dev": (I, [P, I64] + [P] * 7),
        "memb_anchored_predict_f64_dev": (I, [P, I64] + [P] * 6 + [D, D] + [P] * 7),
        "memb_anchored_gate_f64_dev": (I, [P, I64, P, P, D, D, P]),
        "memb_anchored_update_f64_dev": (I, [P, I64] + [P] * 15),
        "memb_compact_sel_dev": (I, [P, I64, P, I, P, P]),
        "memb_mlp_create": (I, [ctypes.POINTER(P), I, I, I, I, I, I]),
        "memb_mlp_destroy": (None, [P]),
        "memb_mlp_last_error": (ctypes.c_char_p, [P]),
        "memb_mlp_set_stream": (I, [P, P]),
        "memb_mlp_num_layers": (I, [P]),
        "memb_mlp_layer_shape": (I, [P, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "memb_mlp_set_layer": (I, [P, I, P, P]),
        "memb_mlp_get_layer": (I, [P, I, P, P]),
        "memb_mlp_set_norm": (I, [P, P, P, P, P]),
        "memb_mlp_init": (I, [P, ctypes.c_uint64]),
        "memb_mlp_forward": (I, [P, I, P, P]),
        "memb_mlp_train_step": (I, [P, I, P, P, P, F, ctypes.POINTER(F)]),
        "memb_mlp_get_grad": (I, [P, I, P, P]),
        "memb_mlp_grad_only": (I, [P, I, P, P, P, ctypes.POINTER(F)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _memb = lib
    return lib


# ---- libfluid.so: the flow update (include/fluid.h) -------------------------------------------
FLUID_PATH = os.path.join(HERE, "libfluid.so")
FLUID_HEADER = os.path.join(ROOT, "include", "fluid.h")
_fluid = None


def declared_fluid_symbols(header: str = FLUID_HEADER):
    """Names of the functions include/fluid.h declares."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fld_[a-z0-9_]+)\s*\(", text)))


def load_fluid() -> ctypes.CDLL:
    global _fluid
    if _fluid is not None:
        return _fluid
    if not os.path.exists(FLUID_PATH):
        raise ImportError(
            f"{FLUID_PATH} is missing: build it with `python -m suguru_amd.build` "
            "(there is no fallback implementation)")
    torch_runtime_first()
    lib = ctypes.CDLL(FLUID_PATH)
    P = ctypes.c_void_p
    D = ctypes.c_double
    I = ctypes.c_int
    sig = {
        "fld_create": (I, [ctypes.POINTER(P), ctypes.POINTER(IbxGrid), D, D, I]),
        "fld_destroy": (None, [P]),
        "fld_last_error": (ctypes.c_char_p, [P]),
        "fld_set_stream": (I, [P, P]),
        "fld_sync": (I, [P]),
        "fld_fftuvw_dev": (I, [P] + [P] * 13),
        "fld_solver1d_dev": (I, [P, P, I]),
        "fld_bounduvw_dev": (I, [P, P, P, P]),
        "fld_boundp_dev": (I, [P, P]),
        "fld_fillps_dev": (I, [P, P, P, P, P]),
        "fld_correc_dev": (I, [P, P, P, P, P]),
        "fld_step_dev": (I, [P] + [P] * 14),
        "fld_chkdiv_dev": (I, [P, P, P, P, ctypes.POINTER(D), ctypes.POINTER(D)]),
        "fld_indicator_div_dev": (I, [P, P, P, P, P]),
        "fld_indicator_finish_dev": (I, [P] + [P] * 5 + [D] * 5),
        "fld_indicator_dev": (I, [P] + [P] * 8 + [D] * 5),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _fluid = lib
    return lib
