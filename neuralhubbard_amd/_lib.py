"""ctypes binding of libnh.so (the C ABI declared in include/nh.h).

The library is built in-tree by ``neuralhubbard_amd/build.py``.  Loading fails loudly when it
is missing, and every device entry point requires a visible GPU: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnh.so")

NH_OK, NH_ERR_ARG, NH_ERR_SHAPE, NH_ERR_WORKSPACE, NH_ERR_HIP, NH_ERR_UNSUPPORTED = range(6)
_NAMES = {1: "NH_ERR_ARG", 2: "NH_ERR_SHAPE", 3: "NH_ERR_WORKSPACE", 4: "NH_ERR_HIP", 5: "NH_ERR_UNSUPPORTED"}
NH_KIND_PFAFFIAN, NH_KIND_BCS = 0, 1
NH_PREC_FP32, NH_PREC_MIXED = 0, 1
NH_NET_CONV, NH_NET_RESCONV, NH_NET_RESCONV_HEAD = 0, 1, 2


class NhError(RuntimeError):
    pass


class NhSurrogate(ctypes.Structure):
    _fields_ = [("LX", ctypes.c_int32), ("LY", ctypes.c_int32), ("channels", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("params", ctypes.c_void_p)]


class NhProposal(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ph_neighbors", ctypes.c_int32), ("se_neighbors", ctypes.c_int32)]


MOVE_EXCHANGE, MOVE_PARTICLE_HOP, MOVE_SITE_EXCHANGE, MOVE_MIX = 0, 1, 2, 3


class NhModel(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("LX", ctypes.c_int32),
        ("LY", ctypes.c_int32),
        ("n_up", ctypes.c_int32),
        ("n_dn", ctypes.c_int32),
        ("channels", ctypes.c_int32),
        ("n_layers", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("project", ctypes.c_int32),
        ("network", ctypes.c_int32),
        ("conv_w", ctypes.c_void_p),
        ("conv_b", ctypes.c_void_p),
        ("wt", ctypes.c_void_p),
        ("base", ctypes.c_void_p),
        ("conv_w_lp", ctypes.c_void_p),
        ("conv_w_mfma", ctypes.c_void_p),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double
_S = ctypes.c_size_t
_MP = ctypes.POINTER(NhModel)

# name: (restype, argtypes)
SIGNATURES = {
    "nh_pairing_occ": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "nh_pairing_ud": (_I, [_P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "nh_slogpf": (_I, [_P, _I, _I, _P, _P, _P]),
    "nh_slogpf_workspace_size": (_S, [_I, _I]),
    "nh_slogpf_ws": (_I, [_P, _I, _I, _P, _P, _P, _P, _S]),
    "nh_slogdet": (_I, [_P, _I, _I, _P, _P, _P]),
    "nh_pairing_occ_vjp": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "nh_slogpf_vjp": (_I, [_P, _I, _I, _P, _P, _P]),
    "nh_slogdet_vjp": (_I, [_P, _I, _I, _P, _P, _P]),
    "nh_pairing_ud_vjp": (_I, [_P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "nh_features": (_I, [_P, _MP, _I, _P, _P]),
    "nh_features_workspace_size": (_S, [_MP, _I]),
    "nh_features_ws": (_I, [_P, _MP, _I, _P, _P, _P, _S]),
    "nh_log_psi_workspace_size": (_S, [_MP, _I]),
    "nh_log_psi": (_I, [_P, _MP, _I, _P, _P, _P, _P, _S]),
    "nh_log_psi_noalloc": (_I, [_P, _MP, _I, _P, _P, _P]),
    "nh_log_psi_jacobian_size": (_S, [_MP]),
    "nh_log_psi_jacobian_workspace_size": (_S, [_MP, _I]),
    "nh_log_psi_jacobian": (_I, [_P, _MP, _D, _I, _P, _P, _P, _P, _P, _S]),
    "nh_local_energy_workspace_size": (_S, [_MP, _D, _I]),
    "nh_local_energy_lowrank_workspace_size": (_S, [_MP, _D, _I]),
    "nh_local_energy_lowrank": (_I, [_P, _MP, _D, _D, _D, _I, _P, _P, _P, _P, _I, _P, _S]),
    "nh_local_energy": (_I, [_P, _MP, _D, _D, _D, _I, _P, _P, _P, _P, _I, _P, _S]),
    "nh_hops": (_I, [_P, _I, _I, _D, _D, _I, _I, _P, _I, _P, _P, _P, _P]),
    "nh_sur_nparams": (_I, [_I]),
    "nh_sur_forward": (_I, [_P, ctypes.POINTER(NhSurrogate), _I, _P, _P]),
    "nh_sur_grad_workspace_size": (_S, [ctypes.POINTER(NhSurrogate), _I]),
    "nh_sur_grad": (_I, [_P, ctypes.POINTER(NhSurrogate), _I, _P, _P, ctypes.c_float, _P, _P, _P, _S]),
    "nh_adam": (_I, [_P, _I, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _I]),
    "nh_metropolis_workspace_size": (_S, [_MP, _I]),
    "nh_metropolis_sweep": (_I, [_P, _MP, ctypes.POINTER(NhSurrogate), ctypes.POINTER(NhProposal), _I, _P, _P,
                                 _P, _P, ctypes.c_uint64, ctypes.c_int64, _I, _P, _P, _S]),
    "nh_profile_enable": (_I, [_I]),
    "nh_profile_query": (_I, [_I, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]),
    "nh_profile_reset": (_I, []),
}
PROF_FEATURES, PROF_PAIRING_LINALG, PROF_OTHER = 0, 1, 2

_lib = None


def load():
    """The loaded library with typed entry points (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NhError(f"{LIB_PATH} is missing: build it with neuralhubbard_amd/build.py")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(SIGNATURES)


def check(rc: int, what: str):
    if rc != NH_OK:
        raise NhError(f"{what} failed with {_NAMES.get(rc, rc)}")


def require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise NhError("no GPU visible: the backflow path runs only on the gfx950 HIP kernels")


def stream_handle(stream=None):
    import torch

    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)
