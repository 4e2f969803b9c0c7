"""ctypes wrapper of oracle/nh_cpu.c (libnhcpu32.so: CPU baseline; libnhcpu64.so: FP64 oracle).

TEST AND BASELINE INFRASTRUCTURE ONLY: used by tests/ and by bench.py's cpu_baseline leg.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class CpuModel(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("kind", "LX", "LY", "n_up", "n_dn", "C", "L", "project", "net")] + [
        (n, ctypes.c_void_p) for n in ("w", "b", "wt", "base")]


_libs = {}


def _load(bits):
    if bits not in _libs:
        path = os.path.join(HERE, f"libnhcpu{bits}.so")
        if not os.path.exists(path):
            from . import build as _b

            _b.build()
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.nhcpu_psi.argtypes = [P, ctypes.c_int, P, P, P, ctypes.c_int]
        lib.nhcpu_local_energy.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, P, P,
                                           P, P, ctypes.c_int]
        _libs[bits] = lib
    return _libs[bits]


class CpuPath:
    """The whole path on host cores: psi (projected) and E_loc, in FP32 (bits=32) or FP64."""

    def __init__(self, workload, params, bits=32, project=True, threads=0):
        from neuralhubbard_amd.backflow import build_wt_ud, build_wtilde, pack_conv

        self.w = workload
        self.lib = _load(bits)
        self.threads = threads
        cw, cb = pack_conv(params)
        self._keep = [np.ascontiguousarray(cw, np.float32), np.ascontiguousarray(cb, np.float32)]
        if workload.kind == "pfaffian":
            wt = build_wtilde(params.W, params.theta, workload.LX, workload.LY)
            base = params.F0
        else:
            wt = build_wt_ud(params.W, params.theta, workload.LX, workload.LY)
            base = params.phi0
        self._keep.append(np.ascontiguousarray(wt, np.float32))
        self._keep.append(None if base is None else np.ascontiguousarray(base, np.float32))
        ptrs = [a.ctypes.data if a is not None else 0 for a in self._keep]
        self.m = CpuModel(0 if workload.kind == "pfaffian" else 1, workload.LX, workload.LY, workload.n_up,
                          workload.n_dn, workload.channels, workload.n_layers, 1 if project else 0,
                          (2 if params.head is not None else 1) if getattr(workload, "network", "conv") == "resconv"
                          else 0, *ptrs)

    def psi(self, s):
        s = np.ascontiguousarray(s, np.int8).reshape(-1, 2 * self.w.N)
        B = s.shape[0]
        sign = np.empty(B, np.float32)
        logabs = np.empty(B, np.float64)
        self.lib.nhcpu_psi(ctypes.byref(self.m), B, s.ctypes.data, sign.ctypes.data, logabs.ctypes.data,
                           self.threads)
        return sign, logabs

    def local_energy(self, s, sign, logabs, t=1.0, tprime=None, U=None):
        w = self.w
        s = np.ascontiguousarray(s, np.int8).reshape(-1, 2 * w.N)
        sign = np.ascontiguousarray(sign, np.float32)
        logabs = np.ascontiguousarray(logabs, np.float64)
        e = np.empty(s.shape[0], np.float64)
        self.lib.nhcpu_local_energy(ctypes.byref(self.m), t, w.tprime if tprime is None else tprime,
                                    w.U if U is None else U, s.shape[0], s.ctypes.data, sign.ctypes.data,
                                    logabs.ctypes.data, e.ctypes.data, self.threads)
        return e
