"""State samplers over libompl_amd.

* sample_uniform: RealVectorStateSampler::sampleUniform
  (base/spaces/src/RealVectorStateSpace.cpp:45-52) for many generators at once;
  generator s is std::mt19937 seeded seeds[s] with ompl::RNG::uniformReal
  semantics (util/RandomNumbers.h:72-76).
* sample_valid: UniformValidStateSampler::sample
  (base/samplers/src/UniformValidStateSampler.cpp:46-57) for many samplers at
  once: each call draws until a valid state or `attempts` draws.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ARM_DOF, _ptr, _stream, _torch
from ._lib import call

__all__ = ["sample_uniform", "sample_valid", "Rng", "MAX_VALID_SAMPLE_ATTEMPTS"]

MAX_VALID_SAMPLE_ATTEMPTS = 100  # tools/config/MagicConstants.h:90


def _host(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype))


def sample_uniform(seeds, n_per_stream: int, lo, hi, f32_soa: bool = False, stream=None):
    """Device tensors: states [n_streams, n_per_stream, dim] float64 (and the
    float32 dof-major copy [dim, n_streams * n_per_stream] when f32_soa)."""
    torch = _torch()
    sd = _host(seeds, np.uint32)
    lo, hi = _host(lo, np.float64), _host(hi, np.float64)
    dim = lo.shape[0]
    out = torch.empty((sd.shape[0], n_per_stream, dim), dtype=torch.float64, device="cuda")
    f32 = torch.empty((dim, sd.shape[0] * n_per_stream), dtype=torch.float32, device="cuda") if f32_soa else None
    call("ompl_amd_sample_uniform_dev", sd.ctypes.data_as(C.c_void_p), sd.shape[0], int(n_per_stream), dim,
         lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), _ptr(out), _ptr(f32), _stream(stream))
    return (out, f32) if f32_soa else out


def sample_valid(env, seeds, n_calls: int, lo=None, hi=None, attempts: int = MAX_VALID_SAMPLE_ATTEMPTS, stream=None):
    """Returns device tensors (states [n_streams, n_calls, 7] float64,
    found [n_streams, n_calls] uint8, draws [n_streams, n_calls] int32)."""
    torch = _torch()
    sd = _host(seeds, np.uint32)
    lo = _host(np.full(ARM_DOF, -np.pi) if lo is None else lo, np.float64)
    hi = _host(np.full(ARM_DOF, np.pi) if hi is None else hi, np.float64)
    ns = sd.shape[0]
    out = torch.empty((ns, n_calls, ARM_DOF), dtype=torch.float64, device="cuda")
    found = torch.empty((ns, n_calls), dtype=torch.uint8, device="cuda")
    draws = torch.empty((ns, n_calls), dtype=torch.int32, device="cuda")
    call("ompl_amd_sample_valid_dev", env.handle, sd.ctypes.data_as(C.c_void_p), ns, int(n_calls), int(attempts),
         lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), _ptr(out), _ptr(found), _ptr(draws),
         _stream(stream))
    return out, found, draws


class Rng:
    """Device-resident generators that persist between calls: stream s is
    std::mt19937 seeded seeds[s], and every call continues it exactly as
    successive calls of one reference sampler would."""

    def __init__(self, seeds):
        sd = _host(seeds, np.uint32)
        self.n = int(sd.shape[0])
        h = C.c_void_p()
        call("ompl_amd_rng_create", sd.ctypes.data_as(C.c_void_p), self.n, C.byref(h))
        self.handle = h

    def uniform(self, n_per_stream: int, lo, hi, stream=None):
        torch = _torch()
        lo, hi = _host(lo, np.float64), _host(hi, np.float64)
        out = torch.empty((self.n, n_per_stream, lo.shape[0]), dtype=torch.float64, device="cuda")
        call("ompl_amd_rng_sample_uniform_dev", self.handle, int(n_per_stream), lo.shape[0],
             lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), _ptr(out), None, _stream(stream))
        return out

    def se3(self, n_per_stream: int, lo, hi, stream=None):
        """SE3StateSampler::sampleUniform: device float64 and float32 [n_streams * n_per_stream, 7]
        rows (x, y, z, qx, qy, qz, qw)."""
        torch = _torch()
        lo, hi = _host(lo, np.float64), _host(hi, np.float64)
        out = torch.empty((self.n * n_per_stream, 7), dtype=torch.float64, device="cuda")
        out32 = torch.empty((self.n * n_per_stream, 7), dtype=torch.float32, device="cuda")
        call("ompl_amd_rng_sample_se3_dev", self.handle, int(n_per_stream), lo.ctypes.data_as(C.c_void_p),
             hi.ctypes.data_as(C.c_void_p), _ptr(out), _ptr(out32), _stream(stream))
        return out, out32

    def valid(self, env, n_calls: int, lo=None, hi=None, attempts: int = MAX_VALID_SAMPLE_ATTEMPTS, stream=None):
        torch = _torch()
        lo = _host(np.full(ARM_DOF, -np.pi) if lo is None else lo, np.float64)
        hi = _host(np.full(ARM_DOF, np.pi) if hi is None else hi, np.float64)
        out = torch.empty((self.n, n_calls, ARM_DOF), dtype=torch.float64, device="cuda")
        found = torch.empty((self.n, n_calls), dtype=torch.uint8, device="cuda")
        draws = torch.empty((self.n, n_calls), dtype=torch.int32, device="cuda")
        call("ompl_amd_rng_sample_valid_dev", self.handle, env.handle, int(n_calls), int(attempts),
             lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), _ptr(out), _ptr(found), _ptr(draws),
             _stream(stream))
        return out, found, draws

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            call("ompl_amd_rng_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
