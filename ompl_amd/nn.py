"""Python mirror of ompl::NearestNeighbors<_T> (datastructures/NearestNeighbors.h:46-115)
over the device index of libompl_amd.  Elements are ids 0, 1, 2, ... in
insertion order; their coordinates live in HBM.

Semantics follow NearestNeighborsLinear (datastructures/NearestNeighborsLinear.h):
nearestK / nearestR return results sorted by distance (reportsSortedResults()
is true), nearestR is inclusive, nearest() raises when the structure is empty
(the reference throws ompl::Exception "No elements found ...").  Equal
distances are ordered by lower id.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import METRIC_L2, METRIC_SE3, _ptr, _stream, _torch
from ._lib import OmplAmdError, call, check, load

__all__ = ["NearestNeighborsGPU", "METRIC_L2", "METRIC_SE3"]


class NearestNeighborsGPU:
    def __init__(self, metric: int = METRIC_L2, dim: int = 7):
        self.metric, self.dim = metric, dim
        h = C.c_void_p()
        call("ompl_amd_nn_create", metric, dim, C.byref(h))
        self.handle = h

    # ---- NearestNeighbors<_T> interface
    def reportsSortedResults(self) -> bool:
        return True

    def clear(self) -> None:
        call("ompl_amd_nn_clear", self.handle)

    def add(self, pts) -> int:
        """Adds one point or a batch; returns the first id assigned."""
        first = C.c_int64()
        if hasattr(pts, "is_cuda") and pts.is_cuda:
            t = pts.reshape(-1, self.dim).contiguous()
            call("ompl_amd_nn_add_dev", self.handle, _ptr(t), t.shape[0], C.byref(first), _stream(None))
            _torch().cuda.current_stream().synchronize()
        else:
            a = np.ascontiguousarray(np.asarray(pts, np.float32).reshape(-1, self.dim))
            call("ompl_amd_nn_add", self.handle, a.ctypes.data_as(C.c_void_p), a.shape[0], C.byref(first))
        return first.value

    def remove(self, idx: int) -> bool:
        r = C.c_int()
        call("ompl_amd_nn_remove", self.handle, int(idx), C.byref(r))
        return bool(r.value)

    def size(self) -> int:
        n = C.c_int64()
        call("ompl_amd_nn_size", self.handle, C.byref(n))
        return n.value

    def nearest(self, q) -> int:
        a = np.ascontiguousarray(np.asarray(q, np.float32).reshape(self.dim))
        out = C.c_int64()
        call("ompl_amd_nn_nearest", self.handle, a.ctypes.data_as(C.c_void_p), C.byref(out))
        return out.value

    def nearestK(self, q, k: int):
        idx, dist = self.knn_batch(np.asarray(q, np.float32).reshape(1, self.dim), k)
        m = idx[0] >= 0
        return idx[0][m], dist[0][m]

    def nearestR(self, q, r: float):
        off, idx, dist = self.radius_batch(np.asarray(q, np.float32).reshape(1, self.dim), r)
        return idx, dist

    # ---- batched forms
    def knn_dev(self, q_dev, k: int, out_i=None, out_d=None, stream=None):
        torch = _torch()
        nq = q_dev.shape[0]
        if out_i is None:
            out_i = torch.empty((nq, k), dtype=torch.int32, device=q_dev.device)
        if out_d is None:
            out_d = torch.empty((nq, k), dtype=torch.float32, device=q_dev.device)
        call("ompl_amd_nn_knn_dev", self.handle, _ptr(q_dev), nq, int(k), _ptr(out_i), _ptr(out_d),
             _stream(stream))
        return out_i, out_d

    def knn_batch(self, q, k: int):
        torch = _torch()
        qd = torch.from_numpy(np.ascontiguousarray(q, np.float32)).cuda()
        i, d = self.knn_dev(qd, k)
        return i.cpu().numpy(), d.cpu().numpy()

    def radius_dev(self, q_dev, r: float, stream=None):
        """CSR radius results (offsets [nq+1], ids, distances) on the device.

        One library call when the outputs, sized from the last call's result
        count per query (at least 16 per query), hold every result; otherwise
        the call reports the total and a second call writes into outputs of
        exactly that size."""
        torch = _torch()
        nq = q_dev.shape[0]
        off = torch.empty(nq + 1, dtype=torch.int64, device=q_dev.device)
        total = C.c_int64()
        per_q = max(16, getattr(self, "_radius_per_query", 16))
        cap = max(1, min(nq * per_q, 1 << 26))
        for _ in range(2):
            idx = torch.empty(cap, dtype=torch.int32, device=q_dev.device)
            dist = torch.empty(cap, dtype=torch.float32, device=q_dev.device)
            code = load().ompl_amd_nn_radius_dev(self.handle, _ptr(q_dev), nq, float(r), _ptr(off), _ptr(idx),
                                                 _ptr(dist), cap, C.byref(total), _stream(stream))
            if code == 0 or total.value <= cap:
                check(code)
                break
            cap = total.value  # outputs too small: offsets and total are written
        else:
            check(code)
        if nq:
            self._radius_per_query = -(-int(total.value) // nq)
        return off, idx[:total.value], dist[:total.value]

    def knn_packed_dev(self, q_dev, k_dev, kmax: int, stream=None, trim: bool = True):
        """Packed batch: each query with its own k (int32 device array,
        clamped to [0, kmax]).  Returns CSR (offsets [nq+1], ids, distances);
        with trim=False the id and distance arrays keep their nq * kmax
        capacity and the call does not wait for the device."""
        torch = _torch()
        nq = q_dev.shape[0]
        off = torch.empty(nq + 1, dtype=torch.int64, device=q_dev.device)
        idx = torch.empty(max(nq * kmax, 1), dtype=torch.int32, device=q_dev.device)
        dist = torch.empty(max(nq * kmax, 1), dtype=torch.float32, device=q_dev.device)
        call("ompl_amd_nn_knn_packed_dev", self.handle, _ptr(q_dev), nq, _ptr(k_dev), int(kmax), _ptr(off),
             _ptr(idx), _ptr(dist), _stream(stream))
        if not trim:
            return off, idx, dist
        total = int(off[-1].item()) if nq else 0
        return off, idx[:total], dist[:total]

    def radius_packed_dev(self, q_dev, r_dev, stream=None):
        """Packed batch: each query with its own radius (float32 device
        array).  Returns CSR (offsets [nq+1], ids, distances)."""
        torch = _torch()
        nq = q_dev.shape[0]
        off = torch.empty(nq + 1, dtype=torch.int64, device=q_dev.device)
        total = C.c_int64()
        cap = max(1, min(nq * 16, 1 << 26))
        for _ in range(2):
            idx = torch.empty(cap, dtype=torch.int32, device=q_dev.device)
            dist = torch.empty(cap, dtype=torch.float32, device=q_dev.device)
            code = load().ompl_amd_nn_radius_packed_dev(self.handle, _ptr(q_dev), nq, _ptr(r_dev), _ptr(off),
                                                        _ptr(idx), _ptr(dist), cap, C.byref(total), _stream(stream))
            if code == 0 or total.value <= cap:
                check(code)
                break
            cap = total.value  # outputs too small: offsets and total are written
        else:
            check(code)
        return off, idx[:total.value], dist[:total.value]

    def radius_batch(self, q, r: float):
        torch = _torch()
        qd = torch.from_numpy(np.ascontiguousarray(q, np.float32)).cuda()
        off, i, d = self.radius_dev(qd, r)
        return off.cpu().numpy(), i.cpu().numpy(), d.cpu().numpy()

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            call("ompl_amd_nn_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ += ["OmplAmdError"]
