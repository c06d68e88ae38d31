"""MI355X-native triangle counting (static + streaming), host side in Python.

A thin ctypes layer over ``libcustinger_tc.so`` (the C-ABI declared in
``include/custinger_tc.h``).  It mirrors the reference's algorithm surface:

* :class:`cuStinger` -- the dynamic graph store built from a CSR
  (``cuStinger::initializeCuStinger``, tests/cct.cpp:261-277);
* :class:`StaticTriangleCounting` -- Init / Reset / Run / Release
  (``StaticAlgorithm``, include/algs.cuh:20-29) over ``callDeviceAllTriangles``
  (include/static_triangle_counting/cct.hpp:15-17);
* :class:`StreamingTriangleCounting` -- Init / runStatic / batchUpdateInserted /
  batchUpdateDeleted / Release (the streaming class shape of
  include/streaming_katz_centrality/katz.cuh:21-58) over
  ``callDeviceNewTriangles`` (src/streaming_triangle_counting/cct_upd.cu:412-440).

There is no CPU fallback: importing works without a GPU, but every compute
call goes to the HIP library and raises if it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CUSTINGER_TC_LIB") or os.path.join(_HERE, "libcustinger_tc.so")

TC_OK, TC_ERR_INVALID, TC_ERR_OOM, TC_ERR_HIP, TC_ERR_UNSUPPORTED, TC_ERR_CAPACITY = range(6)
FMT_AUTO, FMT_DIMACS, FMT_SNAP, FMT_MTX = range(4)
READ_AS_IS = 0x1
TC_HOST, TC_DEVICE = 0, 1
TC_GRAPH_ASSUME_SYMMETRIC, TC_GRAPH_NO_VALIDATE = 0x1, 0x2
POLICY_FIXED, POLICY_LEARNED, POLICY_MERGE, POLICY_BSEARCH, POLICY_HASH, POLICY_MERGEPATH = range(6)
METHOD_AUTO, METHOD_ORIENTED, METHOD_DIRECT = range(3)
MLP_IN, MLP_HIDDEN, MLP_OUT = 4, 32, 3
MLP_PARAMS = MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN \
    + MLP_HIDDEN * MLP_OUT + MLP_OUT

# every symbol include/custinger_tc.h declares (checked by the CPU tests)
EXPORTS = [
    "tc_last_error", "tc_status_string", "tc_version", "tc_device_alloc", "tc_device_free",
    "tc_memcpy", "tc_memset", "tc_device_synchronize", "tc_device_name",
    "tc_graph_create_csr", "tc_graph_destroy", "tc_graph_set_stream", "tc_graph_get_info",
    "tc_graph_to_csr", "tc_graph_degrees", "tc_schedule_default", "tc_static_count",
    "tc_csr_count_device", "tc_batch_apply", "tc_batch_update_only", "tc_batch_prepare",
    "tc_stream_delta_passes",
    "tc_batched_count", "tc_lb_partition", "tc_lb_degree_bins", "tc_gen_rmat_pairs",
    "tc_gen_er_pairs", "tc_gen_rmat_multi", "tc_canon_csr_device", "tc_collect_timings", "tc_mlp_train_step",
    "tc_mlp_forward", "tc_model_load", "tc_model_table", "tc_autotune", "tc_schedule_set_default",
    "tc_batched_graph_create", "tc_batched_graph_count", "tc_read_graph", "tc_host_free",
    "tc_lb_edge_prefix", "tc_graph_get_view", "tc_batch_update_directed", "tc_graph_set_growth",
    "tc_graph_compare", "tc_csr_count_device32", "tc_csr_count_device_noalloc",
    "tc_strategy_table_set",
]


class TCError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{status}] {msg}")
        self.status = status


class Schedule(C.Structure):
    _fields_ = [("method", C.c_int32), ("policy", C.c_int32), ("small_max", C.c_int32),
                ("mid_max", C.c_int32), ("job_items", C.c_int32), ("ratio_x16", C.c_int32),
                ("persistent", C.c_int32), ("flush_items", C.c_int32),
                ("hub_degree", C.c_int32), ("mpath_min", C.c_int32), ("capture", C.c_int32),
                ("build", C.c_int32), ("reserved", C.c_int32 * 4)]


class Stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_prepare", C.c_double), ("ms_count", C.c_double),
                ("ms_bin", C.c_double * 4), ("bytes_alg_bin", C.c_double * 4),
                ("tasks_bin", C.c_int64 * 4), ("jobs_bin", C.c_int64 * 4),
                ("strategy_tasks", C.c_int64 * 3), ("triangles", C.c_uint64),
                ("method_used", C.c_int32), ("graph_replay", C.c_int32),
                ("items_bin", C.c_double * 3), ("stream_bytes_bin", C.c_double * 3),
                ("ms_long", C.c_double * 2), ("tasks_long", C.c_int64 * 2),
                ("items_long", C.c_double * 2), ("stream_bytes_long", C.c_double * 2),
                ("mergepath_tasks", C.c_int64)]

    def as_dict(self) -> dict:
        return {
            "ms_total": self.ms_total, "ms_prepare": self.ms_prepare, "ms_count": self.ms_count,
            "ms_bin": list(self.ms_bin)[:3], "bytes_alg_bin": list(self.bytes_alg_bin)[:3],
            "tasks_bin": list(self.tasks_bin)[:3], "jobs_bin": list(self.jobs_bin)[:3],
            "items_bin": list(self.items_bin),
            "stream_bytes_bin": list(self.stream_bytes_bin),
            "ms_long": list(self.ms_long), "tasks_long": list(self.tasks_long),
            "items_long": list(self.items_long), "stream_bytes_long": list(self.stream_bytes_long),
            "strategy_tasks": dict(zip(("merge", "bsearch", "hash"), self.strategy_tasks),
                                   mergepath=int(self.mergepath_tasks)),
            "triangles": int(self.triangles), "method_used": int(self.method_used),
            "graph_replay": int(self.graph_replay),
        }


class GraphInfo(C.Structure):
    _fields_ = [("nv", C.c_int32), ("nnz", C.c_int64), ("allocated", C.c_int64),
                ("symmetric", C.c_int32), ("self_loops", C.c_int64), ("max_degree", C.c_int32),
                ("reallocations", C.c_int64)]


class BatchStats(C.Structure):
    _fields_ = [("inserted", C.c_int64), ("deleted", C.c_int64), ("dup_inserts", C.c_int64),
                ("missing_deletes", C.c_int64), ("reallocations", C.c_int64),
                ("ms_update", C.c_double), ("ms_delta", C.c_double),
                ("replayed", C.c_int32), ("reserved", C.c_int32 * 7)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def lib():
    """Load the HIP library; raise loudly when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"HIP extension missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u32, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
        L.tc_last_error.restype = C.c_char_p
        L.tc_status_string.restype = C.c_char_p
        L.tc_status_string.argtypes = [C.c_int]
        L.tc_device_alloc.argtypes = [C.POINTER(vp), u64]
        L.tc_device_free.argtypes = [vp]
        L.tc_memcpy.argtypes = [vp, vp, u64, C.c_int]
        L.tc_memset.argtypes = [vp, C.c_int, u64]
        L.tc_device_name.argtypes = [C.c_char_p, C.c_int]
        L.tc_graph_create_csr.argtypes = [i32, i64, vp, vp, C.c_int, u32, i32, C.POINTER(vp)]
        L.tc_graph_destroy.argtypes = [vp]
        L.tc_graph_set_stream.argtypes = [vp, vp]
        L.tc_graph_get_info.argtypes = [vp, C.POINTER(GraphInfo)]
        L.tc_graph_to_csr.argtypes = [vp, vp, vp, C.c_int]
        L.tc_graph_degrees.argtypes = [vp, vp]
        L.tc_schedule_default.argtypes = [C.POINTER(Schedule)]
        L.tc_static_count.argtypes = [vp, vp, C.POINTER(u64), C.POINTER(Schedule), C.POINTER(Stats)]
        L.tc_csr_count_device.argtypes = [i32, i64, vp, vp, vp, C.POINTER(u64), vp]
        L.tc_csr_count_device32.argtypes = [i32, vp, vp, vp, C.POINTER(u64), vp]
        L.tc_csr_count_device_noalloc.argtypes = [i32, i64, vp, vp, vp, C.POINTER(u64), vp]
        L.tc_batch_apply.argtypes = [vp, vp, vp, vp, i64, vp, C.POINTER(i64), C.POINTER(Schedule),
                                     C.POINTER(BatchStats)]
        L.tc_batch_update_only.argtypes = [vp, vp, vp, vp, i64, C.POINTER(BatchStats)]
        L.tc_batch_update_directed.argtypes = [vp, vp, vp, vp, i64, C.POINTER(BatchStats)]
        L.tc_graph_set_growth.argtypes = [vp, i32, i32]
        L.tc_graph_compare.argtypes = [vp, vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]
        L.tc_stream_delta_passes.argtypes = [vp, i64, vp, vp, vp, vp, vp, i32]
        L.tc_batch_prepare.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp]
        L.tc_batched_count.argtypes = [i32, vp, vp, vp, i64, i32, vp, vp, C.POINTER(Schedule),
                                       C.POINTER(Stats), vp]
        L.tc_lb_partition.argtypes = [vp, vp, i64, vp, C.POINTER(i64)]
        L.tc_lb_degree_bins.argtypes = [vp, vp, i64, C.POINTER(i64)]
        L.tc_gen_rmat_pairs.argtypes = [i32, i64, C.c_double, C.c_double, C.c_double, u64, vp, vp, vp]
        L.tc_gen_er_pairs.argtypes = [i32, i64, u64, vp, vp, vp]
        L.tc_gen_rmat_multi.argtypes = [i32, vp, vp, vp, C.c_double, C.c_double, C.c_double, vp,
                                        vp, vp]
        L.tc_canon_csr_device.argtypes = [i32, i64, vp, vp, C.POINTER(vp), C.POINTER(vp),
                                          C.POINTER(i64), vp]
        L.tc_collect_timings.argtypes = [vp, i64, u64, vp, vp, C.POINTER(i64)]
        L.tc_mlp_train_step.argtypes = [vp, vp, vp, i32, vp, vp, i64, C.c_float, vp,
                                        C.POINTER(C.c_float)]
        L.tc_mlp_forward.argtypes = [vp, vp, i64, vp]
        L.tc_model_load.argtypes = [vp]
        L.tc_model_table.argtypes = [vp, C.POINTER(i32)]
        L.tc_strategy_table_set.argtypes = [vp, i32]
        L.tc_autotune.argtypes = [vp, C.POINTER(Schedule), C.POINTER(C.c_double)]
        L.tc_schedule_set_default.argtypes = [C.POINTER(Schedule)]
        L.tc_batched_graph_create.argtypes = [i32, vp, vp, vp, i64, i32, vp, C.POINTER(vp)]
        L.tc_batched_graph_count.argtypes = [vp, vp, vp, C.POINTER(Schedule), C.POINTER(Stats)]
        L.tc_read_graph.argtypes = [C.c_char_p, C.c_int, u32, C.POINTER(i32), C.POINTER(i64),
                                    C.POINTER(C.POINTER(i64)), C.POINTER(C.POINTER(i32))]
        L.tc_host_free.argtypes = [vp]
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != TC_OK:
        raise TCError(status, lib().tc_last_error().decode(errors="replace"))


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().tc_device_name(buf, 256))
    return buf.value.decode()


def synchronize() -> None:
    check(lib().tc_device_synchronize())


# ------------------------------------------------------------ device arrays
class DeviceArray:
    """A HIP device allocation owned by Python (allocDeviceArray / copyArray*)."""

    def __init__(self, n: int, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(self.n * self.dtype.itemsize, 1)
        p = C.c_void_p()
        check(lib().tc_device_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value
        self._owned = True

    @classmethod
    def wrap(cls, ptr: int, n: int, dtype) -> "DeviceArray":
        obj = cls.__new__(cls)
        obj.dtype = np.dtype(dtype)
        obj.n = int(n)
        obj.nbytes = max(obj.n * obj.dtype.itemsize, 1)
        obj.ptr = ptr
        obj._owned = True
        return obj

    @classmethod
    def from_host(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().tc_memcpy(d.ptr, a.ctypes.data, a.nbytes, 0))
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().tc_memcpy(out.ctypes.data, self.ptr, self.n * self.dtype.itemsize, 1))
        return out

    def zero(self) -> "DeviceArray":
        check(lib().tc_memset(self.ptr, 0, self.nbytes))
        return self

    def free(self) -> None:
        if getattr(self, "_owned", False) and self.ptr:
            lib().tc_device_free(self.ptr)
        self.ptr = None
        self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def default_schedule(**over) -> Schedule:
    s = Schedule()
    check(lib().tc_schedule_default(C.byref(s)))
    for k, v in over.items():
        setattr(s, k, v)
    return s


# ------------------------------------------------------------------ the store
class cuStinger:
    """Dynamic adjacency store on the GPU, created from a CSR."""

    def __init__(self, off, adj, *, assume_symmetric=False, validate=True, slack_pct=0,
                 location=TC_HOST):
        flags = (TC_GRAPH_ASSUME_SYMMETRIC if assume_symmetric else 0) | \
            (0 if validate else TC_GRAPH_NO_VALIDATE)
        h = C.c_void_p()
        if location == TC_HOST:
            off = np.ascontiguousarray(off, np.int64)
            adj = np.ascontiguousarray(adj, np.int32)
            nv, nnz = len(off) - 1, len(adj)
            check(lib().tc_graph_create_csr(nv, nnz, off.ctypes.data,
                                            adj.ctypes.data if nnz else None, TC_HOST, flags,
                                            slack_pct, C.byref(h)))
        else:  # DeviceArray inputs
            nv, nnz = off.n - 1, adj.n
            check(lib().tc_graph_create_csr(nv, nnz, off.ptr, adj.ptr if nnz else None, TC_DEVICE,
                                            flags, slack_pct, C.byref(h)))
        self.h = h.value
        self.nv = nv

    def info(self) -> dict:
        i = GraphInfo()
        check(lib().tc_graph_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def to_csr(self):
        nnz = self.info()["nnz"]
        off = np.zeros(self.nv + 1, np.int64)
        adj = np.zeros(max(nnz, 1), np.int32)
        check(lib().tc_graph_to_csr(self.h, off.ctypes.data, adj.ctypes.data, TC_HOST))
        return off, adj[:nnz]

    def set_growth(self, mul_x16: int = 0, add: int = 0) -> None:
        """Update allocator policy: a row that outgrows its block moves to a
        block of len * mul_x16 / 16 + add entries (0, 0: next power of two)."""
        check(lib().tc_graph_set_growth(self.h, mul_x16, add))

    def compare(self, other: "cuStinger") -> dict:
        """compareCUS: rows whose lengths differ, rows whose entries differ,
        the first differing vertex (-1: equal)."""
        a, b, f = C.c_int64(0), C.c_int64(0), C.c_int32(0)
        check(lib().tc_graph_compare(self.h, other.h, C.byref(a), C.byref(b), C.byref(f)))
        return {"size_mismatch": a.value, "entry_mismatch": b.value, "first_vertex": f.value}

    def degrees(self) -> np.ndarray:
        d = DeviceArray(self.nv, np.int32)
        check(lib().tc_graph_degrees(self.h, d.ptr))
        return d.to_host()

    def freecuStinger(self) -> None:
        if self.h:
            check(lib().tc_graph_destroy(self.h))
        self.h = None

    close = freecuStinger

    def __del__(self):
        try:
            self.freecuStinger()
        except Exception:
            pass


def static_count(g: cuStinger, schedule: Schedule | None = None, per_vertex=True,
                 out: DeviceArray | None = None):
    """One static count -> (per-vertex uint32 array or None, uint64 global, stats dict)."""
    own = None
    if per_vertex and out is None:
        own = out = DeviceArray(g.nv, np.uint32)
    glob = C.c_uint64(0)
    st = Stats()
    sp = C.byref(schedule) if schedule is not None else None
    check(lib().tc_static_count(g.h, out.ptr if per_vertex else None, C.byref(glob), sp,
                                C.byref(st)))
    pv = out.to_host() if (per_vertex and own is not None) else None
    if own is not None:
        own.free()
    return pv, int(glob.value), st.as_dict()


def csr_count_device(off, adj):
    """callDeviceAllTrianglesCSR convention (3T) on a host CSR staged to the device."""
    off = np.ascontiguousarray(off, np.int64)
    adj = np.ascontiguousarray(adj, np.int32)
    doff, dadj = DeviceArray.from_host(off), DeviceArray.from_host(adj)
    out = DeviceArray(len(off) - 1, np.uint32)
    glob = C.c_uint64(0)
    check(lib().tc_csr_count_device(len(off) - 1, len(adj), doff.ptr, dadj.ptr, out.ptr,
                                    C.byref(glob), None))
    return out.to_host(), int(glob.value)


def batch_apply(g: cuStinger, u, v, is_delete=None, delta: DeviceArray | None = None,
                schedule: Schedule | None = None, count=True):
    """Apply one batch of undirected pairs; returns (delta_global, stats dict)."""
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    B = len(u)
    du, dv = DeviceArray.from_host(u), DeviceArray.from_host(v)
    dd = DeviceArray.from_host(np.ascontiguousarray(is_delete, np.uint8)) if is_delete is not None else None
    st = BatchStats()
    if count:
        dg = C.c_int64(0)
        sp = C.byref(schedule) if schedule is not None else None
        check(lib().tc_batch_apply(g.h, du.ptr, dv.ptr, dd.ptr if dd else None, B,
                                   delta.ptr if delta is not None else None, C.byref(dg), sp,
                                   C.byref(st)))
        return int(dg.value), st.as_dict()
    check(lib().tc_batch_update_only(g.h, du.ptr, dv.ptr, dd.ptr if dd else None, B, C.byref(st)))
    return 0, st.as_dict()


def batch_update_directed(g: cuStinger, src, dst, is_delete=None) -> dict:
    """edgeInsertions / edgeDeletions of directed entries (src[i] -> dst[i]),
    as given; returns the batch stats."""
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    ds, dd = DeviceArray.from_host(src), DeviceArray.from_host(dst)
    dl = DeviceArray.from_host(np.ascontiguousarray(is_delete, np.uint8)) \
        if is_delete is not None else None
    st = BatchStats()
    try:
        check(lib().tc_batch_update_directed(g.h, ds.ptr, dd.ptr, dl.ptr if dl else None, len(src),
                                             C.byref(st)))
    finally:
        for a in (ds, dd, dl):
            if a is not None:
                a.free()
    return st.as_dict()


def batch_prepare(g: cuStinger, src, dst):
    """BatchUpdate::sortDeviceBUD plus duplicate flags against the current
    store (tests/cct_upd.cpp:395, cct_upd.cu:252): returns the sorted directed
    entries, dup flags and per-vertex batch offsets as host arrays."""
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    nb = len(src)
    nv = g.nv
    ds, dd = DeviceArray.from_host(src), DeviceArray.from_host(dst)
    s_src, s_dst, dup = (DeviceArray(max(nb, 1), np.int32) for _ in range(3))
    boff = DeviceArray(nv + 1, np.int64)
    check(lib().tc_batch_prepare(g.h, nb, ds.ptr, dd.ptr, s_src.ptr, s_dst.ptr, dup.ptr,
                                 boff.ptr))
    return (s_src.to_host()[:nb], s_dst.to_host()[:nb], dup.to_host()[:nb], boff.to_host())


def lb_partition(g: cuStinger, verts) -> np.ndarray:
    """cusLoadBalance::LoadBalanceArray (include/load_balance.cuh:77-102) on the device."""
    dv = DeviceArray.from_host(np.ascontiguousarray(verts, np.int32))
    parts = DeviceArray(1000, np.int64)
    P = C.c_int64(0)
    check(lib().tc_lb_partition(g.h, dv.ptr, dv.n, parts.ptr, C.byref(P)))
    return parts.to_host()[:P.value]


def lb_degree_bins(g: cuStinger, verts) -> np.ndarray:
    """Vertices per degree bin (>=64, 16..63, <16) as in include/operators.cuh:164-288."""
    dv = DeviceArray.from_host(np.ascontiguousarray(verts, np.int32))
    out = (C.c_int64 * 3)()
    check(lib().tc_lb_degree_bins(g.h, dv.ptr, dv.n, out))
    return np.array(out[:], np.int64)


def batched_count(voff, off, adj, schedule: Schedule | None = None, per_vertex=False):
    """Per-graph global counts of G graphs concatenated with global vertex ids."""
    voff = np.ascontiguousarray(voff, np.int64)
    G = len(voff) - 1
    dvo, doff, dadj = (DeviceArray.from_host(np.ascontiguousarray(x, t))
                       for x, t in ((voff, np.int64), (off, np.int64), (adj, np.int32)))
    pg = DeviceArray(G, np.uint64)
    nv = len(off) - 1
    pv = DeviceArray(nv, np.uint32) if per_vertex else None
    st = Stats()
    sp = C.byref(schedule) if schedule is not None else None
    check(lib().tc_batched_count(G, dvo.ptr, doff.ptr, dadj.ptr, dadj.n, nv, pg.ptr,
                                 pv.ptr if pv else None, sp, C.byref(st), None))
    return pg.to_host(), (pv.to_host() if pv else None), st.as_dict()


def batched_count_device(voff: DeviceArray, off: DeviceArray, adj: DeviceArray,
                         per_graph: DeviceArray, schedule: Schedule | None = None):
    st = Stats()
    sp = C.byref(schedule) if schedule is not None else None
    check(lib().tc_batched_count(voff.n - 1, voff.ptr, off.ptr, adj.ptr, adj.n, off.n - 1,
                                 per_graph.ptr, None, sp, C.byref(st), None))
    return st.as_dict()


class BatchedGraph:
    """G graphs concatenated with global vertex ids (device arrays), validated
    once and counted repeatedly; the arrays must outlive this object."""

    def __init__(self, voff: DeviceArray, off: DeviceArray, adj: DeviceArray):
        self._keep = (voff, off, adj)
        self.G = voff.n - 1
        self.nv = off.n - 1
        h = C.c_void_p()
        check(lib().tc_batched_graph_create(self.G, voff.ptr, off.ptr, adj.ptr, adj.n, self.nv,
                                            None, C.byref(h)))
        self.h = h.value

    def count(self, per_graph: DeviceArray, per_vertex: DeviceArray | None = None,
              schedule: Schedule | None = None) -> dict:
        st = Stats()
        sp = C.byref(schedule) if schedule is not None else None
        check(lib().tc_batched_graph_count(self.h, per_graph.ptr,
                                           per_vertex.ptr if per_vertex is not None else None, sp,
                                           C.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if self.h:
            lib().tc_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_graph(path: str, fmt: int = FMT_AUTO, as_is: bool = False):
    """readGraphDIMACS / readGraphSNAP / readGraphMatrixMarket: a graph file as
    a host CSR (off int64[nv+1], adj int32[nnz]); symmetrised, loop-free,
    deduplicated and sorted unless as_is."""
    nv, nnz = C.c_int32(0), C.c_int64(0)
    po, pa = C.POINTER(C.c_int64)(), C.POINTER(C.c_int32)()
    check(lib().tc_read_graph(os.fsencode(path), fmt, READ_AS_IS if as_is else 0, C.byref(nv),
                              C.byref(nnz), C.byref(po), C.byref(pa)))
    try:
        off = np.ctypeslib.as_array(po, shape=(nv.value + 1,)).copy()
        adj = (np.ctypeslib.as_array(pa, shape=(nnz.value,)).copy() if nnz.value
               else np.zeros(0, np.int32))
    finally:
        lib().tc_host_free(C.cast(po, C.c_void_p))
        lib().tc_host_free(C.cast(pa, C.c_void_p))
    return off, adj


# ------------------------------------------------------------ reference shape
class StaticTriangleCounting:
    """StaticAlgorithm surface (include/algs.cuh:20-29) for triangle counting."""

    def __init__(self):
        self.custing = None
        self.d_out = None
        self.global_count = 0
        self.host_out = None
        self.schedule = None
        self.stats = None

    def Init(self, custing: cuStinger):
        self.custing = custing
        self.d_out = DeviceArray(custing.nv, np.uint32)
        self.Reset()

    def Reset(self):
        self.d_out.zero()
        self.global_count = 0

    def setSchedule(self, schedule: Schedule):
        self.schedule = schedule

    def Run(self, custing: cuStinger | None = None):
        g = custing or self.custing
        _, self.global_count, self.stats = static_count(g, self.schedule, True, self.d_out)

    def SyncHostWithDevice(self):
        self.host_out = self.d_out.to_host()

    def SyncDeviceWithHost(self):
        if self.host_out is not None:
            check(lib().tc_memcpy(self.d_out.ptr, self.host_out.ctypes.data, self.host_out.nbytes, 0))

    def getGlobal(self) -> int:
        return self.global_count

    def copyPerVertexToHost(self) -> np.ndarray:
        return self.d_out.to_host()

    def Release(self):
        if self.d_out is not None:
            self.d_out.free()
        self.d_out = None


class StreamingTriangleCounting:
    """Streaming shape (include/streaming_katz_centrality/katz.cuh:21-58) for TC."""

    def __init__(self):
        self.custing = None
        self.d_out = None
        self.global_count = 0
        self.schedule = None
        self.last = None

    def Init(self, custing: cuStinger):
        self.custing = custing
        self.d_out = DeviceArray(custing.nv, np.uint32).zero()

    def runStatic(self, custing: cuStinger | None = None):
        g = custing or self.custing
        _, self.global_count, _ = static_count(g, self.schedule, True, self.d_out)

    def batchUpdateInserted(self, custing, u, v):
        dg, self.last = batch_apply(custing or self.custing, u, v, None, self.d_out, self.schedule)
        self.global_count += dg
        return dg

    def batchUpdateDeleted(self, custing, u, v):
        dg, self.last = batch_apply(custing or self.custing, u, v, np.ones(len(u), np.uint8),
                                    self.d_out, self.schedule)
        self.global_count += dg
        return dg

    def batchUpdate(self, custing, u, v, is_delete):
        dg, self.last = batch_apply(custing or self.custing, u, v, is_delete, self.d_out,
                                    self.schedule)
        self.global_count += dg
        return dg

    def getGlobal(self) -> int:
        return self.global_count

    def copyPerVertexToHost(self) -> np.ndarray:
        return self.d_out.to_host()

    def Release(self):
        if self.d_out is not None:
            self.d_out.free()
        self.d_out = None


# ------------------------------------------------------------------ generators
def rmat_graph_device(scale: int, edge_factor: int, seed: int, abc=(0.57, 0.19, 0.19),
                      slack_pct=0) -> cuStinger:
    """R-MAT graph generated and canonicalised on the device (same RNG as the oracle)."""
    n = 1 << scale
    m = edge_factor * n
    du, dv = DeviceArray(m, np.int32), DeviceArray(m, np.int32)
    check(lib().tc_gen_rmat_pairs(scale, m, abc[0], abc[1], abc[2], seed, du.ptr, dv.ptr, None))
    return _canon(n, m, du, dv, slack_pct)


def er_graph_device(n: int, avg_deg: int, seed: int, slack_pct=0) -> cuStinger:
    m = n * avg_deg // 2
    du, dv = DeviceArray(m, np.int32), DeviceArray(m, np.int32)
    check(lib().tc_gen_er_pairs(n, m, seed, du.ptr, dv.ptr, None))
    return _canon(n, m, du, dv, slack_pct)


_M64 = (1 << 64) - 1


def _mix64(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def rng(seed: int, ctr: int) -> int:
    """The counter-based generator the device kernels use (keyed splitmix64 finaliser)."""
    key = _mix64((seed + 0x9E3779B97F4A7C15) & _M64)
    return _mix64(key ^ ((ctr * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & _M64))


def batched_spec(G: int = 4096, seed_root: int = 1000, nmin: int = 1024, nmax: int = 65536,
                 efmin: int = 4, efmax: int = 32):
    """Sizes of a batched workload of G R-MAT graphs: graph i has
    n_i = floor(exp(U(ln nmin, ln nmax))) vertices and edge factor
    ef_i ~ U{efmin..efmax}, drawn from rng(seed_root + i, 0 / 1), and seed seed_root + i."""
    import math
    n = np.empty(G, np.int32)
    ef = np.empty(G, np.int32)
    seeds = np.empty(G, np.uint64)
    lo, hi = math.log(nmin), math.log(nmax)
    for i in range(G):
        u01 = (rng(seed_root + i, 0) >> 11) / float(1 << 53)
        n[i] = min(nmax, max(nmin, int(math.floor(math.exp(lo + u01 * (hi - lo))))))
        ef[i] = efmin + rng(seed_root + i, 1) % (efmax - efmin + 1)
        seeds[i] = seed_root + i
    return n, ef, seeds


def batched_graph_device(n, ef, seeds, abc=(0.57, 0.19, 0.19)):
    """Concatenated CSR of G R-MAT graphs built on the device -> (voff, off, adj) DeviceArrays."""
    n = np.ascontiguousarray(n, np.int32)
    ef = np.ascontiguousarray(ef, np.int32)
    seeds = np.ascontiguousarray(seeds, np.uint64)
    m = int((n.astype(np.int64) * ef).sum())
    du, dv = DeviceArray(m, np.int32), DeviceArray(m, np.int32)
    check(lib().tc_gen_rmat_multi(len(n), n.ctypes.data, ef.ctypes.data, seeds.ctypes.data,
                                  abc[0], abc[1], abc[2], du.ptr, dv.ptr, None))
    voff = np.concatenate([[0], np.cumsum(n.astype(np.int64))])
    doff, dadj = canon_csr_device(int(voff[-1]), du, dv)
    du.free()
    dv.free()
    return DeviceArray.from_host(voff), doff, dadj


def canon_csr_device(n: int, du: DeviceArray, dv: DeviceArray):
    po, pa, nnz = C.c_void_p(), C.c_void_p(), C.c_int64()
    check(lib().tc_canon_csr_device(n, du.n, du.ptr, dv.ptr, C.byref(po), C.byref(pa),
                                    C.byref(nnz), None))
    return DeviceArray.wrap(po.value, n + 1, np.int64), DeviceArray.wrap(pa.value, nnz.value,
                                                                            np.int32)


def _canon(n, m, du, dv, slack_pct):
    doff, dadj = canon_csr_device(n, du, dv)
    du.free()
    dv.free()
    g = cuStinger(doff, dadj, assume_symmetric=True, validate=True, slack_pct=slack_pct,
                  location=TC_DEVICE)
    doff.free()
    dadj.free()
    return g
