"""ctypes binding of librlc.so (the C ABI declared in include/rlc.h).

Python is only the test and benchmark driver here: every call goes straight to the native
library, which fails loudly when it cannot find a HIP device. Nothing in this module computes
labels itself.

The option and result structures mirror the reference's RakOptions / RakResult
(inc/rak.hxx:19-97); `run` mirrors rakLowmemStaticCuda<SLOTS>(x, o)
(inc/rakLowmemCuda.hxx:624-632) with an optional initial membership.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librlc.so")

MAX_ITER_STATS = 64
EMPTY_LABEL = 0xFFFFFFFF
MODE_ASYNC, MODE_SYNC = 0, 1
SKETCH_BM, SKETCH_MG8 = 1, 8
TIMING_GRAPH, TIMING_PER_ITERATION = 0, 1

STATUS = {
    0: "RLC_OK", 1: "RLC_ERR_INVALID_ARGUMENT", 2: "RLC_ERR_INVALID_GRAPH", 3: "RLC_ERR_OUT_OF_MEMORY",
    4: "RLC_ERR_DEVICE", 5: "RLC_ERR_NO_DEVICE", 6: "RLC_ERR_CAPACITY", 7: "RLC_ERR_CANCELLED",
    8: "RLC_ERR_NOT_READY", 9: "RLC_ERR_BUSY", 10: "RLC_ERR_INTERNAL",
}


class Options(C.Structure):
    _fields_ = [
        ("repeat", C.c_int32), ("tolerance", C.c_double), ("max_iterations", C.c_int32),
        ("slots", C.c_int32), ("mode", C.c_int32), ("pickstep", C.c_int32),
        ("merge_degree", C.c_int32), ("merge_parts", C.c_int32), ("group_degree", C.c_int32 * 4),
        ("timing", C.c_int32), ("sweeps", C.c_int32), ("no_rescan", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("time_ms", C.c_float), ("marking_ms", C.c_float),
        ("init_ms", C.c_float), ("memory_gb", C.c_float), ("changed_last", C.c_uint32),
        ("traversed_edges", C.c_uint64), ("algorithmic_bytes", C.c_uint64),
        ("active", C.c_uint64 * MAX_ITER_STATS), ("edges", C.c_uint64 * MAX_ITER_STATS),
        ("changed", C.c_uint64 * MAX_ITER_STATS), ("changed_edges", C.c_uint64 * MAX_ITER_STATS),
        ("iter_ms", C.c_float * MAX_ITER_STATS),
    ]


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64), ("max_degree", C.c_uint64),
                ("device_bytes", C.c_uint64), ("unit_weights", C.c_int32), ("offsets_bytes", C.c_int32),
                ("segments", C.c_int32), ("reserved", C.c_int32)]


class Csr(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64), ("offsets", C.POINTER(C.c_uint64)),
                ("edges", C.POINTER(C.c_uint32)), ("weights", C.POINTER(C.c_float))]


# (name, restype, argtypes) of every exported entry point, in include/rlc.h order
_P = C.c_void_p
SIGNATURES = [
    ("rlc_last_error", C.c_char_p, []),
    ("rlc_status_string", C.c_char_p, [C.c_int]),
    ("rlc_abi_version", C.c_int32, []),
    ("rlc_options_default", None, [C.POINTER(Options)]),
    ("rlc_context_create", C.c_int, [C.c_int32, C.POINTER(_P)]),
    ("rlc_context_destroy", None, [_P]),
    ("rlc_device_memory", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("rlc_graph_upload", C.c_int, [_P, C.c_uint64, C.c_uint64, _P, _P, _P, C.POINTER(_P)]),
    ("rlc_graph_upload32", C.c_int, [_P, C.c_uint64, C.c_uint64, _P, _P, _P, C.POINTER(_P)]),
    ("rlc_graph_info_get", C.c_int, [_P, C.POINTER(GraphInfo)]),
    ("rlc_graph_destroy", None, [_P]),
    ("rlc_run", C.c_int, [_P, _P, _P, C.POINTER(Options), _P, C.POINTER(RunStats)]),
    ("rlc_graph_labels_device", _P, [_P]),
    ("rlc_graph_vertex_order", C.c_int, [_P, _P]),
    ("rlc_modularity", C.c_int, [_P, _P, _P, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("rlc_gen_sbm", C.c_int, [C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_uint64, C.POINTER(Csr)]),
    ("rlc_gen_rmat", C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_float,
                               C.c_float, C.c_uint64, C.POINTER(Csr)]),
    ("rlc_gen_grid", C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_uint64, C.POINTER(Csr)]),
    ("rlc_csr_release", None, [C.POINTER(Csr)]),
]

_lib = None


class RlcError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def lib() -> C.CDLL:
    """Load librlc.so (built in-tree by `make`); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RlcError(10, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(st: int) -> None:
    if st != 0:
        raise RlcError(st, (lib().rlc_last_error() or b"").decode())


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- graphs -----------------------------------------------------------------------------------
@dataclass
class Graph:
    """Host CSR: n vertices, offsets[n+1] (uint64), edges[m] (uint32), weights[m] (float32 or None)."""
    offsets: np.ndarray
    edges: np.ndarray
    weights: np.ndarray | None

    @property
    def n(self) -> int:
        return len(self.offsets) - 1

    @property
    def m(self) -> int:
        return len(self.edges)

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets)


def _take_csr(c: Csr, unit: bool) -> Graph:
    n, m = int(c.n), int(c.m)
    off = np.ctypeslib.as_array(c.offsets, shape=(n + 1,)).copy()
    edg = np.ctypeslib.as_array(c.edges, shape=(m,)).copy() if m else np.zeros(0, np.uint32)
    wei = None
    if not unit:
        wei = np.ctypeslib.as_array(c.weights, shape=(m,)).copy() if m else np.zeros(0, np.float32)
    lib().rlc_csr_release(C.byref(c))
    return Graph(off, edg, wei)


def gen_sbm(n: int, blocks: int, degree: float, intra: float, seed: int) -> Graph:
    c = Csr()
    _check(lib().rlc_gen_sbm(n, blocks, degree, intra, seed, C.byref(c)))
    return _take_csr(c, unit=True)


def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, wmin=1.0, wmax=1.0, seed=1) -> Graph:
    out = Csr()
    _check(lib().rlc_gen_rmat(scale, edge_factor, a, b, c, wmin, wmax, seed, C.byref(out)))
    return _take_csr(out, unit=(wmin == wmax == 1.0))


def gen_grid(rows: int, cols: int, drop: float, shortcut: float, seed: int) -> Graph:
    c = Csr()
    _check(lib().rlc_gen_grid(rows, cols, drop, shortcut, seed, C.byref(c)))
    return _take_csr(c, unit=True)


def options(**kw) -> Options:
    o = Options()
    lib().rlc_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        if k == "group_degree":
            v = (C.c_int32 * 4)(*v)
        setattr(o, k, v)
    return o


@dataclass
class RakResult:
    """RakResult<K> (inc/rak.hxx:53-97) plus the per-iteration counters of this library."""
    membership: np.ndarray
    iterations: int
    time: float
    markingTime: float
    initializationTime: float
    memory: float
    stats: RunStats


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().rlc_context_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().rlc_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def memory(self):
        f, t = C.c_uint64(), C.c_uint64()
        _check(lib().rlc_device_memory(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def upload(self, g: Graph, offsets32: bool = False) -> "DeviceGraph":
        h = C.c_void_p()
        edg = np.ascontiguousarray(g.edges, dtype=np.uint32)
        wei = None if g.weights is None else np.ascontiguousarray(g.weights, dtype=np.float32)
        if offsets32:
            off = np.ascontiguousarray(g.offsets, dtype=np.uint32)
            _check(lib().rlc_graph_upload32(self.h, g.n, g.m, _ptr(off), _ptr(edg), _ptr(wei), C.byref(h)))
        else:
            off = np.ascontiguousarray(g.offsets, dtype=np.uint64)
            _check(lib().rlc_graph_upload(self.h, g.n, g.m, _ptr(off), _ptr(edg), _ptr(wei), C.byref(h)))
        return DeviceGraph(self, h, g.n)


class DeviceGraph:
    def __init__(self, ctx: Context, h, n: int):
        self.ctx, self.h, self.n = ctx, h, n

    def close(self):
        if self.h:
            lib().rlc_graph_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self) -> GraphInfo:
        i = GraphInfo()
        _check(lib().rlc_graph_info_get(self.h, C.byref(i)))
        return i

    def run(self, opts: Options | None = None, init: np.ndarray | None = None, labels: bool = True) -> RakResult:
        o = opts if opts is not None else options()
        st = RunStats()
        out = np.empty(self.n, np.uint32) if labels else None
        ini = None if init is None else np.ascontiguousarray(init, dtype=np.uint32)
        _check(lib().rlc_run(self.ctx.h, self.h, _ptr(ini), C.byref(o), _ptr(out), C.byref(st)))
        return RakResult(out, st.iterations, st.time_ms, st.marking_ms, st.init_ms, st.memory_gb, st)

    def vertex_order(self) -> np.ndarray:
        o = np.empty(self.n, np.uint32)
        _check(lib().rlc_graph_vertex_order(self.h, _ptr(o)))
        return o

    def modularity(self, labels: np.ndarray | None = None):
        q, k = C.c_double(), C.c_uint64()
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint32)
        _check(lib().rlc_modularity(self.ctx.h, self.h, _ptr(lab), C.byref(q), C.byref(k)))
        return q.value, k.value


def iteration_stats(st: RunStats):
    L = min(st.iterations, MAX_ITER_STATS)
    return np.array([[st.active[i], st.edges[i], st.changed[i], st.changed_edges[i]] for i in range(L)],
                    dtype=np.uint64).reshape(L, 4)
