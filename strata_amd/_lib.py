"""ctypes binding of include/strata_hip.h (libstrata_hip.so).

The library is built in-tree by strata_amd/build.py. Nothing here falls
back to another implementation: if the library is missing or fails to
load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_int32, c_int64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstrata_hip.so")

# Status codes
OK = 0
CANCELLED = 1
EINVAL = -1
EDEVICE = -2
EIO = -3
ENOMEM = -4

MEM_DEVICE = 0
MEM_HOST = 1

# Ops (enum strata_op)
OP_GRADIENT = 1
OP_SLOPE = 2
OP_ASPECT = 3
OP_HILLSHADE = 4
OP_CURVATURE = 5
OP_SURFACE = 6
OP_CORRELATE = 10
OP_SEPARABLE = 11
OP_MEAN = 12
OP_MIN = 13
OP_MAX = 14
OP_LOOKUP = 20
OP_RECLASS = 21

FLAG_ATAN = 0x1
FLAG_TRIG = 0x2
FLAG_FIT = 0x100
CURV_SHIFT = 4
FLAG_MFMA = 0x200
SURF_DX = 0x1000
SURF_DY = 0x2000
SURF_SLOPE = 0x4000
SURF_ASPECT = 0x8000
SURF_HILLSHADE = 0x10000

MAX_TAPS = 320


class Raster(ctypes.Structure):
    _fields_ = [
        ("data", c_void_p),
        ("width", c_int64),
        ("height", c_int64),
        ("stride", c_int64),
        ("valid", c_void_p),
        ("valid_offset", c_int64),
        ("mem", c_int32),
        ("reserved", c_int32),
    ]


class Kernel(ctypes.Structure):
    _fields_ = [
        ("op", c_int32),
        ("radius", c_int32),
        ("flags", c_int32),
        ("ntaps", c_int32),
        ("k", c_float * 16),
        ("taps", c_void_p),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("tile_width", c_int64),
        ("tile_height", c_int64),
        ("workers", c_int64),
        ("stream", c_void_p),
        ("cancel", c_void_p),
        ("async_", c_int32),
        ("reserved", c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("cells", c_int64),
        ("ideal", c_int64),
        ("tiles", c_int64),
        ("bands", c_int64),
        ("source_read", c_int64),
        ("sink_written", c_int64),
        ("kernel_read", c_int64),
        ("kernel_written", c_int64),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("sweep_rows", c_int32), ("fold_stripe_mb", c_int32), ("unit_kcells", c_int32),
                ("stop_after_units", c_int32), ("reserved", c_int32 * 4)]


READ_FN = ctypes.CFUNCTYPE(c_int, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_int64, c_int64,
                           c_int64)
WRITE_FN = ctypes.CFUNCTYPE(c_int, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_int64, c_int64,
                            c_int64)


class Raw(ctypes.Structure):
    """strata_raw: a raw float32 file the library reads or writes itself."""
    _fields_ = [
        ("fd", c_int32),
        ("has_fill", c_int32),
        ("fill", ctypes.c_float),
        ("reserved", c_int32),
        ("offset", c_int64),
    ]


class Source(ctypes.Structure):
    _fields_ = [
        ("width", c_int64),
        ("height", c_int64),
        ("masked", c_int32),
        ("reserved", c_int32),
        ("read", READ_FN),
        ("user", c_void_p),
        ("resident", POINTER(Raster)),
        ("raw", POINTER(Raw)),
    ]


class Sink(ctypes.Structure):
    """strata_sink; its fields are set below, after Fold."""


class Task(ctypes.Structure):
    _fields_ = [
        ("k", Kernel),
        ("post", Kernel * 2),
        ("npost", c_int32),
        ("reserved", c_int32),
        ("src", Raster),
        ("dst", Raster),
    ]


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("rows", c_int32),
        ("blocks_per_cu", c_int32),
        ("graph", c_int32),
        ("reserved", c_int32),
        ("stream", c_void_p),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("tasks", c_int64),
        ("cells", c_int64),
        ("data_units", c_int64),
        ("validity_units", c_int64),
        ("rows", c_int32),
        ("blocks", c_int32),
        ("graph", c_int32),
        ("reserved", c_int32),
    ]


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("src", Raster),
        ("k", Kernel),
        ("post", Kernel * 2),
        ("npost", c_int32),
        ("slots", c_int32),
        ("hbm_budget", c_int64),
        ("max_batch", c_int32),
        ("rows", c_int32),
        ("auto_release", c_int32),
        ("reserved", c_int32),
    ]


class Window(ctypes.Structure):
    _fields_ = [("x", c_int64), ("y", c_int64), ("w", c_int64), ("h", c_int64)]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("requests", c_int64),
        ("batches", c_int64),
        ("cells", c_int64),
        ("max_in_flight", c_int64),
        ("admission_waits", c_int64),
        ("peak_bytes", c_int64),
        ("busy_ms", ctypes.c_double),
        ("queue_ms", ctypes.c_double),
        ("service_ms", ctypes.c_double),
    ]


REDUCE_COUNT = 0x1
REDUCE_MINMAX = 0x2
REDUCE_SUM = 0x4
REDUCE_MOMENTS = 0x8


class Fold(ctypes.Structure):
    """strata_fold: a reduction's combined partial, the exact sums as
    two's complement 64-bit limbs, least significant first."""
    _fields_ = [
        ("count", c_int64),
        ("nan", c_int64),
        ("pos_inf", c_int64),
        ("neg_inf", c_int64),
        ("not_neg_zero", c_int32),
        ("min", ctypes.c_float),
        ("max", ctypes.c_float),
        ("reserved", c_int32),
        ("sum", ctypes.c_uint64 * 6),
        ("sumsq", ctypes.c_uint64 * 10),
    ]


Sink._fields_ = [
    ("width", c_int64),
    ("height", c_int64),
    ("masked", c_int32),
    ("fold_what", c_int32),
    ("write", WRITE_FN),
    ("user", c_void_p),
    ("resident", POINTER(Raster)),
    ("raw", POINTER(Raw)),
    ("fold", POINTER(Fold)),
]


def limbs_to_int(limbs) -> int:
    """The signed integer of two's complement 64-bit limbs."""
    v = 0
    for i, w in enumerate(limbs):
        v |= int(w) << (64 * i)
    n = 64 * len(limbs)
    return v - (1 << n) if v >> (n - 1) else v


# The symbols include/strata_hip.h declares, for the export check.
EXPORTS = (
    "strata_init",
    "strata_last_error",
    "strata_version",
    "strata_trim",
    "strata_process",
    "strata_process_chunked",
    "strata_plan_stats",
    "strata_process_fanout",
    "strata_process_fanout_chunked",
    "strata_plan_stats_fanout",
    "strata_process_chain",
    "strata_process_chain_chunked",
    "strata_plan_stats_chain",
    "strata_reduce",
    "strata_reduce_chunked",
    "strata_reduce_noalloc",
    "strata_get_tuning",
    "strata_set_tuning",
    "strata_selfcheck",
    "strata_timing_begin",
    "strata_timing_end",
    "strata_correlate_error",
    "strata_batch_create",
    "strata_batch_run",
    "strata_batch_run_per_tile",
    "strata_batch_tune",
    "strata_batch_sync",
    "strata_batch_info_get",
    "strata_batch_destroy",
    "strata_serve_create",
    "strata_serve_submit",
    "strata_serve_wait",
    "strata_serve_fetch",
    "strata_serve_release",
    "strata_serve_replay",
    "strata_serve_stats_get",
    "strata_serve_destroy",
)

_lib = None


def lib() -> ctypes.CDLL:
    """Load the library once. Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m strata_amd.build` "
            "(the HIP library is the only implementation)")
    L = ctypes.CDLL(LIB_PATH)
    L.strata_init.argtypes = [c_int]
    L.strata_init.restype = c_int
    L.strata_last_error.argtypes = []
    L.strata_last_error.restype = c_char_p
    L.strata_trim.argtypes = []
    L.strata_trim.restype = c_int
    L.strata_version.argtypes = []
    L.strata_version.restype = c_char_p
    L.strata_process.argtypes = [POINTER(Kernel), POINTER(Raster), c_int, POINTER(Raster), c_int,
                                 POINTER(Options), POINTER(Stats)]
    L.strata_process.restype = c_int
    L.strata_process_chunked.argtypes = [POINTER(Kernel), POINTER(Sink), c_int, POINTER(Source), c_int,
                                         POINTER(Options), POINTER(Stats)]
    L.strata_process_chunked.restype = c_int
    L.strata_plan_stats.argtypes = [POINTER(Kernel), c_int, c_int, c_int64, c_int64, POINTER(Options), c_int,
                                    POINTER(Stats)]
    L.strata_plan_stats.restype = c_int
    L.strata_process_fanout.argtypes = [POINTER(Kernel), c_int, POINTER(Raster), POINTER(Raster),
                                        POINTER(Options), POINTER(Stats)]
    L.strata_process_fanout.restype = c_int
    L.strata_process_fanout_chunked.argtypes = [POINTER(Kernel), c_int, POINTER(Sink), POINTER(Source),
                                                POINTER(Options), POINTER(Stats)]
    L.strata_process_fanout_chunked.restype = c_int
    L.strata_plan_stats_fanout.argtypes = [POINTER(Kernel), c_int, c_int64, c_int64, POINTER(Options), c_int,
                                           POINTER(Stats)]
    L.strata_plan_stats_fanout.restype = c_int
    L.strata_process_chain.argtypes = [POINTER(Kernel), c_int, POINTER(Raster), POINTER(Raster),
                                       POINTER(Options), POINTER(Stats)]
    L.strata_process_chain.restype = c_int
    L.strata_process_chain_chunked.argtypes = [POINTER(Kernel), c_int, POINTER(Sink), POINTER(Source),
                                               POINTER(Options), POINTER(Stats)]
    L.strata_process_chain_chunked.restype = c_int
    L.strata_plan_stats_chain.argtypes = [POINTER(Kernel), c_int, c_int64, c_int64, POINTER(Options), c_int,
                                          POINTER(Stats)]
    L.strata_plan_stats_chain.restype = c_int
    L.strata_reduce.argtypes = [c_int, POINTER(Raster), POINTER(Options), POINTER(Fold), POINTER(Stats)]
    L.strata_reduce.restype = c_int
    L.strata_reduce_chunked.argtypes = [c_int, POINTER(Source), POINTER(Options), POINTER(Fold), POINTER(Stats)]
    L.strata_reduce_chunked.restype = c_int
    L.strata_reduce_noalloc.argtypes = [c_int, POINTER(Raster), POINTER(Options), POINTER(Fold), POINTER(Stats)]
    L.strata_reduce_noalloc.restype = c_int
    L.strata_get_tuning.argtypes = [POINTER(Tuning)]
    L.strata_get_tuning.restype = c_int
    L.strata_set_tuning.argtypes = [POINTER(Tuning)]
    L.strata_set_tuning.restype = c_int
    L.strata_selfcheck.argtypes = [c_int, ctypes.c_uint64, ctypes.c_uint64, POINTER(ctypes.c_uint64),
                                   POINTER(ctypes.c_uint64), POINTER(ctypes.c_uint64)]
    L.strata_selfcheck.restype = c_int
    L.strata_timing_begin.argtypes = []
    L.strata_timing_begin.restype = c_int
    L.strata_timing_end.argtypes = [POINTER(ctypes.c_double), POINTER(c_int64), POINTER(ctypes.c_double),
                                    POINTER(c_int64)]
    L.strata_timing_end.restype = c_int
    _bind_optional(L)
    _lib = L
    return L


def _bind_optional(L) -> None:
    """Bind the batch and serving entry points."""
    for name, argtypes, restype in _OPTIONAL:
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = restype


_OPTIONAL: list = [
    ("strata_correlate_error", [POINTER(Kernel), POINTER(Raster), POINTER(Raster), POINTER(ctypes.c_double),
                                POINTER(ctypes.c_double), POINTER(ctypes.c_double), POINTER(c_int64)], c_int),
    ("strata_batch_create", [POINTER(Task), c_int, POINTER(BatchOptions), POINTER(c_void_p)], c_int),
    ("strata_batch_run", [c_void_p], c_int),
    ("strata_batch_run_per_tile", [c_void_p], c_int),
    ("strata_batch_tune", [c_void_p, c_int, POINTER(BatchOptions)], c_int),
    ("strata_batch_sync", [c_void_p], c_int),
    ("strata_batch_info_get", [c_void_p, POINTER(BatchInfo)], c_int),
    ("strata_batch_destroy", [c_void_p], None),
    ("strata_serve_create", [POINTER(ServeConfig), POINTER(c_void_p)], c_int),
    ("strata_serve_submit", [c_void_p, POINTER(Window), POINTER(c_int64)], c_int),
    ("strata_serve_wait", [c_void_p, c_int64, POINTER(ctypes.c_double)], c_int),
    ("strata_serve_fetch", [c_void_p, c_int64, c_void_p, c_void_p], c_int),
    ("strata_serve_release", [c_void_p, c_int64], c_int),
    ("strata_serve_replay", [c_void_p, POINTER(Window), POINTER(ctypes.c_double), c_int,
                             POINTER(ctypes.c_double)], c_int),
    ("strata_serve_stats_get", [c_void_p, POINTER(ServeStats)], c_int),
    ("strata_serve_destroy", [c_void_p], None),
]


class KernelTiming:
    """Sums of device kernel durations recorded by the library between
    begin() and end(): HIP events around each launch on its own stream."""

    def begin(self) -> "KernelTiming":
        lib().strata_timing_begin()
        return self

    def end(self):
        dms, vms = ctypes.c_double(), ctypes.c_double()
        dn, vn = c_int64(), c_int64()
        rc = lib().strata_timing_end(ctypes.byref(dms), ctypes.byref(dn), ctypes.byref(vms), ctypes.byref(vn))
        if rc != OK:
            raise RuntimeError(last_error())
        self.data_ms, self.data_launches = dms.value, dn.value
        self.validity_ms, self.validity_launches = vms.value, vn.value
        return self


def last_error() -> str:
    return lib().strata_last_error().decode("utf-8", "replace")
