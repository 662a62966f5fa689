"""Python binding of the C ABI in include/chx.h (ctypes over libchx.so).

This mirrors the reference's operator interface for the correlator path --
the OpenCL program's three kernels `offsetAccumulateElements`, `preseed` and
`corr` (offset_accumulator.cl:7, preseed_multifreq_UT.cl:19,
pairwise_correlator_UT.cl:22) and the host loop that chains them
(main_wrapper.c:691-778) -- with the same argument meaning: packed 4-bit
input [time][freq][element], block-order int32 output, the
upper_triangle_convention flag and time_accum.

There is no CPU fallback: if libchx.so is missing, lib() raises.

Processes that also use torch must import torch BEFORE the first lib()
call: torch ships its own HIP runtime (libamdhip64.so without a SONAME);
loaded first, it becomes the runtime libchx binds to, so torch's device
pointers and streams are valid arguments of the C ABI.  Loaded after,
the process would hold two HIP runtimes.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_PATH = os.environ.get("CHX_LIB", os.path.join(HERE, "libchx.so"))
HEADER = os.path.join(REPO, "include", "chx.h")

CONFERENCE, HIGHLY_PACKED, INT8_MFMA, FP6_MFMA = 0, 1, 2, 3
METHOD_NAMES = {CONFERENCE: "conference", HIGHLY_PACKED: "highly_packed", INT8_MFMA: "int8_mfma",
                FP6_MFMA: "fp6_mfma"}

SUCCESS = 0
ERR_INVALID_ARGUMENT, ERR_SHAPE, ERR_TIME_ACCUM, ERR_NO_DEVICE = 1, 2, 3, 4
ERR_OUT_OF_MEMORY, ERR_HIP, ERR_BUDGET, ERR_BUSY, ERR_NOT_FOUND = 5, 6, 7, 8, 9


class ChxError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        detail = lib().chx_last_error_detail().decode()
        msg = f"{where}: {lib().chx_strerror(status).decode()} (status {status})"
        if detail and (status == ERR_HIP or where == "chx_correlate_noalloc"):
            msg += f" [{detail}]"
        super().__init__(msg)


class Config(ctypes.Structure):
    _fields_ = [
        ("num_elements", ctypes.c_int32),
        ("num_freq", ctypes.c_int32),
        ("time_steps", ctypes.c_int64),
        ("time_accum", ctypes.c_int32),
        ("upper_triangle_convention", ctypes.c_int32),
        ("kernel_batch", ctypes.c_int32),
        ("device", ctypes.c_int32),
    ]


class JobDesc(ctypes.Structure):
    _fields_ = [
        ("num_elements", ctypes.c_int32),
        ("num_freq", ctypes.c_int32),
        ("time_steps", ctypes.c_int64),
        ("upper_triangle_convention", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("d_input", ctypes.c_void_p),
        ("d_vis", ctypes.c_void_p),
    ]


class BatchTuning(ctypes.Structure):
    _fields_ = [
        ("max_item_stages", ctypes.c_int32),
        ("workgroups", ctypes.c_int32),
        ("workspace_bytes", ctypes.c_int64),
        ("method", ctypes.c_int32),
        ("small_tile_elements", ctypes.c_int32),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("num_jobs", ctypes.c_int32),
        ("rounds", ctypes.c_int32),
        ("chunk_samples", ctypes.c_int64),
        ("items", ctypes.c_int64),
        ("workgroups", ctypes.c_int32),
        ("max_item_stages", ctypes.c_int32),
        ("workspace_bytes", ctypes.c_size_t),
        ("cmacs", ctypes.c_int64),
        ("method", ctypes.c_int32),
        ("small_tile_jobs", ctypes.c_int32),
    ]


BATCH_EAGER = 1


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("frame_slots", ctypes.c_int32),
        ("slot_bytes", ctypes.c_int64),
        ("workspace_bytes", ctypes.c_int64),
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_batch_frames", ctypes.c_int32),
        ("max_item_stages", ctypes.c_int32),
        ("method", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("frames_submitted", ctypes.c_int64),
        ("frames_done", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("streams_open", ctypes.c_int64),
        ("bytes_in_use", ctypes.c_int64),
        ("wall_seconds", ctypes.c_double),
        ("cmacs", ctypes.c_double),
        ("mean_batch_frames", ctypes.c_double),
        ("latency_p50_ms", ctypes.c_double),
        ("latency_p99_ms", ctypes.c_double),
        ("latency_max_ms", ctypes.c_double),
        ("latency_mean_ms", ctypes.c_double),
    ]

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `make -C chime-x_amd` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sz = ctypes.POINTER(ctypes.c_size_t)
    L.chx_plan_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
    L.chx_plan_destroy.argtypes = [vp]
    L.chx_plan_get_config.argtypes = [vp, ctypes.POINTER(Config)]
    L.chx_plan_sizes.argtypes = [vp, sz, sz, sz]
    for name in ("chx_offset_accumulate", "chx_preseed", "chx_corr", "chx_correlate"):
        getattr(L, name).argtypes = [vp, vp, vp, vp]
    L.chx_correlate_host.argtypes = [vp, vp, vp]
    L.chx_correlate_noalloc.argtypes = [ctypes.POINTER(Config), vp, vp, vp]
    L.chx_device_arch.argtypes = [i32, ctypes.c_char_p, ctypes.c_size_t]
    L.chx_plan_profile.argtypes = [vp, i32]
    L.chx_plan_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int64)]
    L.chx_num_blocks.argtypes = [i32]
    L.chx_num_blocks.restype = i64
    L.chx_block_maps.argtypes = [i32, vp, vp]
    L.chx_blocks_to_triangle.argtypes = [i32, i32, vp, vp]
    L.chx_blocks_to_triangle_device.argtypes = [i32, i32, vp, vp, vp]
    L.chx_blocks_to_full.argtypes = [i32, i32, vp, vp]
    L.chx_blocks_to_full_device.argtypes = [i32, i32, vp, vp, vp]
    L.chx_split_paired_feeds.argtypes = [i32, i32, vp]
    L.chx_feed16_to_triangle.argtypes = [i32, i32, vp, vp]
    L.chx_strerror.argtypes = [ctypes.c_int]
    L.chx_strerror.restype = ctypes.c_char_p
    L.chx_last_error_detail.restype = ctypes.c_char_p
    L.chx_max_time_accum.argtypes = [i32]
    L.chx_max_time_accum.restype = i32
    L.chx_version.restype = ctypes.c_char_p
    L.chx_batch_create.argtypes = [i32, ctypes.POINTER(JobDesc), i32,
                                   ctypes.POINTER(BatchTuning), ctypes.POINTER(vp)]
    L.chx_batch_plan.argtypes = [ctypes.POINTER(JobDesc), i32, ctypes.POINTER(BatchTuning),
                                 ctypes.POINTER(BatchInfo)]
    L.chx_batch_run.argtypes = [vp, i32, vp]
    L.chx_batch_get_info.argtypes = [vp, ctypes.POINTER(BatchInfo)]
    L.chx_batch_destroy.argtypes = [vp]
    L.chx_generate_device.argtypes = [i32] * 7 + [i64, i32, i32, i32, vp, vp]
    L.chx_rand_host.argtypes = [i32, i64, i64, vp]
    pi64 = ctypes.POINTER(ctypes.c_int64)
    L.chx_serve_create.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(vp)]
    L.chx_serve_destroy.argtypes = [vp]
    L.chx_serve_open.argtypes = [vp, i32, i32, i32, i32, i64, pi64]
    L.chx_serve_submit.argtypes = [vp, i64, vp, pi64]
    L.chx_serve_submit_device.argtypes = [vp, i64, vp, pi64]
    L.chx_serve_wait_frame.argtypes = [vp, i64]
    L.chx_serve_read.argtypes = [vp, i64, vp]
    L.chx_serve_read_device.argtypes = [vp, i64, vp]
    L.chx_serve_close.argtypes = [vp, i64]
    L.chx_serve_get_stats.argtypes = [vp, ctypes.POINTER(ServeStats)]
    L.chx_serve_reset_stats.argtypes = [vp]
    L.chx_serve_latencies.argtypes = [vp, vp, vp, i64, pi64]
    # test hook (not in the public header): the next launch fails on the host
    L.chx_serve_debug_fail_next_launch.argtypes = [vp]
    _lib = L
    return L


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chx_[a-z0-9_]+)\s*\(", text)))


def _check(rc: int, where: str) -> None:
    if rc != SUCCESS:
        raise ChxError(rc, where)


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def num_blocks(N: int) -> int:
    return int(lib().chx_num_blocks(N))


def block_maps(N: int):
    nb = num_blocks(N)
    x = np.empty(nb, np.uint32)
    y = np.empty(nb, np.uint32)
    _check(lib().chx_block_maps(N, _p(x), _p(y)), "chx_block_maps")
    return x, y


def blocks_to_triangle(blocks: np.ndarray, N: int) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, np.int32)
    F = blocks.size // (num_blocks(N) * 2048)
    tri = np.empty((F, N * (N + 1) // 2, 2), np.int32)
    _check(lib().chx_blocks_to_triangle(F, N, _p(blocks), _p(tri)), "chx_blocks_to_triangle")
    return tri


def blocks_to_full(blocks: np.ndarray, N: int) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, np.int32)
    F = blocks.size // (num_blocks(N) * 2048)
    full = np.empty((F, N, N, 2), np.int32)
    _check(lib().chx_blocks_to_full(F, N, _p(blocks), _p(full)), "chx_blocks_to_full")
    return full


def blocks_to_full_device(num_freq: int, N: int, d_blocks: int, d_full: int,
                          stream: int = 0) -> None:
    _check(lib().chx_blocks_to_full_device(num_freq, N, d_blocks, d_full, stream or None),
           "chx_blocks_to_full_device")


def split_paired_feeds(full: np.ndarray, num_freq: int, E: int = 16) -> np.ndarray:
    """16-feed unpacking (chx_split_paired_feeds): [F//2][2E][2E][2] full
    matrices of a paired run -> [2*(F//2)][E][E][2]."""
    work = np.array(full, np.int32, order="C", copy=True)
    fin = num_freq // 2
    if work.size < fin * (2 * E) ** 2 * 2:
        raise ValueError("full holds fewer than num_freq//2 matrices of 2E x 2E")
    _check(lib().chx_split_paired_feeds(num_freq, E, _p(work)), "chx_split_paired_feeds")
    return work.reshape(-1)[: fin * 2 * E * E * 2].reshape(2 * fin, E, E, 2).copy()


def feed16_to_triangle(full16: np.ndarray, num_freq_out: int | None = None) -> np.ndarray:
    """[F][16][16][2] -> [num_freq_out][136][2] (chx_feed16_to_triangle),
    zero-padded past F."""
    full16 = np.ascontiguousarray(full16, np.int32)
    F = full16.size // 512
    F_out = F if num_freq_out is None else num_freq_out
    out = np.empty((F_out, 136, 2), np.int32)
    _check(lib().chx_feed16_to_triangle(F_out, F, _p(full16), _p(out)),
           "chx_feed16_to_triangle")
    return out


def device_count() -> int:
    return int(lib().chx_device_count())


class Plan:
    """One correlator configuration on one device (chx_plan)."""

    def __init__(self, num_elements: int, num_freq: int, time_steps: int,
                 time_accum: int = 256, upper_triangle_convention: int = 1,
                 kernel_batch: int = INT8_MFMA, device: int = 0):
        self.cfg = Config(num_elements, num_freq, time_steps, time_accum,
                          upper_triangle_convention, kernel_batch, device)
        h = ctypes.c_void_p()
        _check(lib().chx_plan_create(ctypes.byref(self.cfg), ctypes.byref(h)),
               "chx_plan_create")
        self._h = h
        ib, ob, sb = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().chx_plan_sizes(h, ctypes.byref(ib), ctypes.byref(ob), ctypes.byref(sb)),
               "chx_plan_sizes")
        self.input_bytes, self.output_bytes, self.sums_bytes = ib.value, ob.value, sb.value

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def output_shape(self):
        return (self.cfg.num_freq, num_blocks(self.cfg.num_elements), 32, 32, 2)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().chx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- host pointers ------------------------------------------------------
    def correlate(self, data: np.ndarray) -> np.ndarray:
        """[T][F][N] uint8 -> block-order int32 [F][NB][32][32][2]."""
        data = np.ascontiguousarray(data, np.uint8)
        if data.nbytes != self.input_bytes:
            raise ValueError(f"input has {data.nbytes} bytes, plan expects {self.input_bytes}")
        out = np.empty(self.output_shape, np.int32)
        _check(lib().chx_correlate_host(self._h, _p(data), _p(out)), "chx_correlate_host")
        return out

    # -- device pointers (ints: e.g. torch tensor .data_ptr()) -------------
    def correlate_device(self, d_input: int, d_vis: int, stream: int = 0) -> None:
        _check(lib().chx_correlate(self._h, d_input, d_vis, stream or None), "chx_correlate")

    def offset_accumulate(self, d_input: int, d_sums: int, stream: int = 0) -> None:
        _check(lib().chx_offset_accumulate(self._h, d_input, d_sums, stream or None),
               "chx_offset_accumulate")

    def preseed(self, d_sums: int, d_vis: int, stream: int = 0) -> None:
        _check(lib().chx_preseed(self._h, d_sums, d_vis, stream or None), "chx_preseed")

    def corr(self, d_input: int, d_vis: int, stream: int = 0) -> None:
        _check(lib().chx_corr(self._h, d_input, d_vis, stream or None), "chx_corr")

    # -- kernel timing ------------------------------------------------------
    def profile(self, enable: bool = True) -> None:
        _check(lib().chx_plan_profile(self._h, 1 if enable else 0), "chx_plan_profile")

    def kernel_times(self):
        """(unpack ms, GEMM ms, chunk launches) summed since the last call."""
        u, g, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _check(lib().chx_plan_kernel_times(self._h, ctypes.byref(u), ctypes.byref(g),
                                           ctypes.byref(n)), "chx_plan_kernel_times")
        return u.value, g.value, n.value


NOALLOC_SCRATCH_BYTES = 2 << 20  # CHX_NOALLOC_SCRATCH_BYTES


def correlate_noalloc(num_elements: int, num_freq: int, time_steps: int, d_input: int,
                      d_vis: int, time_accum: int = 256, upper_triangle_convention: int = 1,
                      kernel_batch: int = CONFERENCE, device: int = 0, stream: int = 0) -> None:
    """chx_correlate_noalloc: one correlation of device input into device
    output with no plan and no allocation (packed methods only; scratch in a
    fixed per-GPU device array shared by every call on that GPU)."""
    cfg = Config(num_elements, num_freq, time_steps, time_accum, upper_triangle_convention,
                 kernel_batch, device)
    _check(lib().chx_correlate_noalloc(ctypes.byref(cfg), d_input, d_vis, stream or None),
           "chx_correlate_noalloc")


def blocks_to_triangle_device(num_freq: int, N: int, d_blocks: int, d_tri: int,
                              stream: int = 0) -> None:
    _check(lib().chx_blocks_to_triangle_device(num_freq, N, d_blocks, d_tri, stream or None),
           "chx_blocks_to_triangle_device")


def _job_array(jobs):
    arr = (JobDesc * len(jobs))()
    for k, j in enumerate(jobs):
        arr[k] = JobDesc(j["num_elements"], j["num_freq"], j["time_steps"],
                         j.get("upper_triangle_convention", 1), 0,
                         j["d_input"], j["d_vis"])
    return arr


def _tuning(tuning):
    if not tuning:
        return None
    return BatchTuning(tuning.get("max_item_stages", 0), tuning.get("workgroups", 0),
                       tuning.get("workspace_bytes", 0), tuning.get("method", 0),
                       tuning.get("small_tile_elements", 0))


def batch_plan(jobs, tuning: dict | None = None) -> dict:
    """Host-only planning of a batch (chx_batch_plan): rounds, items and
    workspace, without a GPU.  Device pointers are only checked for NULL."""
    arr = _job_array(jobs)
    t = _tuning(tuning)
    inf = BatchInfo()
    _check(lib().chx_batch_plan(arr, len(jobs), ctypes.byref(t) if t else None,
                                ctypes.byref(inf)), "chx_batch_plan")
    return {name: getattr(inf, name) for name, _ in BatchInfo._fields_}


TUNING_CACHE = os.path.join(HERE, "tuning.json")


def device_arch(device: int = 0) -> str | None:
    """The GPU's target name ("gfx950"), or None without a GPU."""
    buf = ctypes.create_string_buffer(64)
    if lib().chx_device_arch(device, buf, len(buf)) != SUCCESS:
        return None
    return buf.value.decode()


def job_signature(jobs, arch: str | None = None) -> str:
    """Key of a job list's shapes on a GPU target in the autotuner's cache:
    "<arch>/<hash of the sorted shapes>" (arch: the device's, "gfx950" when
    none is visible)."""
    shapes = sorted((j["num_elements"], j["num_freq"], j["time_steps"],
                     j.get("upper_triangle_convention", 1)) for j in jobs)
    arch = arch or device_arch(0) or "gfx950"
    return f"{arch}/" + hashlib.sha1(json.dumps(shapes).encode()).hexdigest()[:16]


def cached_tuning(jobs, path: str = TUNING_CACHE, arch: str | None = None):
    """The autotuned launch shape of this job list's shapes on this GPU
    target, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        ent = json.load(f).get(job_signature(jobs, arch))
    return ent["best"]["tuning"] if ent else None


class Batch:
    """Many matrix-core correlations in one captured graph (chx_batch).

    jobs: sequence of dicts with num_elements, num_freq, time_steps,
    upper_triangle_convention, d_input, d_vis (device pointers as ints).
    tuning: None (defaults), "auto" (the autotuner's cached shape for these
    job shapes, defaults if none), or a dict with any of max_item_stages,
    workgroups, workspace_bytes, method (INT8_MFMA or FP6_MFMA; default FP6).
    """

    def __init__(self, jobs, tuning=None, device: int = 0):
        if tuning == "auto":
            tuning = cached_tuning(jobs)
        arr = _job_array(jobs)
        self._jobs = arr
        self.tuning = _tuning(tuning)
        tp = ctypes.byref(self.tuning) if self.tuning else None
        h = ctypes.c_void_p()
        _check(lib().chx_batch_create(device, arr, len(jobs), tp, ctypes.byref(h)),
               "chx_batch_create")
        self._h = h

    def run(self, stream: int = 0, eager: bool = False) -> None:
        _check(lib().chx_batch_run(self._h, BATCH_EAGER if eager else 0, stream or None),
               "chx_batch_run")

    def info(self) -> dict:
        inf = BatchInfo()
        _check(lib().chx_batch_get_info(self._h, ctypes.byref(inf)), "chx_batch_get_info")
        return {name: getattr(inf, name) for name, _ in BatchInfo._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().chx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Serve:
    """Streaming engine with continuous batching (chx_serve).

    Streams are integrations with device-resident accumulators; frames of
    any open stream are submitted from host memory (copied) or device memory
    (read in place) and integrated in shared launches.  submit/submit_device
    return None instead of raising when no frame slot is free (ERR_BUSY),
    so callers can retry; open returns None when the HBM budget is full.
    """

    def __init__(self, frame_slots: int, slot_bytes: int, workspace_bytes: int = 0,
                 hbm_budget_bytes: int = 0, max_batch_frames: int = 0,
                 max_item_stages: int = 0, device: int = 0, method: int = 0):
        self.cfg = ServeConfig(device, frame_slots, slot_bytes, workspace_bytes,
                               hbm_budget_bytes, max_batch_frames, max_item_stages, method, 0)
        h = ctypes.c_void_p()
        _check(lib().chx_serve_create(ctypes.byref(self.cfg), ctypes.byref(h)),
               "chx_serve_create")
        self._h = h
        self.shapes = {}

    def open(self, num_elements: int, num_freq: int, frame_samples: int,
             upper_triangle_convention: int = 1, integration_frames: int = 0):
        sid = ctypes.c_int64()
        rc = lib().chx_serve_open(self._h, num_elements, num_freq, frame_samples,
                                  upper_triangle_convention, integration_frames,
                                  ctypes.byref(sid))
        if rc == ERR_BUDGET:
            return None
        _check(rc, "chx_serve_open")
        self.shapes[sid.value] = (num_freq, num_blocks(num_elements), 32, 32, 2)
        return sid.value

    def _submit(self, fn, sid: int, ptr: int):
        fid = ctypes.c_int64()
        rc = fn(self._h, sid, ptr, ctypes.byref(fid))
        if rc == ERR_BUSY:
            return None
        _check(rc, fn.__name__)
        return fid.value

    def submit(self, sid: int, h_frame) -> int | None:
        """h_frame: a host address (int) or a C-contiguous uint8 ndarray that
        the caller keeps unchanged until the frame is done."""
        ptr = _p(h_frame) if isinstance(h_frame, np.ndarray) else h_frame
        return self._submit(lib().chx_serve_submit, sid, ptr)

    def submit_device(self, sid: int, d_frame: int) -> int | None:
        return self._submit(lib().chx_serve_submit_device, sid, d_frame)

    def wait_frame(self, fid: int) -> None:
        _check(lib().chx_serve_wait_frame(self._h, fid), "chx_serve_wait_frame")

    def read(self, sid: int) -> np.ndarray:
        out = np.empty(self.shapes[sid], np.int32)
        _check(lib().chx_serve_read(self._h, sid, _p(out)), "chx_serve_read")
        return out

    def read_device(self, sid: int, d_vis: int) -> None:
        _check(lib().chx_serve_read_device(self._h, sid, d_vis), "chx_serve_read_device")

    def close_stream(self, sid: int) -> None:
        _check(lib().chx_serve_close(self._h, sid), "chx_serve_close")
        self.shapes.pop(sid, None)

    def stats(self) -> dict:
        st = ServeStats()
        _check(lib().chx_serve_get_stats(self._h, ctypes.byref(st)), "chx_serve_get_stats")
        return {name: getattr(st, name) for name, _ in ServeStats._fields_}

    def reset_stats(self) -> None:
        _check(lib().chx_serve_reset_stats(self._h), "chx_serve_reset_stats")

    def latencies(self):
        """(frame ids, submit-to-done ms) of the completed frames."""
        n = ctypes.c_int64()
        _check(lib().chx_serve_latencies(self._h, None, None, 0, ctypes.byref(n)),
               "chx_serve_latencies")
        ids = np.empty(n.value, np.int64)
        ms = np.empty(n.value, np.float64)
        _check(lib().chx_serve_latencies(self._h, _p(ids), _p(ms), n.value, ctypes.byref(n)),
               "chx_serve_latencies")
        return ids[:n.value], ms[:n.value]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().chx_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


GEN_CONSTANT, GEN_RAMP_UP, GEN_RAMP_DOWN, GEN_RANDOM = 1, 2, 3, 4


def generate_device(d_data: int, T: int, F: int, N: int, gen_type: int = GEN_RANDOM,
                    seed: int = 42, default_real: int = 0, default_imag: int = 0,
                    initial_real: int = 0, initial_imag: int = 0, single_freq: int = -1,
                    no_repeat: bool = False, stream: int = 0) -> None:
    """The reference generator (input_generator.c) on the device into d_data."""
    _check(lib().chx_generate_device(gen_type, seed, default_real, default_imag, initial_real,
                                     initial_imag, single_freq, T, F, N, int(no_repeat),
                                     d_data, stream or None), "chx_generate_device")


def rand_host(seed: int, k0: int, n: int) -> np.ndarray:
    """glibc rand() outputs k0 .. k0+n-1 after srand(seed) (jump-ahead)."""
    out = np.empty(n, np.int32)
    _check(lib().chx_rand_host(seed, k0, n, _p(out)), "chx_rand_host")
    return out
