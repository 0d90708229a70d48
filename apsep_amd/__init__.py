"""apsep_amd -- MI355X-native APSEP / ANSV-left engine (host bindings).

The product is the gfx950 HIP library ``apsep_amd/libapsep.so`` and its C ABI
(``include/apsep/apsep.h``); C++ callers use the reference-named shims in
``include/apsep/apsep.hpp``.  This module binds the same C ABI with ctypes
for Python callers, tests and the benchmark.  PyTorch is used only for
device memory and streams: every solve runs the HIP kernels, and the module
raises if the library is missing -- there is no CPU fallback.

Entry points mirror the reference's host API (src/apsep.cuh:2503-2605,
3001-3076):

    Scratch(algo, n_cap, incl)      ~ allocWSTLScratch / allocSPTScratch
    Scratch.close()                 ~ freeWSTLScratch / freeSPTScratch
    Scratch.run(d_in, d_out, n)     ~ runWSTL / runSPT (async on a stream)
    launch(algo, d_in, d_out, n)    ~ launchWSTL / launchSPT (sync)
    solve(x)                        convenience: allocate the output, solve
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libapsep.so")

OK, ERR_INVALID_ARG, ERR_ALIGN, ERR_CAPACITY, ERR_ALLOC, ERR_LAUNCH, \
    ERR_DEVICE, ERR_TOO_LARGE, ERR_UNSUPPORTED, ERR_BUDGET = range(10)
ALGO_WSTL, ALGO_SPT, ALGO_LOOKBACK, ALGO_AUTO = range(4)
ALGOS = {"wstl": ALGO_WSTL, "spt": ALGO_SPT, "lookback": ALGO_LOOKBACK,
         "auto": ALGO_AUTO}

# every symbol include/apsep/apsep.h declares
EXPORTS = [
    "apsep_status_string", "apsep_version", "apsep_last_error", "apsep_scratch_create",
    "apsep_scratch_destroy", "apsep_scratch_bytes", "apsep_run",
    "apsep_launch", "apsep_run_mode", "apsep_run_noalloc", "apsep_scratch_query",
    "apsep_scratch_check", "apsep_scratch_set_persistent_blocks", "apsep_time_kernels",
    "apsep_scratch_set_cooperative", "apsep_scratch_cooperative_refusals",
    "apsep_time_kernels_rotating",
    "apsep_batch_plan_create", "apsep_batch_plan_destroy", "apsep_run_segmented",
    "apsep_batch_plan_use_graph", "apsep_batch_plan_tune",
    "apsep_batch_plan_set_tile_lanes", "apsep_batch_plan_tile_lanes",
    "apsep_batch_plan_tile_elems", "apsep_batch_plan_job_tiles", "apsep_run_jobs",
    "apsep_batch_plan_tuned", "apsep_batch_plan_set_persistent", "apsep_batch_plan_check",
    "apsep_batch_plan_tuned_persistent",
    "apsep_server_create", "apsep_server_destroy", "apsep_server_submit",
    "apsep_server_poll", "apsep_server_job_info", "apsep_server_release",
    "apsep_server_drain", "apsep_server_submit_many", "apsep_server_done_upto",
    "apsep_server_job_times", "apsep_server_release_upto",
]

_lib = None


class ScratchInfo(ctypes.Structure):
    """apsep_scratch_info (include/apsep/apsep.h)."""
    _fields_ = [("n_cap", ctypes.c_int64), ("bytes", ctypes.c_int64),
                ("algo", ctypes.c_int32), ("incl", ctypes.c_int32),
                ("device", ctypes.c_int32), ("tile_elems", ctypes.c_int32),
                ("tiles", ctypes.c_int32), ("levels", ctypes.c_int32),
                ("persistent_blocks", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class KernelTimes(ctypes.Structure):
    """apsep_kernel_times (include/apsep/apsep.h): summed device time per kernel."""
    _fields_ = [("runs", ctypes.c_int64), ("tile_ms", ctypes.c_double),
                ("levels_ms", ctypes.c_double), ("fix_ms", ctypes.c_double)]


class JobInfo(ctypes.Structure):
    """apsep_job_info (include/apsep/apsep.h)."""
    _fields_ = [("state", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("n", ctypes.c_int64), ("submit_us", ctypes.c_double),
                ("done_us", ctypes.c_double), ("d_answers", ctypes.c_void_p)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2


class ApsepError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = lib().apsep_status_string(status).decode()
        detail = lib().apsep_last_error().decode()
        super().__init__("%s%s (status %d; %s)" % (what + ": " if what else "", msg, status,
                                                   detail))


def lib() -> ctypes.CDLL:
    """Load libapsep.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("apsep_amd/libapsep.so is not built; run "
                              "`python -m apsep_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int
        L.apsep_status_string.argtypes = [ci]
        L.apsep_status_string.restype = ctypes.c_char_p
        L.apsep_version.restype = ctypes.c_char_p
        L.apsep_last_error.restype = ctypes.c_char_p
        L.apsep_scratch_create.argtypes = [ci, i64, ci, ctypes.POINTER(vp)]
        L.apsep_scratch_destroy.argtypes = [vp]
        L.apsep_scratch_bytes.argtypes = [ci, i64, ci, ctypes.POINTER(i64)]
        L.apsep_run.argtypes = [vp, vp, vp, i64, vp]
        L.apsep_launch.argtypes = [ci, vp, vp, i64, ci]
        L.apsep_run_mode.argtypes = [vp, vp, vp, i64, ci, vp]
        L.apsep_run_noalloc.argtypes = [vp, vp, i64, ci, vp]
        L.apsep_scratch_query.argtypes = [vp, ctypes.POINTER(ScratchInfo)]
        L.apsep_scratch_check.argtypes = [vp, ctypes.POINTER(i32)]
        L.apsep_scratch_set_persistent_blocks.argtypes = [vp, i32]
        L.apsep_scratch_set_cooperative.argtypes = [vp, i32]
        L.apsep_scratch_cooperative_refusals.argtypes = [vp, ctypes.POINTER(i64)]
        L.apsep_time_kernels.argtypes = [vp, vp, vp, i64, ci, ci, vp,
                                         ctypes.POINTER(KernelTimes)]
        L.apsep_time_kernels_rotating.argtypes = [vp, vp, vp, i32, i64, ci, ci, vp,
                                                  ctypes.POINTER(KernelTimes)]
        L.apsep_batch_plan_create.argtypes = [i64, i32, ci, ctypes.POINTER(vp)]
        L.apsep_batch_plan_destroy.argtypes = [vp]
        L.apsep_run_segmented.argtypes = [vp, vp, vp, vp, i32, i64, vp]
        L.apsep_batch_plan_use_graph.argtypes = [vp, ci]
        L.apsep_batch_plan_set_tile_lanes.argtypes = [vp, i32]
        L.apsep_batch_plan_tile_lanes.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.apsep_batch_plan_tile_elems.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.apsep_batch_plan_job_tiles.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.apsep_run_jobs.argtypes = [vp, vp, i32, i64, vp]
        L.apsep_batch_plan_tuned.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                             ctypes.POINTER(ctypes.c_float)]
        L.apsep_batch_plan_set_persistent.argtypes = [vp, ci]
        L.apsep_batch_plan_check.argtypes = [vp, ctypes.POINTER(i32)]
        L.apsep_batch_plan_tuned_persistent.argtypes = [vp, i32, ctypes.POINTER(i32)]
        L.apsep_batch_plan_tune.argtypes = [vp, vp, vp, vp, i32, i64, vp, i32,
                                            ctypes.POINTER(ctypes.c_float),
                                            ctypes.POINTER(ctypes.c_float)]
        L.apsep_server_create.argtypes = [i64, i64, i32, ci, ctypes.POINTER(vp)]
        L.apsep_server_destroy.argtypes = [vp]
        L.apsep_server_submit.argtypes = [vp, vp, i64, ctypes.POINTER(i64)]
        L.apsep_server_poll.argtypes = [vp, ctypes.POINTER(i32)]
        L.apsep_server_job_info.argtypes = [vp, i64, ctypes.POINTER(JobInfo)]
        L.apsep_server_release.argtypes = [vp, i64]
        L.apsep_server_drain.argtypes = [vp]
        L.apsep_server_submit_many.argtypes = [vp, vp, vp, i32, ctypes.POINTER(i64),
                                               ctypes.POINTER(i32)]
        L.apsep_server_done_upto.argtypes = [vp, ctypes.POINTER(i64)]
        L.apsep_server_job_times.argtypes = [vp, i64, i32, vp, vp]
        L.apsep_server_release_upto.argtypes = [vp, i64]
        for f in EXPORTS:
            if f not in ("apsep_status_string", "apsep_version", "apsep_last_error"):
                getattr(L, f).restype = ci
        _lib = L
    return _lib


def _check(st: int, what: str = ""):
    if st != OK:
        raise ApsepError(st, what)


def _algo(a) -> int:
    if isinstance(a, str):
        return ALGOS[a.lower()]
    return int(a)


def scratch_bytes(algo, n_cap: int, incl: bool = False) -> int:
    b = ctypes.c_int64(0)
    _check(lib().apsep_scratch_bytes(_algo(algo), n_cap, int(incl), ctypes.byref(b)),
           "apsep_scratch_bytes")
    return b.value


def _device_index(t) -> int:
    return t.device.index if t.device.index is not None else 0


def check_single_args(x, out, n: int, n_cap: int, device: int):
    """Validate a solve of ``n`` keys of ``x`` into ``out`` before any launch:
    int32 dtypes, 0 <= n <= both sizes and the capacity, contiguous device
    tensors on the scratch's device.  Raises TypeError / ValueError; touches
    no device."""
    import torch
    if x.dtype != torch.int32 or out.dtype != torch.int32:
        raise TypeError("apsep: int32 tensors required")
    if n < 0:
        raise ValueError("apsep: n = %d is negative" % n)
    if n > x.numel() or n > out.numel():
        raise ValueError("apsep: n = %d exceeds the input (%d) or output (%d) size"
                         % (n, x.numel(), out.numel()))
    if n > n_cap:
        raise ValueError("apsep: n = %d exceeds the scratch capacity %d" % (n, n_cap))
    if not (x.is_contiguous() and out.is_contiguous()):
        raise ValueError("apsep: contiguous tensors required")
    if not (x.is_cuda and out.is_cuda):
        raise ValueError("apsep: device tensors required")
    if _device_index(x) != device or _device_index(out) != device:
        raise ValueError("apsep: tensors on device %d/%d, scratch on device %d"
                         % (_device_index(x), _device_index(out), device))


def check_segmented_args(x, offsets, out, total: int, total_cap: int, nseg_cap: int,
                         device: int):
    """Validate a segmented solve before any launch: int32 keys and answers,
    a 1-D int64 offsets tensor of nseg + 1 entries with nseg <= nseg_cap,
    0 <= total <= both sizes and the capacity, contiguous device tensors on
    the plan's device.  Raises TypeError / ValueError; touches no device."""
    import torch
    if x.dtype != torch.int32 or out.dtype != torch.int32 or offsets.dtype != torch.int64:
        raise TypeError("apsep: int32 keys/answers and int64 offsets required")
    if offsets.dim() != 1 or offsets.numel() < 1:
        raise ValueError("apsep: offsets must be 1-D with nseg + 1 >= 1 entries")
    nseg = offsets.numel() - 1
    if nseg > nseg_cap:
        raise ValueError("apsep: %d segments exceed the plan's %d" % (nseg, nseg_cap))
    if total < 0:
        raise ValueError("apsep: total = %d is negative" % total)
    if total > x.numel() or total > out.numel():
        raise ValueError("apsep: total = %d exceeds the keys (%d) or answers (%d) size"
                         % (total, x.numel(), out.numel()))
    if total > total_cap:
        raise ValueError("apsep: total = %d exceeds the plan capacity %d" % (total, total_cap))
    if not (x.is_contiguous() and out.is_contiguous() and offsets.is_contiguous()):
        raise ValueError("apsep: contiguous tensors required")
    if not (x.is_cuda and out.is_cuda and offsets.is_cuda):
        raise ValueError("apsep: device tensors required")
    devs = {_device_index(x), _device_index(out), _device_index(offsets)}
    if devs != {device}:
        raise ValueError("apsep: tensors on devices %s, plan on device %d"
                         % (sorted(devs), device))


class Scratch:
    """Device scratch for solves of up to ``n_cap`` elements."""

    def __init__(self, algo="auto", n_cap: int = 0, incl: bool = False):
        self._h = ctypes.c_void_p(None)
        self.n_cap = n_cap
        self.incl = bool(incl)
        _check(lib().apsep_scratch_create(_algo(algo), n_cap, int(incl),
                                          ctypes.byref(self._h)),
               "apsep_scratch_create")
        self.info = ScratchInfo()
        _check(lib().apsep_scratch_query(self._h, ctypes.byref(self.info)))
        self.algo = self.info.algo
        self.device = self.info.device

    def run_ptr(self, d_in: int, d_out: int, n: int, stream: int = 0,
                incl: bool | None = None):
        """Raw-pointer solve (async on ``stream``)."""
        if incl is None:
            _check(lib().apsep_run(self._h, d_in, d_out, n, stream), "apsep_run")
        else:
            _check(lib().apsep_run_mode(self._h, d_in, d_out, n, int(incl), stream),
                   "apsep_run_mode")

    def run(self, x, out, n: int | None = None, stream=None, incl: bool | None = None):
        """Solve ``x`` (int32 CUDA tensor) into ``out`` on ``stream`` (default:
        torch's current stream)."""
        import torch
        if n is None:
            n = x.numel()
        check_single_args(x, out, n, self.n_cap, self.device)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        self.run_ptr(x.data_ptr(), out.data_ptr(), n, stream.cuda_stream, incl)
        return out

    def set_persistent_blocks(self, blocks: int):
        """SPT persistent grid of later runs (0: the default, occupancy x CUs)."""
        _check(lib().apsep_scratch_set_persistent_blocks(self._h, int(blocks)),
               "apsep_scratch_set_persistent_blocks")

    def set_cooperative(self, enable: bool = True):
        """SPT: launch later eager runs with hipLaunchCooperativeKernel (True)
        or normally (False, the default)."""
        _check(lib().apsep_scratch_set_cooperative(self._h, int(enable)),
               "apsep_scratch_set_cooperative")

    @property
    def cooperative_refusals(self) -> int:
        """Cooperative launches the runtime refused (those runs launched
        normally)."""
        v = ctypes.c_int64(0)
        _check(lib().apsep_scratch_cooperative_refusals(self._h, ctypes.byref(v)),
               "apsep_scratch_cooperative_refusals")
        return v.value

    def check(self):
        """Blocking check of the device-side error word (raises on error)."""
        e = ctypes.c_int32(0)
        _check(lib().apsep_scratch_check(self._h, ctypes.byref(e)), "apsep_scratch_check")
        if e.value:
            raise ApsepError(ERR_LAUNCH, "device-side error %d (grid barrier timeout)"
                             % e.value)

    def time_kernels(self, x, out, reps: int, n: int | None = None, stream=None,
                     incl: bool | None = None) -> KernelTimes:
        """Launch each kernel of the solve of ``x`` into ``out`` ``reps`` times
        back to back and return the summed device time per kernel (blocks)."""
        import torch
        if n is None:
            n = x.numel()
        check_single_args(x, out, n, self.n_cap, self.device)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        t = KernelTimes()
        _check(lib().apsep_time_kernels(self._h, x.data_ptr(), out.data_ptr(), n,
                                        int(self.incl if incl is None else incl), reps,
                                        stream.cuda_stream, ctypes.byref(t)),
               "apsep_time_kernels")
        return t

    def close(self):
        if self._h is not None and self._h.value:
            lib().apsep_scratch_destroy(self._h)
        self._h = ctypes.c_void_p(None)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPlan:
    """Segmented (CSR) batches of up to ``total_cap`` elements in up to
    ``nseg_cap`` segments (apsep_batch_plan_*)."""

    def __init__(self, total_cap: int, nseg_cap: int, incl: bool = False,
                 graph: bool = False):
        self._h = ctypes.c_void_p(None)
        self.incl = bool(incl)
        self.total_cap = total_cap
        self.nseg_cap = nseg_cap
        _check(lib().apsep_batch_plan_create(total_cap, nseg_cap, int(incl),
                                             ctypes.byref(self._h)),
               "apsep_batch_plan_create")
        import torch
        self.device = torch.cuda.current_device()  # the plan binds to it
        if graph:
            self.use_graph(True)

    def use_graph(self, enable: bool = True):
        _check(lib().apsep_batch_plan_use_graph(self._h, int(enable)),
               "apsep_batch_plan_use_graph")

    def set_persistent(self, enable: bool = True):
        """Solve each batch in one persistent launch (True) or by the kernel
        sequence (False, the default)."""
        _check(lib().apsep_batch_plan_set_persistent(self._h, int(enable)),
               "apsep_batch_plan_set_persistent")

    def check(self):
        """Blocking check of the persistent launches' device-side error word
        (raises on error)."""
        e = ctypes.c_int32(0)
        _check(lib().apsep_batch_plan_check(self._h, ctypes.byref(e)), "apsep_batch_plan_check")
        if e.value:
            raise ApsepError(ERR_LAUNCH, "device-side error %d (grid barrier timeout)"
                             % e.value)

    def tuned_persistent(self, size_class: int) -> bool:
        """Whether the tuner chose one persistent launch for the class."""
        v = ctypes.c_int32(0)
        _check(lib().apsep_batch_plan_tuned_persistent(self._h, size_class, ctypes.byref(v)),
               "apsep_batch_plan_tuned_persistent")
        return bool(v.value)

    def tuned(self, size_class: int):
        """(tile lanes, graph, ms) the tuner recorded for batches of
        2^(c-1) + 1 .. 2^c keys; lanes = 0 when the class is untuned."""
        ln, g, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_float(0)
        _check(lib().apsep_batch_plan_tuned(self._h, size_class, ctypes.byref(ln),
                                            ctypes.byref(g), ctypes.byref(ms)),
               "apsep_batch_plan_tuned")
        return ln.value, g.value, ms.value

    def set_tile_lanes(self, lanes: int):
        """Tile shape of this plan's batches: 128, 256 or 512 lanes x 16 keys."""
        _check(lib().apsep_batch_plan_set_tile_lanes(self._h, int(lanes)),
               "apsep_batch_plan_set_tile_lanes")

    @property
    def tile_lanes(self) -> int:
        v = ctypes.c_int32(0)
        _check(lib().apsep_batch_plan_tile_lanes(self._h, ctypes.byref(v)),
               "apsep_batch_plan_tile_lanes")
        return v.value

    def tune(self, x, offsets, out, total: int | None = None, stream=None, reps: int = 3):
        """Time every tile shape and launch sequence (kernel sequence or one
        persistent launch), each by direct launches and by graph replay, on
        this batch and keep the fastest; returns the chosen shape and
        sequence's (eager_ms, graph_ms)."""
        import torch
        if total is None:
            total = x.numel()
        check_segmented_args(x, offsets, out, total, self.total_cap, self.nseg_cap,
                             self.device)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        e, g = ctypes.c_float(0), ctypes.c_float(0)
        _check(lib().apsep_batch_plan_tune(self._h, x.data_ptr(), offsets.data_ptr(),
                                           out.data_ptr(), offsets.numel() - 1, total,
                                           stream.cuda_stream, reps, ctypes.byref(e),
                                           ctypes.byref(g)), "apsep_batch_plan_tune")
        return e.value, g.value

    @property
    def tile_elems(self) -> int:
        v = ctypes.c_int32(0)
        _check(lib().apsep_batch_plan_tile_elems(self._h, ctypes.byref(v)),
               "apsep_batch_plan_tile_elems")
        return v.value

    @property
    def job_tiles(self) -> int:
        v = ctypes.c_int64(0)
        _check(lib().apsep_batch_plan_job_tiles(self._h, ctypes.byref(v)),
               "apsep_batch_plan_job_tiles")
        return v.value

    def job_table(self, xs, outs):
        """The apsep_job table (int64 tensor of 4 words per job, on the
        device) and the total tile count for solving xs[k] into outs[k]
        in place with this plan's current tile shape."""
        import torch
        if len(xs) != len(outs):
            raise ValueError("apsep: one answer tensor per key tensor")
        te = self.tile_elems
        rows, first = [], 0
        for x, o in zip(xs, outs):
            n = x.numel()
            if x.dtype != torch.int32 or o.dtype != torch.int32:
                raise TypeError("apsep: int32 tensors required")
            if o.numel() < n or not (x.is_contiguous() and o.is_contiguous()):
                raise ValueError("apsep: contiguous answers of at least the keys' size required")
            if not (x.is_cuda and o.is_cuda) or {_device_index(x), _device_index(o)} != {self.device}:
                raise ValueError("apsep: device tensors on the plan's device required")
            if n and (x.data_ptr() % 16 or o.data_ptr() % 16):
                raise ValueError("apsep: job keys and answers must be 16-byte aligned")
            rows.append([x.data_ptr(), o.data_ptr(), n, first])
            first += (n + te - 1) // te
        table = torch.tensor(rows, dtype=torch.int64).reshape(-1, 4).to(f"cuda:{self.device}")
        return table, first

    def run_jobs(self, xs, outs, stream=None, table=None, tiles: int | None = None):
        """Solve each xs[k] (int32 device tensor) into outs[k] in place, all in
        one launch sequence (apsep_run_jobs).  Returns the job table used,
        which must stay alive until the solve has run."""
        import torch
        if table is None:
            table, tiles = self.job_table(xs, outs)
        if stream is None:
            stream = torch.cuda.current_stream()
        _check(lib().apsep_run_jobs(self._h, table.data_ptr(), table.shape[0], tiles,
                                    stream.cuda_stream), "apsep_run_jobs")
        return table

    def run_ptr(self, d_in: int, d_offsets: int, d_out: int, nseg: int, total: int,
                stream: int = 0):
        _check(lib().apsep_run_segmented(self._h, d_in, d_offsets, d_out, nseg, total,
                                         stream), "apsep_run_segmented")

    def run(self, x, offsets, out, total: int | None = None, stream=None):
        """Solve every segment of ``x`` (int32) given ``offsets`` (int64, nseg + 1,
        on the device); answers relative to each segment start go to ``out``."""
        import torch
        if total is None:
            total = x.numel()
        check_segmented_args(x, offsets, out, total, self.total_cap, self.nseg_cap,
                             self.device)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        self.run_ptr(x.data_ptr(), offsets.data_ptr(), out.data_ptr(),
                     offsets.numel() - 1, total, stream.cuda_stream)
        return out

    def close(self):
        if self._h is not None and self._h.value:
            lib().apsep_batch_plan_destroy(self._h)
        self._h = ctypes.c_void_p(None)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hiplib = None


def _hip():
    """The HIP runtime (for device-to-device copies out of a server's pool)."""
    global _hiplib
    if _hiplib is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int, ctypes.c_void_p]
        h.hipMemcpyAsync.restype = ctypes.c_int
        _hiplib = h
    return _hiplib


class Server:
    """Serving engine (apsep_server_*): a ``pool_bytes`` device pool for the
    answers, budget admission and continuous batching into in-place job
    batches on two streams (keys are read where the caller holds them)."""

    def __init__(self, pool_bytes: int, max_batch: int, max_batch_jobs: int = 4096,
                 incl: bool = False):
        import torch
        self._h = ctypes.c_void_p(None)
        # the server is bound to the current device (as apsep_server_create)
        self.device = torch.cuda.current_device()
        _check(lib().apsep_server_create(pool_bytes, max_batch, max_batch_jobs, int(incl),
                                         ctypes.byref(self._h)), "apsep_server_create")

    def submit(self, d_keys: int, n: int):
        """Job id, or None when the pool has no room (admission refused)."""
        jid = ctypes.c_int64(-1)
        st = lib().apsep_server_submit(self._h, d_keys, n, ctypes.byref(jid))
        if st == ERR_BUDGET:
            return None
        _check(st, "apsep_server_submit")
        return jid.value

    def poll(self) -> int:
        k = ctypes.c_int32(0)
        _check(lib().apsep_server_poll(self._h, ctypes.byref(k)), "apsep_server_poll")
        return k.value

    def info(self, job_id: int) -> JobInfo:
        ji = JobInfo()
        _check(lib().apsep_server_job_info(self._h, job_id, ctypes.byref(ji)),
               "apsep_server_job_info")
        return ji

    def answers(self, job_id: int, out=None):
        """Copy a done job's answers into ``out`` (int32 device tensor of at
        least n; allocated when None) on torch's current stream, and wait for
        the copy: the job's slot may be released (and reused by the server's
        next batch) as soon as this returns."""
        import torch
        info = self.info(job_id)
        if info.state != JOB_DONE:
            raise ApsepError(ERR_INVALID_ARG, "job %d is not done" % job_id)
        if out is None:
            out = torch.empty(info.n, dtype=torch.int32, device=f"cuda:{self.device}")
        if out.dtype != torch.int32 or not out.is_cuda or not out.is_contiguous() \
                or out.numel() < info.n or _device_index(out) != self.device:
            raise ValueError("apsep: answers need a contiguous int32 tensor of >= %d "
                             "elements on the server's device" % info.n)
        if info.n:
            cs = torch.cuda.current_stream(out.device)
            if _hip().hipMemcpyAsync(out.data_ptr(), info.d_answers, info.n * 4,
                                     3, cs.cuda_stream) != 0:  # hipMemcpyDeviceToDevice
                raise ApsepError(ERR_LAUNCH, "hipMemcpyAsync")
            cs.synchronize()
        return out

    def release(self, job_id: int):
        _check(lib().apsep_server_release(self._h, job_id), "apsep_server_release")

    def drain(self):
        _check(lib().apsep_server_drain(self._h), "apsep_server_drain")

    def submit_many(self, ptrs, lengths):
        """Submit jobs in order (numpy uint64 pointers / int64 lengths);
        returns (first id, number accepted)."""
        import numpy as np
        p = np.ascontiguousarray(ptrs, dtype=np.uint64)
        n = np.ascontiguousarray(lengths, dtype=np.int64)
        first, acc = ctypes.c_int64(0), ctypes.c_int32(0)
        _check(lib().apsep_server_submit_many(self._h, p.ctypes.data, n.ctypes.data, p.size,
                                              ctypes.byref(first), ctypes.byref(acc)),
               "apsep_server_submit_many")
        return first.value, acc.value

    def done_upto(self) -> int:
        d = ctypes.c_int64(0)
        _check(lib().apsep_server_done_upto(self._h, ctypes.byref(d)), "apsep_server_done_upto")
        return d.value

    def job_times(self, first_id: int, count: int):
        import numpy as np
        sub = np.zeros(count, dtype=np.float64)
        done = np.zeros(count, dtype=np.float64)
        _check(lib().apsep_server_job_times(self._h, first_id, count, sub.ctypes.data,
                                            done.ctypes.data), "apsep_server_job_times")
        return sub, done

    def release_upto(self, job_id: int):
        _check(lib().apsep_server_release_upto(self._h, job_id), "apsep_server_release_upto")

    def close(self):
        if self._h is not None and self._h.value:
            lib().apsep_server_destroy(self._h)
        self._h = ctypes.c_void_p(None)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NOALLOC_MAX_N = 1 << 26  # APSEP_NOALLOC_MAX_N (include/apsep/apsep.h)


def run_noalloc(x, out, n: int | None = None, incl: bool = False, stream=None):
    """Solve ``x`` into ``out`` without any device allocation
    (apsep_run_noalloc: the scratch is the library's fixed per-GPU array, n
    <= NOALLOC_MAX_N; calls that overlap on one GPU race)."""
    import torch
    if n is None:
        n = x.numel()
    check_single_args(x, out, n, NOALLOC_MAX_N, _device_index(x))
    if stream is None:
        stream = torch.cuda.current_stream(x.device)
    _check(lib().apsep_run_noalloc(x.data_ptr(), out.data_ptr(), n, int(incl),
                                   stream.cuda_stream), "apsep_run_noalloc")
    return out


def time_kernels_rotating(scratches, xs, outs, reps: int, n: int | None = None,
                          stream=None, incl: bool = False) -> KernelTimes:
    """Each kernel of the solve launched ``reps`` times back to back, launch
    r solving array r % len(xs) with its own scratch, input and output
    (apsep_time_kernels_rotating); returns the summed device time per
    kernel (blocks)."""
    import torch
    m = len(xs)
    if not (len(scratches) == len(outs) == m) or not 1 <= m <= 8:
        raise ValueError("apsep: 1..8 (scratch, input, output) triples required")
    if n is None:
        n = xs[0].numel()
    for sc, x, o in zip(scratches, xs, outs):
        check_single_args(x, o, n, sc.n_cap, sc.device)
    if stream is None:
        stream = torch.cuda.current_stream(xs[0].device)
    P = ctypes.c_void_p * m
    hs = P(*[sc._h.value for sc in scratches])
    ins = P(*[x.data_ptr() for x in xs])
    ous = P(*[o.data_ptr() for o in outs])
    t = KernelTimes()
    _check(lib().apsep_time_kernels_rotating(hs, ins, ous, m, n, int(incl), reps,
                                             stream.cuda_stream, ctypes.byref(t)),
           "apsep_time_kernels_rotating")
    return t


def launch(algo, d_in: int, d_out: int, n: int, incl: bool = False):
    """Synchronous solve with internal scratch (launchWSTL / launchSPT)."""
    _check(lib().apsep_launch(_algo(algo), d_in, d_out, n, int(incl)), "apsep_launch")


def solve(x, algo="auto", incl: bool = False, out=None):
    """Convenience: allocate ``out`` (int32, like ``x``) and solve on torch's
    current stream."""
    import torch
    if out is None:
        out = torch.empty_like(x, dtype=torch.int32)
    with Scratch(algo, x.numel(), incl) as s:
        s.run(x, out)
        torch.cuda.current_stream(x.device).synchronize()
        s.check()
    return out
