"""gpumpi for MI355X: Python bindings of libgpumpi.so (ctypes over the C ABI of
include/gpumpi/gpumpi.h).

The product path is the HIP library; this module only marshals arguments.  Loading
fails loudly when libgpumpi.so is missing -- there is no CPU fallback.

Device buffers are passed as integer device addresses (for example
``torch.Tensor.data_ptr()``) and streams as integer hipStream_t handles.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpumpi.so")

_lib = None

# gpumpi_status
OK, E_ARG, E_HIP, E_RESIDENCY, E_PROGRAM, E_DEVICE, E_NOMEM = 0, -1, -2, -3, -4, -5, -6

# MPI_Datatype enum order (include/gpumpi/mpi.h)
DATATYPES = [
    "MPI_CHAR", "MPI_SHORT", "MPI_INT", "MPI_LONG", "MPI_LONG_LONG_INT", "MPI_LONG_LONG",
    "MPI_SIGNED_CHAR", "MPI_UNSIGNED_CHAR", "MPI_UNSIGNED_SHORT", "MPI_UNSIGNED",
    "MPI_UNSIGNED_LONG", "MPI_UNSIGNED_LONG_LONG", "MPI_FLOAT", "MPI_DOUBLE",
    "MPI_LONG_DOUBLE", "MPI_WCHAR", "MPI_C_BOOL", "MPI_INT8_T", "MPI_INT16_T", "MPI_INT32_T",
    "MPI_INT64_T", "MPI_UINT8_T", "MPI_UINT16_T", "MPI_UINT32_T", "MPI_UINT64_T",
    "MPI_C_COMPLEX", "MPI_C_FLOAT_COMPLEX", "MPI_C_DOUBLE_COMPLEX",
    "MPI_C_LONG_DOUBLE_COMPLEX", "MPI_BYTE",
]
DT = {n: i for i, n in enumerate(DATATYPES)}


class GpumpiError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("num_ranks", ctypes.c_int32), ("ranks_per_block", ctypes.c_int32),
        ("ring_slots", ctypes.c_int32), ("max_requests", ctypes.c_int32),
        ("max_unexpected", ctypes.c_int32), ("bcast_push_min", ctypes.c_int32),
        ("copy_stage_kib", ctypes.c_int32), ("reserved1", ctypes.c_int32),
        ("heap_bytes", ctypes.c_uint64), ("stack_bytes", ctypes.c_uint64),
        ("timeout_s", ctypes.c_double),
    ]


class Job(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("seed", ctypes.c_uint64), ("type", ctypes.c_int32),
                ("ranks", ctypes.c_int32), ("size", ctypes.c_int64)]


class JobResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("ranks", ctypes.c_int32),
                ("dot", ctypes.c_double), ("checksum", ctypes.c_uint64),
                ("t_admit_ticks", ctypes.c_uint64), ("t_done_ticks", ctypes.c_uint64),
                ("bytes", ctypes.c_int64)]


class A2avTune(ctypes.Structure):
    _fields_ = [
        ("block_threads", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
        ("lines_per_lane", ctypes.c_int32), ("nontemporal", ctypes.c_int32),
        ("chunk_rounds", ctypes.c_int32), ("defer_straddle", ctypes.c_int32),
        ("pipeline", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# Exported C-ABI symbols and their ctypes signatures (kept in the order of gpumpi.h).
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_SIGNATURES = {
    "gpumpi_config_default": (None, [ctypes.POINTER(Config)]),
    "gpumpi_runtime_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(_P)]),
    "gpumpi_runtime_destroy": (None, [_P]),
    "gpumpi_run": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                  _P, ctypes.c_size_t, _P, ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_run_async": (ctypes.c_int, [_P, ctypes.c_char_p, _P, ctypes.c_size_t, _P]),
    "gpumpi_launch": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                     _P, ctypes.c_size_t, ctypes.POINTER(Config)]),
    "gpumpi_main": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
    "gpumpi_run_kernel": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                         _P, ctypes.c_size_t, _P, ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_main_kernel": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
    "gpumpi_layout_stamp": (ctypes.c_uint64, []),
    "gpumpi_last_error": (ctypes.c_char_p, [_P]),
    "gpumpi_runtime_copy_stage": (ctypes.c_int, [_P]),
    "gpumpi_last_device_error": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "gpumpi_program_count": (ctypes.c_int, []),
    "gpumpi_program_name": (ctypes.c_char_p, [ctypes.c_int]),
    "gpumpi_heap_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "gpumpi_device_info": (ctypes.c_int, [ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "gpumpi_a2av_plan_create": (ctypes.c_int, [_I32, _I32, _P, _P, _P, _P, _P, _P,
                                               ctypes.POINTER(A2avTune), ctypes.POINTER(_P)]),
    "gpumpi_a2av_plan_destroy": (None, [_P]),
    "gpumpi_a2av_execute": (ctypes.c_int, [_P, _P]),
    "gpumpi_a2av_graph_replay": (ctypes.c_int, [_P, _I32, _P]),
    "gpumpi_a2av_autotune": (ctypes.c_int, [_P, _P, ctypes.POINTER(A2avTune), ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_a2av_phase_times": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_a2av_pipeline_times": (ctypes.c_int, [_P, _I32, _P, ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_a2av_execute_noalloc": (ctypes.c_int, [_I32, _I32, _P, _P, _P, _P, _P, _P, _P]),
    "gpumpi_reduce_plan_create": (ctypes.c_int, [_I32, ctypes.c_int64, _I32, _I32, _I32, _P, _P,
                                                 ctypes.POINTER(_P)]),
    "gpumpi_reduce_plan_destroy": (None, [_P]),
    "gpumpi_reduce_execute": (ctypes.c_int, [_P, _P]),
    "gpumpi_reduce_graph_replay": (ctypes.c_int, [_P, _I32, _P]),
    "gpumpi_reduce_kernel_time": (ctypes.c_int, [_P, _I32, _P, ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_a2av_get_tune": (ctypes.c_int, [_P, ctypes.POINTER(A2avTune), ctypes.POINTER(_I32)]),
    "gpumpi_a2av_payload_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "gpumpi_particles_plan_create": (ctypes.c_int, [_I32, _I32, _I32, ctypes.c_int64, ctypes.c_uint64,
                                                    _I32, ctypes.POINTER(_P)]),
    "gpumpi_particles_plan_destroy": (None, [_P]),
    "gpumpi_particles_init": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P]),
    "gpumpi_particles_prepare": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
    "gpumpi_particles_step": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "gpumpi_particles_graph_replay": (ctypes.c_int, [_P, _P, _P, _P, _P, _I32, _P]),
    "gpumpi_particles_status": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_I32),
                                               ctypes.POINTER(ctypes.c_uint32)]),
    "gpumpi_particles_phase_times": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_particles_exchange_mode": (ctypes.c_int, [_P, _I32]),
    "gpumpi_p2p_plan_create": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_P)]),
    "gpumpi_p2p_plan_destroy": (None, [_P]),
    "gpumpi_p2p_execute": (ctypes.c_int, [_P, _P, ctypes.c_int64, _P]),
    "gpumpi_p2p_graph_replay": (ctypes.c_int, [_P, _P, ctypes.c_int64, _I32, _P]),
    "gpumpi_p2p_kernel_times": (ctypes.c_int, [_P, _P, ctypes.c_int64, _I32, _P, ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_serve_start": (ctypes.c_int, [ctypes.POINTER(Config), _I32, _I32, ctypes.c_double,
                                          ctypes.POINTER(_P)]),
    "gpumpi_serve_submit": (ctypes.c_int, [_P, ctypes.POINTER(Job)]),
    "gpumpi_serve_poll": (ctypes.c_int, [_P, ctypes.POINTER(JobResult), _I32]),
    "gpumpi_serve_stop": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "gpumpi_run_job": (ctypes.c_int, [_P, ctypes.POINTER(Job), _P, _P]),
    "gpumpi_job_heap_bytes": (ctypes.c_int64, [ctypes.POINTER(Job)]),
    "gpumpi_local_sort": (ctypes.c_int, [_P, ctypes.c_int64, _I32, _P, _P]),
    "gpumpi_npb_is_run": (ctypes.c_int, [ctypes.c_char, _I32, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_double)]),
    "gpumpi_gen_keys": (ctypes.c_int, [_P, _I32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32,
                                       _I32, _I32, _P, ctypes.c_int64, _P]),
    "gpumpi_bucket_plan": (ctypes.c_int, [_P, _I32, ctypes.c_int64, _I32, _I32, _P, _P, _P, _P, _P, _P]),
}


def build(verbose=False):
    from importlib import util
    spec = util.spec_from_file_location("gpumpi_build", os.path.join(HERE, "build.py"))
    mod = util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.build(verbose=verbose)


def load_library():
    """Load libgpumpi.so; raises GpumpiError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GpumpiError(f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                          "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def gen_keys(keys_ptr, P, n, seed, step=0, dist="uniform", key_bits=23, cdf_ptr=0, ncdf=0,
             stream=0):
    """Fill keys[P][n] (device uint32) with the workload's seeded keys."""
    d = {"uniform": 0, "zipf": 1, "npb": 2}[dist]
    rc = load_library().gpumpi_gen_keys(keys_ptr, P, n, seed, step, d, key_bits,
                                        cdf_ptr or None, ncdf, stream or None)
    if rc != OK:
        raise GpumpiError(f"gpumpi_gen_keys failed ({rc})")


def bucket_plan(keys_ptr, P, n, key_bits, bucket_bits, kb1_ptr, sc_ptr, sd_ptr, lo_ptr, hi_ptr,
                stream=0):
    """NPB-style bucket plan on device buffers (see gpumpi_bucket_plan)."""
    rc = load_library().gpumpi_bucket_plan(keys_ptr, P, n, key_bits, bucket_bits, kb1_ptr,
                                           sc_ptr, sd_ptr, lo_ptr, hi_ptr, stream or None)
    if rc != OK:
        raise GpumpiError(f"gpumpi_bucket_plan failed ({rc})")


def local_sort(keys_ptr, n, key_bits, out_ptr, stream=0):
    rc = load_library().gpumpi_local_sort(keys_ptr, n, key_bits, out_ptr, stream or None)
    if rc != OK:
        raise GpumpiError(f"gpumpi_local_sort failed ({rc})")


def npb_is(cls, P):
    """NPB IS through the batched exchange; returns (passed, exchange_ms)."""
    passed = ctypes.c_int64()
    ms = ctypes.c_double()
    rc = load_library().gpumpi_npb_is_run(ctypes.c_char(cls.encode()), P, ctypes.byref(passed),
                                          ctypes.byref(ms))
    if rc != OK:
        raise GpumpiError(f"gpumpi_npb_is_run failed ({rc})")
    return passed.value, ms.value


def exported_symbols():
    return list(_SIGNATURES)


def program_names():
    lib = load_library()
    return [lib.gpumpi_program_name(i).decode() for i in range(lib.gpumpi_program_count())]


def device_info():
    lib = load_library()
    cus, khz, waves = _I32(), _I32(), _I32()
    rc = lib.gpumpi_device_info(ctypes.byref(cus), ctypes.byref(khz), ctypes.byref(waves))
    if rc != OK:
        raise GpumpiError("gpumpi_device_info failed")
    return {"cus": cus.value, "wallclock_khz": khz.value, "max_waves_per_cu": waves.value}


def default_config(**kw):
    lib = load_library()
    c = Config()
    lib.gpumpi_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _argv(args):
    arr = (ctypes.c_char_p * (len(args) + 1))()
    for i, a in enumerate(args):
        arr[i] = a.encode() if isinstance(a, str) else a
    arr[len(args)] = None
    return arr


class Runtime:
    """A rank-grid runtime (gpumpi_runtime_create) running built-in rank programs."""

    def __init__(self, num_ranks, **cfg):
        self.lib = load_library()
        self.cfg = default_config(num_ranks=num_ranks, **cfg)
        h = _P()
        rc = self.lib.gpumpi_runtime_create(ctypes.byref(self.cfg), ctypes.byref(h))
        if rc != OK:
            raise GpumpiError(f"gpumpi_runtime_create failed ({rc}): "
                              f"{self.lib.gpumpi_last_error(None).decode()}")
        self.h = h

    def run(self, program, args=None, argv=(), stream=0, check=True):
        """Run `program` on every rank; args is a ctypes.Structure (the argument block).
        Returns (status, elapsed_ms)."""
        ms = ctypes.c_double(0.0)
        blob = ctypes.byref(args) if args is not None else None
        size = ctypes.sizeof(args) if args is not None else 0
        av = _argv(list(argv)) if argv else None
        rc = self.lib.gpumpi_run(self.h, program.encode(), len(argv), av, blob, size,
                                 ctypes.c_void_p(stream) if stream else None, ctypes.byref(ms))
        if check and rc != 0:
            raise GpumpiError(f"{program}: status {rc}: {self.last_error()}")
        return rc, ms.value

    def run_async(self, program, args=None, stream=0):
        blob = ctypes.byref(args) if args is not None else None
        size = ctypes.sizeof(args) if args is not None else 0
        rc = self.lib.gpumpi_run_async(self.h, program.encode(), blob, size,
                                       ctypes.c_void_p(stream) if stream else None)
        if rc != 0:
            raise GpumpiError(f"{program}: status {rc}: {self.last_error()}")

    def last_error(self):
        return self.lib.gpumpi_last_error(self.h).decode()

    def copy_stage_bytes(self):
        """LDS staging bytes per rank that the last launch gave (gpumpi_runtime_copy_stage)."""
        return int(self.lib.gpumpi_runtime_copy_stage(self.h))

    def device_error(self):
        code, rank = _I32(), _I32()
        self.lib.gpumpi_last_device_error(self.h, ctypes.byref(code), ctypes.byref(rank))
        return code.value, rank.value

    def run_job(self, job, result_ptr, stream=0):
        """Plain path: launch one job as its own rank grid (asynchronous)."""
        rc = self.lib.gpumpi_run_job(self.h, ctypes.byref(job), result_ptr,
                                     ctypes.c_void_p(stream) if stream else None)
        if rc != OK:
            raise GpumpiError(f"gpumpi_run_job failed ({rc}): {self.last_error()}")

    def heap_stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.lib.gpumpi_heap_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"in_use": a.value, "peak": b.value, "capacity": c.value}

    def close(self):
        if self.h:
            self.lib.gpumpi_runtime_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Device-resident MPI job server (gpumpi_serve_*)."""

    def __init__(self, slots, queue_capacity=4096, max_job_ranks=256, heap_fraction=0.9, **cfg):
        self.lib = load_library()
        self.cfg = default_config(num_ranks=slots, **cfg)
        h = _P()
        rc = self.lib.gpumpi_serve_start(ctypes.byref(self.cfg), queue_capacity, max_job_ranks,
                                         heap_fraction, ctypes.byref(h))
        if rc != OK:
            raise GpumpiError(f"gpumpi_serve_start failed ({rc}): "
                              f"{self.lib.gpumpi_last_error(None).decode()}")
        self.h = h
        self._buf = (JobResult * 256)()

    def submit(self, job):
        rc = self.lib.gpumpi_serve_submit(self.h, ctypes.byref(job))
        if rc == E_NOMEM:
            return False
        if rc != OK:
            raise GpumpiError(f"gpumpi_serve_submit failed ({rc})")
        return True

    def poll(self):
        out = []
        while True:
            n = self.lib.gpumpi_serve_poll(self.h, self._buf, len(self._buf))
            out.extend(JobResult.from_buffer_copy(self._buf[i]) for i in range(n))
            if n < len(self._buf):
                return out

    def stop(self):
        if not self.h:
            return None
        stats = (ctypes.c_int64 * 4)()
        rc = self.lib.gpumpi_serve_stop(self.h, stats)
        self.h = None
        if rc != OK:
            raise GpumpiError(f"gpumpi_serve_stop: {rc}: {self.lib.gpumpi_last_error(None).decode()}")
        return {"admitted": stats[0], "rejected": stats[1], "max_concurrent": stats[2],
                "backfilled": stats[3]}

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def job_heap_bytes(job):
    return load_library().gpumpi_job_heap_bytes(ctypes.byref(job))


class ReducePlan:
    """Batched MPI_Reduce (root >= 0) / MPI_Allreduce (root = -1) of one step over P ranks
    (gpumpi_reduce_*): sendbufs_ptr / recvbufs_ptr are device arrays of P pointers."""

    def __init__(self, P, count, datatype, op, root, sendbufs_ptr, recvbufs_ptr):
        self.lib = load_library()
        h = _P()
        rc = self.lib.gpumpi_reduce_plan_create(P, count, datatype, op, root, sendbufs_ptr,
                                                recvbufs_ptr, ctypes.byref(h))
        if rc != OK:
            raise GpumpiError(f"gpumpi_reduce_plan_create failed ({rc}): "
                              f"{self.lib.gpumpi_last_error(None).decode()}")
        self.h = h

    def execute(self, stream=0):
        rc = self.lib.gpumpi_reduce_execute(self.h, ctypes.c_void_p(stream) if stream else None)
        if rc != OK:
            raise GpumpiError(f"gpumpi_reduce_execute failed ({rc})")

    def graph_replay(self, steps, stream=0):
        rc = self.lib.gpumpi_reduce_graph_replay(self.h, steps, ctypes.c_void_p(stream) if stream else None)
        if rc != OK:
            raise GpumpiError(f"gpumpi_reduce_graph_replay failed ({rc})")

    def kernel_time(self, reps, stream=0):
        us = ctypes.c_double()
        rc = self.lib.gpumpi_reduce_kernel_time(self.h, reps, ctypes.c_void_p(stream) if stream else None,
                                                ctypes.byref(us))
        if rc != OK:
            raise GpumpiError(f"gpumpi_reduce_kernel_time failed ({rc})")
        return us.value

    def close(self):
        if self.h:
            self.lib.gpumpi_reduce_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class A2avPlan:
    """Batched Alltoallv step: plan, scan and copy kernels, replayed as one hipGraph
    (gpumpi_a2av_*)."""

    def __init__(self, P, elem_bytes, sendbufs_ptr, scounts_ptr, sdispls_ptr, recvbufs_ptr,
                 rcounts_ptr, rdispls_ptr, tune=None):
        self.lib = load_library()
        t = A2avTune(0, -1, 0, -1, 0, -1, -1) if tune is None else A2avTune(*tune)
        if tune is not None and len(tune) < 7:
            t.pipeline = -1
        h = _P()
        rc = self.lib.gpumpi_a2av_plan_create(P, elem_bytes, sendbufs_ptr, scounts_ptr,
                                              sdispls_ptr, recvbufs_ptr, rcounts_ptr,
                                              rdispls_ptr, ctypes.byref(t), ctypes.byref(h))
        if rc != OK:
            raise GpumpiError(f"gpumpi_a2av_plan_create failed ({rc})")
        self.h = h

    def execute(self, stream=0):
        rc = self.lib.gpumpi_a2av_execute(self.h, ctypes.c_void_p(stream) if stream else None)
        if rc != OK:
            raise GpumpiError(f"gpumpi_a2av_execute failed ({rc})")

    def graph_replay(self, steps, stream=0):
        rc = self.lib.gpumpi_a2av_graph_replay(self.h, steps, ctypes.c_void_p(stream) if stream else None)
        if rc != OK:
            raise GpumpiError(f"gpumpi_a2av_graph_replay failed ({rc})")

    def autotune(self, stream=0):
        t = A2avTune()
        us = ctypes.c_double()
        rc = self.lib.gpumpi_a2av_autotune(self.h, ctypes.c_void_p(stream) if stream else None,
                                           ctypes.byref(t), ctypes.byref(us))
        if rc != OK:
            raise GpumpiError(f"gpumpi_a2av_autotune failed ({rc})")
        return t.as_dict(), us.value

    def tune(self):
        t = A2avTune()
        g = _I32()
        self.lib.gpumpi_a2av_get_tune(self.h, ctypes.byref(t), ctypes.byref(g))
        d = t.as_dict()
        d["grid"] = g.value
        return d

    def phase_times(self, enable=None):
        """enable=True makes execute() time its kernels, False stops it; with no argument
        returns [plan_us, scan_us, copy_us] of the last executed step."""
        if enable is not None:
            self.lib.gpumpi_a2av_phase_times(self.h, 1 if enable else 0, None)
            return None
        out = (ctypes.c_double * 3)()
        rc = self.lib.gpumpi_a2av_phase_times(self.h, -1, out)
        if rc != OK:
            raise GpumpiError("phase times not armed")
        return list(out)

    def pipeline_times(self, steps, stream=0):
        """Kernel durations of `steps` pipelined steps (HIP events): {"plan_scan_us":
        first step's plan + scan, "copy_us": average copy kernel that also plans the next
        step, "step_us": the sequence per step}."""
        out = (ctypes.c_double * 3)()
        rc = self.lib.gpumpi_a2av_pipeline_times(self.h, steps, ctypes.c_void_p(stream) if stream else None, out)
        if rc != OK:
            raise GpumpiError(f"gpumpi_a2av_pipeline_times failed ({rc})")
        return {"plan_scan_us": out[0], "copy_us": out[1], "step_us": out[2]}

    def payload_bytes(self):
        b = ctypes.c_int64()
        rc = self.lib.gpumpi_a2av_payload_bytes(self.h, ctypes.byref(b))
        if rc != OK:
            raise GpumpiError(f"gpumpi_a2av step reported a device error ({rc})")
        return b.value

    def close(self):
        if self.h:
            self.lib.gpumpi_a2av_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticlePlan:
    """Batched particle step on a periodic gx*gy*gz rank grid (gpumpi_particles_*): tally,
    place and scatter kernels per step, replayed as one hipGraph.  States are device
    buffers parts[P][cap][8] (float64) and counts[P] (int64)."""

    def __init__(self, dims, cap, seed, nontemporal=True):
        self.lib = load_library()
        gx, gy, gz = dims
        h = _P()
        rc = self.lib.gpumpi_particles_plan_create(gx, gy, gz, cap, seed, 1 if nontemporal else 0,
                                                   ctypes.byref(h))
        if rc != OK:
            raise GpumpiError(f"gpumpi_particles_plan_create failed ({rc})")
        self.h = h

    @staticmethod
    def _s(stream):
        return ctypes.c_void_p(stream) if stream else None

    def _check(self, rc, what):
        if rc != OK:
            raise GpumpiError(f"gpumpi_particles_{what} failed ({rc})")

    def init(self, parts_ptr, counts_ptr, n0, stream=0):
        self._check(self.lib.gpumpi_particles_init(self.h, parts_ptr, counts_ptr, n0, self._s(stream)), "init")

    def prepare(self, parts_ptr, counts_ptr, step, stream=0):
        self._check(self.lib.gpumpi_particles_prepare(self.h, parts_ptr, counts_ptr, step, self._s(stream)),
                    "prepare")

    def step(self, in_ptr, cin_ptr, out_ptr, cout_ptr, stream=0):
        self._check(self.lib.gpumpi_particles_step(self.h, in_ptr, cin_ptr, out_ptr, cout_ptr,
                                                   self._s(stream)), "step")

    def graph_replay(self, a_ptr, ca_ptr, b_ptr, cb_ptr, steps, stream=0):
        self._check(self.lib.gpumpi_particles_graph_replay(self.h, a_ptr, ca_ptr, b_ptr, cb_ptr, steps,
                                                           self._s(stream)), "graph_replay")

    def status(self):
        """(migrants since prepare, sticky overflow flag, index of the next step)."""
        m, o, s = ctypes.c_int64(), _I32(), ctypes.c_uint32()
        self._check(self.lib.gpumpi_particles_status(self.h, ctypes.byref(m), ctypes.byref(o),
                                                     ctypes.byref(s)), "status")
        return m.value, o.value, s.value

    def phase_times(self, enable=None):
        """enable=True/False arms/disarms event timing of step(); with no argument returns
        [tally_us, place_us, scatter_us, p2p_us] of the last timed step (p2p_us: the batched
        point-to-point step of exchange mode 1, 0 in mode 0)."""
        if enable is not None:
            self.lib.gpumpi_particles_phase_times(self.h, 1 if enable else 0, None)
            return None
        out = (ctypes.c_double * 4)()
        self._check(self.lib.gpumpi_particles_phase_times(self.h, -1, out), "phase_times")
        return list(out)

    def exchange_mode(self, mode):
        """0: the scatter writes final slots; 1: neighbour messages through the batched
        point-to-point step (gpumpi_particles_exchange_mode)."""
        self._check(self.lib.gpumpi_particles_exchange_mode(self.h, int(mode)), "exchange_mode")

    def close(self):
        if self.h:
            self.lib.gpumpi_particles_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class P2pPair(ctypes.Structure):
    """gpumpi_p2p_pair: one matched send/receive of a batched point-to-point step."""
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("bytes", ctypes.c_uint64)]


class P2pPlan:
    """Batched point-to-point step (gpumpi_p2p_*): the pairs of a step are a device array of
    gpumpi_p2p_pair (24 bytes each: src, dst, bytes), read when the step runs."""

    def __init__(self, max_pairs):
        self.lib = load_library()
        h = _P()
        rc = self.lib.gpumpi_p2p_plan_create(int(max_pairs), ctypes.byref(h))
        if rc != OK:
            raise GpumpiError(f"gpumpi_p2p_plan_create failed ({rc})")
        self.h = h

    @staticmethod
    def _s(stream):
        return ctypes.c_void_p(stream) if stream else None

    def _check(self, rc, what):
        if rc != OK:
            raise GpumpiError(f"gpumpi_p2p_{what} failed ({rc})")

    def execute(self, pairs_ptr, npairs, stream=0):
        self._check(self.lib.gpumpi_p2p_execute(self.h, pairs_ptr, int(npairs), self._s(stream)), "execute")

    def graph_replay(self, pairs_ptr, npairs, steps, stream=0):
        self._check(self.lib.gpumpi_p2p_graph_replay(self.h, pairs_ptr, int(npairs), int(steps),
                                                     self._s(stream)), "graph_replay")

    def kernel_times(self, pairs_ptr, npairs, reps=10, stream=0):
        """Average [scan_us, copy_us] over `reps` timed steps."""
        out = (ctypes.c_double * 2)()
        self._check(self.lib.gpumpi_p2p_kernel_times(self.h, pairs_ptr, int(npairs), int(reps),
                                                     self._s(stream), out), "kernel_times")
        return list(out)

    def close(self):
        if self.h:
            self.lib.gpumpi_p2p_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from . import programs  # noqa: E402,F401  (argument blocks of the rank programs)
