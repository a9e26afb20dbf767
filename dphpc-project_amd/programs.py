"""ctypes mirrors of the argument blocks of the built-in rank programs
(csrc/programs/*.h).  Field order and types match the C structs exactly."""
import ctypes

_P = ctypes.c_void_p


class OkArgs(ctypes.Structure):
    _fields_ = [("ok", _P)]


class McpiArgs(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("n", ctypes.c_int64), ("per_rank", _P),
                ("total", _P), ("seconds", _P)]


class P2pArgs(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("mode", ctypes.c_int32), ("iters", ctypes.c_int32),
                ("warmup", ctypes.c_int32), ("pad", ctypes.c_int32), ("bytes", ctypes.c_uint64),
                ("sendbase", _P), ("recvbase", _P), ("seconds", _P), ("mismatches", _P)]


class CollArgs(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("kind", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("op", ctypes.c_int32), ("root", ctypes.c_int32), ("count", ctypes.c_int64),
                ("iters", ctypes.c_int32), ("warmup", ctypes.c_int32), ("inbase", _P),
                ("outbase", _P), ("in_stride", ctypes.c_int64), ("out_stride", ctypes.c_int64),
                ("seconds", _P), ("no_fill", ctypes.c_int32), ("pad", ctypes.c_int32)]


class ChurnArgs(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("ops", ctypes.c_int64), ("coll_every", ctypes.c_int32),
                ("dump", ctypes.c_int32), ("errors", _P), ("peak_live", _P), ("live_ptrs", _P),
                ("live_sizes", _P), ("seconds", _P), ("failed", _P)]


class ExchArgs(ctypes.Structure):
    _fields_ = [("phase", ctypes.c_int32), ("elem", ctypes.c_int32), ("sendbufs", _P),
                ("scounts", _P), ("sdispls", _P), ("recvbufs", _P), ("rcounts", _P),
                ("rdispls", _P)]


class ParticleArgs(ctypes.Structure):
    """Particle exchange on a periodic gx*gy*gz rank grid (8 FP64 fields per particle);
    mode 0 = Alltoall(counts) + Alltoallv, mode 1 = neighbour Isend/Irecv."""
    _fields_ = [("seed", ctypes.c_uint64), ("gx", ctypes.c_int32), ("gy", ctypes.c_int32),
                ("gz", ctypes.c_int32), ("mode", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("first_step", ctypes.c_int32), ("n0", ctypes.c_int64), ("cap", ctypes.c_int64),
                ("buf", _P * 2), ("sendbuf", _P), ("counts", _P), ("cnt", _P), ("moved", _P),
                ("seconds", _P)]


class BucketSortArgs(ctypes.Structure):
    """Bucket sort / NPB IS rank program: dist 0 uniform, 1 Zipf (cdf), 2 NPB IS class `cls`."""
    _fields_ = [("dist", ctypes.c_int32), ("cls", ctypes.c_int32), ("key_bits", ctypes.c_int32),
                ("bucket_bits", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("n", ctypes.c_int64), ("seed", ctypes.c_uint64),
                ("cdf", _P), ("ncdf", ctypes.c_int64), ("kb2_cap", ctypes.c_int64),
                ("count_cap", ctypes.c_int64), ("key_array", _P), ("key_buff1", _P),
                ("key_buff2", _P), ("counting", _P), ("bsize", _P), ("btot", _P), ("cnt", _P),
                ("nlocal", _P), ("passed", _P), ("total", _P), ("seconds", _P)]


class OpCase(ctypes.Structure):
    """One (datatype, operator) case of t_op_table: op is the built-in operator index
    (MPI_MAX = 0 .. MPI_MINLOC = 11), -1 the int64 user operator, -2 the double one."""
    _fields_ = [("dtype", ctypes.c_int32), ("op", ctypes.c_int32)]


class OpTableArgs(ctypes.Structure):
    _fields_ = [("cases", _P), ("ncases", ctypes.c_int32), ("count", ctypes.c_int32),
                ("root", ctypes.c_int32), ("stride", ctypes.c_int32), ("inp", _P),
                ("out_all", _P), ("out_root", _P), ("status", _P)]


class LdsArgs(ctypes.Structure):
    _fields_ = [("ok", _P)]


COLL_KINDS = {"bcast": 0, "reduce": 1, "allreduce": 2, "alltoall": 3, "allgather": 4,
              "gather": 5, "scatter": 6}

REFERENCE_TESTS = {
    # program: ranks (as the reference's TestRunner(n) uses)
    "t_single_int": 2, "t_array": 2, "t_sendrecv": 2, "t_repeat": 2, "t_flood": 3,
    "t_p2p_all_to_all": 10, "t_reduce_prod_int16": 5, "t_allreduce_sum_double": 7,
    "t_allgather": 7, "t_scatter": 5, "t_scatterv": 5, "t_comm_rank": 5, "t_comm_dup": 2,
    "t_comm_split": 10,
}
