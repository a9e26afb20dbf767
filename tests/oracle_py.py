"""Python-side helpers around the oracle's C restatement (test infrastructure only)."""
import ctypes

import numpy as np


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class NpbClass(ctypes.Structure):
    _fields_ = [("total_keys_log2", ctypes.c_int), ("max_key_log2", ctypes.c_int),
                ("num_buckets_log2", ctypes.c_int), ("cls", ctypes.c_char)]


def npb_class(orc, cls):
    c = NpbClass()
    assert orc.orc_npb_class_params(ctypes.c_char(cls.encode()), ctypes.byref(c)) == 0
    return c


def alltoallv(orc, sendbufs, counts, sdispls, elem, recv_totals, rdispls=None):
    """Oracle Alltoallv: returns the list of receive buffers (numpy uint8 views)."""
    P = len(sendbufs)
    rcounts = np.ascontiguousarray(counts.T).astype(np.int32)
    if rdispls is None:
        rdispls = np.zeros_like(rcounts)
        rdispls[:, 1:] = np.cumsum(rcounts, axis=1)[:, :-1]
    recv = [np.zeros(max(1, int(n)) * elem, dtype=np.uint8) for n in recv_totals]
    sb = (ctypes.c_void_p * P)(*[b.ctypes.data for b in sendbufs])
    rb = (ctypes.c_void_p * P)(*[b.ctypes.data for b in recv])
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    sdispls = np.ascontiguousarray(sdispls, dtype=np.int32)
    orc.orc_alltoallv(P, elem, elem, sb, ptr(counts), ptr(sdispls), rb, ptr(rcounts), ptr(rdispls))
    return recv, rcounts, rdispls


def reduce(orc, datatype, op, inputs):
    """Oracle reduction in the reference's fold order over a list of numpy arrays."""
    P = len(inputs)
    out = np.zeros_like(inputs[0])
    ins = (ctypes.c_void_p * P)(*[x.ctypes.data for x in inputs])
    count = inputs[0].size if inputs[0].dtype.kind != "c" else inputs[0].size
    rc = orc.orc_reduce(P, datatype, op, ins, ctypes.c_int64(count), ptr(out))
    return rc, out


def fill_int(orc, seed, rank, n, it=0):
    out = np.zeros(n, dtype=np.int32)
    orc.orc_fill_int(ctypes.c_uint64(seed), rank, ctypes.c_uint32(it), ptr(out), ctypes.c_int64(n))
    return out


def fill_normal(orc, seed, rank, n, it=0):
    out = np.zeros(n, dtype=np.float64)
    orc.orc_fill_normal(ctypes.c_uint64(seed), rank, ctypes.c_uint32(it), ptr(out), ctypes.c_int64(n))
    return out


def fill_bytes(orc, seed, rank, n, it=0):
    out = np.zeros(max(n, 1), dtype=np.uint8)
    orc.orc_fill_bytes(ctypes.c_uint64(seed), rank, ctypes.c_uint32(it), ptr(out), ctypes.c_int64(n))
    return out[:n]


def keys_uniform(orc, seed, rank, n, key_bits=23, step=0):
    out = np.zeros(n, dtype=np.uint32)
    orc.orc_fill_keys_uniform(ctypes.c_uint64(seed), rank, ctypes.c_uint32(step), ptr(out),
                              ctypes.c_int64(n), key_bits)
    return out


def keys_zipf(orc, seed, rank, n, cdf, step=0):
    out = np.zeros(n, dtype=np.uint32)
    orc.orc_fill_keys_zipf(ctypes.c_uint64(seed), rank, ctypes.c_uint32(step), ptr(cdf),
                           ctypes.c_int64(cdf.size), ptr(out), ctypes.c_int64(n))
    return out


def zipf_cdf(nkeys, s=1.1):
    """CDF table of Zipf(s) over key ranks 1..nkeys (key = rank-1), sequential sums."""
    w = np.arange(1, nkeys + 1, dtype=np.float64) ** (-s)
    c = np.cumsum(w)
    return c / c[-1]


def bucket_plan(orc, keys, key_bits, bucket_bits):
    P, n = keys.shape
    kb1 = np.zeros_like(keys)
    sc = np.zeros((P, P), dtype=np.int32)
    sd = np.zeros((P, P), dtype=np.int32)
    lo = np.zeros(P, dtype=np.int32)
    hi = np.zeros(P, dtype=np.int32)
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    orc.orc_bucket_plan(P, ctypes.c_int64(n), key_bits, bucket_bits, ptr(keys), ptr(kb1), ptr(sc),
                        ptr(sd), ptr(lo), ptr(hi))
    return kb1, sc, sd, lo, hi


def job_dot(orc, seed, n, P):
    """Dot-product job restated: lane-strided partial sums (64 lanes), butterfly over
    lanes, then MPI_Allreduce SUM in rank order (out = in_0; out = in_r + out)."""
    x = fill_normal(orc, seed, 0, n)
    y = fill_normal(orc, seed, 1, n)
    vals = []
    for r in range(P):
        lo, hi = n * r // P, n * (r + 1) // P
        p = x[lo:hi] * y[lo:hi]
        pad = (-p.size) % 64
        p = np.concatenate([p, np.zeros(pad)]).reshape(-1, 64)
        acc = np.zeros(64)
        for row in p:
            acc = acc + row
        lanes = np.arange(64)
        for o in (32, 16, 8, 4, 2, 1):
            acc = acc + acc[lanes ^ o]
        vals.append(acc[0])
    out = vals[0]
    for v in vals[1:]:
        out = v + out
    return float(out)


def job_alltoall_checksum(orc, seed, m, P):
    """Alltoall job restated: rank d receives block d of every rank's seeded payload; the
    checksum is sum_i w_i (2i+1) mod 2^64 over 8-byte words, summed over ranks."""
    send = [fill_bytes(orc, seed, s, P * m) for s in range(P)]
    total = np.uint64(0)
    for d in range(P):
        rb = np.concatenate([send[s][d * m:(d + 1) * m] for s in range(P)])
        pad = (-rb.size) % 8
        w = np.concatenate([rb, np.zeros(pad, dtype=np.uint8)]).view(np.uint64)
        k = (2 * np.arange(w.size, dtype=np.uint64) + np.uint64(1))
        with np.errstate(over="ignore"):
            total = total + np.sum(w * k, dtype=np.uint64)
    return int(total)


PART_FIELDS = 8


def particles_init(orc, dims, n0, seed, cap):
    """Oracle initial particle state: ([P, cap, 8] float64, counts[P])."""
    gx, gy, gz = dims
    P = gx * gy * gz
    parts = np.zeros((P, cap, PART_FIELDS), dtype=np.float64)
    counts = np.zeros(P, dtype=np.int64)
    orc.orc_particles_init(gx, gy, gz, ctypes.c_int64(n0), ctypes.c_uint64(seed),
                           ctypes.c_int64(cap), ptr(parts), ptr(counts))
    return parts, counts


def particles_step(orc, dims, seed, step, parts, counts, allow_overflow=False):
    """One oracle exchange step; returns (parts, counts, migrants), migrants = -1 when a
    rank exceeded its capacity (only with allow_overflow)."""
    gx, gy, gz = dims
    cap = parts.shape[1]
    out = np.zeros_like(parts)
    cout = np.zeros_like(counts)
    orc.orc_particles_step.restype = ctypes.c_int64
    moved = orc.orc_particles_step(gx, gy, gz, ctypes.c_uint64(seed), ctypes.c_uint32(step),
                                   ctypes.c_int64(cap), ptr(parts), ptr(counts), ptr(out), ptr(cout))
    assert moved >= 0 or allow_overflow, "a rank exceeded its particle capacity"
    return out, cout, moved
