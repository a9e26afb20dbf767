"""Pins the oracle's C restatement (oracle/gpumpi_oracle.c) against known answers:
Random123 KATs and rocRAND for Philox, NPB IS class S/W/A verification, the reference
tests' expected values, brute force and closed forms.  CPU only."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle_py import (alltoallv, bucket_plan, fill_bytes, fill_int, fill_normal, keys_uniform,
                       keys_zipf, npb_class, ptr, reduce, zipf_cdf)

MPI_SHORT, MPI_INT, MPI_LONG = 1, 2, 3
MPI_FLOAT, MPI_DOUBLE = 12, 13
MPI_C_BOOL, MPI_INT16_T, MPI_INT64_T, MPI_UINT8_T, MPI_BYTE = 16, 18, 20, 21, 29
MPI_C_DOUBLE_COMPLEX = 27
OP_MAX, OP_MIN, OP_SUM, OP_PROD, OP_LAND, OP_BAND, OP_LOR, OP_BOR, OP_LXOR, OP_BXOR, OP_MAXLOC = range(11)


def philox(orc, ctr, key):
    out = (ctypes.c_uint32 * 4)()
    orc.orc_philox4x32_10((ctypes.c_uint32 * 4)(*ctr), (ctypes.c_uint32 * 2)(*key), out)
    return list(out)


# Random123 known-answer vectors for philox4x32 with 10 rounds
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_philox_known_answers(oracle, ctr, key, expected):
    assert philox(oracle, ctr, key) == list(expected)


@pytest.mark.skipif(not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"),
                    reason="rocRAND header not present")
def test_philox_matches_rocrand(oracle, tmp_path):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not present")
    exe = tmp_path / "rr"
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-o", str(exe),
                    os.path.join(ROOT, "tests", "golden", "rocrand_philox.cpp")], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    lines = [l.split() for l in out.strip().splitlines()]
    assert len(lines) == 15
    for seed, i, *v in lines:
        seed, i = int(seed), int(i)
        got = philox(oracle, (i, 0, 0, 0), (seed & 0xffffffff, seed >> 32))
        assert got == [int(x) for x in v]


def test_mcpi_brute_force(oracle):
    """Hit count recomputed from raw Philox outputs in Python (exact FP64 squares)."""
    seed, n = 0x5eed, 3000
    for rank in (0, 5, 63):
        hits = 0
        for i in range(n):
            o = philox(oracle, (i, 0, 0, 0), (seed & 0xffffffff, (seed >> 32) ^ rank))
            x = (((o[0] << 32) | o[1]) >> 11) * 2.0 ** -53
            y = (((o[2] << 32) | o[3]) >> 11) * 2.0 ** -53
            hits += (x * x + y * y) <= 1.0
        assert oracle.orc_mcpi_hits(ctypes.c_uint64(seed), rank, ctypes.c_int64(n)) == hits


def test_mcpi_estimates_pi(oracle):
    n = 1 << 16
    hits = sum(oracle.orc_mcpi_hits(ctypes.c_uint64(0x5eed), r, ctypes.c_int64(n)) for r in range(4))
    est = 4.0 * hits / (4 * n)
    sigma = 4.0 * np.sqrt(np.pi / 4 * (1 - np.pi / 4) / (4 * n))
    assert abs(est - np.pi) < 5 * sigma


def test_generators(oracle):
    x = fill_int(oracle, 1, 3, 100000)
    assert x.min() >= -(1 << 20) and x.max() <= (1 << 20)
    assert abs(x.mean()) < 0.02 * (1 << 20)
    z = fill_normal(oracle, 1, 3, 50000)
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1.0) < 0.03
    assert np.all(np.abs(z) <= 8.58)          # sqrt(-2 ln 2^-53)
    b = fill_bytes(oracle, 9, 2, 37)
    o = philox(oracle, (0, 1, 0, 0x0B17E500), (9, 2))
    assert list(b[16:32]) == [(o[k >> 2] >> (8 * (k & 3))) & 0xFF for k in range(16)]
    k = keys_uniform(oracle, 5, 0, 100000, 23)
    assert k.max() < (1 << 23) and abs(k.mean() / (1 << 22) - 1) < 0.02


def test_zipf_keys(oracle):
    cdf = zipf_cdf(1 << 16)
    k = keys_zipf(oracle, 5, 1, 200000, cdf)
    assert k.max() < (1 << 16)
    h = np.arange(1, (1 << 16) + 1, dtype=np.float64) ** -1.1
    p0 = 1.0 / h.sum()
    assert abs((k == 0).mean() - p0) < 4 * np.sqrt(p0 * (1 - p0) / k.size)
    # inverse-CDF property: P(key <= 9) matches the CDF
    assert abs((k <= 9).mean() - cdf[9]) < 0.005


@pytest.mark.parametrize("cls,P", [("S", 1), ("S", 4), ("S", 32), ("W", 4), ("W", 16), ("A", 8)])
def test_npb_is_verification(oracle, cls, P):
    """NPB IS partial verification known answers (is.c:264-287, 849-905) and full verify:
    passed == 5 * MAX_ITERATIONS + P (is.c:1137-1138)."""
    c = npb_class(oracle, cls)
    assert oracle.orc_npb_is_run(ctypes.byref(c), P) == 5 * 10 + P


def test_npb_keys_rank_independent(oracle):
    """find_my_seed jumps the LCG: the keys of P ranks concatenate to the keys of 1 rank."""
    c = npb_class(oracle, "S")
    n = 1 << 16
    one = np.zeros(n, dtype=np.int32)
    oracle.orc_npb_create_keys(ctypes.byref(c), 1, 0, ptr(one))
    parts = []
    for r in range(8):
        a = np.zeros(n // 8, dtype=np.int32)
        oracle.orc_npb_create_keys(ctypes.byref(c), 8, r, ptr(a))
        parts.append(a)
    assert np.array_equal(np.concatenate(parts), one)
    assert one.max() < (1 << 11) and one.min() >= 0


def test_reduce_reference_known_answers(oracle):
    # tests/collectives_test.cu:6-54: PROD of (r+1)(i+1) over 5 ranks in int16 (wraps)
    ins = [np.array([(r + 1) * (i + 1) for i in range(14)], dtype=np.int16) for r in range(5)]
    rc, out = reduce(oracle, MPI_INT16_T, OP_PROD, ins)
    exp = np.ones(14, dtype=np.int64)
    for r in range(5):
        exp = exp * np.array([(r + 1) * (i + 1) for i in range(14)])
    exp = ((exp + 2 ** 15) % 2 ** 16 - 2 ** 15).astype(np.int16)
    assert rc == 0 and np.array_equal(out, exp)
    # tests/collectives_test.cu:57-98: SUM double over 7 ranks
    ins = [np.array([(r + 1) * (i + 1) for i in range(17)], dtype=np.float64) for r in range(7)]
    rc, out = reduce(oracle, MPI_DOUBLE, OP_SUM, ins)
    assert rc == 0 and np.allclose(out, 28 * np.arange(1, 18), atol=1e-5)


def test_reduce_fold_order(oracle):
    """FP64 SUM follows recv = in_r + recv for r = 1..P-1 (mpi.cu:209-225): a sequential
    left-to-right sum, which differs from a pairwise sum on cancellation-heavy input."""
    vals = [1e16, 1.0, -1e16, 1.0]
    ins = [np.array([v]) for v in vals]
    _, out = reduce(oracle, MPI_DOUBLE, OP_SUM, ins)
    acc = 0.0
    for v in vals:
        acc = v + acc
    assert out[0] == acc == 1.0


def test_reduce_ops_and_legality(oracle):
    rng = np.random.default_rng(0)
    ins = [rng.integers(-5, 6, size=32).astype(np.int32) for _ in range(6)]
    for op, f in [(OP_MAX, np.maximum), (OP_MIN, np.minimum), (OP_SUM, np.add),
                  (OP_BAND, np.bitwise_and), (OP_BOR, np.bitwise_or), (OP_BXOR, np.bitwise_xor)]:
        _, out = reduce(oracle, MPI_INT, op, ins)
        exp = ins[0].copy()
        for x in ins[1:]:
            exp = f(x, exp)
        assert np.array_equal(out, exp), op
    assert oracle.orc_op_allowed(MPI_DOUBLE, OP_BAND) == 0
    assert oracle.orc_op_allowed(MPI_BYTE, OP_BXOR) == 1
    assert oracle.orc_op_allowed(MPI_BYTE, OP_SUM) == 0
    assert oracle.orc_op_allowed(MPI_C_BOOL, OP_LAND) == 1
    assert oracle.orc_op_allowed(MPI_INT, OP_MAXLOC) == 0
    assert oracle.orc_op_allowed(MPI_C_DOUBLE_COMPLEX, OP_PROD) == 1
    rc, _ = reduce(oracle, MPI_DOUBLE, OP_BAND, [np.zeros(2)] * 3)
    assert rc == -1


def test_reduce_complex_product(oracle):
    a = np.array([1 + 2j, -3 + 0.5j], dtype=np.complex128)
    b = np.array([2 - 1j, 4 + 4j], dtype=np.complex128)
    _, out = reduce(oracle, MPI_C_DOUBLE_COMPLEX, OP_PROD, [a, b])
    assert np.array_equal(out, b * a)


def test_comm_split_reference_table(oracle):
    """tests/communicator_test.cu:93-152."""
    P = 10
    color = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2, 2], dtype=np.int32)
    key = np.array([0, 1, 2, 3, 2, 1, 1, 1, 1, 0], dtype=np.int32)
    nr = np.zeros(P, dtype=np.int32)
    ns = np.zeros(P, dtype=np.int32)
    oracle.orc_comm_split(P, ptr(color), ptr(key), ptr(nr), ptr(ns))
    exp_r = [0, 1, 2, 2, 1, 0, 1, 2, 3, 0]
    exp_s = [3, 3, 3, 3, 3, 3, 4, 4, 4, 4]
    assert nr.tolist() == exp_r and ns.tolist() == exp_s


def test_alltoallv_is_the_block_permutation(oracle):
    """Independent pure-Python restatement of recv_d[rdispls[d][s]:] = send_s[sdispls[s][d]:]."""
    rng = np.random.default_rng(1)
    P, elem = 6, 4
    counts = rng.integers(0, 9, size=(P, P)).astype(np.int32)
    sd = np.zeros_like(counts)
    sd[:, 1:] = np.cumsum(counts, axis=1)[:, :-1]
    send = [rng.integers(0, 256, size=max(1, counts[s].sum()) * elem, dtype=np.uint8) for s in range(P)]
    got, rc, rd = alltoallv(oracle, send, counts, sd, elem, counts.sum(axis=0))
    for d in range(P):
        exp = b"".join(bytes(send[s][sd[s, d] * elem:(sd[s, d] + counts[s, d]) * elem]) for s in range(P))
        assert bytes(got[d][:len(exp)]) == exp
    assert np.array_equal(rc, counts.T)


@pytest.mark.parametrize("dist", ["uniform", "zipf"])
def test_bucket_plan_properties(oracle, dist):
    P, n, kb, bb = 16, 2048, 16, 8
    if dist == "uniform":
        keys = np.stack([keys_uniform(oracle, 3, r, n, kb) for r in range(P)])
    else:
        cdf = zipf_cdf(1 << kb)
        keys = np.stack([keys_zipf(oracle, 3, r, n, cdf) for r in range(P)])
    kb1, sc, sd, lo, hi = bucket_plan(oracle, keys, kb, bb)
    shift = kb - bb
    assert np.array_equal(sc.sum(axis=1), np.full(P, n))
    for s in range(P):
        # stable bucketing: bucket order, original order within a bucket
        b = kb1[s] >> shift
        assert np.all(np.diff(b.astype(np.int64)) >= 0)
        order = np.argsort(keys[s] >> shift, kind="stable")
        assert np.array_equal(kb1[s], keys[s][order])
    got, _, _ = alltoallv(oracle, [kb1[s].view(np.uint8) for s in range(P)], sc, sd, 4, sc.sum(axis=0))
    allk = []
    for d in range(P):
        k = got[d][: sc[:, d].sum() * 4].view(np.uint32)
        if k.size:
            assert (k >> shift).min() >= lo[d] and (k >> shift).max() <= hi[d]
        allk.append(np.sort(k))
    assert np.array_equal(np.concatenate(allk), np.sort(keys.ravel()))
    # NPB splitter rule (is.c:721-751): rank j's range closes at the first bucket where the
    # global running total reaches (j+1)*n
    tot = np.bincount((keys.ravel() >> shift), minlength=1 << bb).cumsum()
    j = 0
    for i in range(1 << bb):
        if tot[i] >= (j + 1) * n:
            assert hi[j] == i
            j += 1


def test_heap_checker(oracle):
    p = np.array([0x1000, 0x1010, 0x1100], dtype=np.uint64)
    s = np.array([16, 32, 8], dtype=np.uint64)
    assert oracle.orc_heap_check(ptr(p), ptr(s), 3, ctypes.c_uint64(0x1000), ctypes.c_uint64(0x1000)) == 0
    s2 = np.array([32, 32, 8], dtype=np.uint64)   # first block overlaps the second
    assert oracle.orc_heap_check(ptr(p), ptr(s2), 3, ctypes.c_uint64(0x1000), ctypes.c_uint64(0x1000)) == 2
    p3 = np.array([0x1008], dtype=np.uint64)       # misaligned 16-byte block
    s3 = np.array([16], dtype=np.uint64)
    assert oracle.orc_heap_check(ptr(p3), ptr(s3), 1, ctypes.c_uint64(0x1000), ctypes.c_uint64(0x1000)) == 1


def _det(oracle, name):
    f = getattr(oracle, name)
    f.restype, f.argtypes = ctypes.c_double, [ctypes.c_double]
    return f


def test_generator_math_against_libm(oracle):
    """The N(0,1) generator's own ln, sqrt and cos(2 pi u) (+, -, *, / only, so the device
    computes the same bits) against libm: ln within 3 ulps on (0, 1] and at its reduction
    edges, sqrt within 1 ulp on [0, 75], cos(2 pi u) within 1e-15 on [0, 1) including the
    quadrant edges."""
    ln, sq, cs = (_det(oracle, n) for n in ("orc_det_log", "orc_det_sqrt", "orc_det_cos_2pi"))
    rng = np.random.default_rng(3)
    u = np.concatenate([rng.random(20000), 2.0 ** -rng.uniform(0, 53, 5000),
                        [2.0 ** -53, 0.5, 0.7071067811865475, 0.7071067811865476, 1.0 - 2.0 ** -53]])
    got = np.array([ln(x) for x in u])
    ref = np.log(u)
    assert np.all(np.abs(got - ref) <= 3 * np.spacing(np.abs(ref)))
    assert ln(1.0) == 0.0
    x = np.concatenate([rng.uniform(0, 75, 20000), [1e-300, 2.0, 4.0, 73.5]])
    got = np.array([sq(v) for v in x])
    assert np.all(np.abs(got - np.sqrt(x)) <= np.spacing(np.sqrt(x)))
    assert sq(0.0) == 0.0 and sq(4.0) == 2.0
    u2 = np.concatenate([rng.random(20000), [0.0, 0.25, 0.5, 0.75, 0.125, 1.0 - 2.0 ** -53]])
    got = np.array([cs(v) for v in u2])
    assert np.all(np.abs(got - np.cos(2 * np.pi * u2)) <= 1e-15)
    assert cs(0.0) == 1.0


def test_normal_is_box_muller_on_philox(oracle):
    """Element i is sqrt(-2 ln u1) cos(2 pi u2) of Philox block i of the normal stream
    (counter (iter, i, i >> 32, stream), key (seed, seed >> 32 ^ rank)), u1 = words 0-1 as
    53 bits plus one ulp (never 0), u2 = words 2-3: recomputed here with numpy's libm (a
    swapped word pair, a sine for the cosine or a dropped factor fails), and the whole
    stream is N(0,1) by Kolmogorov-Smirnov."""
    from scipy import stats
    seed, rank, it = 0x1234, 5, 2
    z = fill_normal(oracle, seed, rank, 4096, it)
    for i in list(range(32)) + [4095]:
        o = philox(oracle, (it, i, 0, 0x0B17E502), (seed & 0xFFFFFFFF, (seed >> 32) ^ rank))
        u1 = (((o[0] << 32 | o[1]) >> 11) + 1) * 2.0 ** -53
        u2 = ((o[2] << 32 | o[3]) >> 11) * 2.0 ** -53
        ref = np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)
        assert abs(z[i] - ref) <= 4e-15 * max(1.0, abs(ref)), i
    big = fill_normal(oracle, 7, 0, 200000)
    assert stats.kstest(big, "norm").pvalue > 1e-3
    assert abs(np.mean(big ** 4) - 3.0) < 0.06      # normal kurtosis (Irwin-Hall(12): 2.9)
