"""Pins of oracle exports beyond the known-answer tests in test_oracle.py:

  orc_local_sort        against numpy.sort;
  orc_counts_transpose  against a transpose and an exclusive prefix sum;
  orc_npb_is_iteration  against NPB's partial-verification known answers: consecutive
                        rank() iterations 1..10 on one key array each pass all five test
                        keys (is.c:849-905), 5 * 10 + P with the full verification
                        (orc_npb_is_run);
  orc_alltoallv         against MPICH's MPI_Alltoallv byte for byte (oracle/mpi/a2av_check.c);
  orc_reduce            against MPICH's MPI_Reduce/MPI_Allreduce: integers and FP64 MAX
                        equal, FP64 SUM within 1e-12 * sum |x| per element
                        (oracle/mpi/reduce_check.c).
Each check is one a plausible slip fails: an unstable or descending sort, a transposed
or inclusive scan, an iteration that forgets the previous iterations' key changes, a
dropped or misplaced Alltoallv block, a dropped term in the fold."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle_py import npb_class, ptr

MPIRUN = shutil.which("mpirun") or ("/opt/conda/bin/mpirun" if os.path.exists("/opt/conda/bin/mpirun") else None)
BUILD = os.path.join(ROOT, "oracle", "_build")


@pytest.mark.parametrize("n", [0, 1, 7, 1000, 65536])
def test_local_sort_is_numpy_sort(n, oracle):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 23, size=n).astype(np.uint32)
    if n > 10:
        keys[::7] = keys[3]  # duplicates
    out = np.zeros(max(n, 1), dtype=np.uint32)
    oracle.orc_local_sort(ctypes.c_int64(n), ptr(keys), ptr(out))
    assert np.array_equal(out[:n], np.sort(keys))


@pytest.mark.parametrize("P", [1, 5, 64])
def test_counts_transpose(P, oracle):
    rng = np.random.default_rng(P)
    sc = rng.integers(0, 1000, size=(P, P)).astype(np.int32)
    rc = np.zeros_like(sc)
    rd = np.zeros_like(sc)
    oracle.orc_counts_transpose(P, ptr(sc), ptr(rc), ptr(rd))
    assert np.array_equal(rc, sc.T)
    exp = np.zeros_like(sc)
    exp[:, 1:] = np.cumsum(sc.T, axis=1)[:, :-1]
    assert np.array_equal(rd, exp)


@pytest.mark.parametrize("cls,P", [("S", 4), ("S", 16), ("W", 8)])
def test_npb_is_iteration_known_answers(cls, P, oracle):
    c = npb_class(oracle, cls)
    nk = (1 << c.total_keys_log2) // P
    keys = np.zeros(P * nk, dtype=np.int32)
    for r in range(P):
        oracle.orc_npb_create_keys(ctypes.byref(c), P, r, ptr(keys[r * nk:]))
    oracle.orc_npb_is_iteration(ctypes.byref(c), P, 1, ptr(keys))  # the untimed rank(1)
    passed = [oracle.orc_npb_is_iteration(ctypes.byref(c), P, it, ptr(keys)) for it in range(1, 11)]
    assert passed == [5] * 10
    assert oracle.orc_npb_is_run(ctypes.byref(c), P) == 5 * 10 + P
    # the iterations depend on the keys the earlier ones changed (is.c:650-654)
    fresh = np.zeros(P * nk, dtype=np.int32)
    for r in range(P):
        oracle.orc_npb_create_keys(ctypes.byref(c), P, r, ptr(fresh[r * nk:]))
    assert oracle.orc_npb_is_iteration(ctypes.byref(c), P, 10, ptr(fresh)) < 5


def _mpirun(n, prog, *args):
    exe = os.path.join(BUILD, prog)
    if not MPIRUN or not os.path.exists(exe):
        pytest.skip("host MPI or the oracle's MPI programs missing")
    env = dict(os.environ, MPICH_CC="gcc")
    r = subprocess.run([MPIRUN, "-n", str(n), exe, *map(str, args)], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith(prog)]
    assert line, r.stdout
    return dict(kv.split("=") for kv in line[0].split()[1:])


@pytest.mark.parametrize("P,maxc,elem", [(8, 300, 4), (13, 50, 1), (6, 2000, 8), (3, 0, 4)])
def test_alltoallv_against_mpich(P, maxc, elem):
    kv = _mpirun(P, "a2av_check", hex(0x5A + P), maxc, elem)
    assert int(kv["ranks"]) == P
    assert int(kv["mismatches"]) == 0


@pytest.mark.parametrize("P,n", [(16, 4097), (7, 1)])
def test_reduce_against_mpich(P, n):
    kv = _mpirun(P, "reduce_check", "0x1234", n)
    assert int(kv["int_mismatch"]) == 0
    assert int(kv["dmax_mismatch"]) == 0
    assert float(kv["dsum_worst"]) <= 1.0
    assert int(kv["allreduce_mismatch"]) == 0
