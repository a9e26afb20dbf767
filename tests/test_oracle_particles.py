"""Particle-exchange oracle: invariants that pin it, and host MPICH bit-for-bit."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle_py import particles_init, particles_step

SEED = 0x5EED
MPIRUN = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
MPICC = shutil.which("mpicc") or "/opt/conda/bin/mpicc"
BIN = os.path.join(ROOT, "oracle", "_build")


def _cells(parts, n, dims):
    gx, gy, gz = dims
    q = parts[:n]
    cx, cy, cz = np.floor(q[:, 0]), np.floor(q[:, 1]), np.floor(q[:, 2])
    return (cx + gx * (cy + gy * cz)).astype(np.int64)


@pytest.mark.parametrize("dims", [(4, 4, 4), (2, 1, 3)])
def test_particle_step_invariants(oracle, dims):
    gx, gy, gz = dims
    P, n0, cap, steps = gx * gy * gz, 1500, 4096, 3
    parts, counts = particles_init(oracle, dims, n0, SEED, cap)
    total_moved = 0
    for s in range(steps):
        before = {}
        for r in range(P):
            for q in parts[r, :counts[r]]:
                before[int(q[7])] = q.copy()
        parts, counts, moved = particles_step(oracle, dims, SEED, s, parts, counts)
        total_moved += moved
        assert counts.sum() == P * n0
        ids = np.concatenate([parts[r, :counts[r], 7] for r in range(P)]).astype(np.int64)
        assert np.array_equal(np.sort(ids), np.arange(P * n0))
        nmoved = 0
        for r in range(P):
            n = int(counts[r])
            # every particle lies in its owner's cell
            assert np.all(_cells(parts[r], n, dims) == r)
            for q in parts[r, :n]:
                old = before[int(q[7])]
                assert np.array_equal(q[3:], old[3:])
                d = q[:3] - old[:3]
                changed = np.nonzero(d)[0]
                assert len(changed) <= 1
                if len(changed) == 1:
                    nmoved += 1
                    ax = changed[0]
                    g = dims[ax]
                    assert (d[ax] % g) in (1.0, g - 1.0)
                # a particle that moved along an axis of size 1 keeps its coordinates
                own = oracle.orc_particle_dir(ctypes.c_uint64(SEED), ctypes.c_uint32(s),
                                              ctypes.c_uint64(int(q[7])))
                if own >= 0 and dims[own >> 1] > 1:
                    assert len(changed) == 1 and changed[0] == own >> 1
                    sign = 1.0 if own % 2 == 0 else -1.0
                    assert (q[own >> 1] - old[own >> 1] - sign) % dims[own >> 1] == 0
                if own < 0:
                    assert len(changed) == 0
        assert nmoved <= moved
    p =214748365 / 2.0 ** 32
    N = P * n0 * steps
    sigma = np.sqrt(N * p * (1 - p))
    assert abs(total_moved - N * p) < 6 * sigma


def test_particle_stayers_keep_order(oracle):
    dims = (2, 2, 2)
    parts, counts = particles_init(oracle, dims, 800, SEED, 2048)
    out, cout, _ = particles_step(oracle, dims, SEED, 0, parts, counts)
    for r in range(8):
        stay = [q[7] for q in parts[r, :counts[r]]
                if oracle.orc_particle_dir(ctypes.c_uint64(SEED), ctypes.c_uint32(0),
                                           ctypes.c_uint64(int(q[7]))) < 0]
        assert np.array_equal(out[r, :len(stay), 7], np.array(stay))
        # arrivals follow, grouped by ascending source rank
        src = [int(i) // 800 for i in out[r, len(stay):cout[r], 7]]
        assert src == sorted(src)


@pytest.fixture(scope="module")
def mpi_particles():
    if not (os.path.exists(MPIRUN) and os.path.exists(MPICC)):
        pytest.skip("host MPI (mpicc/mpirun) not available")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mpi"], check=True)
    return os.path.join(BIN, "particles")


@pytest.mark.parametrize("dims", [(2, 2, 2), (1, 2, 3)])
def test_particles_match_host_mpich(oracle, mpi_particles, dims, tmp_path):
    gx, gy, gz = dims
    P, n0, steps = gx * gy * gz, 700, 4
    prefix = str(tmp_path / "state")
    res = subprocess.run([MPIRUN, "-n", str(P), mpi_particles, str(gx), str(gy), str(gz), str(n0),
                          hex(SEED), str(steps), prefix], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    parts, counts = particles_init(oracle, dims, n0, SEED, 4 * n0)
    moved = 0
    for s in range(steps):
        parts, counts, m = particles_step(oracle, dims, SEED, s, parts, counts)
        moved += m
    assert f"moved={moved} " in res.stdout
    for r in range(P):
        raw = open(f"{prefix}.{r}", "rb").read()
        n = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
        got = np.frombuffer(raw[8:], dtype=np.float64).reshape(n, 8)
        assert n == counts[r]
        assert got.tobytes() == parts[r, :n].tobytes()
