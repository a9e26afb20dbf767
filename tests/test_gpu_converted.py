"""C MPI programs converted by converter.py and built by gpumpicc.py (in build()), run
through their own ---gpumpi command line like the reference's toolchain tests
(toolchain_tests/pi/test.sh, global_var/test.sh).  Each program checks itself and exits
with status 0 when every rank's checks hold."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
BUILD = os.path.join(ROOT, "tests", "programs", "_build")


def run(name, args, timeout=120, stdin_text=None):
    exe = os.path.join(BUILD, name)
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} was not built: run __graft_entry__.build()")
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=timeout,
                          input=stdin_text)


@pytest.mark.parametrize("ranks", [1, 7, 64])
def test_quadrature(ranks):
    """Bcast of a double array, Allreduce and Gather in a converted C program: the total is
    pi within 1e-10, and the gathered blocks added in rank order give the reduced total's
    exact bits (MPI_Allreduce folds in rank order here)."""
    print(f"[gpu] converted quad on {ranks} ranks", flush=True)
    r = run("quad", ["2000", "---gpumpi", "-n", str(ranks), "-p", str(64 << 20), "-t", "60"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"quad: {ranks} blocks x 2000 panels, integral 3.14159265358" in r.stdout
    assert ", ok" in r.stdout
    line = [l for l in r.stdout.splitlines() if l.startswith("quad: reduced")][0].split()
    assert line[2] == line[4], line


@pytest.mark.parametrize("ranks", [1, 16, 64])
def test_globals_and_libc(ranks):
    print(f"[gpu] converted globals on {ranks} ranks", flush=True)
    r = run("globals", ["---gpumpi", "-n", str(ranks), "-p", str(64 << 20), "-t", "60"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"globals: ok on {ranks} ranks" in r.stdout
    assert "rank check failed" not in r.stdout


@pytest.mark.parametrize("ranks", [2, 8, 33])
def test_coalesced_alltoall(ranks):
    print(f"[gpu] converted coalesce on {ranks} ranks", flush=True)
    r = run("coalesce", ["---gpumpi", "-n", str(ranks), "-p", str(64 << 20), "-t", "60"])
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("ranks", [2, 9])
def test_two_file_ring(ranks):
    print(f"[gpu] converted ring on {ranks} ranks", flush=True)
    r = run("ring", ["---gpumpi", "-n", str(ranks), "-p", str(64 << 20), "-t", "60"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"ring: ok after {ranks} hops" in r.stdout


def test_exit_ends_ranks_with_status():
    print("[gpu] converted exit", flush=True)
    r = run("exits", ["---gpumpi", "-n", "4", "-p", str(64 << 20), "-t", "60"])
    assert r.returncode == 1, r.stdout + r.stderr
    assert "before exit" in r.stdout
    assert "not reached" not in r.stdout


@pytest.mark.parametrize("ranks", [1, 8, 40])
def test_file_io_through_the_host(ranks, tmp_path):
    """Ranks write and read host files (device-to-host delegation); the files they wrote
    are checked here byte for byte."""
    import numpy as np
    print(f"[gpu] converted fileio on {ranks} ranks", flush=True)
    r = run("fileio", [str(tmp_path), "stdin", "---gpumpi", "-n", str(ranks), "-p", str(64 << 20),
                       "-t", "60"], stdin_text="4242\n")
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"fileio: ok on {ranks} ranks" in r.stdout
    for k in range(ranks):
        data = (tmp_path / f"rank{k}.txt").read_bytes()
        head = f"rank {k} of {ranks} pi=3.141593\nline two\n".encode()
        assert data.startswith(head)
        ints = np.frombuffer(data[len(head):], dtype=np.int32)
        assert np.array_equal(ints, k * 100000 + 7 * np.arange(1500, dtype=np.int32))
