"""Converted programs (the converter's output contract: __gpu_main, __gpu_global,
__gpu_max_threads) built as executables against the runtime and run through their own
---gpumpi command line, as the reference's toolchain tests do (toolchain_tests/pi/test.sh)."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "programs", "_build", "pi_converted")


def run(args, timeout=120):
    if not os.path.exists(EXE):
        raise RuntimeError(f"{EXE} was not built: run __graft_entry__.build()")
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("ranks", [1, 4, 64, 257])
def test_converted_pi(ranks):
    print(f"[gpu] converted pi on {ranks} ranks", flush=True)
    r = run(["5", "---gpumpi", "-n", str(ranks), "-p", str(64 << 20)])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "pi is approximately 3.14159265" in r.stdout
    assert f"5000 intervals, {ranks} ranks" in r.stdout


def test_converted_program_rejects_too_many_ranks():
    """More ranks than the per-rank global arrays hold (the program is compiled for 512):
    the launch fails cleanly (main.cu:113-122 behaviour) instead of corrupting memory."""
    r = run(["5", "---gpumpi", "-n", "600", "-p", str(64 << 20), "-t", "5"])
    assert r.returncode != 0
    assert "GPU_MPI_MAX_RANKS" in (r.stdout + r.stderr) or "device error" in (r.stdout + r.stderr)


def test_rank_grid_beyond_half_the_wave_slots_is_refused():
    """A grid that needs more than half of the wave slots its kernel can occupy is refused
    before launch (the reference's co-residency check, main.cu:189-209, with headroom)."""
    r = run(["5", "---gpumpi", "-n", "1100", "-p", str(64 << 20), "-t", "5"])
    assert r.returncode != 0
    assert "exceeds half of the" in r.stderr, r.stderr


def test_cli_requires_gpumpi_separator():
    r = run(["5"])
    assert r.returncode == 1
    assert "---gpumpi" in r.stderr
