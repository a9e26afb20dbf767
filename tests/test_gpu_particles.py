"""Particle exchange on the rank grid against the oracle, bit for bit (positions are
exact multiples of 2^-20, so every FP64 field is compared as bytes)."""
import numpy as np
import pytest

from oracle_py import particles_init, particles_step

pytestmark = pytest.mark.gpu

SEED = 0x5EED


def run_particles(pkg, torch, dims, n0, steps, mode, cap, first_step=0, state=None):
    from dphpc_project_amd import programs as pr
    gx, gy, gz = dims
    P = gx * gy * gz
    bufs = [torch.zeros(P * cap * 8, dtype=torch.float64, device="cuda") for _ in range(2)]
    send = torch.zeros(P * cap * 8, dtype=torch.float64, device="cuda")
    counts = torch.zeros(2 * P, dtype=torch.int64, device="cuda")
    cnt = torch.zeros(4 * P * P, dtype=torch.int32, device="cuda")
    moved = torch.zeros(P, dtype=torch.int64, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    if state is not None:
        parts, c = state
        bufs[0].copy_(torch.from_numpy(parts.reshape(-1)))
        counts[:P].copy_(torch.from_numpy(c))
    args = pr.ParticleArgs(SEED, gx, gy, gz, mode, steps, first_step, 0 if state is not None else n0,
                           cap, (pr._P * 2)(bufs[0].data_ptr(), bufs[1].data_ptr()), send.data_ptr(),
                           counts.data_ptr(), cnt.data_ptr(), moved.data_ptr(), sec.data_ptr())
    print(f"[gpu] particles dims={dims} n0={n0} steps={steps} mode={mode}", flush=True)
    with pkg.Runtime(P, heap_bytes=64 << 20, timeout_s=30.0) as rt:
        rt.run("particles", args)
    torch.cuda.synchronize()
    fin = steps & 1
    out = bufs[fin].cpu().numpy().reshape(P, cap, 8)
    c = counts[fin * P:(fin + 1) * P].cpu().numpy()
    return out, c, moved.cpu().numpy()


def oracle_run(oracle, dims, n0, steps, cap, first_step=0):
    parts, counts = particles_init(oracle, dims, n0, SEED, cap)
    moved = 0
    for s in range(first_step, first_step + steps):
        parts, counts, m = particles_step(oracle, dims, SEED, s, parts, counts)
        moved += m
    return parts, counts, moved


def assert_same_state(P, got, gc, exp, ec):
    assert np.array_equal(gc, ec)
    for r in range(P):
        n = int(ec[r])
        assert got[r, :n].tobytes() == exp[r, :n].tobytes(), f"rank {r}"


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 4, 4), (1, 2, 3)])
def test_particles_match_oracle(pkg, oracle, torch_gpu, dims, mode):
    P = dims[0] * dims[1] * dims[2]
    n0, steps, cap = 1000, 4, 2048
    got, gc, gm = run_particles(pkg, torch_gpu, dims, n0, steps, mode, cap)
    exp, ec, em = oracle_run(oracle, dims, n0, steps, cap)
    assert int(gm.sum()) == em
    assert_same_state(P, got, gc, exp, ec)


def test_particles_initial_state(pkg, oracle, torch_gpu):
    dims, n0, cap = (2, 2, 2), 777, 1024
    got, gc, _ = run_particles(pkg, torch_gpu, dims, n0, 0, 0, cap)
    exp, ec = particles_init(oracle, dims, n0, SEED, cap)
    assert_same_state(8, got, gc, exp, ec)


def test_particles_resume_from_state(pkg, oracle, torch_gpu):
    """A launch that starts from a given state at a later step index continues the run."""
    dims, n0, cap = (2, 2, 2), 600, 1536
    exp2, ec2, _ = oracle_run(oracle, dims, n0, 2, cap)
    got, gc, _ = run_particles(pkg, torch_gpu, dims, n0, 3, 0, cap, first_step=2, state=(exp2, ec2))
    exp, ec, _ = oracle_run(oracle, dims, n0, 5, cap)
    assert_same_state(8, got, gc, exp, ec)


def test_particles_full_grid(pkg, oracle, torch_gpu):
    """The 8x8x8 grid of the workload (512 ranks) at 4096 particles per rank, 3 steps."""
    dims, n0, cap = (8, 8, 8), 4096, 6144
    got, gc, gm = run_particles(pkg, torch_gpu, dims, n0, 3, 0, cap)
    exp, ec, em = oracle_run(oracle, dims, n0, 3, cap)
    assert int(gm.sum()) == em
    assert_same_state(512, got, gc, exp, ec)
