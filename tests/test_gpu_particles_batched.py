"""Batched particle step (gpumpi_particles_*: tally, place and scatter kernels; hipGraph
replay) against the oracle's orc_particles_step, bit for bit: every FP64 field of every
particle, the counts, the migrant totals and the overflow path.  At the workload's size
(2^24 particles, 8x8x8 grid) against the rank-grid MPI program and by conservation.
Exchange mode 1 (the neighbour messages through the batched point-to-point step) must
give the same states as mode 0 in every test it runs in."""
import numpy as np
import pytest

from oracle_py import particles_init, particles_step

pytestmark = pytest.mark.gpu

SEED = 0x5EED


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).cuda()


def host_state(torch, parts, counts, P, cap):
    torch.cuda.synchronize()
    return parts.cpu().numpy().reshape(P, cap, 8), counts.cpu().numpy()


def assert_same_state(P, got, gc, exp, ec, whole=False):
    assert np.array_equal(gc, ec), f"counts differ: {np.nonzero(gc != ec)[0][:8]}"
    for r in range(P):
        n = exp.shape[1] if whole else int(ec[r])
        assert got[r, :n].tobytes() == exp[r, :n].tobytes(), f"rank {r}"


class Chain:
    """Two device state buffers stepped by one plan."""

    def __init__(self, pkg, torch, dims, cap, nontemporal=True, exchange=0):
        self.torch = torch
        self.dims, self.cap = dims, cap
        self.P = dims[0] * dims[1] * dims[2]
        self.plan = pkg.ParticlePlan(dims, cap, SEED, nontemporal)
        self.plan.exchange_mode(exchange)
        self.bufs = [torch.zeros(self.P * cap * 8, dtype=torch.float64, device="cuda") for _ in range(2)]
        self.cnts = [torch.zeros(self.P, dtype=torch.int64, device="cuda") for _ in range(2)]
        self.cur = 0

    def ptrs(self, i):
        return self.bufs[i].data_ptr(), self.cnts[i].data_ptr()

    def load(self, parts, counts, step):
        self.bufs[self.cur].copy_(dev(self.torch, parts))
        self.cnts[self.cur].copy_(dev(self.torch, counts))
        self.plan.prepare(*self.ptrs(self.cur), step)

    def step(self):
        self.plan.step(*self.ptrs(self.cur), *self.ptrs(self.cur ^ 1))
        self.cur ^= 1

    def state(self):
        return host_state(self.torch, self.bufs[self.cur], self.cnts[self.cur], self.P, self.cap)


@pytest.mark.parametrize("exchange", [0, 1])
@pytest.mark.parametrize("dims,cap", [((2, 2, 2), 2000), ((4, 4, 4), 2000), ((1, 2, 3), 2000),
                                      ((1, 1, 1), 2000), ((3, 1, 5), 1001), ((2, 3, 1), 1203)])
def test_batched_steps_match_oracle(pkg, oracle, torch_gpu, dims, cap, exchange):
    """Every step equals the oracle's, including grids whose axes of length 1 or 2 send
    several directions to one rank or back to the sender, and slabs whose size is not a
    multiple of the 256-slot chunk or of 4 (direction bytes read at any alignment)."""
    n0, steps = min(1000, cap * 4 // 5), 4
    print(f"[gpu] batched particles dims={dims} cap={cap} exchange={exchange}", flush=True)
    parts, counts = particles_init(oracle, dims, n0, SEED, cap)
    ch = Chain(pkg, torch_gpu, dims, cap, exchange=exchange)
    ch.load(parts, counts, 0)
    total = 0
    for s in range(steps):
        parts, counts, m = particles_step(oracle, dims, SEED, s, parts, counts)
        total += m
        ch.step()
        got, gc = ch.state()
        assert_same_state(ch.P, got, gc, parts, counts)
    moved, overflow, nxt = ch.plan.status()
    assert (moved, overflow, nxt) == (total, 0, steps)


def test_batched_init_matches_oracle(pkg, oracle, torch_gpu):
    dims, n0, cap = (2, 3, 2), 777, 1024
    ch = Chain(pkg, torch_gpu, dims, cap)
    ch.plan.init(*ch.ptrs(0), n0)
    got, gc = ch.state()
    exp, ec = particles_init(oracle, dims, n0, SEED, cap)
    assert_same_state(ch.P, got, gc, exp, ec, whole=True)


@pytest.mark.parametrize("exchange", [0, 1])
def test_batched_graph_replay_continues_steps(pkg, oracle, torch_gpu, exchange):
    """Graph replays carry the step index on: 3 replays of a 2-step graph, then a plain
    step, equal 7 oracle steps."""
    print(f"[gpu] batched particles graph replay exchange={exchange}", flush=True)
    dims, n0, cap = (4, 4, 4), 1500, 3000
    parts, counts = particles_init(oracle, dims, n0, SEED, cap)
    ch = Chain(pkg, torch_gpu, dims, cap, exchange=exchange)
    ch.load(parts, counts, 0)
    for rep in range(3):
        ch.plan.graph_replay(*ch.ptrs(0), *ch.ptrs(1), 2)
        for s in (2 * rep, 2 * rep + 1):
            parts, counts, _ = particles_step(oracle, dims, SEED, s, parts, counts)
        got, gc = ch.state()
        assert_same_state(ch.P, got, gc, parts, counts)
    ch.step()
    parts, counts, _ = particles_step(oracle, dims, SEED, 6, parts, counts)
    got, gc = ch.state()
    assert_same_state(ch.P, got, gc, parts, counts)
    assert ch.plan.status()[2] == 7


def test_batched_resume_at_later_step(pkg, oracle, torch_gpu):
    dims, n0, cap = (2, 2, 2), 600, 1536
    parts, counts = particles_init(oracle, dims, n0, SEED, cap)
    for s in range(2):
        parts, counts, _ = particles_step(oracle, dims, SEED, s, parts, counts)
    ch = Chain(pkg, torch_gpu, dims, cap)
    ch.load(parts, counts, 2)
    for s in range(2, 5):
        parts, counts, _ = particles_step(oracle, dims, SEED, s, parts, counts)
        ch.step()
    got, gc = ch.state()
    assert_same_state(ch.P, got, gc, parts, counts)


@pytest.mark.parametrize("exchange", [0, 1])
def test_batched_overflow_keeps_stayers(pkg, oracle, torch_gpu, exchange):
    """A rank that would receive more than cap particles: like the MPI program, the step
    writes the stayers and the counts only, and the overflow flag is raised (exchange 1:
    no message moves, so nothing lands past a slab's capacity)."""
    print(f"[gpu] batched particles overflow exchange={exchange}", flush=True)
    dims, cap = (2, 2, 2), 512
    parts, counts = particles_init(oracle, dims, cap, SEED, cap)   # every slab full
    exp, ec, m = particles_step(oracle, dims, SEED, 0, parts, counts, allow_overflow=True)
    assert m == -1
    ch = Chain(pkg, torch_gpu, dims, cap, exchange=exchange)
    ch.load(parts, counts, 0)
    ch.step()
    got, gc = ch.state()
    assert_same_state(ch.P, got, gc, exp, ec, whole=True)
    assert ch.plan.status()[1] == 1


@pytest.mark.parametrize("nontemporal,exchange", [(True, 0), (False, 0), (True, 1)])
def test_batched_full_grid_against_oracle(pkg, oracle, torch_gpu, nontemporal, exchange):
    print(f"[gpu] batched particles 8x8x8 nontemporal={nontemporal} exchange={exchange}", flush=True)
    dims, n0, cap = (8, 8, 8), 4096, 6144
    parts, counts = particles_init(oracle, dims, n0, SEED, cap)
    ch = Chain(pkg, torch_gpu, dims, cap, nontemporal, exchange)
    ch.load(parts, counts, 0)
    for s in range(3):
        parts, counts, _ = particles_step(oracle, dims, SEED, s, parts, counts)
        ch.step()
    got, gc = ch.state()
    assert_same_state(512, got, gc, parts, counts)


@pytest.mark.parametrize("exchange", [0, 1])
def test_batched_workload_size_matches_rank_grid_program(pkg, torch_gpu, exchange):
    """2^24 particles on the 8x8x8 grid, 3 steps: the batched step and the rank-grid MPI
    program (itself checked against the oracle; neighbour mode for exchange 1) end in the
    same state; ids stay a permutation and every particle lies in its owner's cell."""
    torch = torch_gpu
    from dphpc_project_amd import programs as pr
    dims, n0, cap, steps = (8, 8, 8), 32768, 40960, 3
    P = 512
    ch = Chain(pkg, torch, dims, cap, exchange=exchange)
    ch.plan.init(*ch.ptrs(0), n0)
    ch.plan.prepare(*ch.ptrs(0), 0)
    for _ in range(steps):
        ch.step()
    torch.cuda.synchronize()
    bufs = [torch.zeros(P * cap * 8, dtype=torch.float64, device="cuda") for _ in range(2)]
    send = torch.zeros(P * cap * 8, dtype=torch.float64, device="cuda")
    counts = torch.zeros(2 * P, dtype=torch.int64, device="cuda")
    cnt = torch.zeros(4 * P * P, dtype=torch.int32, device="cuda")
    moved = torch.zeros(P, dtype=torch.int64, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    args = pr.ParticleArgs(SEED, 8, 8, 8, exchange, steps, 0, n0, cap,
                           (pr._P * 2)(bufs[0].data_ptr(), bufs[1].data_ptr()), send.data_ptr(),
                           counts.data_ptr(), cnt.data_ptr(), moved.data_ptr(), sec.data_ptr())
    print(f"[gpu] rank-grid particles 8x8x8 n0=32768 steps=3 mode={exchange}", flush=True)
    with pkg.Runtime(P, heap_bytes=64 << 20, timeout_s=60.0) as rt:
        rt.run("particles", args)
    torch.cuda.synchronize()
    fin = steps & 1
    assert torch.equal(counts[fin * P:(fin + 1) * P], ch.cnts[ch.cur])
    c = ch.cnts[ch.cur].cpu().numpy()
    mask = (torch.arange(cap, device="cuda")[None, :] < ch.cnts[ch.cur][:, None])
    a = ch.bufs[ch.cur].view(P, cap, 8)[mask]
    b = bufs[fin].view(P, cap, 8)[mask]
    assert torch.equal(a.view(torch.int64), b.view(torch.int64))
    assert int(c.sum()) == P * n0
    ids = torch.sort(a[:, 7].to(torch.int64)).values
    assert torch.equal(ids, torch.arange(P * n0, device="cuda"))
    owner = torch.repeat_interleave(torch.arange(P, device="cuda"), ch.cnts[ch.cur])
    cell = a[:, 0].floor().to(torch.int64) + 8 * (a[:, 1].floor().to(torch.int64) +
                                                    8 * a[:, 2].floor().to(torch.int64))
    assert torch.equal(cell, owner)
    mv, ov, nxt = ch.plan.status()
    assert ov == 0 and nxt == steps and mv == int(moved.sum().item())
