"""Batched MPI_Reduce / MPI_Allreduce (gpumpi_reduce_*) against the oracle.

Bar: bit-exact with orc_reduce, which folds in the reference's rank order
(gpu_mpi/mpi.cu:209-225: recv = in_0; recv = op(in_r, recv), r = 1..P-1) -- FP64 included,
since the batched kernel keeps that order per element -- and with the rank grid's
MPI_Allreduce (the same fold) at the workload's 1 MiB per rank."""
import numpy as np
import pytest

from oracle_py import reduce

pytestmark = pytest.mark.gpu

MPI_INT, MPI_UNSIGNED, MPI_FLOAT, MPI_DOUBLE, MPI_INT64_T = 2, 9, 12, 13, 20
OP_MAX, OP_MIN, OP_SUM, OP_PROD, OP_BAND = 0, 1, 2, 3, 5
NP = {MPI_INT: np.int32, MPI_UNSIGNED: np.uint32, MPI_FLOAT: np.float32, MPI_DOUBLE: np.float64,
      MPI_INT64_T: np.int64}


def _inputs(rng, dt, op, P, n):
    t = NP[dt]
    if np.issubdtype(t, np.floating):
        return [rng.normal(size=n).astype(t) for _ in range(P)]
    if op == OP_PROD:
        return [rng.integers(-1 if t != np.uint32 else 0, 3, size=n).astype(t) for _ in range(P)]
    return [rng.integers(-(1 << 20) if t != np.uint32 else 0, 1 << 20, size=n).astype(t) for _ in range(P)]


def _run(pkg, torch, ins, dt, op, root, graph_steps=0):
    dev = torch.device("cuda")
    P, n = len(ins), ins[0].size
    es = ins[0].itemsize
    d_in = [torch.from_numpy(x.view(np.uint8).copy()).to(dev) for x in ins]
    d_out = [torch.full((n * es + 16,), 0xAB, dtype=torch.uint8, device=dev) for _ in range(P)]
    sp = torch.tensor([t.data_ptr() for t in d_in], dtype=torch.int64, device=dev)
    rp = torch.tensor([t.data_ptr() for t in d_out], dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    plan = pkg.ReducePlan(P, n, dt, op, root, sp.data_ptr(), rp.data_ptr())
    try:
        if graph_steps:
            plan.graph_replay(graph_steps)
        else:
            plan.execute()
        torch.cuda.synchronize()
    finally:
        plan.close()
    outs = [t.cpu().numpy() for t in d_out]
    for o in outs:
        assert np.all(o[n * es:] == 0xAB), "write past the receive buffer"
    return [o[:n * es] for o in outs]


@pytest.mark.parametrize("dt", [MPI_INT, MPI_DOUBLE, MPI_INT64_T, MPI_FLOAT, MPI_UNSIGNED])
@pytest.mark.parametrize("op", [OP_SUM, OP_MAX, OP_MIN, OP_PROD])
@pytest.mark.parametrize("P,n", [(1, 7), (7, 1), (9, 1000), (64, 4097), (130, 333)])
def test_batched_reduce_matches_oracle(dt, op, P, n, pkg, oracle, torch_gpu):
    print(f"[gpu] batched reduce dt={dt} op={op} P={P} n={n}", flush=True)
    rng = np.random.default_rng(P * 7919 + n + dt * 31 + op)
    ins = _inputs(rng, dt, op, P, n)
    rc, exp = reduce(oracle, dt, op, ins)
    assert rc == 0
    eb = exp.view(np.uint8)
    outs = _run(pkg, torch_gpu, ins, dt, op, -1)
    for r in range(P):
        assert np.array_equal(outs[r], eb), ("allreduce", r)
    root = P // 2
    outs = _run(pkg, torch_gpu, ins, dt, op, root)
    assert np.array_equal(outs[root], eb)
    for r in range(P):
        if r != root:
            assert np.all(outs[r] == 0xAB), "MPI_Reduce wrote a non-root buffer"


def test_batched_allreduce_1024_ranks_graph(pkg, oracle, torch_gpu):
    """1024 ranks x 2^15 doubles, three steps replayed from a graph: still the oracle's
    rank-order fold."""
    rng = np.random.default_rng(1)
    P, n = 1024, 1 << 15
    ins = [rng.normal(size=n) for _ in range(P)]
    rc, exp = reduce(oracle, MPI_DOUBLE, OP_SUM, ins)
    assert rc == 0
    outs = _run(pkg, torch_gpu, ins, MPI_DOUBLE, OP_SUM, -1, graph_steps=3)
    eb = exp.view(np.uint8)
    for r in range(0, P, 37):
        assert np.array_equal(outs[r], eb), r
    # against an independent order: within 1e-12 normwise
    ref = np.sum(np.stack(ins), axis=0)
    norm = np.sum(np.abs(np.stack(ins)), axis=0)
    assert np.all(np.abs(outs[0].view(np.float64) - ref) <= 1e-12 * norm)


@pytest.mark.parametrize("dt", [MPI_DOUBLE, MPI_INT])
def test_batched_equals_rank_grid_allreduce(dt, pkg, torch_gpu):
    """The workload's size, 1024 ranks x 1 MiB: the batched step and the rank grid's
    MPI_Allreduce (one wave per rank, the coll program) give the same bytes."""
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    P = 1024
    n = (1 << 20) // (8 if dt == MPI_DOUBLE else 4)
    es = 8 if dt == MPI_DOUBLE else 4
    inb = torch.zeros((P, n * es), dtype=torch.uint8, device="cuda")
    outb = torch.zeros((P, n * es), dtype=torch.uint8, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    args = pr.CollArgs(0xBEEF, pr.COLL_KINDS["allreduce"], dt, 0, 0, n, 1, 0, inb.data_ptr(),
                       outb.data_ptr(), n * es, n * es, sec.data_ptr())
    print(f"[gpu] rank-grid allreduce P={P} n={n} dt={dt}", flush=True)
    with pkg.Runtime(P, ranks_per_block=4, heap_bytes=256 << 20, timeout_s=60.0) as rt:
        rt.run("coll", args)
    torch.cuda.synchronize()
    grid_out = outb[0].clone()
    bout = torch.zeros((P, n * es), dtype=torch.uint8, device="cuda")
    sp = torch.tensor([inb[r].data_ptr() for r in range(P)], dtype=torch.int64, device="cuda")
    rp = torch.tensor([bout[r].data_ptr() for r in range(P)], dtype=torch.int64, device="cuda")
    plan = pkg.ReducePlan(P, n, dt, OP_SUM, -1, sp.data_ptr(), rp.data_ptr())
    try:
        plan.graph_replay(2)
        torch.cuda.synchronize()
    finally:
        plan.close()
    assert bool(torch.equal(bout[0], grid_out))
    for r0 in range(0, P, 128):
        assert bool((bout[r0:r0 + 128] == grid_out[None, :]).all()), r0


def test_batched_reduce_rejects_pairs_off_its_path(pkg, torch_gpu):
    torch = torch_gpu
    sp = torch.zeros(4, dtype=torch.int64, device="cuda")
    with pytest.raises(pkg.GpumpiError, match="not on the batched path"):
        pkg.ReducePlan(4, 16, MPI_INT, OP_BAND, -1, sp.data_ptr(), sp.data_ptr())
