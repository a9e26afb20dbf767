"""Rank-grid workload programs against the oracle (bit-exact for integer, byte and
index results; FP64 reductions bit-exact because the device folds in the reference's
rank order 0..P-1, gpu_mpi/mpi.cu:209-225)."""
import ctypes

import numpy as np
import pytest

from oracle_py import fill_bytes, fill_int, fill_normal, reduce

pytestmark = pytest.mark.gpu

MPI_INT, MPI_DOUBLE = 2, 13
OP_MAX, OP_SUM = 0, 2


def test_mcpi(pkg, oracle, torch_gpu):
    """Config 1 shape: 64 ranks, 2^20 pairs each; int64 count bit-exact."""
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    P, n, seed = 64, 1 << 20, 0x5eed
    print(f"[gpu] mcpi P={P} n={n}", flush=True)
    per = torch.zeros(P, dtype=torch.int64, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    with pkg.Runtime(P, heap_bytes=64 << 20, timeout_s=20.0) as rt:
        rt.run("mcpi", pr.McpiArgs(seed, n, per.data_ptr(), tot.data_ptr(), sec.data_ptr()))
    got = per.cpu().numpy()
    # oracle on a subset of ranks at full n, all ranks at small n below
    for r in (0, 1, 31, 63):
        assert got[r] == oracle.orc_mcpi_hits(ctypes.c_uint64(seed), r, ctypes.c_int64(n)), r
    assert tot.item() == int(got.sum())
    est = 4.0 * tot.item() / (P * n)
    assert abs(est - np.pi) < 1e-3


def test_mcpi_all_ranks_small(pkg, oracle, torch_gpu):
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    P, n, seed = 64, 4096, 0xabcdef
    per = torch.zeros(P, dtype=torch.int64, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    with pkg.Runtime(P, heap_bytes=64 << 20) as rt:
        rt.run("mcpi", pr.McpiArgs(seed, n, per.data_ptr(), tot.data_ptr(), sec.data_ptr()))
    exp = [oracle.orc_mcpi_hits(ctypes.c_uint64(seed), r, ctypes.c_int64(n)) for r in range(P)]
    assert per.cpu().tolist() == exp
    assert tot.item() == sum(exp)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("nbytes", [8, 80, 81, 4096, 1 << 20])
def test_p2p_payloads(mode, nbytes, pkg, oracle, torch_gpu):
    """Config 2 shapes at reduced size: ping-pong / ring, blocking / non-blocking; every
    received byte equal to the sender's Philox payload (and to the oracle's bytes)."""
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    P, seed = 16, 0x77
    print(f"[gpu] p2p mode={mode} bytes={nbytes}", flush=True)
    sb = torch.zeros(P * nbytes, dtype=torch.uint8, device="cuda")
    rb = torch.zeros(P * nbytes, dtype=torch.uint8, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    mis = torch.full((P,), -1, dtype=torch.int64, device="cuda")
    with pkg.Runtime(P, heap_bytes=64 << 20, timeout_s=20.0) as rt:
        rt.run("p2p", pr.P2pArgs(seed, mode, 3, 1, 0, nbytes, sb.data_ptr(), rb.data_ptr(),
                                 sec.data_ptr(), mis.data_ptr()))
    assert mis.cpu().tolist() == [0] * P
    recv = rb.cpu().numpy().reshape(P, nbytes)
    send = sb.cpu().numpy().reshape(P, nbytes)
    for r in range(P):
        assert np.array_equal(send[r], fill_bytes(oracle, seed, r, nbytes))
        origin = (r if r % 2 == 0 else r ^ 1) if mode < 2 else (r - 1) % P
        assert np.array_equal(recv[r], fill_bytes(oracle, seed, origin, nbytes)), r


@pytest.mark.parametrize("mode", [0, 3])
@pytest.mark.parametrize("nbytes", [8, 4 << 20, 64 << 20])
def test_p2p_256_ranks(mode, nbytes, pkg, oracle, torch_gpu):
    """Config 2 at its own shape: 256 ranks (128 ping-pong pairs, or the ring), up to 64 MiB
    per message.  Every rank checks every received byte against the Philox generator on
    the device; the host compares sampled ranks' buffers with the oracle's bytes."""
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    P, seed = 256, 0x2561
    print(f"[gpu] p2p P={P} mode={mode} bytes={nbytes}", flush=True)
    sb = torch.zeros(P * nbytes, dtype=torch.uint8, device="cuda")
    rb = torch.zeros(P * nbytes, dtype=torch.uint8, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    mis = torch.full((P,), -1, dtype=torch.int64, device="cuda")
    with pkg.Runtime(P, ranks_per_block=4, heap_bytes=256 << 20, timeout_s=60.0) as rt:
        rt.run("p2p", pr.P2pArgs(seed, mode, 2, 1, 0, nbytes, sb.data_ptr(), rb.data_ptr(),
                                 sec.data_ptr(), mis.data_ptr()))
    assert mis.cpu().tolist() == [0] * P
    for r in (0, 1, 128, 255):
        origin = (r if r % 2 == 0 else r ^ 1) if mode < 2 else (r - 1) % P
        got = rb[r * nbytes:(r + 1) * nbytes].cpu().numpy()
        assert np.array_equal(got, fill_bytes(oracle, seed, origin, nbytes)), r


@pytest.mark.parametrize("stage_kib", [-1, 16, 32, 64])
@pytest.mark.parametrize("offset,nbytes", [(0, 1 << 20), (5, 3 * 32768 + 16 * 7), (0, 65536 + 16),
                                           (3, 16384 + 80)])
@pytest.mark.parametrize("mode", [0, 3])
def test_p2p_lds_staged_copies(mode, offset, nbytes, stage_kib, pkg, oracle, torch_gpu):
    """Large messages through the LDS-staged copy (copy_stage_kib 16 / 32 / 64) and through
    registers only (-1), byte-exact against the oracle.  Buffers at offset 5 give every
    message an 11-byte head before the 16-byte aligned part and a byte tail after it;
    3 x 32 KiB + 112 leaves a part batch after the whole stage batches; 16 KiB + 80 at
    offset 3 is one whole 16 KiB batch, or none of 32 KiB, around a 13-byte head."""
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    P, seed = 16, 0x5ba
    print(f"[gpu] p2p staged mode={mode} offset={offset} bytes={nbytes} stage={stage_kib}", flush=True)
    sb = torch.zeros(P * nbytes + 16, dtype=torch.uint8, device="cuda")
    rb = torch.zeros(P * nbytes + 16, dtype=torch.uint8, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    mis = torch.full((P,), -1, dtype=torch.int64, device="cuda")
    with pkg.Runtime(P, heap_bytes=64 << 20, timeout_s=20.0, copy_stage_kib=stage_kib) as rt:
        rt.run("p2p", pr.P2pArgs(seed, mode, 2, 1, 0, nbytes, sb.data_ptr() + offset,
                                 rb.data_ptr() + offset, sec.data_ptr(), mis.data_ptr()))
        spb = torch.cuda.get_device_properties(0).shared_memory_per_block
        fits = stage_kib > 0 and (stage_kib << 10) + 1024 <= spb
        assert rt.copy_stage_bytes() == (stage_kib << 10 if fits else 0)
    assert mis.cpu().tolist() == [0] * P
    recv = rb.cpu().numpy()[offset:offset + P * nbytes].reshape(P, nbytes)
    for r in range(P):
        origin = (r if r % 2 == 0 else r ^ 1) if mode < 2 else (r - 1) % P
        assert np.array_equal(recv[r], fill_bytes(oracle, seed, origin, nbytes)), r


def test_copy_stage_follows_residency(pkg, torch_gpu):
    """The automatic stage is the largest of 64 / 32 / 16 KiB per rank with which the grid
    keeps its residency headroom (half of the co-resident wave slots) and the workgroup's
    LDS fits the device's per-workgroup limit.  On 256 CUs with 160 KiB of LDS each and a
    program of two waves per SIMD: 256 ranks take 64 KiB (2 workgroups per CU), 512 ranks
    32 KiB (5 per CU), 1024 ranks 16 KiB (8 per CU); 4 ranks per workgroup take 32 KiB
    when a workgroup may hold 128 KiB.  Where the device allows less per workgroup, the
    next size down."""
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    spb = torch.cuda.get_device_properties(0).shared_memory_per_block
    slack = 1024   # the program's own static LDS (a few pointers)
    cases = ((256, 1, 64 << 10 if spb >= (64 << 10) + slack else 32 << 10),
             (512, 1, 32 << 10), (1024, 1, 16 << 10),
             (256, 4, 32 << 10 if spb >= (128 << 10) + slack else 16 << 10))
    print(f"[gpu] LDS per workgroup {spb}", flush=True)
    for P, rpb, want in cases:
        print(f"[gpu] copy stage P={P} rpb={rpb}", flush=True)
        sb = torch.zeros(P * 64, dtype=torch.uint8, device="cuda")
        rb = torch.zeros_like(sb)
        sec = torch.zeros(P, dtype=torch.float64, device="cuda")
        mis = torch.full((P,), -1, dtype=torch.int64, device="cuda")
        with pkg.Runtime(P, ranks_per_block=rpb, heap_bytes=64 << 20, timeout_s=20.0) as rt:
            rt.run("p2p", pr.P2pArgs(0x9, 0, 1, 0, 0, 64, sb.data_ptr(), rb.data_ptr(),
                                     sec.data_ptr(), mis.data_ptr()))
            assert rt.copy_stage_bytes() == want, (P, rpb, rt.copy_stage_bytes())
        assert mis.cpu().tolist() == [0] * P


def run_coll(pkg, torch, P, kind, dtype, op, count, root=0, seed=0x1234, **cfg):
    from dphpc_project_amd import programs as pr
    es = 8 if dtype == MPI_DOUBLE else 4
    in_elems = count * (P if kind == "scatter" else 1)
    out_elems = count * (P if kind in ("allgather", "gather") else 1)
    inb = torch.zeros(P * in_elems * es + 16, dtype=torch.uint8, device="cuda")
    outb = torch.zeros(P * out_elems * es + 16, dtype=torch.uint8, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    prog_op = 0 if op == OP_SUM else 1  # the program's own encoding: 0 = MPI_SUM, 1 = MPI_MAX
    args = pr.CollArgs(seed, pr.COLL_KINDS[kind], dtype, prog_op, root, count, 1, 0,
                       inb.data_ptr(), outb.data_ptr(), in_elems * es, out_elems * es, sec.data_ptr())
    with pkg.Runtime(P, heap_bytes=128 << 20, timeout_s=20.0, **cfg) as rt:
        rt.run("coll", args)
    npdt = np.float64 if dtype == MPI_DOUBLE else np.int32
    out = outb.cpu().numpy()[:P * out_elems * es].view(npdt).reshape(P, out_elems)
    return out


def gen(oracle, dtype, seed, r, n):
    return fill_normal(oracle, seed, r, n) if dtype == MPI_DOUBLE else fill_int(oracle, seed, r, n)


@pytest.mark.parametrize("dtype", [MPI_INT, MPI_DOUBLE])
@pytest.mark.parametrize("op", [OP_SUM, OP_MAX])
@pytest.mark.parametrize("P,count", [(7, 1), (64, 1000), (33, 4097)])
def test_reduce_allreduce(dtype, op, P, count, pkg, oracle, torch_gpu):
    """Config 3: Reduce and Allreduce bit-exact with the oracle in reference fold order."""
    print(f"[gpu] reduce/allreduce dtype={dtype} op={op} P={P} count={count}", flush=True)
    seed = 0x1234
    ins = [gen(oracle, dtype, seed, r, count) for r in range(P)]
    rc, exp = reduce(oracle, dtype, op, ins)
    assert rc == 0
    out = run_coll(pkg, torch_gpu, P, "reduce", dtype, op, count, root=P // 2, seed=seed)
    assert np.array_equal(out[P // 2], exp)
    out = run_coll(pkg, torch_gpu, P, "allreduce", dtype, op, count, seed=seed)
    for r in range(P):
        assert np.array_equal(out[r], exp), r
    if dtype == MPI_DOUBLE and op == OP_SUM:
        # against a different (pairwise) order: within 1e-12 normwise
        ref = np.sum(np.stack(ins), axis=0)
        norm = np.sum(np.abs(np.stack(ins)), axis=0)
        assert np.all(np.abs(out[0] - ref) <= 1e-12 * norm)


@pytest.mark.parametrize("dtype", [MPI_INT, MPI_DOUBLE])
@pytest.mark.parametrize("P,count", [(5, 3), (64, 256), (128, 1000)])
def test_bcast_alltoall_gather_scatter(dtype, P, count, pkg, oracle, torch_gpu):
    print(f"[gpu] bcast/alltoall/gather/scatter dtype={dtype} P={P} count={count}", flush=True)
    seed = 0x99
    ins = [gen(oracle, dtype, seed, r, count) for r in range(P)]
    root = P - 1
    out = run_coll(pkg, torch_gpu, P, "bcast", dtype, OP_SUM, count, root=root, seed=seed)
    for r in range(P):
        assert np.array_equal(out[r], ins[root])
    # Alltoall: per-rank total `count`, per pair count // P
    per = count // P
    if per > 0:
        out = run_coll(pkg, torch_gpu, P, "alltoall", dtype, OP_SUM, count, seed=seed)
        for d in range(P):
            exp = np.concatenate([ins[s][d * per:(d + 1) * per] for s in range(P)])
            assert np.array_equal(out[d][:P * per], exp)
    out = run_coll(pkg, torch_gpu, P, "allgather", dtype, OP_SUM, count, seed=seed)
    exp = np.concatenate(ins)
    for r in range(P):
        assert np.array_equal(out[r], exp)
    out = run_coll(pkg, torch_gpu, P, "gather", dtype, OP_SUM, count, root=1 % P, seed=seed)
    assert np.array_equal(out[1 % P], exp)
    out = run_coll(pkg, torch_gpu, P, "scatter", dtype, OP_SUM, count, root=0, seed=seed)
    src = gen(oracle, dtype, seed, 0, count * P)
    for r in range(P):
        assert np.array_equal(out[r], src[r * count:(r + 1) * count])


@pytest.mark.parametrize("dtype,P,count,root", [
    (MPI_INT, 3, 1027, 0),          # 4108 B: 2 KiB slices, a 12-byte tail on rank 2
    (MPI_DOUBLE, 64, 100003, 63),   # 13 KiB slices (1 KiB passes), an 8-byte tail; root pushes too
    (MPI_DOUBLE, 5, 25001, 2),      # 40 KiB slices: two 16 KiB passes + 1 KiB passes + a tail
    (MPI_INT, 257, (1 << 20) + 3, 128),  # full 16 KiB passes
    (MPI_DOUBLE, 1024, 1 << 18, 5),
])
def test_bcast_push_and_pull(dtype, P, count, root, pkg, oracle, torch_gpu):
    """MPI_Bcast under both schedules (bcast_push_min 1: every size pushes; -1: never) and
    the default choice: every rank's buffer equals the root's input."""
    print(f"[gpu] bcast push/pull dtype={dtype} P={P} count={count} root={root}", flush=True)
    exp = gen(oracle, dtype, 0x77, root, count)
    for pmin in (1, -1, 0):
        out = run_coll(pkg, torch_gpu, P, "bcast", dtype, OP_SUM, count, root=root, seed=0x77,
                       bcast_push_min=pmin)
        bad = [r for r in range(P) if not np.array_equal(out[r], exp)]
        assert not bad, (pmin, bad[:8])


def test_heap_churn_on_a_full_arena(pkg, oracle, torch_gpu):
    """The churn on an arena too small for its live set (64 ranks, 32 MiB, ~26 MiB live at
    peak): the page bump runs out, so allocations are served by other ranks' free stacks
    and by splitting larger blocks, and some fail with nullptr.  No block may be handed out
    twice (stamps intact, live blocks disjoint and aligned) and the heap ends empty."""
    from dphpc_project_amd import programs as pr
    from oracle_py import ptr
    torch = torch_gpu
    P, ops, heap = 64, 5000, 32 << 20
    print(f"[gpu] heap churn on a full arena P={P} ops={ops} heap={heap >> 20} MiB", flush=True)
    err = torch.full((P,), -1, dtype=torch.int64, device="cuda")
    failed = torch.full((P,), -1, dtype=torch.int64, device="cuda")
    peak = torch.zeros(P, dtype=torch.int64, device="cuda")
    lp = torch.zeros(P * 64, dtype=torch.int64, device="cuda")
    ls = torch.zeros(P * 64, dtype=torch.int64, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    with pkg.Runtime(P, heap_bytes=heap, timeout_s=60.0) as rt:
        rt.run("churn", pr.ChurnArgs(0x5eed, ops, 100, 1, err.data_ptr(), peak.data_ptr(),
                                     lp.data_ptr(), ls.data_ptr(), sec.data_ptr(), failed.data_ptr()))
        stats = rt.heap_stats()
    assert err.cpu().tolist() == [0] * P
    assert int(failed.sum().item()) > 0, "the arena was expected to run out"
    assert stats["in_use"] == 0
    p = lp.cpu().numpy().view(np.uint64)
    s = ls.cpu().numpy().view(np.uint64)
    keep = p != 0
    p, s = np.ascontiguousarray(p[keep]), np.ascontiguousarray(s[keep])
    base = int(p.min()) & ~0xFFFF
    assert oracle.orc_heap_check(ptr(p), ptr(s), ctypes.c_int64(p.size), ctypes.c_uint64(base),
                                 ctypes.c_uint64(heap + (1 << 20))) == 0


def test_heap_churn(pkg, oracle, torch_gpu):
    """Config 3 allocator churn at reduced length: stamps survive, live blocks are disjoint
    and aligned, the heap returns to empty, with an Allreduce interleaved."""
    from dphpc_project_amd import programs as pr
    from oracle_py import ptr
    torch = torch_gpu
    P, ops = 256, 20000
    print(f"[gpu] heap churn P={P} ops={ops}", flush=True)
    err = torch.full((P,), -1, dtype=torch.int64, device="cuda")
    peak = torch.zeros(P, dtype=torch.int64, device="cuda")
    lp = torch.zeros(P * 64, dtype=torch.int64, device="cuda")
    ls = torch.zeros(P * 64, dtype=torch.int64, device="cuda")
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    heap = 1 << 30
    with pkg.Runtime(P, heap_bytes=heap, timeout_s=60.0) as rt:
        rt.run("churn", pr.ChurnArgs(0x5eed, ops, 1000, 1, err.data_ptr(), peak.data_ptr(),
                                     lp.data_ptr(), ls.data_ptr(), sec.data_ptr()))
        stats = rt.heap_stats()
    assert err.cpu().tolist() == [0] * P
    assert stats["in_use"] == 0
    p = lp.cpu().numpy().view(np.uint64)
    s = ls.cpu().numpy().view(np.uint64)
    keep = p != 0
    p, s = np.ascontiguousarray(p[keep]), np.ascontiguousarray(s[keep])
    base = int(p.min()) & ~0xFFFF
    assert oracle.orc_heap_check(ptr(p), ptr(s), ctypes.c_int64(p.size), ctypes.c_uint64(base),
                                 ctypes.c_uint64(heap + (1 << 20))) == 0
    assert 0 < peak.cpu().max() <= 64
