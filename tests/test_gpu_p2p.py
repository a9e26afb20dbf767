"""Batched point-to-point step (gpumpi_p2p_*): every pair's bytes land in its destination
exactly, and nothing else in the destination buffer changes.  Each pair is a matched
MPI_Isend/MPI_Irecv (gpu_libc/cuda_mpi.cu:65-118, gpu_mpi/mpi.cu:617-645).  The expected
buffer is built on the host with numpy.

The pair sizes straddle every boundary of the copy: zero bytes, pieces below one 16-byte
vector, the 8 KiB chunk edges, and multi-MiB messages.  Source and destination addresses
are misaligned at random.  Destinations get random gaps between them, which must keep
their sentinel bytes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SENTINEL = 0xA5


def build(torch, rng, sizes, src_len=None):
    """Device src/dst buffers and a pair tensor for `sizes`; returns the host expectation."""
    sizes = [int(x) for x in sizes]
    src_len = src_len or max(sum(sizes) // 2, max(sizes) + 64, 64)
    src_h = rng.integers(0, 256, size=src_len, dtype=np.uint8)
    offs, cur = [], int(rng.integers(0, 16))
    for n in sizes:
        offs.append(cur)
        cur += n + int(rng.integers(0, 40))
    dst_len = cur + 64
    src = torch.from_numpy(src_h).cuda()
    dst = torch.full((dst_len,), SENTINEL, dtype=torch.uint8, device="cuda")
    exp = np.full(dst_len, SENTINEL, dtype=np.uint8)
    rows = []
    for n, o in zip(sizes, offs):
        so = int(rng.integers(0, src_len - n + 1))
        exp[o:o + n] = src_h[so:so + n]
        rows.append((src.data_ptr() + so, dst.data_ptr() + o, n))
    pairs = torch.tensor(rows, dtype=torch.int64, device="cuda") if rows else \
        torch.zeros((0, 3), dtype=torch.int64, device="cuda")
    return src, dst, exp, pairs


def check(torch, dst, exp):
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    bad = np.nonzero(got != exp)[0]
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


EDGE_SIZES = [0, 1, 3, 15, 16, 17, 31, 100, 1023, 4095, 4096, 8191, 8192, 8193, 12345,
              16384, 24577, 65541, 1 << 20, (1 << 20) + 13, 3 * 8192 + 1]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_edge_sizes_misaligned(pkg, torch_gpu, seed):
    torch = torch_gpu
    rng = np.random.default_rng(seed)
    sizes = rng.permutation(EDGE_SIZES * 3)
    print(f"[gpu] p2p edge sizes seed={seed} pairs={len(sizes)}", flush=True)
    src, dst, exp, pairs = build(torch, rng, sizes)
    plan = pkg.P2pPlan(len(sizes))
    plan.execute(pairs.data_ptr(), len(sizes))
    check(torch, dst, exp)
    plan.close()


@pytest.mark.parametrize("npairs", [5000, 40001])
def test_many_small_pairs(npairs, pkg, torch_gpu):
    """Pairs of 0-300 bytes: 5000 fit one scan pass (7168 pairs), 40001 take six with a
    part pass and a part thread at the end; copy waves' chunk ranges start in the middle
    of the list."""
    torch = torch_gpu
    rng = np.random.default_rng(7)
    sizes = rng.integers(0, 301, size=npairs)
    print(f"[gpu] p2p {npairs} small pairs", flush=True)
    src, dst, exp, pairs = build(torch, rng, sizes)
    plan = pkg.P2pPlan(npairs + 100)
    plan.execute(pairs.data_ptr(), len(sizes))
    check(torch, dst, exp)


def test_large_messages(pkg, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(11)
    sizes = [64 << 20, (64 << 20) + 3, 8 << 20, 5]
    print("[gpu] p2p large messages", flush=True)
    src, dst, exp, pairs = build(torch, rng, sizes, src_len=(64 << 20) + 4096)
    plan = pkg.P2pPlan(4)
    plan.execute(pairs.data_ptr(), len(sizes))
    check(torch, dst, exp)


def test_pairs_rewritten_on_the_device_between_steps(pkg, torch_gpu):
    """The list is read when the step runs: a device-side rewrite of the sizes and
    destinations (torch kernels on the same stream) takes effect in the next step, with no
    new plan and no host copy of the list."""
    torch = torch_gpu
    rng = np.random.default_rng(5)
    sizes = rng.integers(1, 20000, size=64)
    src, dst, exp, pairs = build(torch, rng, sizes)
    plan = pkg.P2pPlan(64)
    plan.execute(pairs.data_ptr(), 64)
    check(torch, dst, exp)
    # step 2: halve every size and move every destination into a second buffer
    dst2 = torch.full_like(dst, SENTINEL)
    pairs[:, 2] //= 2
    pairs[:, 1] += dst2.data_ptr() - dst.data_ptr()
    plan.execute(pairs.data_ptr(), 64)
    rows = pairs.cpu().numpy()
    src_h = src.cpu().numpy()
    exp2 = np.full(dst2.numel(), SENTINEL, dtype=np.uint8)
    for s, d, n in rows:
        so, do = s - src.data_ptr(), d - dst2.data_ptr()
        exp2[do:do + n] = src_h[so:so + n]
    check(torch, dst2, exp2)
    check(torch, dst, exp)   # the first step's destination is untouched


def test_graph_replay_and_limits(pkg, torch_gpu):
    from dphpc_project_amd import GpumpiError
    torch = torch_gpu
    rng = np.random.default_rng(9)
    sizes = rng.permutation(EDGE_SIZES)
    src, dst, exp, pairs = build(torch, rng, sizes)
    plan = pkg.P2pPlan(len(sizes))
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(2):   # the second call replays the captured graph
        dst.fill_(SENTINEL)
        plan.graph_replay(pairs.data_ptr(), len(sizes), 3, stream)
        check(torch, dst, exp)
    scan_us, copy_us = plan.kernel_times(pairs.data_ptr(), len(sizes), 3, stream)
    assert scan_us > 0 and copy_us > 0
    check(torch, dst, exp)
    plan.execute(pairs.data_ptr(), 0)      # an empty step is a no-op
    with pytest.raises(GpumpiError):
        plan.execute(pairs.data_ptr(), len(sizes) + 1)
