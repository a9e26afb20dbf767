"""Config 3 at its own shape: 1024 ranks on one GPU (the co-residency limit of the rank
grid is half of the wave slots a program can occupy; the rank programs hold two waves per
SIMD, so 1024 ranks fit with room to spare).

Bcast, Reduce and Allreduce (SUM and MAX) and Alltoall on MPI_INT and MPI_DOUBLE:
  * counts 1 and 4097: every output element against the oracle (orc_reduce folds in the
    reference's rank order, gpu_mpi/mpi.cu:209-225, so FP64 sums are bit-exact);
  * 2^22 elements per rank (the sweep's largest, 16 / 32 MiB per rank): every element
    checked on the device against closed forms (Bcast = the root's input, Allreduce
    replicated on every rank, integer SUM/MAX in any order, Alltoall = the block
    transpose of the inputs), FP64 SUM within 1e-12 normwise of an independent order, and
    sampled elements -- both ends of every rank's reduction slice plus random ones --
    bit-exact against the oracle's fold of the oracle's own inputs."""
import ctypes

import numpy as np
import pytest

from oracle_py import fill_int, fill_normal, ptr, reduce

pytestmark = pytest.mark.gpu

MPI_INT, MPI_DOUBLE = 2, 13
OP_MAX, OP_SUM = 0, 2
P = 1024
SEED = 0xC011


def _runtime(pkg):
    return pkg.Runtime(P, ranks_per_block=4, heap_bytes=256 << 20, timeout_s=60.0)


def _run(pkg, rt, torch, kind, dtype, op, count, root, inb, outb):
    from dphpc_project_amd import programs as pr
    es = 8 if dtype == MPI_DOUBLE else 4
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    prog_op = 0 if op == OP_SUM else 1  # the program's own encoding: 0 = MPI_SUM, 1 = MPI_MAX
    args = pr.CollArgs(SEED, pr.COLL_KINDS[kind], dtype, prog_op, root, count, 1, 0,
                       inb.data_ptr(), outb.data_ptr(), count * es, count * es, sec.data_ptr())
    print(f"[gpu] coll P={P} kind={kind} dtype={dtype} op={op} count={count}", flush=True)
    rt.run("coll", args)
    torch.cuda.synchronize()


def _gen(oracle, dtype, r, n):
    return fill_normal(oracle, SEED, r, n) if dtype == MPI_DOUBLE else fill_int(oracle, SEED, r, n)


def _gen_at(oracle, dtype, r, idx):
    if dtype == MPI_DOUBLE:
        out = np.zeros(idx.size, dtype=np.float64)
        oracle.orc_fill_normal_at(ctypes.c_uint64(SEED), r, ctypes.c_uint32(0), ptr(idx),
                                  ctypes.c_int64(idx.size), ptr(out))
    else:
        out = np.zeros(idx.size, dtype=np.int32)
        oracle.orc_fill_int_at(ctypes.c_uint64(SEED), r, ctypes.c_uint32(0), ptr(idx),
                               ctypes.c_int64(idx.size), ptr(out))
    return out


@pytest.mark.parametrize("dtype", [MPI_INT, MPI_DOUBLE])
@pytest.mark.parametrize("count", [1, 4097])
def test_collectives_1024_ranks_against_the_oracle(dtype, count, pkg, oracle, torch_gpu):
    torch = torch_gpu
    tdt = torch.float64 if dtype == MPI_DOUBLE else torch.int32
    inb = torch.zeros((P, count), dtype=tdt, device="cuda")
    outb = torch.zeros((P, count), dtype=tdt, device="cuda")
    ins = [_gen(oracle, dtype, r, count) for r in range(P)]
    root = 517
    with _runtime(pkg) as rt:
        _run(pkg, rt, torch, "bcast", dtype, OP_SUM, count, root, inb, outb)
        out = outb.cpu().numpy()
        assert np.array_equal(inb[root].cpu().numpy(), ins[root])
        assert (out == ins[root][None, :]).all()
        for op in (OP_SUM, OP_MAX):
            rc, exp = reduce(oracle, dtype, op, ins)
            assert rc == 0
            outb.zero_()
            _run(pkg, rt, torch, "reduce", dtype, op, count, root, inb, outb)
            assert np.array_equal(outb[root].cpu().numpy(), exp), ("reduce", op)
            outb.zero_()
            _run(pkg, rt, torch, "allreduce", dtype, op, count, 0, inb, outb)
            out = outb.cpu().numpy()
            assert (out == exp[None, :]).all(), ("allreduce", op)
        per = count // P
        if per > 0:
            outb.zero_()
            _run(pkg, rt, torch, "alltoall", dtype, OP_SUM, count, 0, inb, outb)
            out = outb.cpu().numpy()
            for d in range(0, P, 97):
                exp = np.concatenate([ins[s][d * per:(d + 1) * per] for s in range(P)])
                assert np.array_equal(out[d][:P * per], exp), d


@pytest.mark.parametrize("dtype", [MPI_INT, MPI_DOUBLE])
@pytest.mark.parametrize("per", [1, 2, 4])
def test_alltoall_small_blocks_1024_ranks(dtype, per, pkg, oracle, torch_gpu):
    """MPI_Alltoall of 1, 2 and 4 elements per pair (4-32 byte blocks): blocks of up to 16
    bytes are pulled lane by lane (16 sources per lane), 32-byte ones two lanes per source.  Every destination's whole receive buffer equals the transpose of the
    inputs."""
    torch = torch_gpu
    print(f"[gpu] alltoall 1024 ranks dtype={dtype} per={per}", flush=True)
    count = per * P
    tdt = torch.float64 if dtype == MPI_DOUBLE else torch.int32
    inb = torch.zeros((P, count), dtype=tdt, device="cuda")
    outb = torch.zeros((P, count), dtype=tdt, device="cuda")
    ins = np.stack([_gen(oracle, dtype, r, count) for r in range(P)])
    exp = ins.reshape(P, P, per).transpose(1, 0, 2).reshape(P, count)
    with _runtime(pkg) as rt:
        for _ in range(2):   # twice: the second call reuses the slots and counters
            outb.zero_()
            _run(pkg, rt, torch, "alltoall", dtype, OP_SUM, count, 0, inb, outb)
            out = outb.cpu().numpy()
            bad = np.nonzero((out != exp).any(axis=1))[0]
            assert bad.size == 0, f"destinations {bad[:8]} differ"


@pytest.mark.parametrize("dtype,per", [(MPI_INT, 8), (MPI_INT, 12), (MPI_INT, 16), (MPI_INT, 32),
                                       (MPI_INT, 128), (MPI_INT, 256), (MPI_INT, 512), (MPI_INT, 768),
                                       (MPI_INT, 1024), (MPI_INT, 1028), (MPI_DOUBLE, 6),
                                       (MPI_DOUBLE, 8), (MPI_DOUBLE, 3), (MPI_DOUBLE, 128),
                                       (MPI_DOUBLE, 512)])
def test_alltoall_blocks_1024_ranks_transpose(dtype, per, pkg, torch_gpu):
    """MPI_Alltoall of 32 B .. 4 KiB blocks: whole 16-byte vectors up to 4 KiB are pulled
    several sources per pass -- up to 512 B with 2-32 lanes per source (32, 48, 64, 128 and
    512 B here), above with 1, 2 or 4 vectors per lane (8, 4 or 2 sources); 4112-byte and
    24-byte blocks take the one-source-per-pass copy.  Every destination's receive buffer
    equals the block transpose of the inputs (which the program generated on the device),
    compared on the device; twice per runtime."""
    torch = torch_gpu
    print(f"[gpu] alltoall 1024 ranks dtype={dtype} per={per} (transpose)", flush=True)
    count = per * P
    tdt = torch.float64 if dtype == MPI_DOUBLE else torch.int32
    inb = torch.zeros((P, count), dtype=tdt, device="cuda")
    outb = torch.zeros((P, count), dtype=tdt, device="cuda")
    with _runtime(pkg) as rt:
        for _ in range(2):
            outb.zero_()
            _run(pkg, rt, torch, "alltoall", dtype, OP_SUM, count, 0, inb, outb)
            exp = inb.view(P, P, per).transpose(0, 1).contiguous().view(P, count)
            same = (outb == exp).all(dim=1)
            bad = torch.nonzero(~same).flatten()[:8].tolist()
            assert not bad, f"destinations {bad} differ"
            del exp, same
    del inb, outb
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype,count", [(MPI_INT, 8), (MPI_INT, 16), (MPI_INT, 96), (MPI_INT, 256),
                                         (MPI_INT, 768), (MPI_INT, 1024),
                                         (MPI_INT, 1028), (MPI_DOUBLE, 3), (MPI_DOUBLE, 512)])
def test_allgather_1024_ranks(dtype, count, pkg, torch_gpu):
    """MPI_Allgather of 32 B .. 4 KiB per rank: whole 16-byte vectors up to 4 KiB are pulled
    several sources per pass (up to 512 B several sources per row of lanes), 4112-byte and 24-byte blocks one source at a time.  Every
    rank's receive buffer equals all inputs in rank order, compared on the device."""
    torch = torch_gpu
    print(f"[gpu] allgather 1024 ranks dtype={dtype} count={count}", flush=True)
    tdt = torch.float64 if dtype == MPI_DOUBLE else torch.int32
    inb = torch.zeros((P, count), dtype=tdt, device="cuda")
    outb = torch.zeros((P, P * count), dtype=tdt, device="cuda")
    from dphpc_project_amd import programs as pr
    es = 8 if dtype == MPI_DOUBLE else 4
    sec = torch.zeros(P, dtype=torch.float64, device="cuda")
    with _runtime(pkg) as rt:
        for _ in range(2):
            outb.zero_()
            args = pr.CollArgs(SEED, pr.COLL_KINDS["allgather"], dtype, 0, 0, count, 1, 0,
                               inb.data_ptr(), outb.data_ptr(), count * es, P * count * es,
                               sec.data_ptr())
            rt.run("coll", args)
            torch.cuda.synchronize()
            same = (outb == inb.reshape(1, P * count)).all(dim=1)
            bad = torch.nonzero(~same).flatten()[:8].tolist()
            assert not bad, f"ranks {bad} differ"
    del inb, outb
    torch.cuda.empty_cache()


def _sample_indices(count, rng):
    """Both ends of every rank's reduction slice (multiples of 256 elements, the runtime's
    slicing) plus random elements."""
    per = -(-count // P)
    per = (per + 255) & ~255
    ends = []
    for k in range(P):
        lo, hi = k * per, min(count, (k + 1) * per)
        if lo < hi:
            ends += [lo, hi - 1]
    extra = rng.integers(0, count, size=512)
    return np.unique(np.concatenate([np.array(ends, dtype=np.int64), extra.astype(np.int64)]))


@pytest.mark.parametrize("dtype", [MPI_INT, MPI_DOUBLE])
def test_collectives_1024_ranks_full_size(dtype, pkg, oracle, torch_gpu):
    torch = torch_gpu
    count = 1 << 22
    tdt = torch.float64 if dtype == MPI_DOUBLE else torch.int32
    inb = torch.zeros((P, count), dtype=tdt, device="cuda")
    outb = torch.zeros((P, count), dtype=tdt, device="cuda")
    idx = _sample_indices(count, np.random.default_rng(5))
    samples = [_gen_at(oracle, dtype, r, idx) for r in range(P)]
    root = 3
    chunk = 64
    with _runtime(pkg) as rt:
        # Bcast: every rank holds the root's input, which is the oracle's stream
        _run(pkg, rt, torch, "bcast", dtype, OP_SUM, count, root, inb, outb)
        assert np.array_equal(inb[root].cpu().numpy()[idx], samples[root])
        for r0 in range(0, P, chunk):
            assert bool((outb[r0:r0 + chunk] == inb[root][None, :]).all()), ("bcast", r0)
        # inputs of every rank sampled against the oracle
        got = inb[:, torch.from_numpy(idx).cuda()].cpu().numpy()
        for r in range(0, P, 31):
            assert np.array_equal(got[r], samples[r]), ("input", r)
        for op in (OP_SUM, OP_MAX):
            rc, exp_s = reduce(oracle, dtype, op, samples)
            assert rc == 0
            # an independent reduction of the device inputs (exact for integers and MAX)
            if op == OP_MAX:
                ref = inb.amax(dim=0)
            elif dtype == MPI_INT:
                ref = inb.sum(dim=0, dtype=torch.int64).to(torch.int32)
            else:
                ref = inb.sum(dim=0)
                norm = inb.abs().sum(dim=0)
            for kind, rroot in (("reduce", root), ("allreduce", 0)):
                outb.zero_()
                _run(pkg, rt, torch, kind, dtype, op, count, rroot, inb, outb)
                res = outb[rroot]
                assert np.array_equal(res.cpu().numpy()[idx], exp_s), (kind, op)
                if dtype == MPI_DOUBLE and op == OP_SUM:
                    assert bool(((res - ref).abs() <= 1e-12 * norm).all()), (kind, op)
                else:
                    assert bool(torch.equal(res, ref)), (kind, op)
                if kind == "allreduce":
                    for r0 in range(0, P, chunk):
                        assert bool((outb[r0:r0 + chunk] == res[None, :]).all()), ("allreduce", op, r0)
        # Alltoall: out[d] block s == in[s] block d
        per = count // P
        outb.zero_()
        _run(pkg, rt, torch, "alltoall", dtype, OP_SUM, count, 0, inb, outb)
        iv = inb.view(P, P, per)
        ov = outb.view(P, P, per)
        for d0 in range(0, P, chunk):
            exp = iv[:, d0:d0 + chunk, :].transpose(0, 1)
            assert bool(torch.equal(ov[d0:d0 + chunk], exp)), ("alltoall", d0)
