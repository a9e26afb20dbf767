"""The allocation-free Alltoallv step (gpumpi_a2av_execute_noalloc) against the step of a
plan (gpumpi_a2av_plan_create + gpumpi_a2av_execute): the same receive buffers, counts and
displacements, bit for bit, on two calls in a row -- the second with fewer ranks and a
smaller input, so stale tables from the first call would show -- and the oracle's
restatement of MPI_Alltoallv (gpu_mpi/mpi.cu:363-408).  Rank counts above the static
workspace's 512 are refused before any launch."""
import ctypes

import numpy as np
import pytest

from oracle_py import alltoallv

pytestmark = pytest.mark.gpu


def _random_case(rng, P, elem, maxc):
    counts = rng.integers(0, maxc + 1, size=(P, P)).astype(np.int32)
    counts[rng.random((P, P)) < 0.2] = 0
    sdispls = np.zeros_like(counts)
    for s in range(P):
        off = int(rng.integers(0, 5))
        for d in rng.permutation(P):
            sdispls[s, d] = off
            off += int(counts[s, d]) + int(rng.integers(0, 3))
    sizes = [int((sdispls[s] + counts[s]).max()) for s in range(P)]
    send = [rng.integers(0, 256, size=max(1, sizes[s]) * elem, dtype=np.uint8) for s in range(P)]
    return send, counts, sdispls


def _buffers(torch, send, counts, elem):
    dev = torch.device("cuda")
    P = len(send)
    rtot = counts.sum(axis=0)
    d_send = [torch.from_numpy(b.copy()).to(dev) for b in send]
    d_recv = [torch.full((int(n) * elem + 16,), 0xAB, dtype=torch.uint8, device=dev) for n in rtot]
    sptr = torch.tensor([t.data_ptr() for t in d_send], dtype=torch.int64, device=dev)
    rptr = torch.tensor([t.data_ptr() for t in d_recv], dtype=torch.int64, device=dev)
    return d_send, d_recv, sptr, rptr


def _run(pkg, torch, send, counts, sdispls, elem, noalloc):
    lib = pkg.load_library()
    d_send, d_recv, sptr, rptr = _buffers(torch, send, counts, elem)
    P = len(send)
    d_sc = torch.from_numpy(np.ascontiguousarray(counts)).cuda()
    d_sd = torch.from_numpy(np.ascontiguousarray(sdispls)).cuda()
    d_rc = torch.full_like(d_sc, -7)
    d_rd = torch.full_like(d_sc, -7)
    torch.cuda.synchronize()
    if noalloc:
        rc = lib.gpumpi_a2av_execute_noalloc(P, elem, sptr.data_ptr(), d_sc.data_ptr(), d_sd.data_ptr(),
                                             rptr.data_ptr(), d_rc.data_ptr(), d_rd.data_ptr(), None)
        assert rc == 0, lib.gpumpi_last_error(None)
    else:
        plan = pkg.A2avPlan(P, elem, sptr.data_ptr(), d_sc.data_ptr(), d_sd.data_ptr(), rptr.data_ptr(),
                            d_rc.data_ptr(), d_rd.data_ptr())
        plan.execute()
        torch.cuda.synchronize()
        plan.close()
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in d_recv], d_rc.cpu().numpy(), d_rd.cpu().numpy()


@pytest.mark.parametrize("elem", [4, 1])
def test_noalloc_step_matches_the_plan_step(elem, pkg, oracle, torch_gpu):
    rng = np.random.default_rng(900 + elem)
    for P, maxc in ((96, 700), (24, 300)):  # the second call is smaller
        print(f"[gpu] noalloc Alltoallv P={P} elem={elem}", flush=True)
        send, counts, sdispls = _random_case(rng, P, elem, maxc)
        got, rc, rd = _run(pkg, torch_gpu, send, counts, sdispls, elem, True)
        ref, prc, prd = _run(pkg, torch_gpu, send, counts, sdispls, elem, False)
        exp, erc, erd = alltoallv(oracle, send, counts, sdispls, elem, counts.sum(axis=0))
        assert np.array_equal(rc, prc) and np.array_equal(rd, prd)
        assert np.array_equal(rc, erc) and np.array_equal(rd, erd)
        for d in range(P):
            assert np.array_equal(got[d], ref[d]), d  # guard bytes included
            n = int(counts[:, d].sum()) * elem
            assert np.array_equal(got[d][:n], exp[d][:n]), d


def test_noalloc_step_refuses_more_ranks_than_its_workspace(pkg, torch_gpu):
    lib = pkg.load_library()
    z = torch_gpu.zeros(4, dtype=torch_gpu.int64, device="cuda")
    p = z.data_ptr()
    assert lib.gpumpi_a2av_execute_noalloc(513, 4, p, p, p, p, p, p, None) == pkg.E_ARG
    assert b"512" in lib.gpumpi_last_error(None)
