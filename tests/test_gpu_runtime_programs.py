"""Runtime behaviour beyond the reference's own unit tests, on the GPU through the C ABI:
the whole operator table (every legal (datatype, operator) pair of gpu_mpi/operators.cu
:153-200 plus MPI_Op_create user operators) bit-exact against the oracle's fold in the
reference's order, illegal pairs recorded as errors, LDS buffers in point-to-point and
collective calls, and the served-job rings with a queue smaller than the job stream."""
import ctypes
import time

import numpy as np
import pytest

from oracle_py import ptr, reduce

pytestmark = pytest.mark.gpu

# MPI_Datatype enum order (gpu_mpi/datatypes.cuh:12-14 + mpi_types_list.cuh:4-33)
NP_TYPES = {
    0: np.int8, 1: np.int16, 2: np.int32, 3: np.int64, 4: np.int64, 5: np.int64, 6: np.int8,
    7: np.uint8, 8: np.uint16, 9: np.uint32, 10: np.uint64, 11: np.uint64, 12: np.float32,
    13: np.float64, 14: np.float64, 15: np.int32, 16: np.uint8, 17: np.int8, 18: np.int16,
    19: np.int32, 20: np.int64, 21: np.uint8, 22: np.uint16, 23: np.uint32, 24: np.uint64,
    25: np.complex64, 26: np.complex64, 27: np.complex128, 28: np.complex128, 29: np.uint8,
}
OP_MAX, OP_MIN, OP_SUM, OP_PROD, OP_LAND, OP_BAND, OP_LOR, OP_BOR, OP_LXOR, OP_BXOR, \
    OP_MAXLOC, OP_MINLOC = range(12)
MPI_C_BOOL, MPI_INT64_T, MPI_DOUBLE, MPI_INT, MPI_FLOAT = 16, 20, 13, 2, 12


def _values(rng, dtype, op, n):
    t = NP_TYPES[dtype]
    if np.issubdtype(t, np.complexfloating):
        return (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(t)
    if np.issubdtype(t, np.floating):
        return rng.normal(size=n).astype(t)
    if dtype == MPI_C_BOOL:
        return rng.integers(0, 2, size=n).astype(t)
    if op in (OP_LAND, OP_LOR, OP_LXOR):
        return rng.choice(np.array([0, 0, 1, 5], dtype=t), size=n)
    if op == OP_PROD:
        return rng.integers(0, 4, size=n).astype(t) if np.issubdtype(t, np.unsignedinteger) \
            else rng.integers(-3, 4, size=n).astype(t)
    info = np.iinfo(t)
    if op == OP_SUM and np.issubdtype(t, np.signedinteger) and info.bits >= 32:
        # signed overflow is undefined in C on both sides: stay in range
        return rng.integers(-(1 << 20), (1 << 20) + 1, size=n).astype(t)
    return rng.integers(info.min, info.max, size=n, dtype=t, endpoint=True)


def _user_fold(dtype, op, ins):
    acc = ins[0].copy()
    for x in ins[1:]:
        if op == -1:
            acc = (acc.astype(np.uint64) * np.uint64(2) + x.astype(np.uint64)).astype(acc.dtype)
        else:
            acc = 0.5 * acc + x
    return acc


def _run_op_table(pkg, torch, P, cases, count, root, seed):
    from dphpc_project_amd import programs as pr
    rng = np.random.default_rng(seed)
    stride = count * 16
    n = len(cases)
    inp = np.zeros((n, P, stride), dtype=np.uint8)
    ins = []
    for c, (dt, op) in enumerate(cases):
        vals = [_values(rng, dt, op, count) for _ in range(P)]
        ins.append(vals)
        for r in range(P):
            b = vals[r].view(np.uint8)
            inp[c, r, :b.size] = b
    d_cases = torch.tensor(np.array(cases, dtype=np.int32).reshape(-1), device="cuda")
    d_in = torch.from_numpy(inp.reshape(-1)).cuda()
    d_all = torch.zeros(n * P * stride, dtype=torch.uint8, device="cuda")
    d_root = torch.zeros(n * stride, dtype=torch.uint8, device="cuda")
    d_st = torch.full((n * P,), -1, dtype=torch.int32, device="cuda")
    args = pr.OpTableArgs(d_cases.data_ptr(), n, count, root, stride, d_in.data_ptr(),
                          d_all.data_ptr(), d_root.data_ptr(), d_st.data_ptr())
    with pkg.Runtime(P, heap_bytes=64 << 20, timeout_s=20.0) as rt:
        rc, _ = rt.run("t_op_table", args, check=False)
        err = rt.last_error() if rc != 0 else ""
    return rc, err, ins, d_all.cpu().numpy().reshape(n, P, stride), \
        d_root.cpu().numpy().reshape(n, stride), d_st.cpu().numpy().reshape(n, P)


def test_every_legal_operator_pair_and_user_ops(pkg, oracle, torch_gpu):
    P, count, root = 7, 37, 4
    cases = [(dt, op) for dt in range(30) for op in range(12) if oracle.orc_op_allowed(dt, op)]
    assert len(cases) > 150
    assert not any(op in (OP_MAXLOC, OP_MINLOC) for _, op in cases)  # no legal type
    cases += [(MPI_INT64_T, -1), (MPI_DOUBLE, -2)]
    print(f"[gpu] op table: {len(cases)} cases on {P} ranks", flush=True)
    rc, err, ins, out_all, out_root, status = _run_op_table(pkg, torch_gpu, P, cases, count,
                                                            root, seed=21)
    assert rc == 0, err
    assert (status == 0).all()
    for c, (dt, op) in enumerate(cases):
        if op < 0:
            exp = _user_fold(dt, op, ins[c])
        else:
            rcode, exp = reduce(oracle, dt, op, ins[c])
            assert rcode == 0
        eb = exp.view(np.uint8)
        assert np.array_equal(out_root[c, :eb.size], eb), ("reduce", dt, op)
        for r in range(P):
            assert np.array_equal(out_all[c, r, :eb.size], eb), ("allreduce", dt, op, r)


@pytest.mark.parametrize("dt,op", [(MPI_INT, OP_MAXLOC), (MPI_DOUBLE, OP_MINLOC),
                                   (MPI_FLOAT, OP_BAND), (MPI_DOUBLE, OP_LXOR)])
def test_illegal_operator_pair_is_an_error(dt, op, pkg, oracle, torch_gpu):
    """operators.cu:202-208: an illegal pair is reported (the reference prints and aborts);
    here the first error is recorded, every rank's waits end, and the launch fails."""
    assert not oracle.orc_op_allowed(dt, op)
    rc, err, _, _, _, status = _run_op_table(pkg, torch_gpu, 5, [(dt, op)], 8, 0, seed=3)
    assert rc != 0
    assert "not allowed" in err, err


@pytest.mark.parametrize("rpb,push", [(1, 0), (4, 0), (4, 1)])
def test_lds_buffers(rpb, push, pkg, torch_gpu):
    """Point-to-point and collectives on LDS buffers; push=1 sends the Bcast into LDS
    through the push schedule (a staged buffer stored into by the other ranks)."""
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    P = 10
    print(f"[gpu] lds buffers P={P} rpb={rpb} push={push}", flush=True)
    ok = torch.zeros(P, dtype=torch.int32, device="cuda")
    with pkg.Runtime(P, ranks_per_block=rpb, heap_bytes=64 << 20, timeout_s=20.0,
                     bcast_push_min=1 if push else 0) as rt:
        rt.run("t_lds_buffers", pr.LdsArgs(ok.data_ptr()))
    assert ok.cpu().tolist() == [1] * P


def test_serve_queue_smaller_than_the_stream(pkg, oracle, torch_gpu):
    """A queue of 4 entries and 40 jobs: submit until the ring reports full, poll, repeat.
    Every job completes exactly once with its oracle result (no result overwritten before
    it was polled, no waiting job overwritten in the device's copy of the ring)."""
    from oracle_py import job_alltoall_checksum, job_dot
    jobs = []
    rng = np.random.default_rng(4)
    for i in range(40):
        P = int(rng.choice([8, 16, 32, 64]))
        if i % 3 == 0:
            jobs.append(pkg.Job(i, 500 + i, 1, P, int(rng.integers(1024, 4096))))
        else:
            jobs.append(pkg.Job(i, 500 + i, 0, P, int(rng.integers(1 << 10, 1 << 16))))
    print(f"[gpu] serve with queue_capacity=4, {len(jobs)} jobs", flush=True)
    srv = pkg.Server(256, queue_capacity=4, max_job_ranks=64, heap_fraction=0.9,
                     ranks_per_block=4, heap_bytes=1 << 30, timeout_s=30.0)
    got = {}
    nxt = 0
    full = 0
    t0 = time.time()
    try:
        while len(got) < len(jobs) and time.time() - t0 < 60:
            while nxt < len(jobs):
                if not srv.submit(jobs[nxt]):
                    full += 1
                    break
                nxt += 1
            for r in srv.poll():
                assert r.id not in got, f"job {r.id} reported twice"
                got[r.id] = r
    finally:
        srv.stop()
    assert full > 0, "the ring never filled"
    assert sorted(got) == list(range(len(jobs)))
    for j in jobs:
        r = got[j.id]
        assert r.status == 1, (j.id, r.status)
        if j.type == 0:
            assert r.dot == job_dot(oracle, j.seed, j.size, j.ranks), j.id
        else:
            assert r.checksum == job_alltoall_checksum(oracle, j.seed, j.size, j.ranks), j.id
