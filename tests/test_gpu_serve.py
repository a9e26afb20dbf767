"""Served MPI jobs: every job's result is the same when it runs in the served mix,
alone on the plain path (one launch per job) and in the oracle's restatement."""
import numpy as np
import pytest

from oracle_py import job_alltoall_checksum, job_dot

pytestmark = pytest.mark.gpu


def make_jobs(pkg, n=40, seed=11):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        P = int(2 ** (3 + 5 * rng.random()))
        if i % 2 == 0:
            size = int(np.clip(np.exp(rng.normal(np.log(20000), 1.5)), 1 << 10, 1 << 18))
            jobs.append(pkg.Job(i, 1000 + i, 0, P, size))
        else:
            size = int(2 ** rng.uniform(10, 13)) + int(rng.integers(0, 7))
            jobs.append(pkg.Job(i, 1000 + i, 1, min(P, 64), size))
    return jobs


def test_served_equals_plain_equals_oracle(pkg, oracle, torch_gpu):
    torch = torch_gpu
    jobs = make_jobs(pkg)
    print(f"[gpu] serve {len(jobs)} jobs", flush=True)
    srv = pkg.Server(512, queue_capacity=256, max_job_ranks=256, heap_fraction=0.9,
                     ranks_per_block=4, heap_bytes=2 << 30, timeout_s=30.0)
    for j in jobs:
        assert srv.submit(j)
    import time
    served = {}
    t0 = time.time()
    while len(served) < len(jobs) and time.time() - t0 < 60:
        for r in srv.poll():
            served[r.id] = r
    stats = srv.stop()
    assert len(served) == len(jobs), f"only {len(served)} jobs completed"
    assert stats["admitted"] == len(jobs) and stats["max_concurrent"] > 1
    # plain path: each job alone
    res = torch.zeros(len(jobs) * 64, dtype=torch.uint8, device="cuda")
    rt = pkg.Runtime(256, heap_bytes=2 << 30, timeout_s=30.0)
    stride = 64
    for j in jobs:
        rt.run_job(j, res.data_ptr() + stride * j.id)
    torch.cuda.synchronize()
    raw = res.cpu().numpy().reshape(len(jobs), stride)
    rt.close()
    for j in jobs:
        s = served[j.id]
        p = pkg.JobResult.from_buffer_copy(raw[j.id].tobytes())
        assert s.status == 1 and p.status == 1, (j.id, s.status, p.status)
        if j.type == 0:
            exp = job_dot(oracle, j.seed, j.size, j.ranks)
            assert s.dot == p.dot == exp, (j.id, s.dot, p.dot, exp)
        else:
            exp = job_alltoall_checksum(oracle, j.seed, j.size, j.ranks)
            assert s.checksum == p.checksum == exp, (j.id, s.checksum, p.checksum, exp)


def test_server_rejects_what_never_fits(pkg, torch_gpu):
    srv = pkg.Server(128, queue_capacity=16, max_job_ranks=64, heap_bytes=64 << 20)
    big = pkg.Job(1, 1, 1, 64, 1 << 20)     # 2 * 64 * 64 MiB: never fits a 64 MiB heap
    ok = pkg.Job(2, 2, 0, 8, 4096)
    assert srv.submit(big) and srv.submit(ok)
    import time
    got = {}
    t0 = time.time()
    while len(got) < 2 and time.time() - t0 < 30:
        for r in srv.poll():
            got[r.id] = r
    stats = srv.stop()
    assert got[1].status == -1 and got[2].status == 1
    assert stats["rejected"] == 1


def test_full_size_jobs_served_plain_and_oracle(pkg, oracle, torch_gpu):
    """Config-5 jobs at the stream's extremes: a dot product of 2^22 FP64 elements on 256
    and on 8 ranks, and Alltoalls of 16 MiB per pair on 8 ranks, 256 KiB on 64 and 4 KiB
    on 256.  Served, plain and oracle results agree (dots bit-exact: the same rank-order
    Allreduce)."""
    torch = torch_gpu
    jobs = [pkg.Job(0, 71, 0, 256, 1 << 22), pkg.Job(1, 72, 0, 8, 1 << 22),
            pkg.Job(2, 73, 1, 8, 16 << 20), pkg.Job(3, 74, 1, 64, 256 << 10),
            pkg.Job(4, 75, 1, 256, 4 << 10)]
    print(f"[gpu] full-size served jobs: {len(jobs)}", flush=True)
    heap = 12 << 30
    srv = pkg.Server(512, queue_capacity=16, max_job_ranks=256, heap_fraction=0.9,
                     ranks_per_block=4, heap_bytes=heap, timeout_s=60.0)
    import time
    served = {}
    try:
        for j in jobs:
            assert srv.submit(j)
        t0 = time.time()
        while len(served) < len(jobs) and time.time() - t0 < 120:
            for r in srv.poll():
                served[r.id] = r
    finally:
        srv.stop()
    assert sorted(served) == list(range(len(jobs)))
    res = torch.zeros(len(jobs) * 64, dtype=torch.uint8, device="cuda")
    with pkg.Runtime(256, heap_bytes=heap, timeout_s=60.0) as rt:
        for j in jobs:
            rt.run_job(j, res.data_ptr() + 64 * j.id)
        torch.cuda.synchronize()
    raw = res.cpu().numpy().reshape(len(jobs), 64)
    for j in jobs:
        s = served[j.id]
        p = pkg.JobResult.from_buffer_copy(raw[j.id].tobytes())
        assert s.status == 1 and p.status == 1, (j.id, s.status, p.status)
        if j.type == 0:
            exp = job_dot(oracle, j.seed, j.size, j.ranks)
            assert s.dot == p.dot == exp, (j.id, s.dot, p.dot, exp)
        else:
            exp = job_alltoall_checksum(oracle, j.seed, j.size, j.ranks)
            assert s.checksum == p.checksum == exp, (j.id, s.checksum, p.checksum, exp)
