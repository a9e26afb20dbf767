"""Batched persistent-kernel Alltoallv (gpumpi_a2av_*) against the oracle.

Bar: received bytes, receive counts and receive displacements bit-exact with the
oracle's Alltoallv restatement (gpu_mpi/mpi.cu:363-408 semantics) on the same inputs.
"""
import numpy as np
import pytest

from oracle_py import alltoallv, bucket_plan, keys_uniform, keys_zipf, zipf_cdf

pytestmark = pytest.mark.gpu


def run_exchange(pkg, torch, sendbufs, counts, sdispls, elem, tune=None, offsets=None,
                 graph_steps=0, autotune=False):
    """Run the batched step; sendbufs are numpy uint8 arrays (one per rank).
    offsets: per-rank byte offset of the receive buffer inside a padded allocation
    (exercises unaligned destinations)."""
    dev = torch.device("cuda")
    P = len(sendbufs)
    rtot = counts.sum(axis=0)
    offsets = offsets if offsets is not None else [0] * P
    d_send = [torch.from_numpy(b.copy()).to(dev) for b in sendbufs]
    d_recv_alloc = [torch.full((int(n) * elem + offsets[d] + 16,), 0xAB, dtype=torch.uint8, device=dev)
                    for d, n in enumerate(rtot)]
    sptr = torch.tensor([t.data_ptr() for t in d_send], dtype=torch.int64, device=dev)
    rptr = torch.tensor([t.data_ptr() + offsets[d] for d, t in enumerate(d_recv_alloc)],
                        dtype=torch.int64, device=dev)
    d_sc = torch.from_numpy(np.ascontiguousarray(counts, dtype=np.int32)).to(dev)
    d_sd = torch.from_numpy(np.ascontiguousarray(sdispls, dtype=np.int32)).to(dev)
    d_rc = torch.zeros_like(d_sc)
    d_rd = torch.zeros_like(d_sc)
    torch.cuda.synchronize()
    plan = pkg.A2avPlan(P, elem, sptr.data_ptr(), d_sc.data_ptr(), d_sd.data_ptr(),
                        rptr.data_ptr(), d_rc.data_ptr(), d_rd.data_ptr(), tune=tune)
    if autotune:
        plan.autotune()
    if graph_steps:
        plan.graph_replay(graph_steps)
    else:
        plan.execute()
    torch.cuda.synchronize()
    nbytes = plan.payload_bytes()
    plan.close()
    got = []
    for d, t in enumerate(d_recv_alloc):
        a = t.cpu().numpy()
        n = int(rtot[d]) * elem
        # guard bytes around the destination must be untouched
        assert np.all(a[:offsets[d]] == 0xAB) and np.all(a[offsets[d] + n:] == 0xAB), f"rank {d} overrun"
        got.append(a[offsets[d]:offsets[d] + n])
    return got, d_rc.cpu().numpy(), d_rd.cpu().numpy(), nbytes


def random_case(rng, P, elem, maxc, permute=True):
    counts = rng.integers(0, maxc + 1, size=(P, P)).astype(np.int32)
    counts[rng.random((P, P)) < 0.2] = 0
    sdispls = np.zeros_like(counts)
    for s in range(P):
        order = rng.permutation(P) if permute else np.arange(P)
        off = int(rng.integers(0, 5))  # leading gap
        for d in order:
            sdispls[s, d] = off
            off += int(counts[s, d]) + int(rng.integers(0, 3))  # gaps between blocks
    sizes = [int((sdispls[s] + counts[s]).max()) if P else 0 for s in range(P)]
    send = [rng.integers(0, 256, size=max(1, sizes[s]) * elem, dtype=np.uint8) for s in range(P)]
    return send, counts, sdispls


@pytest.mark.parametrize("defer", [0, 1])
@pytest.mark.parametrize("P,elem,maxc", [(1, 4, 100), (2, 1, 50), (7, 1, 333), (16, 4, 300),
                                         (16, 8, 120), (64, 4, 64), (33, 1, 2000), (130, 1, 3)])
def test_a2av_random(P, elem, maxc, defer, pkg, oracle, torch_gpu):
    print(f"[gpu] a2av random P={P} elem={elem} maxc={maxc} defer={defer}", flush=True)
    rng = np.random.default_rng(P * 1000 + elem)
    send, counts, sdispls = random_case(rng, P, elem, maxc)
    offsets = [int(rng.integers(0, 16)) for _ in range(P)]
    got, rc, rd, nbytes = run_exchange(pkg, torch_gpu, send, counts, sdispls, elem, offsets=offsets,
                                       tune=(0, -1, 0, -1, 0, defer))
    exp, erc, erd = alltoallv(oracle, send, counts, sdispls, elem, counts.sum(axis=0))
    assert nbytes == int(counts.sum()) * elem
    assert np.array_equal(rc, erc) and np.array_equal(rd, erd)
    for d in range(P):
        assert np.array_equal(got[d], exp[d][:got[d].size]), f"rank {d} payload differs"


@pytest.mark.parametrize("tune", [(256, 0, 2, 0, 1, 0), (256, 16, 2, 1, 2, 0), (256, 1, 8, 1, 1, 0),
                                  (256, 4, 4, 0, 4, 0), (256, 0, 4, 1, 8, 0), (256, 0, 2, 0, 1, 1),
                                  (256, 16, 4, 1, 2, 1), (256, 1, 8, 1, 1, 1), (256, 0, 4, 1, 4, 1)])
def test_a2av_launch_shapes(tune, pkg, oracle, torch_gpu):
    print(f"[gpu] a2av launch shape {tune}", flush=True)
    rng = np.random.default_rng(11)
    send, counts, sdispls = random_case(rng, 24, 4, 700)
    got, rc, rd, _ = run_exchange(pkg, torch_gpu, send, counts, sdispls, 4, tune=tune)
    exp, _, _ = alltoallv(oracle, send, counts, sdispls, 4, counts.sum(axis=0))
    for d in range(24):
        assert np.array_equal(got[d], exp[d][:got[d].size])


def test_a2av_empty_and_single_heavy(pkg, oracle, torch_gpu):
    """All-zero rows, an all-zero matrix corner and one receiver getting most bytes."""
    print("[gpu] a2av empty and heavy", flush=True)
    P = 9
    counts = np.zeros((P, P), dtype=np.int32)
    counts[:, 0] = 50000          # everyone sends a lot to rank 0
    counts[3, :] = 0              # rank 3 sends nothing
    counts[5, 7] = 1
    sdispls = np.zeros_like(counts)
    sdispls[:, 1:] = np.cumsum(counts, axis=1)[:, :-1]
    rng = np.random.default_rng(3)
    send = [rng.integers(0, 256, size=max(1, int(counts[s].sum())) * 4, dtype=np.uint8) for s in range(P)]
    exp, _, _ = alltoallv(oracle, send, counts, sdispls, 4, counts.sum(axis=0))
    for defer in (0, 1):
        got, rc, rd, _ = run_exchange(pkg, torch_gpu, send, counts, sdispls, 4,
                                      tune=(0, -1, 0, -1, 0, defer))
        for d in range(P):
            assert np.array_equal(got[d], exp[d][:got[d].size]), f"defer={defer} rank {d}"


def test_a2av_graph_and_autotune(pkg, oracle, torch_gpu):
    print("[gpu] a2av graph replay + autotune", flush=True)
    rng = np.random.default_rng(5)
    send, counts, sdispls = random_case(rng, 32, 4, 2000, permute=False)
    got, _, _, _ = run_exchange(pkg, torch_gpu, send, counts, sdispls, 4, graph_steps=3,
                                autotune=True)
    exp, _, _ = alltoallv(oracle, send, counts, sdispls, 4, counts.sum(axis=0))
    for d in range(32):
        assert np.array_equal(got[d], exp[d][:got[d].size])


@pytest.mark.parametrize("defer", [0, 1])
@pytest.mark.parametrize("pipeline", [1, 0])
@pytest.mark.parametrize("steps", [1, 2, 3, 10, 11])
def test_a2av_pipelined_graphs(steps, pipeline, defer, pkg, oracle, torch_gpu):
    """Graphs of 1..11 steps, with each step's copy kernel planning the next step into the
    other workspace (pipeline 1) or three kernels per step (0): the receive buffers,
    counts and displacements after the graph are the oracle's."""
    print(f"[gpu] a2av pipelined graph steps={steps} pipeline={pipeline} defer={defer}", flush=True)
    rng = np.random.default_rng(100 + steps)
    send, counts, sdispls = random_case(rng, 24, 4, 900)
    got, rc, rd, nbytes = run_exchange(pkg, torch_gpu, send, counts, sdispls, 4,
                                       tune=(0, -1, 0, -1, 0, defer, pipeline), graph_steps=steps)
    exp, erc, erd = alltoallv(oracle, send, counts, sdispls, 4, counts.sum(axis=0))
    assert nbytes == int(counts.sum()) * 4
    assert np.array_equal(rc, erc) and np.array_equal(rd, erd)
    for d in range(24):
        assert np.array_equal(got[d], exp[d][:got[d].size]), f"rank {d} payload differs"


def test_a2av_pipelined_graph_replans_new_inputs(pkg, oracle, torch_gpu):
    """A pipelined graph plans its first step from the inputs as they are at launch: new
    counts, displacements and payloads written between two replays give the new
    exchange (buffers sized for both)."""
    torch = torch_gpu
    dev = torch.device("cuda")
    P, elem = 20, 4
    rng = np.random.default_rng(77)
    cases = [random_case(rng, P, elem, 500), random_case(rng, P, elem, 500)]
    cap_s = max(max(b.size for b in c[0]) for c in cases)
    cap_r = max(int(c[1].sum(axis=0).max()) for c in cases) * elem
    d_send = [torch.zeros(cap_s, dtype=torch.uint8, device=dev) for _ in range(P)]
    d_recv = [torch.zeros(cap_r + 16, dtype=torch.uint8, device=dev) for _ in range(P)]
    sptr = torch.tensor([t.data_ptr() for t in d_send], dtype=torch.int64, device=dev)
    rptr = torch.tensor([t.data_ptr() for t in d_recv], dtype=torch.int64, device=dev)
    d_sc = torch.zeros((P, P), dtype=torch.int32, device=dev)
    d_sd = torch.zeros_like(d_sc)
    d_rc = torch.zeros_like(d_sc)
    d_rd = torch.zeros_like(d_sc)
    plan = pkg.A2avPlan(P, elem, sptr.data_ptr(), d_sc.data_ptr(), d_sd.data_ptr(),
                        rptr.data_ptr(), d_rc.data_ptr(), d_rd.data_ptr(),
                        tune=(0, -1, 0, -1, 0, 0, 1))
    try:
        for send, counts, sdispls in cases + cases[:1]:
            for s_ in range(P):
                d_send[s_][:send[s_].size].copy_(torch.from_numpy(send[s_]))
            d_sc.copy_(torch.from_numpy(counts))
            d_sd.copy_(torch.from_numpy(sdispls))
            torch.cuda.synchronize()
            plan.graph_replay(5)
            torch.cuda.synchronize()
            exp, erc, erd = alltoallv(oracle, send, counts, sdispls, elem, counts.sum(axis=0))
            assert np.array_equal(d_rc.cpu().numpy(), erc)
            assert np.array_equal(d_rd.cpu().numpy(), erd)
            for d in range(P):
                n = int(counts[:, d].sum()) * elem
                assert np.array_equal(d_recv[d][:n].cpu().numpy(), exp[d][:n]), d
    finally:
        plan.close()


def test_a2av_pipeline_times(pkg, oracle, torch_gpu):
    torch = torch_gpu
    dev = torch.device("cuda")
    rng = np.random.default_rng(9)
    send, counts, sdispls = random_case(rng, 16, 4, 3000)
    P = 16
    d_send = [torch.from_numpy(b.copy()).to(dev) for b in send]
    rtot = counts.sum(axis=0)
    d_recv = [torch.zeros(int(n) * 4 + 16, dtype=torch.uint8, device=dev) for n in rtot]
    sptr = torch.tensor([t.data_ptr() for t in d_send], dtype=torch.int64, device=dev)
    rptr = torch.tensor([t.data_ptr() for t in d_recv], dtype=torch.int64, device=dev)
    d_sc = torch.from_numpy(counts).to(dev)
    d_sd = torch.from_numpy(sdispls).to(dev)
    d_rc = torch.zeros_like(d_sc)
    d_rd = torch.zeros_like(d_sc)
    plan = pkg.A2avPlan(P, 4, sptr.data_ptr(), d_sc.data_ptr(), d_sd.data_ptr(),
                        rptr.data_ptr(), d_rc.data_ptr(), d_rd.data_ptr())
    try:
        t = plan.pipeline_times(6)
        assert t["plan_scan_us"] > 0 and t["copy_us"] > 0 and t["step_us"] > 0
        torch.cuda.synchronize()
        assert plan.payload_bytes() == int(counts.sum()) * 4
        exp, _, _ = alltoallv(oracle, send, counts, sdispls, 4, rtot)
        for d in range(P):
            n = int(rtot[d]) * 4
            assert np.array_equal(d_recv[d][:n].cpu().numpy(), exp[d][:n]), d
    finally:
        plan.close()


@pytest.mark.parametrize("dist", ["uniform", "zipf"])
def test_bucket_exchange_small(dist, pkg, oracle, torch_gpu):
    """NPB-style bucket exchange of the workload keys at a reduced size."""
    print(f"[gpu] bucket exchange small {dist}", flush=True)
    P, n, key_bits, bucket_bits = 32, 4096, 23, 10
    if dist == "uniform":
        keys = np.stack([keys_uniform(oracle, 0x5eed, r, n, key_bits) for r in range(P)])
    else:
        cdf = zipf_cdf(1 << key_bits)
        keys = np.stack([keys_zipf(oracle, 0x5eed, r, n, cdf) for r in range(P)])
    kb1, sc, sd, lo, hi = bucket_plan(oracle, keys, key_bits, bucket_bits)
    send = [kb1[s].view(np.uint8) for s in range(P)]
    got, rc, rd, _ = run_exchange(pkg, torch_gpu, send, sc, sd, 4)
    exp, _, _ = alltoallv(oracle, send, sc, sd, 4, sc.sum(axis=0))
    allk = []
    for d in range(P):
        assert np.array_equal(got[d], exp[d][:got[d].size])
        allk.append(np.sort(got[d].view(np.uint32)))
    # global order after the local sort equals a plain sort of all keys
    assert np.array_equal(np.concatenate(allk), np.sort(keys.ravel()))


@pytest.mark.parametrize("dist", ["uniform", "zipf", "npb"])
def test_gen_keys_and_bucket_plan_match_oracle(dist, pkg, oracle, torch_gpu):
    """Device key generators and the NPB-style bucket plan are bit-exact with the oracle."""
    import ctypes
    from oracle_py import npb_class, ptr
    torch = torch_gpu
    print(f"[gpu] gen_keys + bucket_plan {dist}", flush=True)
    P, n = 16, 4096
    key_bits, bucket_bits = (11, 9) if dist == "npb" else (23, 10)
    dev = torch.device("cuda")
    keys = torch.empty(P * n, dtype=torch.int32, device=dev)
    cdf_np = zipf_cdf(1 << key_bits) if dist == "zipf" else None
    cdf = torch.from_numpy(cdf_np).to(dev) if cdf_np is not None else None
    pkg.gen_keys(keys.data_ptr(), P, n, 0x5eed, 0, dist, key_bits,
                 cdf.data_ptr() if cdf is not None else 0, cdf.numel() if cdf is not None else 0)
    torch.cuda.synchronize()
    got = keys.cpu().numpy().view(np.uint32).reshape(P, n)
    if dist == "uniform":
        exp = np.stack([keys_uniform(oracle, 0x5eed, r, n, key_bits) for r in range(P)])
    elif dist == "zipf":
        exp = np.stack([keys_zipf(oracle, 0x5eed, r, n, cdf_np) for r in range(P)])
    else:
        c = npb_class(oracle, "S")
        exp = np.zeros((P, n), dtype=np.int32)
        for r in range(P):
            row = np.zeros(n, dtype=np.int32)
            oracle.orc_npb_create_keys(ctypes.byref(c), P, r, ptr(row))
            exp[r] = row
        exp = exp.view(np.uint32)
    assert np.array_equal(got, exp)
    kb1 = torch.empty_like(keys)
    sc = torch.zeros((P, P), dtype=torch.int32, device=dev)
    sd = torch.zeros_like(sc)
    lo = torch.zeros(P, dtype=torch.int32, device=dev)
    hi = torch.zeros_like(lo)
    pkg.bucket_plan(keys.data_ptr(), P, n, key_bits, bucket_bits, kb1.data_ptr(), sc.data_ptr(),
                    sd.data_ptr(), lo.data_ptr(), hi.data_ptr())
    e_kb1, e_sc, e_sd, e_lo, e_hi = bucket_plan(oracle, exp, key_bits, bucket_bits)
    assert np.array_equal(kb1.cpu().numpy().view(np.uint32).reshape(P, n), e_kb1)
    assert np.array_equal(sc.cpu().numpy(), e_sc)
    assert np.array_equal(sd.cpu().numpy(), e_sd)
    assert np.array_equal(lo.cpu().numpy(), e_lo) and np.array_equal(hi.cpu().numpy(), e_hi)


@pytest.mark.parametrize("dist", ["uniform", "zipf"])
def test_full_size_exchange(dist, pkg, oracle, torch_gpu):
    """512 ranks x 2^18 keys: every destination bit-exact against the oracle's Alltoallv."""
    import importlib.util, os
    torch = torch_gpu
    print(f"[gpu] full-size exchange {dist}", flush=True)
    spec = importlib.util.spec_from_file_location("bench", os.path.join(os.path.dirname(__file__), "..", "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    ex = bench.Exchange(pkg, torch, 512, 1 << 18, dist)
    plan = ex.plan(pkg)
    plan.graph_replay(2)
    torch.cuda.synchronize()
    P, n = ex.P, ex.n
    kb1 = ex.kb1.cpu().numpy().view(np.uint32).reshape(P, n)
    send = [kb1[s].view(np.uint8) for s in range(P)]
    exp, erc, erd = alltoallv(oracle, send, ex.counts.astype(np.int32), ex.sd.cpu().numpy(), 4, ex.rtot)
    recv = ex.recv.cpu().numpy().view(np.uint8)
    assert np.array_equal(ex.rc.cpu().numpy(), erc) and np.array_equal(ex.rd.cpu().numpy(), erd)
    for d in range(P):
        o = int(ex.roff[d]) * 4
        assert np.array_equal(recv[o:o + int(ex.rtot[d]) * 4], exp[d][:int(ex.rtot[d]) * 4]), d
    plan.close()


@pytest.mark.parametrize("cls,P", [("S", 1), ("S", 4), ("S", 16), ("S", 64), ("W", 8), ("W", 32),
                                   ("A", 16), ("B", 128), ("C", 512)])
def test_npb_is_known_answers(cls, P, pkg, torch_gpu):
    """NPB IS partial + full verification through the batched exchange: 5*10 + P passes
    (the class known answers of is.c:264-287 with the per-iteration shifts of 849-905).
    Class C on 512 ranks is the exchange workload's size: 2^27 keys < 2^23, 2^18 per rank."""
    print(f"[gpu] NPB IS class {cls} on {P} ranks", flush=True)
    passed, ms = pkg.npb_is(cls, P)
    assert passed == 5 * 10 + P, passed


@pytest.mark.parametrize("dist", ["uniform", "zipf"])
def test_local_sort_gives_global_order(dist, pkg, torch_gpu):
    torch = torch_gpu
    import importlib.util, os
    print(f"[gpu] bucket exchange + local sort {dist}", flush=True)
    spec = importlib.util.spec_from_file_location("bench", os.path.join(os.path.dirname(__file__), "..", "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    ex = bench.Exchange(pkg, torch, 64, 1 << 16, dist)
    plan = ex.plan(pkg)
    plan.execute()
    torch.cuda.synchronize()
    out = torch.zeros_like(ex.recv)
    total = int(ex.rtot.sum())
    pkg.local_sort(ex.recv.data_ptr(), total, 23, out.data_ptr())
    got = out.cpu().numpy().view(np.uint32)[:total]
    assert np.array_equal(got, np.sort(ex.keys.cpu().numpy().view(np.uint32)))
    plan.close()
