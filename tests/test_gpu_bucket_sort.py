"""The bucket-sort rank program (NPB IS rank()/full_verify through the device MPI API):
NPB IS class known answers, and for the exchange workload's keys the global order equal
to a plain sort of all keys, bit for bit."""
import numpy as np
import pytest

from oracle_py import keys_uniform, keys_zipf, zipf_cdf

pytestmark = pytest.mark.gpu

NPB = {"S": (16, 11, 9), "W": (20, 16, 10), "A": (23, 19, 10), "B": (25, 21, 10)}


def run_bucket_sort(pkg, torch, P, dist, n=0, key_bits=23, bucket_bits=10, cls="S", iters=10,
                    seed=0x5EED, cdf=None):
    from dphpc_project_amd import programs as pr
    if dist == 2:
        tk, key_bits, bucket_bits = NPB[cls]
        n = (1 << tk) // P
    NB = 1 << bucket_bits
    # a Zipf receiver can hold most of the keys (bucket 0 goes to rank 0 whole)
    kb2_cap = P * n if dist == 1 else 4 * n + 4096
    count_cap = 1 << key_bits
    dev = dict(device="cuda")
    ka = torch.zeros(P * kb2_cap, dtype=torch.int32, **dev)
    kb1 = torch.zeros(P * n, dtype=torch.int32, **dev)
    kb2 = torch.zeros(P * kb2_cap, dtype=torch.int32, **dev)
    counting = torch.zeros(P * count_cap, dtype=torch.int32, **dev)
    bs = torch.zeros(P * (NB + 5), dtype=torch.int32, **dev)
    bt = torch.zeros(P * (NB + 5), dtype=torch.int32, **dev)
    cnt = torch.zeros(7 * P * P, dtype=torch.int32, **dev)
    nlocal = torch.zeros(P, dtype=torch.int64, **dev)
    passed = torch.zeros(P, dtype=torch.int64, **dev)
    total = torch.zeros(1, dtype=torch.int64, **dev)
    sec = torch.zeros(P, dtype=torch.float64, **dev)
    cdf_t = torch.from_numpy(cdf).to("cuda") if cdf is not None else None
    args = pr.BucketSortArgs(dist, ord(cls), key_bits, bucket_bits, iters, 0, n, seed,
                             cdf_t.data_ptr() if cdf_t is not None else None,
                             len(cdf) if cdf is not None else 0, kb2_cap, count_cap,
                             ka.data_ptr(), kb1.data_ptr(), kb2.data_ptr(), counting.data_ptr(),
                             bs.data_ptr(), bt.data_ptr(), cnt.data_ptr(), nlocal.data_ptr(),
                             passed.data_ptr(), total.data_ptr(), sec.data_ptr())
    print(f"[gpu] bucket_sort P={P} dist={dist} cls={cls} n={n}", flush=True)
    with pkg.Runtime(P, heap_bytes=256 << 20, timeout_s=60.0) as rt:
        rt.run("bucket_sort", args)
    torch.cuda.synchronize()
    nl = nlocal.cpu().numpy()
    rows = ka.cpu().numpy().view(np.uint32).reshape(P, kb2_cap)
    return [rows[r, :nl[r]] for r in range(P)], passed.cpu().numpy(), int(total.item())


@pytest.mark.parametrize("cls,P", [("S", 4), ("S", 16), ("W", 16), ("W", 64), ("A", 64), ("B", 128)])
def test_npb_is_rank_program_verifies(pkg, torch_gpu, cls, P):
    """is.c:1137-1138: passed_verification == 5 * MAX_ITERATIONS + P."""
    _, passed, total = run_bucket_sort(pkg, torch_gpu, P, 2, cls=cls)
    assert total == 5 * 10 + P, (total, passed.tolist())


def test_npb_is_rank_program_matches_oracle_order(pkg, oracle, torch_gpu):
    """Class S on 8 ranks: every rank's sorted keys are the oracle's keys in order."""
    import ctypes
    from oracle_py import npb_class
    P = 8
    rows, _, total = run_bucket_sort(pkg, torch_gpu, P, 2, cls="S")
    assert total == 5 * 10 + P
    c = npb_class(oracle, "S")
    nk = (1 << 16) // P
    keys = np.zeros(P * nk, dtype=np.int32)
    for r in range(P):
        oracle.orc_npb_create_keys(ctypes.byref(c), P, r, keys[r * nk:].ctypes.data_as(ctypes.c_void_p))
    for it in range(1, 11):   # is.c:650-654 alterations of rank 0's keys
        keys[it] = it
        keys[it + 10] = (1 << 11) - it
    assert np.array_equal(np.concatenate(rows).astype(np.int64), np.sort(keys).astype(np.int64))


@pytest.mark.parametrize("dist", [0, 1])
def test_bucket_sort_global_order(pkg, oracle, torch_gpu, dist):
    P, n, kb, bb = 64, 4096, 20, 10
    cdf = zipf_cdf(1 << kb) if dist == 1 else None
    rows, passed, total = run_bucket_sort(pkg, torch_gpu, P, dist, n=n, key_bits=kb, bucket_bits=bb,
                                          iters=2, seed=0x77, cdf=cdf)
    if dist == 0:
        allk = np.concatenate([keys_uniform(oracle, 0x77, r, n, kb) for r in range(P)])
    else:
        allk = np.concatenate([keys_zipf(oracle, 0x77, r, n, cdf) for r in range(P)])
    assert total == P
    assert np.array_equal(np.concatenate(rows), np.sort(allk))
