#!/usr/bin/env python3
"""Benchmark: batched persistent-kernel MPI_Alltoallv on one MI355X.

Workload (the bucket-sort exchange): P = 512 ranks, 2^18 uint32 keys per rank (2^27
keys, 512 MiB) uniform in [0, 2^23), bucketed NPB-style into 1024 buckets (NPB IS
rank(): histogram, splitters, stable bucketing).  One timed step = the batched exchange
kernel: Alltoall of the counts + receive displacements + the skewed Alltoallv of the
keys, one persistent kernel replayed from a hipGraph.  Inputs are resident in HBM
before the timed region.

value  = algorithmic bytes / time = 2 x payload bytes (read once + written once) per
         step / seconds per step, in GB/s.
Also reported: the plain per-call path (one rank-grid kernel launch per MPI call:
MPI_Alltoall then MPI_Alltoallv on 512 wavefront ranks) and the ratio, the Zipf(1.1)
variant, and the host-MPI CPU baseline (the same exchange under MPICH, one rank per
host core, 2^18 keys per rank).
"""
import argparse
import ctypes
import importlib.util
import json
import os
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "dphpc-project_amd")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    name = "dphpc_project_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class Exchange:
    """Device buffers of one bucket-sort exchange (keys, bucketed send buffers, counts)."""

    def __init__(self, pkg, torch, P, n, dist, seed=0x5eed, key_bits=23, bucket_bits=10):
        import numpy as np
        dev = torch.device("cuda")
        self.P, self.n = P, n
        self.keys = torch.empty(P * n, dtype=torch.int32, device=dev)
        self.kb1 = torch.empty(P * n, dtype=torch.int32, device=dev)
        self.sc = torch.zeros((P, P), dtype=torch.int32, device=dev)
        self.sd = torch.zeros((P, P), dtype=torch.int32, device=dev)
        self.rc = torch.zeros((P, P), dtype=torch.int32, device=dev)
        self.rd = torch.zeros((P, P), dtype=torch.int32, device=dev)
        self.lo = torch.zeros(P, dtype=torch.int32, device=dev)
        self.hi = torch.zeros(P, dtype=torch.int32, device=dev)
        cdf = None
        if dist == "zipf":
            w = np.arange(1, (1 << key_bits) + 1, dtype=np.float64) ** -1.1
            c = np.cumsum(w)
            cdf = torch.from_numpy(c / c[-1]).to(dev)
        stream = torch.cuda.current_stream().cuda_stream
        pkg.gen_keys(self.keys.data_ptr(), P, n, seed, 0, dist, key_bits,
                     cdf.data_ptr() if cdf is not None else 0, cdf.numel() if cdf is not None else 0,
                     stream)
        pkg.bucket_plan(self.keys.data_ptr(), P, n, key_bits, bucket_bits, self.kb1.data_ptr(),
                        self.sc.data_ptr(), self.sd.data_ptr(), self.lo.data_ptr(),
                        self.hi.data_ptr(), stream)
        torch.cuda.synchronize()
        counts = self.sc.cpu().numpy().astype(np.int64)
        rtot = counts.sum(axis=0)
        self.payload = int(counts.sum()) * 4
        self.recv = torch.zeros(int(rtot.sum()) + 64, dtype=torch.int32, device=dev)
        roff = np.concatenate([[0], np.cumsum(rtot)[:-1]])
        base = self.recv.data_ptr()
        self.sptr = torch.tensor([self.kb1.data_ptr() + 4 * n * s for s in range(P)],
                                 dtype=torch.int64, device=dev)
        self.rptr = torch.tensor([base + 4 * int(o) for o in roff], dtype=torch.int64, device=dev)
        self.roff, self.rtot, self.counts = roff, rtot, counts
        self.key_bits, self.bucket_bits = key_bits, bucket_bits
        torch.cuda.synchronize()

    def plan(self, pkg, tune=None):
        return pkg.A2avPlan(self.P, 4, self.sptr.data_ptr(), self.sc.data_ptr(), self.sd.data_ptr(),
                            self.rptr.data_ptr(), self.rc.data_ptr(), self.rd.data_ptr(), tune=tune)

    def verify(self, torch, samples=6):
        """Exact check of sampled destinations against the send blocks, and global
        properties of all destinations (key ranges, counts, key sum)."""
        import numpy as np
        P, n = self.P, self.n
        shift = self.key_bits - self.bucket_bits
        rc = self.rc.cpu().numpy()
        assert np.array_equal(rc, self.counts.T), "receive counts differ from the transposed send counts"
        kb1 = self.kb1.cpu().numpy().view(np.uint32).reshape(P, n)
        sd = self.sd.cpu().numpy()
        recv = self.recv.cpu().numpy().view(np.uint32)
        lo, hi = self.lo.cpu().numpy(), self.hi.cpu().numpy()
        rng = np.random.default_rng(1)
        for d in set([0, P - 1] + list(rng.integers(0, P, size=samples))):
            exp = np.concatenate([kb1[s, sd[s, d]:sd[s, d] + self.counts[s, d]] for s in range(P)])
            got = recv[self.roff[d]:self.roff[d] + self.rtot[d]]
            assert np.array_equal(got, exp), f"destination {d} differs"
        b = (recv[:int(self.rtot.sum())] >> shift).astype(np.int64)
        owner = np.repeat(np.arange(P), self.rtot)
        assert np.all(b >= lo[owner]) and np.all(b <= hi[owner]), "a key reached the wrong rank"
        assert int(recv[:int(self.rtot.sum())].astype(np.int64).sum()) == int(kb1.astype(np.int64).sum())


def time_steps(torch, fn, steps):
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps  # ms per step


def cpu_baseline(args):
    """The same exchange under host MPI (MPICH), one rank per host core, timed by the
    oracle's MPI program.  Returns the cpu_baseline object or None."""
    exe = os.path.join(ROOT, "oracle", "_build", "a2av_bucket")
    mpirun = shutil.which("mpirun") or ("/opt/conda/bin/mpirun" if os.path.exists("/opt/conda/bin/mpirun") else None)
    if not os.path.exists(exe) or not mpirun:
        log("cpu baseline skipped: host MPI program or mpirun missing")
        return None
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 64)
    n = 1 << args.keys_log2
    cmd = [mpirun, "-n", str(cores), exe, str(n), "0", "0x5eed", str(args.cpu_iters), "2"]
    env = dict(os.environ, MPICH_CC="gcc")
    t0 = time.time()
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=240, env=env)
    except subprocess.TimeoutExpired:
        log("cpu baseline timed out")
        return None
    line = [l for l in out.stdout.splitlines() if l.startswith("a2av ")]
    if out.returncode != 0 or not line:
        log("cpu baseline failed:", out.stdout[-500:], out.stderr[-500:])
        return None
    kv = dict(x.split("=") for x in line[0].split()[1:])
    secs = float(kv["seconds"])
    payload = int(kv["payload_bytes"])
    log(f"cpu baseline: {cores} ranks, {payload} B payload, {secs*1e3:.3f} ms/exchange "
        f"(wall {time.time()-t0:.1f} s)")
    return {"value": round(2 * payload / secs / 1e9, 3), "unit": "GB/s", "cores": cores,
            "kind": "port",
            "sample": f"MPICH 3.3.2 MPI_Alltoall+MPI_Alltoallv bucket exchange, {cores} ranks x "
                      f"2^{args.keys_log2} uniform uint32 keys (one rank per host core), "
                      f"{args.cpu_iters} timed repetitions, sorted_ok={kv['sorted_ok']}"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ranks", type=int, default=512)
    ap.add_argument("--keys-log2", type=int, default=18)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-zipf", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune", type=str, default="", help="block_threads,blocks_per_cu,lines,nt")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 512 MiB device copy (scale for rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; the exchange of one "
                                   "GPU's ranks is measured (see DESIGN.md for sharding)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    pkg = load_pkg()
    pkg.load_library()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    P, n = args.ranks, 1 << args.keys_log2
    stream = torch.cuda.current_stream().cuda_stream
    log(f"device: {torch.cuda.get_device_name(0)}; info {pkg.device_info()}")

    # ---------------- batched path, uniform keys
    ex = Exchange(pkg, torch, P, n, "uniform")
    log(f"uniform: payload {ex.payload} B, max recv {int(ex.rtot.max())} keys")
    tune = tuple(int(x) for x in args.tune.split(",")) if args.tune else None
    plan = ex.plan(pkg, tune)
    if not args.no_autotune and tune is None:
        chosen, us = plan.autotune(stream)
        log(f"autotune: {chosen} best {us:.1f} us")
    for _ in range(max(1, args.warmup)):
        plan.graph_replay(1, stream)
    torch.cuda.synchronize()
    ex.verify(torch)
    assert plan.payload_bytes() == ex.payload
    ms = time_steps(torch, lambda: plan.graph_replay(1, stream), args.steps)
    algo = 2 * ex.payload
    gbs = algo / (ms * 1e-3) / 1e9
    log(f"batched: {ms*1e3:.1f} us/step, {gbs:.1f} GB/s ({100*gbs/HBM_PEAK_GBS:.1f}% of HBM peak), "
        f"tune {plan.tune()}")
    ex.verify(torch)
    tune_used = plan.tune()

    # ---------------- plain per-call path (one rank-grid launch per MPI call)
    plain = None
    if not args.no_plain:
        from dphpc_project_amd import programs as pr
        rt = pkg.Runtime(P, heap_bytes=256 << 20, timeout_s=60.0)
        a0 = pr.ExchArgs(0, 4, ex.sptr.data_ptr(), ex.sc.data_ptr(), ex.sd.data_ptr(),
                         ex.rptr.data_ptr(), ex.rc.data_ptr(), ex.rd.data_ptr())
        a1 = pr.ExchArgs(1, 4, ex.sptr.data_ptr(), ex.sc.data_ptr(), ex.sd.data_ptr(),
                         ex.rptr.data_ptr(), ex.rc.data_ptr(), ex.rd.data_ptr())
        ex.recv.zero_()

        def plain_step():
            rt.run_async("exchange", a0, stream)
            rt.run_async("exchange", a1, stream)

        plain_step()
        torch.cuda.synchronize()
        ex.verify(torch)
        pms = time_steps(torch, plain_step, args.plain_steps)
        plain = {"ms_per_step": round(pms, 4), "GBps": round(algo / (pms * 1e-3) / 1e9, 2)}
        log(f"plain: {pms*1e3:.1f} us/step, {plain['GBps']} GB/s; batched/plain = {pms/ms:.2f}x")
        rt.close()

    # ---------------- Zipf(1.1) keys
    zipf = None
    if not args.no_zipf:
        del plan
        ez = Exchange(pkg, torch, P, n, "zipf")
        zp = ez.plan(pkg, (tune_used["block_threads"], tune_used["blocks_per_cu"],
                           tune_used["lines_per_lane"], tune_used["nontemporal"]))
        for _ in range(max(1, args.warmup)):
            zp.graph_replay(1, stream)
        torch.cuda.synchronize()
        ez.verify(torch)
        zms = time_steps(torch, lambda: zp.graph_replay(1, stream), args.steps)
        zipf = {"ms_per_step": round(zms, 5), "GBps": round(2 * ez.payload / (zms * 1e-3) / 1e9, 2),
                "max_recv_bytes": int(ez.rtot.max()) * 4, "payload_bytes": ez.payload}
        log(f"zipf: {zms*1e3:.1f} us/step, {zipf['GBps']} GB/s, heaviest receiver "
            f"{zipf['max_recv_bytes']/2**20:.1f} MiB")
        zp.close()
        del ez

    if args.copy_probe:
        a = torch.empty(1 << 27, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    cpu = None if args.no_cpu else cpu_baseline(args)

    traffic = None
    tf = os.path.join(ROOT, "profiles", "a2av_step_uniform_512x2^18_traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    result = {
        "metric": "batched Alltoallv GB/s (% HBM peak) + p50/p99 job latency vs plain launch",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (Philox uniform keys, NPB-style buckets)",
        "config": {"workload": "bucket-sort exchange: 512 ranks x 2^18 uint32 keys (2^27 total), "
                               "uniform [0,2^23), 1024 buckets; one step = Alltoall(counts) + "
                               "Alltoallv(keys) in one persistent kernel, hipGraph replay",
                   "ranks": P, "keys_per_rank": n, "payload_bytes": ex.payload,
                   "tune": tune_used},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "plain_path": plain,
        "batched_vs_plain": round(plain["ms_per_step"] / ms, 2) if plain else None,
        "zipf": zipf,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
