"""Register budget of the rank programs, read from the built gfx950 code object.

The launcher admits a rank grid only if it needs at most half of the wave slots its
kernel can occupy.  A rank program within 256 VGPRs + AGPRs per lane holds two waves per
SIMD (2048 on 256 CUs), so 1024-rank grids (the collective sweep, the heap churn, the job
server) fit; a program above it holds one wave per SIMD and runs at most 512 ranks.  A
change that pushes a runtime function past the budget shows up here, on the CPU, instead
of as a refused launch on the GPU."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import PKG_DIR

LLVM = "/opt/rocm/llvm/bin"
LIB = os.path.join(PKG_DIR, "libgpumpi.so")
# programs that need more than 256 registers for their own loops, and run at most 512
# ranks (the NPB IS rank program and the particle exchange, both on 512 ranks)
HEAVY = {"w_bucket_sort", "w_particles"}


MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _kernels(tmp_path):
    """Kernel name -> VGPRs + AGPRs per lane, over the code objects of every translation
    unit (the library's .hip_fatbin section holds one offload bundle per unit)."""
    tools = [os.path.join(LLVM, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")]
    if not all(os.path.exists(t) for t in tools) or not os.path.exists(LIB):
        pytest.skip("LLVM tools or the built library missing")
    fat = tmp_path / "fatbin.bin"
    subprocess.run([tools[0], "-O", "binary", "--only-section=.hip_fatbin", LIB, str(fat)], check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    out = {}
    for i, a in enumerate(starts):
        piece = tmp_path / f"bundle{i}.bin"
        piece.write_bytes(data[a:starts[i + 1] if i + 1 < len(starts) else len(data)])
        co = tmp_path / f"gfx950_{i}.co"
        subprocess.run([tools[1], "--unbundle", "--type=o", f"--input={piece}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True)
        notes = subprocess.run([tools[2], "--notes", str(co)], capture_output=True, text=True,
                               check=True).stdout
        for item in re.split(r"\n  - ", notes):
            m = re.search(r"\.name:\s+(\S+)", item)
            v = re.search(r"\.vgpr_count:\s+(\d+)", item)
            if m and v:
                out[m.group(1)] = max(out.get(m.group(1), 0), int(v.group(1)))
    return out


def test_rank_programs_fit_two_waves_per_simd(tmp_path):
    ks = _kernels(tmp_path)
    rank = {re.search(r"progs\d+(\w+?)EiPPc", k).group(1): v for k, v in ks.items()
            if "rank_grid_kernel" in k}
    assert len(rank) >= 20, rank
    for name, v in rank.items():
        limit = 512 if name in HEAVY else 256
        assert v <= limit, f"{name}: {v} VGPRs + AGPRs per lane (budget {limit})"
    for name in ("w_coll", "w_churn", "w_serve", "w_job", "w_mcpi", "w_p2p", "w_exchange"):
        assert rank[name] <= 256, name


def test_batched_kernels_keep_their_occupancy(tmp_path):
    """The copy kernels (batched Alltoallv, batched point-to-point) stream with at least
    four waves per SIMD.  The reduction kernels'
    narrow shape (2 columns x 8 ranks in flight per lane, the 1 MiB-per-rank case) keeps
    at least three waves per SIMD (168 registers measured); the wide shape (8 x 4, the
    large-buffer case with fewer, longer waves) stays within one wave's 512; the integer
    split fold (2 x 4 per wave, 4 waves on one tile) keeps four waves per SIMD (86-108)."""
    ks = _kernels(tmp_path)
    n = 0
    for k, v in ks.items():
        if "a2av_copy_kernel" in k or "p2p_copy_kernel" in k:
            assert v <= 128, f"{k}: {v}"
            n += 1
        if "batched_reduce_kernel" in k:
            narrow = "Li2ELi8EEEvNS_9ReduceDevE" in k
            assert v <= (170 if narrow else 512), f"{k}: {v}"
            n += narrow
        if "batched_reduce_split_kernel" in k:
            assert v <= 128, f"{k}: {v}"
            n += 1
    assert n >= 24 + 1 + 24 + 16
