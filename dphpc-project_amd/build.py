"""Build recipe of libgpumpi.so (gfx950) -- compiled in-tree with hipcc.

The device runtime is header-only and is compiled together with the built-in rank
programs and the host C ABI as one translation unit (csrc/gpumpi.hip).  Flags:
  -ffp-contract=off          every FP expression rounds as written (bit-exact FP64
                             reductions and Monte-Carlo hit tests against the oracle)
  -packed-fp32-ops disabled  no v_pk_*_f32 instructions are generated
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "gpumpi.hip")
LIB = os.path.join(HERE, "libgpumpi.so")
ARCH = os.environ.get("GPUMPI_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gpumpi HIP library cannot be built")


def compile_cmd(out):
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-Wno-unused-value",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-o", out, SRC,
    ]


# Programs in the converter's output style, compiled as executables against the runtime
# (each includes gpumpi/gpumpi_main.h) -- used by the tests.
PROGRAMS = [os.path.join(ROOT, "tests", "programs", "pi_converted.hip")]


def program_cmd(src, out):
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
        "-Wno-unused-value", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "include", "gpumpi"),
        "-I", os.path.join(HERE, "csrc"), src, "-L", HERE, "-lgpumpi",
        "-Wl,-rpath," + HERE, "-o", out,
    ]


def build_programs(verbose=False):
    outs = []
    for src in PROGRAMS:
        out_dir = os.path.join(os.path.dirname(src), "_build")
        os.makedirs(out_dir, exist_ok=True)
        out = os.path.join(out_dir, os.path.splitext(os.path.basename(src))[0])
        cmd = program_cmd(src, out)
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout)
            raise RuntimeError(f"hipcc failed to build {src}")
        outs.append(out)
    return outs


# C MPI programs built through the converter and gpumpicc (rank arrays sized for 64 ranks):
# (output name, sources relative to tests/programs/c, extra gpumpicc arguments)
C_PROGRAMS = [
    ("cpi", ["cpi.c"], []),
    ("globals", ["globals.c"], []),
    ("coalesce", ["coalesce.c"], ["--malloc-version", "5"]),
    ("ring", ["proj/main.c", "proj/util.c"], []),
    ("exits", ["exits.c"], []),
    ("fileio", ["fileio.c"], []),
]


def build_c_programs(verbose=False):
    src_dir = os.path.join(ROOT, "tests", "programs", "c")
    out_dir = os.path.join(ROOT, "tests", "programs", "_build")
    os.makedirs(out_dir, exist_ok=True)
    env = dict(os.environ, GPU_MPI_MAX_RANKS="64")
    outs = []
    for name, srcs, extra in C_PROGRAMS:
        out = os.path.join(out_dir, name)
        cmd = [sys.executable, os.path.join(HERE, "gpumpicc.py"), *[os.path.join(src_dir, s) for s in srcs],
               "-o", out, "--build-dir", os.path.join(out_dir, name + ".src"), *extra]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
        if res.returncode != 0:
            sys.stderr.write(res.stdout)
            raise RuntimeError(f"gpumpicc failed to build {name}")
        outs.append(out)
    return outs


def build(verbose=False):
    """Compile libgpumpi.so (always recompiles); atomically replaces the old library."""
    fd, tmp = tempfile.mkstemp(prefix="libgpumpi.", suffix=".so.tmp", dir=HERE)
    os.close(fd)
    cmd = compile_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        os.unlink(tmp)
        sys.stderr.write(res.stdout)
        raise RuntimeError("hipcc failed to build libgpumpi.so")
    os.replace(tmp, LIB)
    build_programs(verbose)
    build_c_programs(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
