"""Build recipe of libgpumpi.so (gfx950) -- compiled in-tree with hipcc.

The device runtime is header-only.  The library is nine translation units compiled in
parallel and linked (csrc/units.h): the host C ABI with the batched steps' host code
(gpumpi.hip), three units of rank programs, and the many-instance batched Alltoallv (one
unit per lines-per-lane value) and reduction kernels.  Flags:
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
UNITS = [os.path.join(HERE, "csrc", u) for u in
         ("gpumpi.hip", "programs_tests.hip", "programs_work.hip", "programs_heavy.hip",
          "kernels_a2av.hip", "kernels_a2av_u2.hip", "kernels_a2av_u4.hip", "kernels_a2av_u8.hip",
          "kernels_reduce.hip")]
LIB = os.path.join(HERE, "libgpumpi.so")
ARCH = os.environ.get("GPUMPI_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gpumpi HIP library cannot be built")


def unit_cmd(src, obj):
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
        "-ffp-contract=off", "-Wno-unused-value",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-o", obj, src,
    ]


def link_cmd(objs, out):
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs]


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
    ("quad", ["quad.c"], []),
    ("globals", ["globals.c"], []),
    ("coalesce", ["coalesce.c"], ["--malloc-version", "5"]),
    ("ring", ["proj/main.c", "proj/util.c"], []),
    ("exits", ["exits.c"], []),
    ("fileio", ["fileio.c"], []),
]


def _build_one_c_program(item, src_dir, out_dir, env, verbose):
    name, srcs, extra = item
    out = os.path.join(out_dir, name)
    cmd = [sys.executable, os.path.join(HERE, "gpumpicc.py"), *[os.path.join(src_dir, s) for s in srcs],
           "-o", out, "--build-dir", os.path.join(out_dir, name + ".src"), *extra]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"gpumpicc failed to build {name}")
    return out


def build_c_programs(verbose=False):
    """The converted C test programs, compiled four at a time."""
    src_dir = os.path.join(ROOT, "tests", "programs", "c")
    out_dir = os.path.join(ROOT, "tests", "programs", "_build")
    os.makedirs(out_dir, exist_ok=True)
    env = dict(os.environ, GPU_MPI_MAX_RANKS="64")
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=4) as pool:
        return list(pool.map(lambda it: _build_one_c_program(it, src_dir, out_dir, env, verbose), C_PROGRAMS))


# NPB IS and DT, the reference's toolchain tests (toolchain_tests/npb/is_test.sh,
# dt_test.sh), converted and compiled from the reference's sources where they are present.
# The sources and the converted text stay in a temporary directory outside the repository;
# only the executables are kept (tests/programs/_build/npb_is_{S,W,A,B,C}, npb_dt_S), so GPU tests
# can run them and parse the benchmarks' own verification lines.
NPB = "/root/reference/toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI"
NPB_COMMON_DEFS = ('#define COMPILETIME "now"\n#define NPBVERSION "3.4.1"\n#define MPICC "gpumpicc"\n'
                   '#define CFLAGS "-O3"\n#define CLINK "hipcc"\n#define CLINKFLAGS ""\n'
                   '#define CMPI_LIB "-lgpumpi"\n#define CMPI_INC ""\n')
NPB_PROGRAMS = [
    # (output, benchmark dir, sources in it, common sources, npbparams.h)
    ("npb_is_S", "IS", ["is.c"], ["c_print_results.c", "c_timers.c"],
     "#define CLASS 'S'\n#define NUM_PROCS 4\n"),
    ("npb_is_W", "IS", ["is.c"], ["c_print_results.c", "c_timers.c"],
     "#define CLASS 'W'\n#define NUM_PROCS 16\n"),
    ("npb_is_A", "IS", ["is.c"], ["c_print_results.c", "c_timers.c"],
     "#define CLASS 'A'\n#define NUM_PROCS 64\n"),
    ("npb_is_B", "IS", ["is.c"], ["c_print_results.c", "c_timers.c"],
     "#define CLASS 'B'\n#define NUM_PROCS 64\n"),
    ("npb_is_C", "IS", ["is.c"], ["c_print_results.c", "c_timers.c"],
     "#define CLASS 'C'\n#define NUM_PROCS 64\n"),
    # class S parameters of NPB's setparams.c (write_dt_info)
    ("npb_dt_S", "DT", ["dt.c", "DGraph.c"], ["c_print_results.c", "c_timers.c", "randdp.c"],
     "#define CLASS 'S'\n#define NUM_PROCS 5\n#define NUM_SAMPLES 1728\n"
     "#define STD_DEVIATION 128\n#define NUM_SOURCES 4\n"),
]


def _build_one_npb(item, verbose):
    import gpumpicc
    name, bdir, srcs, common, params = item
    out_dir = os.path.join(ROOT, "tests", "programs", "_build")
    tmp = tempfile.mkdtemp(prefix="gpumpi_npb_")
    try:
        proj = os.path.join(tmp, "npb")
        os.makedirs(os.path.join(proj, bdir))
        os.makedirs(os.path.join(proj, "common"))
        for f in srcs + (["DGraph.h"] if bdir == "DT" else []):
            shutil.copy(os.path.join(NPB, bdir, f), os.path.join(proj, bdir, f))
        for f in common + ["c_timers.h"]:
            shutil.copy(os.path.join(NPB, "common", f), os.path.join(proj, "common", f))
        with open(os.path.join(proj, bdir, "npbparams.h"), "w") as fh:
            fh.write(params + NPB_COMMON_DEFS)
        files = [os.path.join(proj, bdir, f) for f in srcs] + \
                [os.path.join(proj, "common", f) for f in common]
        exe = gpumpicc.build(files, os.path.join(tmp, name), project=proj,
                             build_dir=os.path.join(tmp, name + ".gpumpi"), ranks=64)
        final = os.path.join(out_dir, name)
        shutil.copy(exe, final)
        if verbose:
            print(f"built {final}", flush=True)
        return final
    except Exception as e:  # a converter gap is reported, the rest of the build goes on
        sys.stderr.write(f"NPB {name}: {e}\n")
        return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def build_npb(verbose=False):
    """The NPB executables, compiled four at a time (each is one large translation unit)."""
    if not os.path.isdir(NPB):
        return []
    sys.path.insert(0, HERE)
    os.makedirs(os.path.join(ROOT, "tests", "programs", "_build"), exist_ok=True)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=4) as pool:
        outs = list(pool.map(lambda it: _build_one_npb(it, verbose), NPB_PROGRAMS))
    return [o for o in outs if o]


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"hipcc failed to build {what}")


def build(verbose=False):
    """Compile libgpumpi.so (always recompiles: the units in parallel, then one link) and
    atomically replace the old library; then the test programs."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = tempfile.mkdtemp(prefix="gpumpi_units_")
    try:
        objs = [os.path.join(objdir, os.path.splitext(os.path.basename(u))[0] + ".o") for u in UNITS]
        with ThreadPoolExecutor(max_workers=8) as pool:
            for f in [pool.submit(_run, unit_cmd(u, o), verbose, os.path.basename(u)) for u, o in zip(UNITS, objs)]:
                f.result()
        fd, tmp = tempfile.mkstemp(prefix="libgpumpi.", suffix=".so.tmp", dir=HERE)
        os.close(fd)
        try:
            _run(link_cmd(objs, tmp), verbose, "libgpumpi.so")
        except Exception:
            os.unlink(tmp)
            raise
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    build_programs(verbose)
    # the converted C programs and the NPB executables: one pool, eight compiles at a time
    with ThreadPoolExecutor(max_workers=2) as pool:
        jobs = [pool.submit(build_c_programs, verbose), pool.submit(build_npb, verbose)]
        for j in jobs:
            j.result()
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
