"""gpumpicc: convert MPI C sources and build them into a gpumpi executable.

The reference's compiler wrapper (scripts/gpumpicc.py:140-211) copies each source to
.cu and each project header to .cuh, runs the converter, appends `.cuh` to every
#include so that the C library headers resolve to its gpu_libc shims, and compiles with
nvcc -rdc against its static libraries.  Here:

  1. every C source goes through the converter (converter.py); the project headers it
     depends on (clang -MM, files under the project directory) are converted or copied
     into a build directory that mirrors the project tree (`foo.h` -> `foo.h.hip`);
  2. includes are redirected: `mpi.h` and the C library headers that have a device
     version (include/gpumpi/libc) to those, project headers to their converted copies;
  3. the converted sources are compiled as one translation unit, followed by
     gpumpi_main.h (the device runtime and the host launcher main), with hipcc for
     gfx950 and linked against libgpumpi.so.

    python dphpc-project_amd/gpumpicc.py prog.c [more.c ...] -o prog [-I dir] [-D X=1] \\
        [--malloc-version V] [--project DIR]
    ./prog <args> ---gpumpi -n 64
"""
import argparse
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIBC = os.path.join(INCLUDE, "gpumpi", "libc")
sys.path.insert(0, HERE)
import converter  # noqa: E402

ARCH = "gfx950"


def libc_headers():
    out = set()
    for dirpath, _, files in os.walk(LIBC):
        for f in files:
            rel = os.path.relpath(os.path.join(dirpath, f), LIBC)
            if rel.endswith(".h") and rel not in ("format.h", "parse.h", "undef.h"):
                out.add(rel)
    return out


INCLUDE_RE = re.compile(rb'^(\s*#\s*include\s*)([<"])([^>"]+)([>"])', re.M)


def project_deps(src, clang_args, project):
    """Project headers a source includes (directly or not), by clang -MM."""
    cmd = [converter.clang_path(), "-x", "c", "-MM", "-I", converter.C_MPI_DIR, *clang_args, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"clang -MM failed on {src}:\n{r.stderr}")
    deps = r.stdout.replace("\\\n", " ").split(":", 1)[1].split()
    out = []
    for d in deps[1:]:
        rp = os.path.realpath(d)
        if rp.startswith(project + os.sep) and rp != os.path.realpath(src):
            out.append(rp)
    return out


def redirect_includes(data, file_real, project, build, libc, converted_headers, inc_dirs):
    def sub(m):
        name = m.group(3).decode()
        if name == "mpi.h":
            return m.group(1) + b"<gpumpi/libc/mpi.h>"
        if m.group(2) == b'"':
            for base in [os.path.dirname(file_real), *inc_dirs]:
                cand = os.path.realpath(os.path.join(base, name))
                if cand in converted_headers:
                    return m.group(1) + b'"' + (name + ".hip").encode() + b'"'
        if name in libc:
            return m.group(1) + ("<gpumpi/libc/%s>" % name).encode()
        return m.group(0)
    return INCLUDE_RE.sub(sub, data)


def build(sources, output, clang_args=(), malloc_version=0, project=None, hipcc_args=(),
          build_dir=None, ranks=None):
    """Converts and compiles; returns the executable's path.  `ranks` overrides
    GPU_MPI_MAX_RANKS (the size of the per-rank global arrays)."""
    sources = [os.path.abspath(s) for s in sources]
    project = os.path.realpath(project or os.environ.get("GPU_MPI_PROJECT") or
                               os.path.commonpath([os.path.dirname(s) for s in sources]))
    build_dir = os.path.abspath(build_dir or output + ".gpumpi")
    os.makedirs(build_dir, exist_ok=True)
    libc = libc_headers()
    inc_dirs = [a[2:] for a in clang_args if a.startswith("-I") and len(a) > 2]
    inc_dirs = [os.path.realpath(d) for d in inc_dirs]

    def mirror(real):
        return os.path.join(build_dir, os.path.relpath(real, project))

    converted = {}   # real path -> bytes
    headers = set()
    for src in sources:
        for h in project_deps(src, clang_args, project):
            headers.add(h)
        out = converter.convert_file(src, project, malloc_version, clang_args, ranks)
        for real, data in out.items():
            if real in converted and real != os.path.realpath(src):
                continue
            converted[real] = data
    for h in headers:
        if h not in converted:
            converted[h] = open(h, "rb").read()
    header_set = {h for h in converted if h not in {os.path.realpath(s) for s in sources}}
    unity = []
    for real, data in converted.items():
        data = redirect_includes(data, real, project, build_dir, libc, header_set, inc_dirs)
        if real in header_set:
            dest = mirror(real) + ".hip"
        else:
            dest = os.path.splitext(mirror(real))[0] + ".hip"
            unity.append(dest)
        os.makedirs(os.path.dirname(dest), exist_ok=True)
        with open(dest, "wb") as fh:
            fh.write(data)
    # keep the order of the command line
    order = {os.path.splitext(mirror(os.path.realpath(s)))[0] + ".hip": i for i, s in enumerate(sources)}
    unity.sort(key=lambda p: order.get(p, len(order)))
    tu = os.path.join(build_dir, "__gpumpi_program.hip")
    with open(tu, "w") as fh:
        for u in unity:
            fh.write('#include "%s"\n' % u)
        fh.write("#include <gpumpi/gpumpi_main.h>\n")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    extra_inc = []
    for d in inc_dirs:
        if d.startswith(project + os.sep) or d == project:
            extra_inc += ["-I", mirror(d)]
        else:
            extra_inc += ["-I", d]
    defines = [a for a in clang_args if a.startswith("-D")]
    cmd = [hipcc, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Wno-unused-value", "-Wno-unused-result", "-Wno-format",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", INCLUDE, "-I", os.path.join(INCLUDE, "gpumpi"), "-I", os.path.join(HERE, "csrc"),
           "-I", build_dir, *extra_inc, *defines, *hipcc_args, tu,
           "-L", HERE, "-lgpumpi", "-Wl,-rpath," + HERE, "-o", os.path.abspath(output)]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=build_dir)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stderr[-6000:])
    return os.path.abspath(output)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("sources", nargs="+")
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("-I", dest="inc", action="append", default=[])
    ap.add_argument("-D", dest="defs", action="append", default=[])
    ap.add_argument("--malloc-version", type=int, default=0)
    ap.add_argument("--project", default=None)
    ap.add_argument("--build-dir", default=None)
    a = ap.parse_args(argv)
    clang_args = ["-I" + i for i in a.inc] + ["-D" + d for d in a.defs]
    out = build(a.sources, a.output, clang_args, a.malloc_version, a.project, build_dir=a.build_dir)
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
