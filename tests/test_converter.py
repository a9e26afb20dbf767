"""The source converter and the device C library on the CPU.

The converter's expected output follows the reference converter's rewrite rules
(source_converter/converter.cpp:59-330) and its golden test's expectations for the
coalescing rewrite (tests/converter/malloc.c: `bool __coalesced = false` on functions that
allocate or exchange outside branches, `__coalesced` passed down, `true` from main,
MPI_Alltoall_coalesce, guarded calls left alone).  The programs are this repository's
own (tests/programs/c); the GPU runs of the built programs are in test_gpu_converted.py.
"""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "dphpc-project_amd"))
import converter  # noqa: E402

C = os.path.join(ROOT, "tests", "programs", "c")


def convert(name, version=0, ranks=4, project=C):
    path = os.path.join(C, name)
    out = converter.convert_file(path, project, version, (), ranks)
    return out[os.path.realpath(path)].decode(), out


def test_coalescing_rewrite():
    text, _ = convert("coalesce.c", version=5)
    assert text.startswith("#define GPUMPI_MALLOC_V5\n#include <gpumpi/global_vars.h>\n")
    for line in [
        "__device__ static int *make_block(int n, bool __coalesced = false) {",
        "int *p = (int *)malloc(n * sizeof(int), __coalesced);",
        "__device__ static int exchange(int *send, int n, int rank, int size, bool __coalesced = false) {",
        "int *recv = (int *)malloc(n * size * sizeof(int), __coalesced);",
        "MPI_Alltoall_coalesce((const void *)send, n, MPI_INT, (void *)recv, n, MPI_INT, MPI_COMM_WORLD, __coalesced);",
        "__device__ int __gpu_main(int argc, char** argv) {",
        "int *send = make_block(n * size, true);",
        "int ok = exchange(send, n, rank, size, true);",
        "ok = exchange(send, n, rank, size) && ok;",   # under a branch: default flag
        "free((void *)send);",
        "__device__ int __gpu_max_threads() {\n    return 4;\n}",
    ]:
        assert line in text, line


def test_malloc_v3_brackets_main(tmp_path):
    src = tmp_path / "m.c"
    src.write_text("#include <stdlib.h>\nint main(void) {\n    void *p = malloc(8);\n"
                   "    free(p);\n    return 0;\n}\n")
    out = converter.convert_file(str(src), str(tmp_path), 3, (), 8)
    text = out[os.path.realpath(str(src))].decode()
    assert "int __gpu_main(int argc, char** argv) {\n    __gpu_init_malloc();" in text
    assert "void *p = malloc(8, true);" in text
    assert "__gpu_clean_malloc();\n    return 0;" in text
    assert text.startswith("#define GPUMPI_MALLOC_V3\n")


def test_globals_rewrite():
    text, _ = convert("globals.c", ranks=4)
    for line in [
        "__device__ extern int counter[4];\n__device__ extern int counter[4];\n__device__ int counter[4] = {-7, -7, -7, -7};",
        "__device__ double weights[4][3] = {{0.5, 1.5, 2.5}, {0.5, 1.5, 2.5}, {0.5, 1.5, 2.5}, {0.5, 1.5, 2.5}};",
        "__device__ int grid[4][2][3] = {{{1, 2, 3}, {4, 5, 6}}, ",
        "__device__ int *slots[4][2] = {{NULL, NULL}, ",
        "__device__ const int kTable[4] = {10, 20, 30, 40};",         # const: one shared copy
        "__device__ int __decl_class[4] = {3, 3, 3, 3};",
        "#define ALIAS_WEIGHTS __gpu_global(weights)",                 # a global named in a macro body
        "__device__ static int calls[4][4] = {{100}, {100}, {100}, {100}};",
        "__gpu_global(calls)[0] += 1;",
        "__device__ int oldstyle(int a, double b) { return a + (int)b; }",
        "__device__ static int cmp_int(const void *x, const void *y) {",
        "CHECK(__gpu_global(counter) == -7);",                        # a global as a macro argument
        "__gpu_global(__decl_class) = rank;",
        "int *buf = (int *)malloc(8 * sizeof(int));",
        "qsort((void *)buf, 8, sizeof(int), cmp_int);",
        "MPI_Allreduce((const void *)&__gpu_global(counter), (void *)&total, 1, MPI_INT, MPI_SUM, MPI_COMM_WORLD);",
        "&& kTable[3] == 40 &&",
    ]:
        assert line in text, line
    assert "__gpu_global(kTable)" not in text


def test_open_array_and_function_pointer_globals(tmp_path):
    src = tmp_path / "g.c"
    src.write_text("static int twice(int x) { return 2 * x; }\n"
                   "const char *names[] = {\"a\", \"b\", \"c\"};\n"
                   "int (*op)(int) = twice;\n"
                   "int main(void) { return op(1) == 2 && names[2][0] == 'c' ? 0 : 1; }\n")
    text = converter.convert_file(str(src), str(tmp_path), 0, (), 2)[os.path.realpath(str(src))].decode()
    assert '__device__ const char *names[2][3] = {{"a", "b", "c"}, {"a", "b", "c"}};' in text
    assert "__device__ int (*op[2])(int) = {twice, twice};" in text
    assert "return __gpu_global(op)(1) == 2 && __gpu_global(names)[2][0] == 'c' ? 0 : 1;" in text


def test_project_header_converted():
    _, out = convert("proj/util.c", ranks=4, project=os.path.join(C, "proj"))
    hdr = out[os.path.realpath(os.path.join(C, "proj", "util.h"))].decode()
    assert "__device__ extern int hops[4];" in hdr
    assert "__device__ int ring_pass(int value, int rank, int size);" in hdr
    util = out[os.path.realpath(os.path.join(C, "proj", "util.c"))].decode()
    assert "__device__ int hops[4] = {0, 0, 0, 0};" in util
    assert "__gpu_global(hops)++;" in util


def test_parse_error_is_reported(tmp_path):
    src = tmp_path / "bad.c"
    src.write_text("int main( { return 0; }\n")
    with pytest.raises(RuntimeError, match="could not parse"):
        converter.convert_file(str(src), str(tmp_path), 0, (), 4)


def test_device_libc_formatter_and_parsers_match_the_c_library(tmp_path):
    """include/gpumpi/libc/format.h and parse.h built for the host and compared with
    snprintf / strtoll / strtod on fixed and 6000 random inputs."""
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = tmp_path / "libc_host_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "libc_host_check.cpp"), "-o", str(exe), "-lm"],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "fails=0" in r.stdout


@pytest.mark.parametrize("name", ["quad", "globals", "coalesce", "ring", "exits"])
def test_converted_programs_built(name):
    """build() converts and compiles every C test program for gfx950."""
    exe = os.path.join(ROOT, "tests", "programs", "_build", name)
    if not os.path.exists(exe):
        pytest.skip("programs not built yet (run __graft_entry__.build())")
    src = os.path.join(ROOT, "tests", "programs", "_build", name + ".src", "__gpumpi_program.hip")
    assert os.path.exists(src)
    assert "#include <gpumpi/gpumpi_main.h>" in open(src).read()


NPB = "/root/reference/toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI"


@pytest.mark.skipif(not os.path.exists(NPB), reason="reference NPB sources not present")
def test_npb_is_converts_and_compiles(tmp_path):
    """A real program: the NPB IS kernel the reference converts (toolchain_tests/npb:
    IS/is.c with common/c_timers.c and c_print_results.c), copied into a scratch project
    with a class S / 4-process npbparams.h written here, through the converter and hipcc
    (three translation units in one build)."""
    proj = tmp_path / "npb"
    (proj / "IS").mkdir(parents=True)
    (proj / "common").mkdir()
    shutil.copy(os.path.join(NPB, "IS", "is.c"), proj / "IS" / "is.c")
    for f in ("c_timers.h", "c_timers.c", "c_print_results.c"):
        shutil.copy(os.path.join(NPB, "common", f), proj / "common" / f)
    (proj / "IS" / "npbparams.h").write_text(
        "#define CLASS 'S'\n#define NUM_PROCS 4\n#define COMPILETIME \"now\"\n"
        "#define NPBVERSION \"3.4.1\"\n#define MPICC \"gpumpicc\"\n#define CFLAGS \"-O3\"\n"
        "#define CLINK \"hipcc\"\n#define CLINKFLAGS \"\"\n#define CMPI_LIB \"-lgpumpi\"\n"
        "#define CMPI_INC \"\"\n")
    sys.path.insert(0, os.path.join(ROOT, "dphpc-project_amd"))
    import gpumpicc
    srcs = [str(proj / "IS" / "is.c"), str(proj / "common" / "c_print_results.c"),
            str(proj / "common" / "c_timers.c")]
    exe = gpumpicc.build(srcs, str(tmp_path / "is_S"), project=str(proj), ranks=8)
    assert os.path.exists(exe)
    converted = (tmp_path / "is_S.gpumpi" / "IS" / "is.hip").read_text()
    assert "__device__ int __gpu_main(int argc, char** argv)" in converted
