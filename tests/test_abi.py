"""The C-ABI library loads and exports every symbol include/gpumpi/gpumpi.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "gpumpi", "gpumpi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gpumpi_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_declared_symbols(pkg):
    lib = pkg.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(names) == set(pkg.exported_symbols())


def test_program_registry(pkg):
    names = pkg.program_names()
    for p in pkg.programs.REFERENCE_TESTS:
        assert p in names
    for p in ["mcpi", "p2p", "coll", "churn", "exchange"]:
        assert p in names


def test_config_defaults(pkg):
    c = pkg.default_config(num_ranks=4)
    assert c.num_ranks == 4 and c.ring_slots == 256 and c.max_requests == 1024
    assert c.heap_bytes == 1 << 30 and abs(c.timeout_s - 20.0) < 1e-9


def test_mpi_header_datatype_order(pkg):
    """MPI_Datatype keeps the reference's enum order (mpi_types_list.cuh:4-33)."""
    src = open(os.path.join(ROOT, "include", "gpumpi", "mpi.h")).read()
    body = re.search(r"enum MPI_Datatype \{(.*?)\};", src, re.S).group(1)
    names = [x.strip() for x in body.replace("\n", " ").split(",") if x.strip()]
    assert names == pkg.DATATYPES


def test_particle_plan_rejects_bad_arguments(pkg):
    """Argument checks of the batched particle step return before any device call."""
    lib = pkg.load_library()
    h = ctypes.c_void_p()
    for dims, cap in [((0, 2, 2), 64), ((2, 2, 2), 0), ((2, 2, 2), 1 << 31)]:
        assert lib.gpumpi_particles_plan_create(*dims, cap, 1, 1, ctypes.byref(h)) == pkg.E_ARG
    assert lib.gpumpi_particles_step(None, None, None, None, None, None) == pkg.E_ARG
    assert lib.gpumpi_particles_graph_replay(None, None, None, None, None, 2, None) == pkg.E_ARG


def test_built_programs_match_the_library_layout(pkg):
    """A converted program compares the runtime layout it was compiled against with the
    library's (gpumpi_layout_stamp) before it launches anything -- a stale build would read
    the runtime's structures at other offsets.  `---gpumpi -h` passes the check and prints
    the launcher's help without a GPU."""
    import subprocess
    import pytest
    assert pkg.load_library().gpumpi_layout_stamp() != 0
    exe = os.path.join(ROOT, "tests", "programs", "_build", "quad")
    if not os.path.exists(exe):
        pytest.skip("converted programs not built")
    r = subprocess.run([exe, "---gpumpi", "-h"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "GPU MPI" in r.stdout
