"""The reference's own device unit tests, run as rank programs on the GPU.

tests/cuda_mpi_test.cu, tests/collectives_test.cu and tests/communicator_test.cu of the
reference assert per-rank `ok` flags with the rank counts used here (TestRunner(n)).
"""
import pytest

from conftest import load_pkg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("program", sorted(load_pkg().programs.REFERENCE_TESTS))
def test_reference_program(program, pkg, torch_gpu):
    from dphpc_project_amd import programs as pr
    torch = torch_gpu
    n = pr.REFERENCE_TESTS[program]
    print(f"[gpu] reference test {program} on {n} ranks", flush=True)
    ok = torch.zeros(n, dtype=torch.int32, device="cuda")
    with pkg.Runtime(n, heap_bytes=64 << 20, timeout_s=10.0) as rt:
        rt.run(program, pr.OkArgs(ok.data_ptr()))
    assert ok.cpu().tolist() == [1] * n
