"""NPB IS (classes S, W, A, B, C on 4, 16, 64, 64, 64 ranks) and DT class S converted by this build's
converter (gpumpicc) from the
reference's toolchain-test sources and run through their own ---gpumpi command line, as
toolchain_tests/npb/is_test.sh and dt_test.sh do; the benchmarks verify themselves and the
test parses their report (tests/golden/npb_converted_expected.json).  The executables are
built by build() where the reference's sources are present; elsewhere the test skips."""
import json
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
EXPECTED = json.load(open(os.path.join(ROOT, "tests", "golden", "npb_converted_expected.json")))


@pytest.mark.parametrize("name", ["npb_is_S", "npb_is_W", "npb_is_A", "npb_is_B", "npb_is_C", "npb_dt_S"])
def test_converted_npb_verifies(name):
    exe = os.path.join(ROOT, "tests", "programs", "_build", name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built (the reference's NPB sources were not present at build time)")
    exp = EXPECTED[name]
    print(f"[gpu] converted {name} {' '.join(exp['args'])}", flush=True)
    r = subprocess.run([exe, *exp["args"]], capture_output=True, text=True, timeout=240)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    lines = r.stdout.splitlines()
    for want in exp["lines"]:
        assert want in lines, (want, out[-3000:])
