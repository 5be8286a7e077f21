"""The batch driver's captured-graph life cycle with its host code under
AddressSanitizer.

tests/cpp/graph_churn.cpp solves one batch with every launch shape the
autotuner tries, three times over (the tile table regrows, captured graphs
are destroyed and captured again, compaction re-captures mid-solve), then a
lone QP on direct launches and cached graphs, then two engines admitting and
retiring QPs of every size.  It links libmprgp_amd_hostasan.so, the library
built with -Xarch_host -fsanitize=address (device code unchanged), so any
use-after-free or overflow in the host code -- the graph cache, the launch
tables, the staging buffers, the engine's slots -- aborts the run with a
report.  It runs in a child process with leak checks off (the HIP runtime
keeps allocations until exit).  After destroying every library object it
ends with _exit(0): the HIP runtime's own static destructors fail an
internal check of the sanitizer's device-memory support at process exit
(runtime memory freed after the device runtime shut down), which is no part
of the code under test."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "graph_churn")


def test_graph_churn_under_host_asan():
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/graph_churn is built by __graft_entry__.build()")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0")
    r = subprocess.run([BIN], capture_output=True, text=True, env=env, timeout=150)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "graph_churn ok" in r.stdout
    assert "AddressSanitizer" not in r.stderr
