"""GPU: the reference join test program's JoinInstance flow (test/join/main.cc) through
the C++ facade, built by __graft_entry__.build() as spid-join_amd/join_example."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "spid-join_amd", "join_example")


@pytest.mark.parametrize("args", [["-s", "1048576", "-t", "2"], ["-s", "262144", "-t", "8", "-z", "150", "-B", "8"],
                                  ["-s", "131072", "-t", "4", "-z", "50", "-a", "2"]])
def test_reference_program_flow(args):
    if not os.path.exists(EXE):
        pytest.fail("join_example not built (run __graft_entry__.build())")
    print("case", args, flush=True)
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "bad 0 repeated 0 missing 0" in r.stdout
    # RecordLogs: one line per node of the plan the design selects
    assert "-COMPOUND_FUNC_" in r.stdout and "-HOST_FUNC_RECV_DATA_OPT-" in r.stdout
    assert ("-COMPOUND_FUNC_BG_BROADCAST-" in r.stdout) == ("-B" in args)


def test_deprecated_constructor_throws():
    """JoinInstance(int, int) is [[deprecated]] in the reference and exits
    there (query.cc:733-736); here it throws std::logic_error."""
    if not os.path.exists(EXE):
        pytest.fail("join_example not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, "-d"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "deprecated constructor refused" in r.stdout
