"""The C++ drop-in (tnl_amd/include/TNL/MI355X.h): the reference's SpMV
known-answer tests and solver example compiled against TNL's class templates
with Devices::MI355X (tests/cpp/test_tnl_api.cpp), run as a program."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_tnl_api")
ROWS_BIN = os.path.join(HERE, "cpp", "test_device_rows")


def test_cpp_dropin_kats_and_solvers():
    assert os.path.exists(BIN), "tests/cpp/test_tnl_api is not built (run __graft_entry__.build())"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(p.stdout[-3000:], p.stderr[-3000:])
    assert p.returncode == 0 and "ALL OK" in p.stdout


def test_cpp_device_rows_and_elements():
    """getView().getRow( i ).setElement inside parallelFor< Devices::MI355X >,
    forAllRows / forAllElements / const traversals, every format, Index int
    and long, against the same matrix built through setElement
    (tests/cpp/test_device_rows.cpp, compiled with hipcc)."""
    assert os.path.exists(ROWS_BIN), "tests/cpp/test_device_rows is not built (run __graft_entry__.build())"
    p = subprocess.run([ROWS_BIN], capture_output=True, text=True, timeout=300)
    print(p.stdout[-3000:], p.stderr[-3000:])
    assert p.returncode == 0 and "ALL OK" in p.stdout
