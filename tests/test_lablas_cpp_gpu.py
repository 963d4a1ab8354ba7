"""The C++ lablas front end (include/gbk/lablas.hpp) driven like the
reference's own drivers: each mode builds a seeded R-MAT through
lablas::Matrix::build_from_csr_arrays, runs on the GPU and checks itself with
a sequential checker, printing "error_count: k/n" (spmv.cpp:102-122 style)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "lablas_driver")


@pytest.mark.parametrize("mode,scale,fmt", [("spmv", 14, 0), ("spmv", 14, 6), ("spmv", 14, 3), ("spmv", 14, 4),
                                            ("vxm", 13, 0), ("vxm", 13, 6), ("masked", 13, 0), ("bfs", 14, 0),
                                            ("sssp", 13, 0), ("cc", 14, 0), ("pr", 13, 0)])
def test_driver(mode, scale, fmt):
    assert os.path.exists(EXE), "build the drivers first (__graft_entry__.build())"
    r = subprocess.run([EXE, mode, str(scale), str(fmt)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "error_count: 0/" in r.stdout, r.stdout
