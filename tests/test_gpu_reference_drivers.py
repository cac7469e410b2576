"""The reference's own example programs (precision_benchmarks/01_ozaki_scheme2_example.cpp and
03_ozaki_scheme1_example.cpp), compiled unmodified against this repository's drop-in C++ API by
`make -C oracle drivers` (oracle/_ref/drivers/), run on the GPU: each builds its 8192 x 8192
inputs at phi = 4 with the reference generator, computes rocBLAS DGEMM through
HipExecutor::get_blas_handle(), then ozaki_scheme2_gemm / ozaki_scheme1_gemm over its slice counts,
and prints the reference's error metrics against the DGEMM result.  The programs must run to the
end, and the errors they print must fall as the slice count grows.  Skipped when the drivers were not
built (no reference checkout at build time)."""
import math
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVERS = os.path.join(ROOT, "oracle", "_ref", "drivers")


def run_driver(name):
    path = os.path.join(DRIVERS, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (make -C oracle drivers needs the reference checkout)")
    print(f"running {name}", flush=True)
    out = subprocess.run([path], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "Done!" in out.stdout
    return out.stdout


def relative_errors(text):
    return [float(x) for x in re.findall(r"Max Relative Error:\s+(\S+)", text)]


@pytest.mark.parametrize("name,slices", [("01_ozaki_scheme2_example", [4, 6, 8, 12, 16]),
                                         ("03_ozaki_scheme1_example", list(range(2, 16)))])
def test_reference_example_runs_on_this_library(name, slices):
    errs = relative_errors(run_driver(name))
    assert len(errs) == len(slices), errs
    assert all(math.isfinite(e) for e in errs), errs
    # more slices, smaller error, until it meets DGEMM's own rounding (the programs compare with DGEMM)
    assert errs[-1] < errs[0] * 1e-3, errs
    for a, b in zip(errs, errs[1:]):
        assert b <= a * 1.5 or b < errs[0] * 1e-3, errs
