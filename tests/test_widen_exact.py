"""The traversal kernels widen far slab distances with one FP32 fused
multiply-add, fmaf(t, 2*gamma(3), t), where bounds3.h:159-180 multiplies by the
double 1.0 + 2.0*gamma(3) and rounds back to FP32. The two agree on every FP32
input: checked here over all 2^32 bit patterns (tests/native/widen_exact.c,
built with gcc; libm's fmaf is correctly rounded)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_fma_widening_equals_fp64_widening_for_all_floats(tmp_path):
    exe = tmp_path / "widen_exact"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), os.path.join(HERE, "native", "widen_exact.c"),
                    "-lm", "-lpthread"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.stdout.strip() == "mismatches 0", out.stdout + out.stderr
    assert out.returncode == 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_widening_check_catches_a_wrong_factor(tmp_path):
    """The same scan with the factor taken one FP64 ulp away must report
    mismatches: the check can fail."""
    src = open(os.path.join(HERE, "native", "widen_exact.c")).read()
    src = src.replace("factor = 1.0 + 2.0 * (double)gamma_n(3);",
                      "factor = nextafter(1.0 + 2.0 * (double)gamma_n(3), 2.0);")
    src = src.replace("if ((double)delta != factor - 1.0) {", "if (0) {")
    c = tmp_path / "wrong.c"
    c.write_text(src)
    exe = tmp_path / "wrong"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm", "-lpthread"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 1 and out.stdout.startswith("mismatches ") and out.stdout.strip() != "mismatches 0"
