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


def test_sah_arithmetic_in_fp32_equals_the_double_forms():
    """The build computes Bounds3::surface_area's 2.0 * (FP32 sum) and the SAH
    cost's 0.125 + (FP32 x) in FP32 (box_area, sah_cost): doubling is exact in
    both formats, and 0.125 + x is exact in double below 2^50 (above, 0.125 is lost
    in both), so one FP32 rounding gives the double form's value. Checked on
    random and extreme FP32 values, including overflow to infinity."""
    import numpy as np

    rng = np.random.default_rng(7)
    bits = rng.integers(0, 0x7F800000, 2_000_000, dtype=np.int64).astype(np.uint32)
    x = bits.view(np.float32)
    extra = np.array([0.0, np.float32(np.finfo(np.float32).max), np.float32(np.finfo(np.float32).max) / 2,
                      np.nextafter(np.float32(np.finfo(np.float32).max) / 2, np.float32(np.inf)),
                      np.float32(2.0 ** 50), np.float32(2.0 ** 49), np.float32(1e-45), np.float32(0.875)],
                     np.float32)
    x = np.concatenate([x, extra])
    with np.errstate(over="ignore"):
        area_double = (2.0 * x.astype(np.float64)).astype(np.float32)
        area_single = np.float32(2.0) * x
        cost_double = (0.125 + x.astype(np.float64)).astype(np.float32)
        cost_single = np.float32(0.125) + x
    assert np.array_equal(area_double.view(np.uint32), area_single.view(np.uint32))
    assert np.array_equal(cost_double.view(np.uint32), cost_single.view(np.uint32))
