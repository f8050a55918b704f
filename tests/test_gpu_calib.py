"""bench.py's vector-ALU calibration (libpt_calib.so) runs on the device and
reports rates in the range a wave64 VALU can issue: below two instructions per
clock per CU at 2.4 GHz (4.8 per ns per CU), and ordered as the instruction
forms' costs are (an FMA on VGPRs issues faster than a transcendental)."""
import pytest

pytestmark = pytest.mark.gpu


def test_calibration_kernels_time_each_form(ctx):
    import bench
    r = bench.valu_calibration()
    for form in ("fma_vgpr", "fma_sgpr", "add_vgpr", "rsq", "fma_f64"):
        assert 0.05 < r[form] < 4.8, (form, r[form])
    assert r["rsq"] < r["fma_vgpr"]
    assert r["fma_tflops"] == pytest.approx(r["fma_vgpr"] * 256 * 128 / 1e3)
