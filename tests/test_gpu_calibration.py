"""The measurement-only VALU issue calibration: every kind runs and issues
the instruction count its loop implies at a plausible rate (a 16-lane SIMD
issues at most one wave64 FP32 instruction per 4 cycles without packed
FP32); an unknown kind is rejected on the host."""
import pytest

import measure
from forcepho_amd import _lib

pytestmark = pytest.mark.gpu

# kind -> VALU instructions per wave of the timed loop (32768 iterations of
# 16 FMAs, multiplies or adds; 8192 iterations of 8 FMA + 8 exp, or of 16 exp)
PER_WAVE = {0: 16 * 32768, 1: 16 * 8192, 2: 16 * 32768, 3: 16 * 8192, 4: 16 * 32768,
            5: 16 * 32768, 6: 16 * 32768}


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4, 5, 6])
def test_issue_calibration_rates(hip, kind):
    ms, wi, clk = measure.issue_calibration(kind)
    cus = _lib.cu_count()
    waves = 8 * cus * 4
    assert wi == waves * PER_WAVE[kind]
    rate = wi / (4 * cus * ms * 1e-3 * clk * 1e3)   # per SIMD per cycle
    print(f"kind {kind}: {ms:.3f} ms, {rate:.3f} wave64 VALU per SIMD per cycle", flush=True)
    assert 0.02 < rate < 1.1


def test_issue_calibration_rejects_unknown_kind(hip):
    with pytest.raises(_lib.FphoError, match="kind"):
        measure.issue_calibration(7)
