"""dfrg_valu_peak: the measured vector FMA rate that confirms, on the device, the published vector
peak the integrands' roofline is quoted against (MI355X: 78.6 TFLOP/s FP64; plain FP32 FMAs issue at
the same instruction rate, the published 157.3 TFLOP/s FP32 figure needs packed FMAs, which this
build does not emit)."""
import pytest

pytestmark = pytest.mark.gpu

PUBLISHED_FP64 = 78.6e12


def test_fma_chain_rate_near_published_peak(gpu):
    props = gpu.device_props()
    # the published figure scaled to this device's CU count and clock (256 CUs at 2.4 GHz)
    expected = PUBLISHED_FP64 * props["cu_count"] / 256 * (props["clock_khz"] / 2.4e6 if props["clock_khz"] else 1.0)
    print(f"[case] FMA-chain probe on {props['name']}: {props['cu_count']} CUs, {props['clock_khz']} kHz", flush=True)
    f64 = gpu.valu_peak(True)
    f32 = gpu.valu_peak(False)
    print(f"  fp64 {f64 / 1e12:.1f} TFLOP/s, fp32 {f32 / 1e12:.1f} TFLOP/s, expected {expected / 1e12:.1f}", flush=True)
    assert 0.5 * expected <= f64 <= 1.05 * expected
    assert 0.5 * expected <= f32 <= 1.05 * 2 * expected
    with pytest.raises(gpu.DfrgInvalidArgument):
        gpu.check(gpu.lib().dfrg_valu_peak(2, None))
