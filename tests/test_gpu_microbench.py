"""The HIP microbenchmarks give physically ordered figures on the GPU.

HBM streams and compute peaks (microbenches/membw.cu, roofline.cu) and the
memory-hierarchy sweep (microbenches/instruction_rates.cu): results must sit
below the MI355X datasheet peaks, and a working set that fits a cache level
must be served faster than one that only fits HBM.
"""
import pytest

from gpu_util import note

pytestmark = pytest.mark.gpu


def test_stream_and_peaks_below_datasheet(gpu):
    note("microbench: read stream 1 GiB, FP64 matrix-core peak")
    rd = gpu.mb_stream("read", 1 << 30, 2)
    assert 2000.0 < rd < 8000.0 * 1.02, rd
    fp = gpu.mb_peak("fp64_mfma", 1)
    assert 10.0 < fp < 78.6 * 1.02, fp


def test_hierarchy_is_ordered(gpu):
    note("microbench: latency 16 KiB / 2 MiB / 1 GiB, re-read 8 MiB / 4 GiB, LDS")
    l1 = gpu.mb_cache("latency", 16 << 10, 1)
    l2 = gpu.mb_cache("latency", 2 << 20, 1)
    hbm = gpu.mb_cache("latency", 1 << 30, 1)
    assert 0 < l1 < l2 < hbm, (l1, l2, hbm)
    assert hbm < 10000.0, hbm        # a dependent load takes well under 10 us
    bw_cache = gpu.mb_cache("bandwidth", 8 << 20, 1)
    bw_hbm = gpu.mb_cache("bandwidth", 4 << 30, 1)
    assert bw_cache > bw_hbm > 1000.0, (bw_cache, bw_hbm)
    lds = gpu.mb_cache("lds_bandwidth", 0, 1)
    assert lds > bw_hbm, (lds, bw_hbm)
    lds_lat = gpu.mb_cache("lds_latency", 0, 1)
    assert 0 < lds_lat < l1, (lds_lat, l1)


def test_bad_arguments_raise(gpu):
    with pytest.raises(gpu.SkinnyError):
        gpu.mb_cache("latency", 100, 1)


def test_grid_barrier_cost(gpu):
    note("microbench: grid barrier")
    us = gpu.mb_barrier(1, 1)
    assert us is not None and 0.1 < us < 1000.0, us
    assert gpu.mb_barrier(1000, 1) is None      # cannot be resident: refused, not launched


def test_fp64_matrix_core_forms_and_clock(gpu):
    """The four-block 4x4x4 FP64 form and the shader clock: every peak stays
    below the datasheet figure scaled to the measured clock (78.6 TFLOP/s at
    2400 MHz), and the clock readings lie between 100 MHz and the 2400 MHz
    maximum engine clock."""
    note("microbench: fp64 4x4x4 peak and clocks")
    q4 = gpu.mb_peak("fp64_mfma_4x4x4", 1)
    assert 10.0 < q4 < 78.6 * 1.02, q4
    for kind in ("idle", "fp64_fma", "fp64_mfma", "fp64_mfma_4x4x4", "int8_mfma"):
        mhz = gpu.mb_clock(kind)
        assert 100.0 < mhz < 2400.0 * 1.05, (kind, mhz)
    mhz = gpu.mb_clock("fp64_mfma_4x4x4")
    assert q4 < 78.6 * (mhz / 2400.0) * 1.05, (q4, mhz)


def test_fp32_matrix_core_forms(gpu):
    """FP32 matrix-core peaks (16x16x4, 32x32x2, 16-block 4x4x1): each below
    the 157.3 TFLOP/s dense datasheet figure and above the FP64 vector rate;
    the clock read during each loop lies between 100 and 2400 MHz.  (The
    clock is read by one workgroup at its start and end, which for the
    32x32x2 loop does not cover the whole run, so the peaks are not scaled
    by it.)"""
    fma = gpu.mb_peak("fp64_fma", 1)
    for kind in ("fp32_mfma", "fp32_mfma_32x32x2", "fp32_mfma_4x4x1"):
        note(f"microbench: {kind}")
        tf = gpu.mb_peak(kind, 1)
        mhz = gpu.mb_clock(kind)
        assert 100.0 < mhz < 2400.0 * 1.05, (kind, mhz)
        assert fma < tf < 157.3 * 1.02, (kind, tf, mhz)
