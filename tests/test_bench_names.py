"""bench.py names the kernel of each single-kernel stage the way rocprofv3
does, so that its HIP-event figure and the committed trace summaries refer to
the same instantiation (CPU only: no kernel runs)."""
import os
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_contraction_names_follow_the_padded_descriptor_length():
    t = {"gpr_rows": 128, "waves_fwd": 1, "waves_bwd": 2}
    assert bench.stage_kernel("gpr_contraction", True, t, True, 128) == "k_gpr_mixed<false, 128, 64, 2>"
    # above 128 features: 32-point chunks and 64-row tiles only
    assert bench.stage_kernel("gpr_contraction", True, t, True, 256) == "k_gpr_mixed<false, 256, 32, 1>"
    assert bench.stage_kernel("gpr_contraction", True, t, True, 66) == "k_gpr_mixed<false, 96, 64, 2>"
    assert bench.stage_kernel("gpr_contraction", False, t, True, 256) == "k_gpr_f64<256>"


def test_descriptor_names():
    t = {"gpr_rows": 64, "waves_fwd": 4, "waves_bwd": 2}
    assert bench.stage_kernel("descriptors_fwd", True, t, True, 128) == "k_desc_fwd_grid<float, 4, 2, true>"
    assert bench.stage_kernel("descriptors_bwd", False, t, False, 128) == "k_desc_bwd<double, 2, 4, false>"
