"""bench.py names the kernel of each single-kernel stage the way rocprofv3
does, with its grid in threads, so that its HIP-event figure and the
committed trace summaries refer to the same instantiation and launch (CPU
only: no kernel runs)."""
import os
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def name(*a, **k):
    return bench.stage_kernel(*a, **k)[0]


def test_contraction_names_follow_the_padded_descriptor_length():
    t = {"gpr_rows": 128, "waves_fwd": 1, "waves_bwd": 2, "gpr_blocks": 512}
    assert name("gpr_contraction", True, t, True, 128, 4096, 4000) == "k_gpr_mixed<false, 128, 2, 4>"
    assert name("gpr_contraction", True, dict(t, gpr_rows=64), True, 128, 4096, 4000) == \
        "k_gpr_mixed<false, 128, 1, 4>"
    # above 128 features: eight waves of one 16-row block, whatever the tuning says
    assert name("gpr_contraction", True, dict(t, gpr_rows=64), True, 256, 4096, 4000) == \
        "k_gpr_mixed<false, 256, 1, 8>"
    assert name("gpr_contraction", True, t, True, 66, 4096, 4000) == "k_gpr_mixed<false, 96, 2, 4>"
    assert name("gpr_contraction", False, t, True, 256, 4096, 4000) == "k_gpr_f64<256>"
    # a model whose mixed runs contract in FP64
    assert name("gpr_contraction", True, t, True, 66, 64, 129, c64=True) == "k_gpr_f64<96>"


def test_descriptor_names_and_grids():
    t = {"gpr_rows": 64, "waves_fwd": 4, "waves_bwd": 2, "gpr_blocks": 512}
    assert bench.stage_kernel("descriptors_fwd", True, t, True, 128, 1000, 4000) == \
        ("k_desc_fwd_grid<float, 4, 2, true, false>", 1000 * 128)
    assert bench.stage_kernel("descriptors_bwd", False, t, False, 128, 1000, 4000) == \
        ("k_desc_bwd<double, 2, 4, false, false>", 1000 * 128)


def test_contraction_grid_follows_the_host_split():
    # 4096 atoms, D = 128, M = 4000, 128-row tiles, 512 workgroups targeted:
    # 32 tiles, 16 splits of 256 points (Mp = 4032 in 64-point chunks)
    t = {"gpr_rows": 128, "gpr_blocks": 512}
    assert bench.contraction_grid(4096, 128, 4000, True, t) == 32 * 16 * 256
    # the FP64 kernel always takes 64-row tiles
    assert bench.contraction_grid(4096, 128, 4000, False, t) == 64 * 8 * 256
    # 32768 atoms, D = 256, M = 8000: 128-row tiles of 512 threads, 32-point
    # chunks, 1024 workgroups targeted: 256 tiles, 4 splits of 2016 points
    t = {"gpr_rows": 64, "gpr_blocks": 1024}
    assert bench.contraction_grid(32768, 256, 8000, True, t) == 256 * 4 * 512
