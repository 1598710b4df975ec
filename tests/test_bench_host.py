"""Host-side arithmetic of bench.py (no GPU): profile folding, the tracer
kernel form read from a profile, the roofline figures of the dominant
kernel, and the refusal of multi-GPU runs."""
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_merge_profile_folds_tile0_and_tracer_forms():
    prof = {"inv_z_bs": (3.0, 3), "inv_z_bs_t0": (0.6, 3), "interp_ab_col": (20.0, 3), "fwd_z": (1.8, 3)}
    m = bench.merge_profile(prof)
    assert m["inv_z_bs"] == (3.6, 3)
    assert "inv_z_bs_t0" not in m and "interp_ab_col" not in m
    assert m["interp_ab"] == (20.0, 3)
    assert m["fwd_z"] == (1.8, 3)
    assert bench.merge_profile({"interp_ab_gather": (5.0, 2)})["interp_ab"] == (5.0, 2)


def test_tracer_form_from_profile_names():
    assert bench.tracer_form({"interp_ab_col": (1, 1)}) == "column walk"
    assert bench.tracer_form({"interp_ab_gather": (1, 1)}) == "global gather"
    assert bench.tracer_form({"interp_ab": (1, 1)}) == "bricks"
    assert bench.tracer_form({"inv_y": (1, 1)}) is None


def test_roofline_of_a_field_pass():
    n, es = 512, 4
    r = bench.roofline({"inv_y": (3.0, 3), "fwd_z": (1.0, 3)}, n, es, 0, 6, False, 3)
    assert r["kernel"] == "inv_y" and r["bound"] == "hbm"
    assert r["bytes_per_launch"] == 12 * n ** 3 * es
    assert r["achieved"] == pytest.approx(12 * n ** 3 * es / 1e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["launches_per_step"] == 1.0 and r["traffic"] is None


def test_roofline_of_the_tracer_loop():
    n, es, np_, nb = 512, 4, 1 << 22, 6
    ms = 6.5
    r = bench.roofline({"interp_ab_col": (3 * ms, 3), "inv_y": (3.9, 3)}, n, es, np_, nb, False, 3)
    assert r["kernel"] == "interp_ab" and r["form"] == "column walk"
    w = 2 * nb + 2
    # tracer state (positions, AB4 history) and one read of the velocity
    assert r["bytes_per_launch"] == np_ * (48 + 24 * 3 + 24) + 3 * n ** 3 * es
    # stencil reads dominate the LDS bytes; the column windows add the halo
    stencil = np_ * 3 * w ** 3 * es
    assert stencil < r["lds"]["bytes_per_launch"] < 1.2 * stencil
    # the reference's FP64 arithmetic: 2 flops of weight product per point
    # and 2 per component
    assert r["fp64"]["flops_per_launch"] == np_ * w ** 3 * 8
    assert r["fp64"]["achieved"] == pytest.approx(np_ * w ** 3 * 8 / (ms * 1e-3) / 1e12)
    assert r["fp64"]["peak"] == pytest.approx(78.6)
    assert 0 < r["fp64"]["frac"] < 1


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
