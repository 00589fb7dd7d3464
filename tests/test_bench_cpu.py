"""CPU checks of bench.py's measurement plumbing: the committed PMC summaries carry the
traffic lines the bench reads, the roofline record has the fields and the arithmetic
its JSON line promises, and the arguments it refuses are refused before any GPU
call."""
import importlib.util
import os
import subprocess
import sys

import pytest

from lpm_harness import ROOT


def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("kernel", ["k_cg_spmv", "k_slice_mv", "k_j2_particle"])
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_committed_traffic_lines(kernel, form):
    """Every profiled kernel's PMC summary under profiles/ holds a machine-readable
    traffic line per form, at least the kernel's algorithmic bytes at 1M."""
    b = bench()
    t = b.measured_traffic(kernel, form)
    assert t is not None and t > 1e9, (kernel, form, t)


def test_roofline_record_fields_and_arithmetic():
    b = bench()
    by = 2047315200.0
    r = b.roofline((7.37, 10, by), "mixed", "k_j2_particle")
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in r
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["avg_launch_ms"] - 0.737) < 1e-12
    # achieved = algorithmic bytes per launch / average launch time, in GB/s
    assert abs(r["achieved"] - by / 0.737e-3 / 1e9) <= 1e-12 * r["achieved"]
    assert abs(r["frac"] - r["achieved"] / r["peak"]) <= 1e-15
    assert r["traffic"] == b.measured_traffic("k_j2_particle", "mixed")
    assert b.roofline((0.0, 0, by), "mixed", "k_j2_particle") is None


def test_dominant_kernel_comes_first():
    b = bench()
    prof = {"k_slice_mv": (5.0, 10, 2.4e9), "k_j2_particle": (7.0, 10, 2.0e9), "k_cg_spmv": (1.0, 2, 2.4e9)}
    rl = b.form_rooflines({"prof": prof}, "mixed")
    assert [r["kernel"] for r in rl] == ["k_j2_particle", "k_slice_mv", "k_cg_spmv"]
    assert rl[0]["total_ms"] == 7.0


@pytest.mark.parametrize("args,code,text", [(["--gpus", "2"], 2, "one GPU only"),
                                            (["--warmup", "45", "--steps", "10"], 1, "--warmup + --steps")])
def test_refused_arguments(args, code, text):
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == code and text in p.stderr, (p.returncode, p.stderr)
