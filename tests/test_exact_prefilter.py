"""Exact shortcuts of cms-cuda_amd/csrc/cms_device.h against the definitions
they replace, compiled for the host from the same header (hipcc, no GPU):
the squared-distance prefilter of the triangle tests (norm_lt) against the
square-root comparison, and the closed-form subdivision count against its
defining loop."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run_check(tmp_path, name):
    hipcc = os.environ.get("HIPCC") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / name)
    src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
    inc = os.path.join(ROOT, "cms-cuda_amd", "csrc")
    cmd = [hipcc if os.path.exists(hipcc) else "hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", inc, src, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    fields = out.stdout.split()
    assert int(fields[fields.index("mismatches") + 1]) == 0
    return fields


def test_norm_lt_matches_sqrt(tmp_path):
    fields = _run_check(tmp_path, "norm_lt_check")
    assert int(fields[fields.index("near-band") + 1]) > 1000   # the exact branch was exercised


def test_subdiv_count_matches_loop(tmp_path):
    fields = _run_check(tmp_path, "subdiv_count_check")
    assert int(fields[fields.index("on-point") + 1]) > 100000   # angles exactly on a subdivision point
