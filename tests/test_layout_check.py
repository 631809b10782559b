"""Every address the sweep loops issue, checked on the CPU.

tests/cpp/layout_check.cpp is built with hipcc for the host only and calls
the very __host__ __device__ addressing functions the kernels inline
(sat_engine.hpp: SweepAddr, store_offset, WaveLayout, SegLayout;
sat_device.hpp: wave_geom, wave_index; sat_phases.hpp: wave_decode) for grid
widths 3..512, five heights, 1/2/4/8 waves per pair and the four sweep
directions.  It asserts that every load lies inside its plane and inside the
contiguous run of its step (idle lanes load what the last active lane
loads), that live stores stay below the raw buffer's num_records and hit each
slot once, that dropped stores lie at or past it, that segmented-layout idle
lanes store into the trash row, and that wave_decode inverts wave_index.
A run against planes one element short must fail.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip("hipcc is not available")
    exe = str(tmp_path_factory.mktemp("layout") / "layout_check")
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++20", "-O2", "-Wall", "-Werror",
           "-Wno-unused-function", os.path.join(ROOT, "tests", "cpp", "layout_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_every_sweep_address_is_in_bounds(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=600)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
    assert int(r.stdout.split()[1]) > 2000  # shapes checked


def test_a_plane_one_element_short_is_caught(checker):
    r = subprocess.run([checker, "--plant"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and r.stdout.startswith("FAIL"), r.stdout
