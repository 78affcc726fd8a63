"""The C++ host layer (include/gpurt.hpp: gpurt::Mesh / gpurt::Accel, shaped like the
reference's VK::Mesh / VK::Accel) end to end, through the prebuilt accel_demo."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "gpu-rt_amd", "_build", "accel_demo")


def test_demo_is_built():
    assert os.path.exists(DEMO), "build gpu-rt_amd (make -C gpu-rt_amd) first"


@pytest.mark.gpu
def test_cpp_accel_demo_runs():
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "accel_demo ok" in r.stdout


TRACER_DEMO = os.path.join(ROOT, "gpu-rt_amd", "_build", "tracer_demo")


def test_tracer_demo_is_built():
    assert os.path.exists(TRACER_DEMO), "build gpu-rt_amd (make -C gpu-rt_amd) first"


@pytest.mark.gpu
def test_cpp_tracer_demo_runs():
    """gpurt::Tlas + gpurt::RTPipe (the reference's RTPipe shape) with exact answers."""
    r = subprocess.run([TRACER_DEMO], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "tracer_demo ok" in r.stdout
