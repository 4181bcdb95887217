"""Opt-in pytest plugin: runs the GPU test files against the CPU build of libkcore_hip (the
library's host code over the wave64 emulator), to rehearse them without a GPU. Nothing it runs is a
GPU result, and it only acts when named explicitly:

    python tests/hostemu/cpu_lib.py /tmp/libkcore_hip_cpu.so
    KCEMU_LIB=/tmp/libkcore_hip_cpu.so KCEMU_WALL_KHZ=10000000 PYTHONPATH=tests/hostemu \\
        python -m pytest -p cpu_lib_plugin tests/test_gpu_serve.py -m gpu

Full-size cases (R-MAT scale 24) take far too long on the CPU; select the others with -k."""
import os

import pytest

from kcoregpu_amd import _native as N


def pytest_configure(config):
    path = os.environ.get("KCEMU_LIB")
    if not path or not os.path.exists(path):
        raise pytest.UsageError("cpu_lib_plugin: KCEMU_LIB must name a CPU build of libkcore_hip (cpu_lib.py)")
    N.HIP_LIB = path


def pytest_report_header(config):
    return f"libkcore_hip: CPU emulation build {N.HIP_LIB} -- not a GPU run"
