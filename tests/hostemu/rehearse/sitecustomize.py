"""Opt-in rehearsal of whole programs (bench.py with its child processes, scripts) on the CPU build
of libkcore_hip (tests/hostemu/cpu_lib.py): with this directory on PYTHONPATH and KCEMU_LIB naming
the CPU build, every Python process started points the package at that library. Nothing such a run
prints is a GPU result; a banner on stderr says so. Test infrastructure only.

    python tests/hostemu/cpu_lib.py /tmp/libkcore_hip_cpu.so
    PYTHONPATH=tests/hostemu/rehearse KCEMU_LIB=/tmp/libkcore_hip_cpu.so KCEMU_WALL_KHZ=10000000 \\
        python bench.py --scale 12 --batch-graphs 8 --only single"""
import os
import sys

_lib = os.environ.get("KCEMU_LIB")
if _lib:
    try:
        from kcoregpu_amd import _native

        _native.HIP_LIB = _lib
        print(f"[rehearsal] libkcore_hip is the CPU emulation build {_lib}: not a GPU run", file=sys.stderr,
              flush=True)
    except ImportError:
        pass
