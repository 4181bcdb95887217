"""Code generation of the gfx950 kernels, read from the compiler's own resource report
(-Rpass-analysis=kernel-resource-usage, a device-only compile here, no GPU): no kernel may use
scratch memory or spill vector registers. Scratch is what an out-of-line callee taking the kernel
arguments by reference costs (the arguments are copied to scratch and reloaded from there), so
this keeps every call on the kernels' paths inlined."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "kcoregpu_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="needs hipcc")


def resource_report(source, tmp_path):
    hipcc = HIPCC if os.path.exists(HIPCC) else "hipcc"
    r = subprocess.run([hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "--cuda-device-only",
                        "-c", os.path.join(CSRC, source), "-o", str(tmp_path / (source + ".o")),
                        "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-2000:]
    rows, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            rows[cur] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs Spill|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and cur:
            rows[cur][m.group(1).split(" [")[0]] = int(m.group(2))
    return rows


@pytest.mark.parametrize("source,min_kernels", [("kcore_hip.hip", 25), ("serve.hip", 3)])
def test_no_scratch_and_no_spills(tmp_path, source, min_kernels):
    rows = resource_report(source, tmp_path)
    assert len(rows) >= min_kernels, sorted(rows)
    for name, r in rows.items():
        assert r.get("ScratchSize") == 0, (name, r)
        assert r.get("VGPRs Spill") == 0, (name, r)  # SGPR spills go to VGPR lanes, not to memory
    peel = {n: r for n, r in rows.items() if "peel_" in n}
    if source == "kcore_hip.hip":
        # every single-graph kernel (sub-round, cascade-queue, hand-off) and the batch kernel
        assert len(peel) == 21, sorted(peel)
        assert min(r["Occupancy"] for r in peel.values()) >= 5, peel
