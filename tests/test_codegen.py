"""Code generation of the gfx950 kernels, read from the compiler's own resource report
(-Rpass-analysis=kernel-resource-usage, a device-only compile here, no GPU): no kernel may use
scratch memory or spill vector registers. Scratch is what an out-of-line callee taking the kernel
arguments by reference costs (the arguments are copied to scratch and reloaded from there), so
this keeps every call on the kernels' paths inlined. From the same compile's assembly: the degree
decrements of the peel loops overlap, i.e. a group of four atomics is not separated by full memory
waits (each arc's load and atomic under its own branch made the compiler wait for everything
before every one)."""
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
    """Per-kernel resource rows, and the device assembly (compiled here from this repository)."""
    hipcc = HIPCC if os.path.exists(HIPCC) else "hipcc"
    asm = tmp_path / (source + ".s")
    r = subprocess.run([hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "--cuda-device-only",
                        "-S", os.path.join(CSRC, source), "-o", str(asm),
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
    return rows, asm.read_text()


def kernel_bodies(asm):
    """name -> instruction lines of every function in the assembly."""
    out, name = {}, None
    for line in asm.splitlines():
        m = re.match(r"^([_A-Za-z][\w.$]*):", line)
        if m and not line.startswith("."):
            name = m.group(1)
            out[name] = []
        elif name and line.startswith("\t") and not line.strip().startswith(";"):
            out[name].append(line.strip())
    return out


def decrement_groups(body):
    """Each run of four consecutive degree-decrement atomics (within 40 instructions), with the
    number of full memory waits between its first and last atomic."""
    idx = [i for i, ins in enumerate(body) if ins.startswith("global_atomic_sub ")]
    groups = []
    i = 0
    while i + 3 < len(idx):
        if idx[i + 3] - idx[i] <= 40:
            waits = sum(1 for ins in body[idx[i]:idx[i + 3]] if ins.startswith("s_waitcnt") and "vmcnt(0)" in ins)
            groups.append(waits)
            i += 4
        else:
            i += 1
    return groups


_COMPILED = {}


@pytest.fixture(scope="module")
def compiled(tmp_path_factory):
    """One compile per source for the whole module."""
    def get(source):
        if source not in _COMPILED:
            _COMPILED[source] = resource_report(source, tmp_path_factory.mktemp("codegen"))
        return _COMPILED[source]
    return get


@pytest.mark.parametrize("source,min_kernels", [("kcore_hip.hip", 25), ("serve.hip", 3)])
def test_no_scratch_and_no_spills(compiled, source, min_kernels):
    rows, asm = compiled(source)
    assert len(rows) >= min_kernels, sorted(rows)
    for name, r in rows.items():
        assert r.get("ScratchSize") == 0, (name, r)
        assert r.get("VGPRs Spill") == 0, (name, r)  # SGPR spills go to VGPR lanes, not to memory
    peel = {n: r for n, r in rows.items() if "peel_" in n}
    if source == "kcore_hip.hip":
        # every single-graph kernel (sub-round, cascade-queue, hand-off) and the batch kernel
        assert len(peel) == 21, sorted(peel)
        assert min(r["Occupancy"] for r in peel.values()) >= 5, peel


@pytest.mark.parametrize("source", ["kcore_hip.hip", "serve.hip"])
def test_decrements_overlap(compiled, source):
    _, asm = compiled(source)
    bodies = kernel_bodies(asm)
    kernels = {n: b for n, b in bodies.items() if re.search(r"peel_persistent|peel_async|peel_batch|serve_kernel", n)}
    assert kernels
    for name, body in kernels.items():
        groups = decrement_groups(body)
        assert groups, name  # every peel kernel decrements four arcs per lane at a time
        assert max(groups) <= 1, (name, groups)
