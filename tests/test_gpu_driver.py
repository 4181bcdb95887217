"""The `kcore <edgelist>` driver (reference main.cu:44-91) on the GPU: the reference's stdout lines
in order, the serialized cache on the second run, and core numbers / JSON / peel order equal to
the oracle's."""
import os
import re
import subprocess

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O
from kcoregpu_amd import _native as N

pytestmark = pytest.mark.gpu
BIN = os.path.join(N.BIN_DIR, "kcore")
LABELS = ["Ours: ", "SM:  ", "VP: ", "BC: ", "BC + SM: ", "BC + VP: ", "EC: ", "EC + SM: ", "EC + VP: "]


def run(cwd, *args):
    r = subprocess.run([BIN, *args], cwd=cwd, capture_output=True, text=True, timeout=180)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.returncode
    return r.stdout.split("\n")


def test_driver_lines_cache_and_outputs(tmp_path):
    g = K.Graph.erdos_renyi(3000, 15000, 11)
    g.write_edgelist(str(tmp_path / "er.txt"))
    os.makedirs(tmp_path / "output")
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)

    out = run(tmp_path, "er.txt", "--core-out", "core.u32", "--core-text", "core.txt", "--json", "--order-out",
              "order.u32")
    assert out[0] == "Graph loading Started... "
    assert out[1] == "serialized-er.txt : File couldn't open"
    assert out[2] == "Reading normal file... "
    assert re.fullmatch(r"File Loaded in: \d+", out[3])
    assert out[4] == "******************  er.txt ****************** "
    assert out[5] == f"V: {g.V}" and out[6] == f"E: {g.arcs}"
    i = 7
    for k, lab in enumerate(LABELS):
        assert re.fullmatch(re.escape(lab) + r"\d+ ", out[i]), out[i]
        i += 1
        if k == 0:
            assert out[i] == f"Kmax: {kmax}"
            i += 1
    assert out[i] == "Deallocated... " and out[i + 1] == ""

    got = np.fromfile(tmp_path / "core.u32", np.uint32)
    assert (got == core).all()
    assert (tmp_path / "core.txt").read_text() == "".join(f"{c}\n" for c in core)
    order = np.fromfile(tmp_path / "order.u32", np.uint32)
    assert O.check_certificate(g.neighbors_offset, g.neighbors, got, order) == 0
    shells = np.fromfile(tmp_path / "order.u32.shells", np.uint32)
    assert shells[0] == 0 and shells[-1] == g.V and len(shells) == kmax + 2
    # the JSON of non-zero cores in the reference's byte format
    want = "{ " + "".join(f'"{v}": {c}, \n' for v, c in enumerate(core) if c)
    want = want[:-3] + " }\n" if want != "{ " else want
    assert (tmp_path / "output" / "pkc-kcore-er.txt").read_text() == want

    # second run: the serialized cache written by the first run is read instead of the text
    out = run(tmp_path, "er.txt", "--variant", "lds", "--core-out", "core2.u32")
    assert out[:2] == ["Graph loading Started... ", "Reading serialized file... "]
    assert out[3] == f"V: {g.V}" and out[4] == f"E: {g.arcs}"
    assert re.fullmatch(r"lds: \d+ ", out[5]) and out[6] == f"Kmax: {kmax}"
    assert (np.fromfile(tmp_path / "core2.u32", np.uint32) == core).all()


@pytest.mark.parametrize("variant", ["async", "stepwise"])
def test_driver_other_variants(tmp_path, variant):
    g = K.Graph.erdos_renyi(3000, 15000, 12)
    g.write_edgelist(str(tmp_path / "er.txt"))
    os.makedirs(tmp_path / "output")
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    out = run(tmp_path, "er.txt", "--variant", variant, "--core-out", "core.u32")
    assert out[5] == f"V: {g.V}" and out[6] == f"E: {g.arcs}"
    assert re.fullmatch(re.escape(variant) + r": \d+ ", out[7]) and out[8] == f"Kmax: {kmax}"
    assert (np.fromfile(tmp_path / "core.u32", np.uint32) == core).all()


def test_driver_binary_cache(tmp_path):
    g = K.Graph.chung_lu(20000, 7)
    g.write_edgelist(str(tmp_path / "cl.txt"))
    os.makedirs(tmp_path / "output")
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    out = run(tmp_path, "cl.txt", "--binary-cache", "--variant", "auto", "--core-out", "c1.u32")
    assert out[1] == "Reading normal file... " and re.fullmatch(r"File Loaded in: \d+", out[2])
    out = run(tmp_path, "cl.txt", "--binary-cache", "--variant", "auto", "--core-out", "c2.u32")
    assert out[1] == "Reading binary cache... "
    assert out[3] == f"V: {g.V}" and out[4] == f"E: {g.arcs}" and out[6] == f"Kmax: {kmax}"
    for f in ("c1.u32", "c2.u32"):
        assert (np.fromfile(tmp_path / f, np.uint32) == core).all()
