"""Host-side logic that runs without a GPU: the serving arena's first-fit allocator (compiled from
its header with g++ and checked against a brute-force model), the registry that defers frees while
a serving kernel is resident, the served-workload request generator's distribution, and the host
library under sanitizers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import kcoregpu_amd as K

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_arena_first_fit_against_model(tmp_path):
    exe = tmp_path / "arena_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fsanitize=address,undefined", "-o", str(exe),
                    os.path.join(HERE, "cpp", "arena_test.cpp")], check=True)
    r = subprocess.run([str(exe), "30"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")


@pytest.mark.parametrize("sanitizer", ["address,undefined", "thread"])
def test_deferred_free_registry(tmp_path, sanitizer):
    """hipFree / hipHostFree synchronise the whole device and would wait for a resident serving
    kernel: the registry that defers the library's frees meanwhile (resident.h), with a fake
    backend, single-threaded and under eight threads, with ASan+UBSan and with TSan."""
    exe = tmp_path / "resident_test"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread", f"-fsanitize={sanitizer}", "-o", str(exe),
                    os.path.join(HERE, "cpp", "resident_test.cpp")], check=True)
    r = subprocess.run([str(exe), "20000" if sanitizer != "thread" else "5000"], capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0 and "WARNING: ThreadSanitizer" not in r.stderr, r.stdout + r.stderr[-3000:]
    assert "ok host_allocs" in r.stdout


def test_request_shape_distribution():
    """Edge counts: lognormal around the median, truncated to [mmin, mmax] by redrawing; the
    family is a fair coin; everything is a pure function of (seed, i)."""
    N = 20000
    ms, kinds = zip(*[K.Graph.request_shape(5, i, median_m=50000.0, sigma=1.0) for i in range(N)])
    ms, kinds = np.array(ms), np.array(kinds)
    assert ms.min() >= 1000 and ms.max() <= 50_000_000
    assert abs(np.median(ms) / 50000.0 - 1) < 0.05
    logm = np.log(ms)
    assert abs(logm.std() - 1.0) < 0.05  # sigma (truncation at 1K/50M cuts < 0.1% here)
    assert abs(kinds.mean() - 0.5) < 0.02
    assert K.Graph.request_shape(5, 17) == K.Graph.request_shape(5, 17)
    # a tight truncation window is respected
    tight = [K.Graph.request_shape(9, i, median_m=50000.0, sigma=3.0, mmin=40000, mmax=60000)[0]
             for i in range(500)]
    assert min(tight) >= 40000 and max(tight) <= 60000


def test_request_graphs_match_their_shape():
    for i in range(12):
        m, kind = K.Graph.request_shape(5, i)
        g = K.Graph.request(5, i)
        assert g.kind == kind
        assert 0.5 * m <= g.m <= m  # dedup and loops drop some of the m draws
        if kind == 1:
            assert g.V <= (2 * m + 11) // 12 + 1
        else:
            assert g.V <= 1 << max(6, int(np.ceil(np.log2(m / 16))))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_library_under_sanitizers(tmp_path):
    src = os.path.join(os.path.dirname(HERE), "kcoregpu_amd", "csrc")
    exe = tmp_path / "host_sanitize"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fopenmp", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-o", str(exe),
                    os.path.join(HERE, "cpp", "host_sanitize.cpp")] +
                   [os.path.join(src, f) for f in ("graph.cpp", "generators.cpp", "graph_capi.cpp", "bincache.cpp")],
                   check=True)
    fixtures = sorted(os.path.join(HERE, "golden", "loader", f)
                      for f in os.listdir(os.path.join(HERE, "golden", "loader")) if f.endswith(".txt"))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", TMPDIR=str(tmp_path))
    r = subprocess.run([str(exe), *fixtures], capture_output=True, text=True, timeout=300, env=env, cwd=tmp_path)
    assert r.returncode == 0 and "runtime error" not in r.stderr, r.stderr[-3000:]
    assert r.stdout.startswith("ok")
