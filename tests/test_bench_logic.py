"""bench.py's own logic without a GPU: the argument contract, the served leg's arrival-rate rule
(0.7 x the measured served capacity unless a rate is given) and sweep, and the JSON keys it reports,
with kcore_serve replaced by a stand-in that answers from its arguments."""
import argparse
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def serve_args(**kw):
    a = dict(serve_pool=8, serve_seed=5, serve_seconds=1.0, serve_rate=0.0, serve_closed=100,
             serve_sweep=[], plain_rate=150.0, plain_closed=20)
    a.update(kw)
    return argparse.Namespace(**a)


def fake_run_serve(calls):
    def run(argv, limit):
        calls.append(list(argv))
        mode = argv[argv.index("--mode") + 1]
        closed = "--closed" in argv
        rate = float(argv[argv.index("--rate") + 1]) if "--rate" in argv else 0.0
        gps = {("served", True): 10000.0, ("plain", True): 200.0}.get((mode, closed), rate)
        return {"offered_rate": rate, "seconds": 1.0, "closed": 100 if closed else 0, "requests": 100,
                "completed": 100, "unfinished": 0, "dropped": 0, "overloaded": False, "graphs_per_s": gps,
                "edges_per_s": gps * 5e4, "p50_ms": 1.0, "p90_ms": 2.0, "p99_ms": 3.0, "max_ms": 4.0,
                "mismatches": 0, "checked_against": "expected core numbers"}
    return run


def test_served_rate_is_seventy_percent_of_capacity(monkeypatch):
    calls = []
    monkeypatch.setattr(bench, "run_serve", fake_run_serve(calls))
    res = bench.bench_serve(serve_args(serve_sweep=[2000.0, 16000.0]), None, "/x/expect.bin")
    assert calls[0][calls[0].index("--mode") + 1] == "served" and "--closed" in calls[0]
    served = calls[1]
    assert served[served.index("--rate") + 1] == "7000"
    assert res["rate_rule"] == "0.7 x served_capacity: 7000/s"
    assert res["capacity_vs_plain_capacity"] == 10000.0 / 200.0
    assert {"served_capacity", "served", "served_at_plain_rate", "plain", "plain_capacity",
            "served_at_2000", "served_at_16000"} <= set(res)
    for c in calls:
        assert c[c.index("--expect") + 1] == "/x/expect.bin"


def test_served_rate_given(monkeypatch):
    calls = []
    monkeypatch.setattr(bench, "run_serve", fake_run_serve(calls))
    res = bench.bench_serve(serve_args(serve_rate=3000.0), None, None)
    served = calls[1]
    assert float(served[served.index("--rate") + 1]) == 3000.0
    assert res["rate_rule"].startswith("--serve-rate")
    assert all("--expect" not in c for c in calls)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_child_leg_failures_are_reported(monkeypatch):
    """The tuning and the batched leg run as child processes: one that dies, prints nothing or
    overruns its limit comes back as an error string, never as an exception."""
    args = argparse.Namespace()

    def dead(argv, **kw):
        assert argv[-2:] == ["--leg", "batch"] or "--tune-cache" in argv
        return subprocess.CompletedProcess(argv, -11, stdout="partial output\n")

    monkeypatch.setattr(bench.subprocess, "run", dead)
    assert bench.run_leg(args, "batch", 10) == (None, "child process exited -11")

    def slow(argv, **kw):
        raise subprocess.TimeoutExpired(argv, kw.get("timeout"))

    monkeypatch.setattr(bench.subprocess, "run", slow)
    assert bench.run_leg(args, "tune", 5, tune_cache="/x/t.txt") == (None, "timed out after 5 s")

    def ok(argv, **kw):
        assert argv[argv.index("--tune-cache") + 1] == "/x/t.txt"
        return subprocess.CompletedProcess(argv, 0, stdout='noise\n{"tuned": {"variant": 2}}\n')

    monkeypatch.setattr(bench.subprocess, "run", ok)
    assert bench.run_leg(args, "tune", 5, tune_cache="/x/t.txt") == ({"tuned": {"variant": 2}}, None)
