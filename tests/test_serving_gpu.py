"""Serving stream with continuous batching: every request converges, and each
request's result is the same alone and in the mix (P1 requests bit for bit --
their vectors never interact; P2 requests up to the order of floating-point
atomic sums)."""
import numpy as np
import pytest

from hyteg_amd import serving

pytestmark = pytest.mark.gpu


def test_stream_converges_and_results_do_not_depend_on_the_mix(ctx):
    kinds = np.array([0, 0, 1, 0, 3, 0, 1, 2, 0, 0, 3, 0, 1, 0, 0, 0], dtype=np.int32)
    levels = np.array([3, 4, 3, 3, 3, 5, 3, 2, 4, 3, 4, 3, 4, 6, 3, 4], dtype=np.int32)
    seeds = np.arange(100, 100 + len(kinds), dtype=np.uint64)
    arrivals = np.linspace(0.0, 0.3, len(kinds))
    print("mixed stream", flush=True)
    mix = serving.serve_stream(kinds, levels, seeds, arrivals, tol=1e-8, slots=8)
    assert (mix["rel_residual"] <= 1e-8).all(), mix["rel_residual"]
    assert (mix["latency_s"] > 0).all()
    for i in (0, 2, 4, 7, 13):
        print(f"alone: request {i} kind {kinds[i]} level {levels[i]}", flush=True)
        one = serving.serve_stream(kinds[i:i + 1], levels[i:i + 1], seeds[i:i + 1], np.zeros(1), tol=1e-8, slots=8)
        assert one["cycles"][0] == mix["cycles"][i]
        if kinds[i] in (serving.P1_CUBE, serving.P1_SHELL):
            assert one["checksum"][0] == mix["checksum"][i]
        else:
            assert abs(one["checksum"][0] - mix["checksum"][i]) <= 1e-10 * abs(mix["checksum"][i])


def test_graph_replayed_cycles_equal_eager_cycles(ctx, monkeypatch):
    """P2 requests (shell level 2, cube level 3, six at once on six workers): the
    V-cycle captured once per request and replayed (14 times, with the residual
    of the eager code between replays) gives the cycle counts and, up to the
    order of floating-point atomic sums, the results of eager cycles; so do P1
    requests (one capture per worker, bit for bit)"""
    kinds = np.array([2, 2, 2, 3, 3, 3, 0, 1], dtype=np.int32)
    levels = np.array([2, 2, 2, 3, 3, 3, 4, 3], dtype=np.int32)
    seeds = np.arange(7, 7 + len(kinds), dtype=np.uint64)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HG_SERVE_NO_GRAPHS", mode)
        print(f"burst, graphs {'off' if mode == '1' else 'on'}", flush=True)
        out[mode] = serving.serve_stream(kinds, levels, seeds, np.zeros(len(kinds)), tol=1e-8, slots=8)
        assert (out[mode]["rel_residual"] <= 1e-8).all(), out[mode]["rel_residual"]
    assert np.array_equal(out["0"]["cycles"], out["1"]["cycles"])
    p1 = kinds <= 1
    assert np.array_equal(out["0"]["checksum"][p1], out["1"]["checksum"][p1])
    assert np.all(np.abs(out["0"]["checksum"] - out["1"]["checksum"]) <= 1e-10 * np.abs(out["1"]["checksum"]))
