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
