"""Serving-stream driver: seeded Poisson-arrival request streams for
hg_serve_stream and their latency / throughput summary.

    python -m hyteg_amd.serving [--rates 10,50,200,500] [--requests 2000] [--tol 1e-8]

Request mix (kind, levels, share): P1 Poisson on the unit cube, levels 3-7,
40 %; P1 Poisson on the shell macro mesh, levels 3-5, 25 %; P2 div(k grad) on
the cube, levels 3-5, 20 %; P2 div(k grad) on the blended shell, levels 2-3,
15 %.  Coefficients and right-hand sides come from each request's seed.  Kind 4
(P2-P1 Stokes on the shell) is served too; --stokes-share mixes it in.
"""
import argparse
import ctypes as C
import json
import os

# The classes of a stream run on one stream each; HIP maps streams onto its
# hardware queues (4 by default), and classes sharing a queue run one after the
# other.  16 queues let the classes' kernels overlap (50 requests/s: p99 2.9 s
# against 16-18 s with 4 on one device).  Read when the HIP runtime loads, so it
# only takes effect when this module is imported before the library is loaded;
# a smaller setting (HIP's default is 4) is raised to 16.
if not os.environ.get("GPU_MAX_HW_QUEUES", "").isdigit() or int(os.environ["GPU_MAX_HW_QUEUES"]) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np  # noqa: E402

from ._lib import HygError, lib  # noqa: E402

P1_CUBE, P1_SHELL, P2_SHELL, P2_CUBE, STOKES_SHELL = 0, 1, 2, 3, 4
MIX = [  # (kind, levels, level weights, share)
    (P1_CUBE, (3, 4, 5, 6, 7), (0.25, 0.25, 0.25, 0.15, 0.10), 0.40),
    (P1_SHELL, (3, 4, 5), (0.4, 0.4, 0.2), 0.25),
    (P2_CUBE, (3, 4, 5), (0.4, 0.4, 0.2), 0.20),
    (P2_SHELL, (2, 3), (0.6, 0.4), 0.15),
]


def make_stream(n, rate, seed=0, stokes_share=0.0, stokes_levels=(2,)):
    """n requests, exponential inter-arrival times of mean 1/rate (seeded); with
    stokes_share > 0 that share of the requests are Stokes solves on the shell"""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals -= arrivals[0]
    mix = MIX + ([(STOKES_SHELL, tuple(stokes_levels), (1.0,) * len(stokes_levels), stokes_share)]
                 if stokes_share > 0 else [])
    shares = np.array([m[3] for m in mix])
    which = rng.choice(len(mix), n, p=shares / shares.sum())
    kinds = np.array([mix[w][0] for w in which], dtype=np.int32)
    levels = np.array([rng.choice(mix[w][1], p=np.array(mix[w][2]) / sum(mix[w][2])) for w in which], dtype=np.int32)
    seeds = rng.integers(0, 2 ** 62, n, dtype=np.uint64)
    return kinds, levels, seeds, arrivals


# macro-primitive counts (vertices, edges, faces, cells) of the two meshes
MESH_COUNTS = {"cube": (8, 19, 18, 6), "shell": (486, 2724, 4160, 1920)}


def _lattice_dofs(mesh, level):
    nv, ne, nf, nc = MESH_COUNTS[mesh]
    m = 2 ** level
    tri = lambda k: k * (k + 1) // 2 if k > 0 else 0  # noqa: E731
    tet = lambda k: k * (k + 1) * (k + 2) // 6 if k > 0 else 0  # noqa: E731
    return nv + ne * (m - 1) + nf * tri(m - 2) + nc * tet(m - 3)


def ndofs(kind, level):
    """canonical DoFs of a request's solution: P1 DoFs of `level`, P2 DoFs of
    `level` (= P1 DoFs of level + 1), or for Stokes three P2 velocity components
    and the P1 pressure"""
    mesh = "cube" if kind in (P1_CUBE, P2_CUBE) else "shell"
    if kind == STOKES_SHELL:
        return 3 * _lattice_dofs(mesh, level + 1) + _lattice_dofs(mesh, level)
    return _lattice_dofs(mesh, level + (1 if kind in (P2_SHELL, P2_CUBE) else 0))


def request_coefficient(seed):
    """(a, omega, phi) of k(x) = exp(sum_m a_m sin(omega_m . x + phi_m)) of a diffusion request"""
    a, om, ph = np.zeros(16), np.zeros(48), np.zeros(16)
    p = lambda v: v.ctypes.data_as(C.c_void_p)  # noqa: E731
    if lib().hg_serve_request_coefficient(C.c_uint64(int(seed)), p(a), p(om), p(ph)) != 0:
        raise HygError("hg_serve_request_coefficient failed")
    return a, om.reshape(16, 3), ph


def request_blobs(seed):
    """(12, 3) centres of the Gaussian temperature blobs of a Stokes request"""
    c = np.zeros(36)
    if lib().hg_serve_request_blobs(C.c_uint64(int(seed)), c.ctypes.data_as(C.c_void_p)) != 0:
        raise HygError("hg_serve_request_blobs failed")
    return c.reshape(12, 3)


def serve_stream(kinds, levels, seeds, arrivals, tol=1e-8, slots=32, mem_budget_gb=200.0, device=0,
                 solutions=False):
    """runs the stream; with solutions=True the result also holds every request's
    solution (canonical FP64 array) under the key solutions"""
    n = len(kinds)
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    levels = np.ascontiguousarray(levels, dtype=np.int32)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    arrivals = np.ascontiguousarray(arrivals, dtype=np.float64)
    lat, rel, chk = np.zeros(n), np.zeros(n), np.zeros(n)
    cyc = np.zeros(n, dtype=np.int32)
    wall = C.c_double()
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    sols, ptrs = None, None
    if solutions:
        sols = [np.zeros(ndofs(int(k), int(lv))) for k, lv in zip(kinds, levels)]
        ptrs = (C.c_void_p * n)(*[s.ctypes.data for s in sols])
    st = lib().hg_serve_stream_solutions(device, n, p(kinds), p(levels), p(seeds), p(arrivals), tol, slots,
                                         mem_budget_gb, p(lat), p(rel), p(chk), p(cyc), C.byref(wall), ptrs)
    if st != 0:
        raise HygError(f"hg_serve_stream: status {st}: {lib().hg_serve_last_error().decode()}")
    out = {"latency_s": lat, "rel_residual": rel, "checksum": chk, "cycles": cyc, "wall_s": wall.value}
    if solutions:
        out["solutions"] = sols
    return out


KIND_NAMES = {P1_CUBE: "p1_cube", P1_SHELL: "p1_shell", P2_SHELL: "p2_shell", P2_CUBE: "p2_cube",
              STOKES_SHELL: "stokes_shell"}


def summary(out, n, kinds=None):
    """throughput and latency percentiles, overall and (given kinds) per request kind"""
    lat = out["latency_s"]
    res = {"requests": n, "wall_s": out["wall_s"], "solves_per_s": n / out["wall_s"],
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "max_rel_residual": float(out["rel_residual"].max())}
    if kinds is not None:
        res["by_kind"] = {KIND_NAMES[k]: {"requests": int((kinds == k).sum()),
                                          "p50_latency_s": float(np.percentile(lat[kinds == k], 50)),
                                          "p99_latency_s": float(np.percentile(lat[kinds == k], 99))}
                          for k in sorted(set(int(v) for v in kinds))}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="10,50,200,500")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--slots", type=int, default=32)
    ap.add_argument("--kind", choices=list(KIND_NAMES.values()), default=None,
                    help="instead of the mixed streams: a burst of --requests requests of this kind at t = 0")
    ap.add_argument("--level", type=int, default=3, help="level of the --kind burst")
    ap.add_argument("--stokes-share", type=float, default=0.0, help="share of Stokes requests in the mixed streams")
    args = ap.parse_args()
    if args.kind is not None:
        kind = {v: k for k, v in KIND_NAMES.items()}[args.kind]
        n = args.requests
        kinds = np.full(n, kind, dtype=np.int32)
        out = serve_stream(kinds, np.full(n, args.level, dtype=np.int32), np.arange(1, n + 1, dtype=np.uint64),
                           np.zeros(n), tol=args.tol, slots=args.slots)
        cycles = sorted(set(int(c) for c in out["cycles"]))
        print(json.dumps({"kind": args.kind, "level": args.level, "cycles": cycles, **summary(out, n)}), flush=True)
        return
    for rate in [float(r) for r in args.rates.split(",")]:
        stream = make_stream(args.requests, rate, seed=0, stokes_share=args.stokes_share)
        out = serve_stream(*stream, tol=args.tol, slots=args.slots)
        print(json.dumps({"rate_per_s": rate, **summary(out, args.requests, stream[0])}), flush=True)


if __name__ == "__main__":
    main()
