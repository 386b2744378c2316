"""Benchmark of the matrix-free operator / multigrid hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload p2_shell|p1_cube] [--level L]

Default step: one application of the P2 variable-coefficient diffusion
operator div(k grad) (P2ElementwiseOperator with the blended
p2_div_k_grad_blending_q4 form, FP64, exact reference formulation, k stored at
the 11 quadrature points of every micro-cell) on the thick spherical shell
r in [0.55, 1] (ntan 5, 3 radial layers, 1920 macro-cells) at level 6:
673,710,594 DoFs, inputs resident in HBM.  Prints ONE JSON line with DoF/s per
apply, the HBM roofline of the dominant kernel (HIP-event timed on the
library's stream; algorithmic bytes: src, dst and one coefficient value per
DoF), the CPU oracle timed on this host, and secondary results: the apply with
k evaluated in the kernel, the low-precision (reduced-form) apply, the P1
apply and Gauss-Seidel sweep and the P2 transfers with their HBM fractions,
FP64 and mixed-precision GMG time-to-1e-8 (cfg2), the cfg1 GMG, 64 batched P1
solves (cfg4), the Stokes Uzawa GMG (cfg3), the serving stream (cfg5), and
CPU-oracle baselines of cfg2 (small level), cfg4 and cfg1.
"""
import argparse
import json
import os
import sys
import time

# hardware queues of the process (read when the HIP runtime loads): the serving
# leg runs one stream per request class, and with HIP's default of 4 queues
# classes sharing a queue serialise (a smaller setting is raised to 16); the
# other legs use one stream
if not os.environ.get("GPU_MAX_HW_QUEUES", "").isdigit() or int(os.environ["GPU_MAX_HW_QUEUES"]) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# FP64 operations of one micro-cell action in kP2TileApplyStored<double>, counted
# from its instruction stream (FMA = 2): 11 quadrature points x (43 FMA + 21 MUL +
# 5 ADD + 2 reciprocal / reciprocal-sqrt seeds) + about 230 for the gradient set-up
# and the result = 1484
P2_FLOPS_PER_ELEMENT = 11 * (2 * 43 + 21 + 5 + 2) + 230
METRIC = "DoF/s per apply; time-to-1e-8 residual mixed vs FP64; solves/s + p99 latency"


def tet(n):
    return 0 if n <= 0 else n * (n + 1) * (n + 2) // 6


def load_traffic(key):
    try:
        return json.load(open(os.path.join(ROOT, "profiles", "traffic.json"))).get(key)
    except (OSError, ValueError):
        return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_applies(ctx, apply, steps, warmup):
    for _ in range(warmup):
        apply()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        apply()
    ctx.synchronize()
    return (time.perf_counter() - t0) / steps


def kernel_ms(ctx, apply, reps):
    ctx.event_timing(True)
    ms = []
    for _ in range(reps):
        apply()
        ms.append(ctx.last_ms())
    ctx.event_timing(False)
    return float(np.mean(ms))


P2_STATE = {}


def p2_shell(hg, ctx, L, steps, warmup):
    from hyteg_amd import problems
    st = hg.PrimitiveStorage.shell(ctx, **problems.SHELL)
    a, om, ph = problems.shell_coefficient(0)
    A = hg.P2DivKGradOperator(st, L, L, a, om, ph)
    n = st.p2_ndofs(L)
    src = hg.P2Function("src", st, L, L)
    dst = hg.P2Function("dst", st, L, L)
    src.interpolate_random(1, -1.0, 1.0, L)
    log(f"P2 shell level {L}: {n} DoFs")
    apply = lambda: A.apply(src, dst, L, hg.ALL)  # noqa: E731
    sec = timed_applies(ctx, apply, steps, warmup)
    kms = kernel_ms(ctx, apply, max(3, min(steps, 5)))
    nelem = 1920 * 8 ** L
    alg = 3 * 8 * n  # algorithmic HBM bytes: src, dst and one coefficient value per P2 DoF, FP64
    roof = {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "algorithmic_bytes_per_launch": alg,
            "note": "algorithmic bytes = 24 per P2 DoF (src, dst, one k value, FP64); the exact "
                    "p2_div_k_grad_blending_q4 element integration issues ~1.5 kflop FP64 per micro-cell, so "
                    "the kernel is bound by FP64 VALU issue and latency, not by HBM (compute_fp64 below)",
            "compute_fp64": {"achieved_tflops": P2_FLOPS_PER_ELEMENT * nelem / (kms * 1e-3) / 1e12,
                             "peak_tflops": FP64_PEAK_TFLOPS,
                             "frac": P2_FLOPS_PER_ELEMENT * nelem / (kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                             "flops_per_launch": P2_FLOPS_PER_ELEMENT * nelem}}
    P2_STATE.update(st=st, A=A, src=src, dst=dst, L=L, n=n)
    return n, sec, kms, roof, {"workload": f"P2 div(k grad) apply (blended shell map, 11-point quadrature, "
                                           f"k = exp(16 sine modes) stored at the quadrature points -- 11 FP64 "
                                           f"values per micro-cell, a 53 GB table built once per level -- FP64), "
                                           f"spherical shell r 0.55..1, ntan 5, 3 layers (1920 macro-cells), "
                                           f"level {L} ({n} DoFs)",
                               "level": L, "dofs": n, "mesh": "meshSphericalShell(5, 3, 0.55, 1.0)",
                               "coefficients": "stored at the quadrature points"}, \
        "kP2TileApplyStored<double> + kP2TileApplyDiag<double> (interior and last-layer tiles of one apply)", \
        f"p2_apply_shell_L{L}_f64"


def p2_apply_variants(hg, ctx):
    """the headline operator with k evaluated in the kernel (the reference's default
    apply) and the reduced-form low-precision apply of the mixed-precision V-cycle"""
    st, A, L, n = P2_STATE["st"], P2_STATE["A"], P2_STATE["L"], P2_STATE["n"]
    src, dst = P2_STATE["src"], P2_STATE["dst"]
    out = {}
    A.store_coefficients(False)
    kms = kernel_ms(ctx, lambda: A.apply(src, dst, L, hg.ALL), 3)
    A.store_coefficients(True)
    out["k_in_kernel_f64"] = {"kernel": "kP2TileApplyEval<double>", "kernel_ms": kms, "dof_per_s": n / (kms * 1e-3),
                              "hbm_frac_algorithmic": 16 * n / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    for dt, name, s in ((hg.F16, "reduced_f16", 2), (hg.F32, "reduced_f32", 4)):
        fs = hg.P2Function("s", st, L, L, dtype=dt)
        fd = hg.P2Function("d", st, L, L, dtype=dt)
        fs.interpolate_random(1, -1.0, 1.0, L)
        apply = lambda: A.apply(fs, fd, L, hg.ALL)  # noqa: E731
        kms = kernel_ms(ctx, apply, 5)
        sec = timed_applies(ctx, apply, 5, 1)
        # src, dst, and one FP16 k value per micro-cell (about 0.75 per DoF)
        alg = (2 * s + 2 * 0.75) * n
        out[name] = {"kernel": f"kP2TileApplyReduced<{'__half' if dt == hg.F16 else 'float'}>",
                     "kernel_ms": kms, "ms_per_apply": 1e3 * sec, "dof_per_s": n / sec,
                     "hbm_frac_algorithmic": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        log(f"P2 reduced apply {name}: kernel {kms:.3f} ms, {1e3 * sec:.3f} ms per apply")
        fs.destroy()
        fd.destroy()
    return out


def p2_transfers(hg, ctx):
    """P2 quadratic restriction and prolongation between shell levels 6 and 5 (FP64),
    cell kernels HIP-event timed; algorithmic bytes 9 per fine DoF"""
    st, L = P2_STATE["st"], P2_STATE["L"]
    f = hg.P2Function("t", st, L - 1, L)
    f.interpolate_random(3, -1.0, 1.0, L)
    f.interpolate_random(4, -1.0, 1.0, L - 1)
    nf = st.p2_ndofs(L)
    r = kernel_ms(ctx, lambda: hg.restrict(f, L, hg.GMG_FLAG), 3)
    p = kernel_ms(ctx, lambda: hg.prolongate(f, L - 1, hg.GMG_FLAG, hg.ADD), 3)
    f.destroy()
    alg = 9 * nf
    log(f"P2 transfers level {L}<->{L - 1}: restrict {r:.3f} ms, prolongate {p:.3f} ms")
    return {"fine_dofs": nf, "restrict_kernel_ms": r, "prolongate_kernel_ms": p,
            "restrict_hbm_frac": alg / (r * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "prolongate_hbm_frac": alg / (p * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "note": "kP2RestrictCellList / kP2ProlongChunks; algorithmic bytes 9 per fine DoF (8 fine + 8/8 coarse)"}


def p1_cube(hg, ctx, L, steps, warmup):
    st = hg.PrimitiveStorage.cuboid(ctx)
    A = hg.P1ConstantLaplaceOperator(st, L, L)
    n = st.ndofs(L)
    src = hg.P1Function("src", st, L, L)
    dst = hg.P1Function("dst", st, L, L)
    src.interpolate_random(1, -1.0, 1.0, L)
    log(f"P1 cube level {L}: {n} DoFs")
    apply = lambda: A.apply(src, dst, L, hg.ALL)  # noqa: E731
    sec = timed_applies(ctx, apply, steps, warmup)
    kms = kernel_ms(ctx, apply, max(5, min(steps, 20)))
    alg = 2 * 8 * 6 * tet(2 ** L - 3)
    roof = {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "algorithmic_bytes_per_launch": alg}
    # one Gauss-Seidel sweep (multicolour: owned boundary DoFs, then 8 parity colours
    # of the cell interiors), timed from its first to its last kernel
    b = hg.P1Function("b", st, L, L)
    b.interpolate_random(2, -1.0, 1.0, L)
    gs = kernel_ms(ctx, lambda: A.smooth_gs(src, b, L, hg.GMG_FLAG), 3)
    roof["gauss_seidel_sweep"] = {"ms": gs, "dofs": n, "algorithmic_bytes": 3 * 8 * n,
                                  "achieved_gbs": 3 * 8 * n / (gs * 1e-3) / 1e9,
                                  "frac": 3 * 8 * n / (gs * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"P1 level {L}: apply kernel {kms:.3f} ms, Gauss-Seidel sweep {gs:.3f} ms")
    b.destroy()
    return n, sec, kms, roof, {"workload": f"P1ConstantLaplaceOperator apply, unit cube 6 macro-tets, level {L} "
                                          f"({n} DoFs), FP64", "level": L, "dofs": n, "mesh": "meshCuboid 1x1x1"}, \
        "kCellApplyBricks<double>", f"p1_apply_cube_L{L}_f64"


def p2_const_cube(hg, ctx, L=8, steps=10, warmup=2):
    """P2ConstantLaplaceOperator apply on the unit cube (6 macro-tets), level L, FP64:
    one 10x10 element product per micro-cell"""
    st = hg.PrimitiveStorage.cuboid(ctx)
    A = hg.P2ConstantLaplaceOperator(st, L, L)
    n = st.p2_ndofs(L)
    src = hg.P2Function("src", st, L, L)
    dst = hg.P2Function("dst", st, L, L)
    src.interpolate_random(1, -1.0, 1.0, L)
    apply = lambda: A.apply(src, dst, L, hg.ALL)  # noqa: E731
    sec = timed_applies(ctx, apply, steps, warmup)
    kms = kernel_ms(ctx, apply, 5)
    log(f"P2 constant apply cube level {L}: {n} DoFs, {1e3 * sec:.3f} ms/apply, kernel {kms:.3f} ms")
    for f in (src, dst):
        f.destroy()
    A.destroy()
    return {"dofs": n, "ms_per_apply": 1e3 * sec, "dof_per_s": n / sec, "kernel_ms": kms,
            "kernel_gbs_algorithmic": 16.0 * n / (kms * 1e-3) / 1e9}


def cpu_baseline(workload, seconds):
    """oracle (C restatement of the reference's CPU path), all host cores, bounded sample"""
    from oracle import pyoracle as po
    from hyteg_amd import problems
    if workload == "p2_shell":
        L = 3
        m = po.Mesh.shell(**problems.SHELL)
        op = po.P2Op(m, L, po.KCoeff(*problems.shell_coefficient(0)))
        n = m.ndofs(L + 1)
        desc = f"oracle P2 div(k grad) apply, shell level {L} ({n} DoFs)"
    else:
        L = 8
        m = po.Mesh.cuboid()
        op = po.P1Op(m, L)
        n = m.ndofs(L)
        desc = f"oracle P1 apply, cube level {L} ({n} DoFs)"
    x = np.random.default_rng(1).uniform(-1, 1, n)
    y = np.zeros(n)
    op.apply(x, y, po.ALL)
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        op.apply(x, y, po.ALL)
        reps += 1
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n * reps / dt, "unit": "DoF/s", "cores": cores, "kind": "port",
            "sample": f"{desc}, {reps} applies in {dt:.1f} s, OpenMP over macro-cells"}


def gmg_cfg1(hg, ctx):
    """cube L5, FP64 GMG V(3,3), GS smoother, CG at L0, to 1e-8 relative residual"""
    from oracle import pyoracle as po
    L = 5
    st = hg.PrimitiveStorage.cuboid(ctx)
    A = hg.P1ConstantLaplaceOperator(st, 0, L)
    fl = st.flags(L)
    b = np.where(fl == hg.INNER, np.random.default_rng(0).uniform(-1, 1, len(fl)), 0.0)
    x = hg.P1Function("x", st, 0, L)
    fb = hg.P1Function("b", st, 0, L)
    fb.upload(L, b)
    hg.gmg_solve(A, x, fb, 0, L, 3, 3, cycles=2)
    x.interpolate(0.0, L)
    ctx.synchronize()
    t0 = time.perf_counter()
    k, hist = hg.gmg_solve(A, x, fb, 0, L, 3, 3, cycles=100, tol=1e-8)
    ctx.synchronize()
    t_gpu = time.perf_counter() - t0
    om = po.Mesh.cuboid()
    xo = np.zeros(len(b))
    t0 = time.perf_counter()
    ko, ho = om.gmg(0, L, xo, b, 3, 3, cycles=100, tol=1e-8)
    t_cpu = time.perf_counter() - t0
    err = float(np.abs(x.download(L) - xo).max() / np.abs(xo).max())
    return {"workload": "P1 cube level 5 (35,937 DoFs), FP64 GMG V(3,3) GS, CG coarse, to 1e-8",
            "cycles": k, "time_s": t_gpu, "cpu_oracle_cycles": ko, "cpu_oracle_time_s": t_cpu,
            "mean_factor": float((hist[-1] / hist[0]) ** (1.0 / max(k, 1))),
            "rel_diff_vs_oracle_solution": err}


def gmg_p2_shell(hg, ctx, L):
    """FP64 GMG V(2,2), Chebyshev order 3 (radius from 20 power iterations on the finest
    level), P2 quadratic transfer, Jacobi-PCG to 1e-14 on level 0, to 1e-8 relative residual"""
    from hyteg_amd import problems
    st = hg.PrimitiveStorage.shell(ctx, **problems.SHELL)
    A = hg.P2DivKGradOperator(st, 0, L, *problems.shell_coefficient(0))
    x = hg.P2Function("x", st, 0, L)
    b = hg.P2Function("b", st, 0, L)
    b.interpolate_random(0, -1.0, 1.0, L, hg.INNER)
    t = hg.P2Function("t", st, L, L)
    t.interpolate_random(1, -1.0, 1.0, L, hg.ALL)
    ctx.synchronize()
    t0 = time.perf_counter()
    rho = hg.p2_estimate_radius(A, t, L, 20)
    hg.p2_gmg_solve(A, x, b, 0, L, rho, cycles=1)  # builds the per-level data once
    ctx.synchronize()
    setup = time.perf_counter() - t0
    x.interpolate(0.0, L)
    log(f"P2 GMG level {L}: radius {rho:.4f}, set-up {setup:.2f} s")
    ctx.synchronize()
    t0 = time.perf_counter()
    k, hist = hg.p2_gmg_solve(A, x, b, 0, L, rho, pre=2, post=2, order=3, cycles=60, tol=1e-8)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    log(f"P2 GMG level {L}: {k} cycles, {dt:.2f} s")
    out = {"workload": f"P2 div(k grad) shell level {L}, FP64 GMG V(2,2) Chebyshev(3), Jacobi-PCG to 1e-14 on "
                       f"level 0, to 1e-8 relative residual",
           "cycles": k, "time_s": dt, "setup_s": setup, "spectral_radius": rho,
           "factors": [float(v) for v in hist[1:] / hist[:-1]],
           "relative_residual": float(hist[-1] / hist[0])}
    t.destroy()
    # mixed precision: FP64 iterative refinement around one V-cycle per step in FP32
    # arithmetic on FP16 (or FP32) vectors; the low-precision applies use the reduced
    # element form (affine micro-cells, k constant per micro-cell, FP32 accumulation)
    for low, key, what in ((hg.F16, "mixed_fp32_compute_fp16_storage", "FP16 vectors"),
                           (hg.F32, "mixed_fp32_compute_fp32_storage", "FP32 vectors")):
        xm = hg.P2Function("xm", st, 0, L, dtype=hg.F64)
        hg.p2_ir_solve(A, xm, b, 0, L, rho, low_dtype=low, max_outer=1)  # builds the low-precision tables once
        xm.interpolate(0.0, L)
        ctx.synchronize()
        t0 = time.perf_counter()
        km, hm = hg.p2_ir_solve(A, xm, b, 0, L, rho, low_dtype=low, inner_cycles=1, max_outer=60, tol=1e-8)
        ctx.synchronize()
        dtm = time.perf_counter() - t0
        # || x_mixed - x_fp64 ||_2 / || x_fp64 ||_2 on the device
        xm.assign([1.0, -1.0], [xm, x], L, hg.ALL)
        err = (xm.dot(xm, L, hg.ALL) / x.dot(x, L, hg.ALL)) ** 0.5
        log(f"P2 mixed IR ({what}) level {L}: {km} outer steps, {dtm:.2f} s, speed-up {dt / dtm:.2f}, "
            f"error vs FP64 {err:.2e}")
        out[key] = {
            "workload": "FP64 iterative refinement (FP64 residual, FP64 PCG on level 0), one V(2,2) Chebyshev(3) "
                        f"cycle per step in FP32 arithmetic with {what}, FP16 coefficients and FP32 accumulation "
                        "(reduced element form on levels >= 2), to 1e-8",
            "outer_steps": km, "time_s": dtm, "speedup_vs_fp64": dt / dtm,
            "relative_residual": float(hm[-1] / hm[0]),
            "factors": [float(v) for v in hm[1:] / hm[:-1]],
            "rel_l2_difference_to_fp64_solution": err}
        xm.destroy()
    for f in (x, b):
        f.destroy()
    A.destroy()
    return out


def batched_cfg4(hg, ctx, L=6, nvec=64):
    """64 P1 Laplace systems on the cube at level 6, RHS uniform[-1,1] on inner DoFs from
    seeds 0..63: one batched FP16-storage V(3,3) Gauss-Seidel hierarchy inside FP64
    iterative refinement (interleaved vectors) versus 64 sequential FP64 GMG solves,
    all to 1e-10 relative residual"""
    st = hg.PrimitiveStorage.cuboid(ctx)
    A = hg.P1ConstantLaplaceOperator(st, 0, L)
    xb = hg.P1Function("xb", st, 0, L, nvec=nvec)
    bb = hg.P1Function("bb", st, 0, L, nvec=nvec)
    bb.interpolate_random(0, -1.0, 1.0, L, hg.INNER)
    hg.batched_ir_solve(A, xb, bb, 0, L, low_dtype=hg.F16, max_outer=1)
    xb.interpolate(0.0, L)
    ctx.synchronize()
    t0 = time.perf_counter()
    kir, hir = hg.batched_ir_solve(A, xb, bb, 0, L, low_dtype=hg.F16, max_outer=60, tol=1e-10)
    ctx.synchronize()
    t_ir = time.perf_counter() - t0
    log(f"batched FP16 refinement: {kir} steps, {t_ir:.3f} s for {nvec} systems")
    # 64 sequential FP64 solves, each from its own seed
    x = hg.P1Function("x", st, 0, L)
    b = hg.P1Function("b", st, 0, L)
    e = hg.P1Function("e", st, L, L)
    errs, cycles, t_seq = [], [], 0.0
    xs = xb.download(L)
    for k in range(nvec):
        b.interpolate_random(k, -1.0, 1.0, L, hg.INNER)
        x.interpolate(0.0, L)
        ctx.synchronize()
        t0 = time.perf_counter()
        kk, _ = hg.gmg_solve(A, x, b, 0, L, 3, 3, cycles=60, tol=1e-10)
        ctx.synchronize()
        t_seq += time.perf_counter() - t0
        cycles.append(kk)
        xk = x.download(L)
        errs.append(float(np.abs(xs[:, k] - xk).max() / np.abs(xk).max()))
    log(f"sequential FP64: {t_seq:.3f} s for {nvec} systems")
    return {"workload": f"{nvec} x P1 Laplace, unit cube level {L} ({st.ndofs(L)} DoFs each), RHS seeds 0..{nvec - 1}, "
                        "to 1e-10", "batched_fp16_ir": {"outer_steps": kir, "time_s": t_ir, "solves_per_s": nvec / t_ir},
            "sequential_fp64_gmg": {"cycles_per_solve": cycles[0], "time_s": t_seq, "solves_per_s": nvec / t_seq},
            "speedup": t_seq / t_ir, "max_rel_error_vs_fp64": max(errs)}


def serving_cfg5(n=400, rate=200.0):
    """seeded Poisson stream of mixed P1/P2 solve requests (to 1e-8) through the
    continuous-batching server; latency = completion - arrival"""
    from hyteg_amd import serving
    stream = serving.make_stream(n, rate, seed=0)
    out = serving.serve_stream(*stream, tol=1e-8, slots=32)
    res = serving.summary(out, n, stream[0])
    log(f"serving {n} requests at {rate:g}/s: {res['solves_per_s']:.1f} solves/s, p50 {res['p50_latency_s'] * 1e3:.1f} ms, "
        f"p99 {res['p99_latency_s'] * 1e3:.1f} ms")
    return {"workload": f"{n} requests, Poisson arrivals at {rate:g}/s, mix P1 cube L3-7 40%, P1 shell L3-5 25%, "
                        "P2 cube L3-5 20%, P2 shell L2-3 15%, each solved to 1e-8", "rate_per_s": rate,
            "hip_hardware_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), **res}


def stokes_cfg3(hg, ctx, L=5, cycles=120, tol=1e-7):
    """P2-P1 Stokes on the shell at level L, eta = exp(ln(1e4)(1-r)/0.45), nodal
    buoyancy of 12 Gaussian blobs, free-slip on the inner and no-slip on the outer
    surface (free-slip operator wrapper, normal projections): the all-at-once
    Uzawa-smoothed GMG of the free-slip solver template (V(3,3) + 2 per coarser
    level, omega 0.3, Chebyshev(3) velocity smoothing, MINRES(10) on level 0),
    FP64, until the residual drops by `tol`"""
    from hyteg_amd import problems
    st = hg.PrimitiveStorage.shell(ctx, **problems.SHELL, inner_flag=hg.FREESLIP, outer_flag=hg.DIRICHLET)
    A = hg.P2P1StokesOperator(st, 0, L, *problems.STOKES_VISCOSITY, freeslip=True)
    P = hg.P2ProjectNormalOperator(st)
    X = hg.P2P1TaylorHoodFunction("x", st, 0, L)
    B = hg.P2P1TaylorHoodFunction("b", st, 0, L)
    n2, n1 = st.p2_ndofs(L), st.ndofs(L)
    t0 = time.perf_counter()
    xyz = st.coords(L + 1)
    f = problems.buoyancy(xyz)
    fl = st.flags(L + 1)
    live = (fl == hg.INNER) | (fl == hg.FREESLIP)
    del xyz
    B.upload(L, np.concatenate([np.where(live, f[:, c], 0.0) for c in range(3)]), np.zeros(n1))
    del f, fl, live
    P.project(B, L)
    for c in range(3):
        X.uvw[c].interpolate_random(7 + c, -1.0, 1.0, L, hg.INNER)
    T = hg.P2P1TaylorHoodFunction("t", st, L, L)
    rho = hg.stokes_estimate_radius(A, X, T, L, 25)
    T.destroy()
    ctx.synchronize()
    setup = time.perf_counter() - t0
    log(f"Stokes level {L}: {3 * n2 + n1} DoFs, radius {rho:.4f}, set-up {setup:.2f} s")
    Y = hg.P2P1TaylorHoodFunction("y", st, L, L)
    kms = kernel_ms(ctx, lambda: A.apply(B, Y, L, hg.GMG_FLAG), 3)
    Y.destroy()
    X.interpolate(0.0, L)
    ctx.synchronize()
    t0 = time.perf_counter()
    k, hist = hg.stokes_gmg_solve(A, X, B, 0, L, rho, cycles=cycles, tol=tol)
    ctx.synchronize()
    t_uz = time.perf_counter() - t0
    rel = float(hist[-1] / hist[0])
    log(f"Stokes Uzawa GMG: {k} cycles {t_uz:.2f} s, relative residual {rel:.2e}")
    for fn in (X, B):
        fn.destroy()
    A.destroy()
    return {"workload": f"P2-P1 Stokes, shell level {L} ({3 * n2} velocity + {n1} pressure DoFs), "
                        "eta = exp(ln(1e4)(1-r)/0.45), 12-blob buoyancy, free-slip inner / no-slip outer",
            "dofs": 3 * n2 + n1, "setup_s": setup, "spectral_radius": rho,
            "stokes_apply_kernel_ms": kms, "stokes_apply_dof_per_s": (3 * n2 + n1) / (kms * 1e-3),
            "uzawa_gmg_f64": {"solver": "Uzawa GMG V(3,3)+2, omega 0.3, Chebyshev(3), MINRES(10) on level 0",
                              "tol": tol, "cycles": k, "time_s": t_uz, "s_per_cycle": t_uz / max(k, 1),
                              "relative_residual": rel, "reached_tol": bool(rel <= tol),
                              "factors_first": [float(v) for v in (hist[1:] / hist[:-1])[:5]],
                              "mean_factor": float(rel ** (1.0 / max(k, 1)))}}


def cpu_extra_baselines():
    """oracle (C restatement, all host cores) solves of cfg1, cfg2 (at a level the
    oracle finishes in bounded time) and cfg4"""
    from oracle import pyoracle as po
    from hyteg_amd import problems
    out = {}
    # cfg2 at shell level 2 (174,114 DoFs): FP64 GMG V(2,2) Chebyshev(3) to 1e-8
    m = po.Mesh.shell(**problems.SHELL)
    k = po.KCoeff(*problems.shell_coefficient(0))
    L = 2
    fl = m.flags(L + 1)
    b = np.where(fl == po.INNER, problems.canonical_uniform(len(fl), 0), 0.0)
    t0 = time.perf_counter()
    op = po.P2Op(m, L, k)
    rho = op.power_iteration(op.inv_diag(), problems.canonical_uniform(len(fl), 1), 20)
    x = np.zeros(len(fl))
    kk, hist = m.p2_gmg(0, L, k, x, b, 2, 2, 3, rho, cycles=60, tol=1e-8)
    out["cfg2_oracle_shell_L2"] = {"dofs": len(fl), "cycles": kk, "time_s": time.perf_counter() - t0,
                                   "note": "power iteration (20) + FP64 GMG V(2,2) Chebyshev(3) to 1e-8"}
    # cfg4: 64 cube level-6 systems, FP64 GMG V(3,3) lexicographic GS to 1e-10, OpenMP over the systems
    mc = po.Mesh.cuboid()
    L = 6
    fl = mc.flags(L)
    B = np.stack([np.where(fl == po.INNER, problems.canonical_uniform(len(fl), s), 0.0) for s in range(64)])
    t0 = time.perf_counter()
    X, done = mc.gmg_many(0, L, np.zeros_like(B), B, cycles=60, tol=1e-10)
    dt = time.perf_counter() - t0
    out["cfg4_oracle_64_solves"] = {"time_s": dt, "solves_per_s": 64 / dt, "cycles": int(done.max()),
                                    "note": "oracle GMG, one system per OpenMP thread"}
    out["cores"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="p2_shell", choices=["p2_shell", "p1_cube"])
    ap.add_argument("--level", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gmg", action="store_true", help="skip the P2 GMG solve")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched 64-system P1 solve")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-stokes", action="store_true", help="skip the Stokes solve")
    ap.add_argument("--only-apply", action="store_true",
                    help="run the apply measurement only (for counter collection)")
    ap.add_argument("--only", default=None,
                    help="the apply measurement and the named legs only (comma separated: variants, transfers, "
                         "p1, p2const, gmg, cfg1, cfg4, stokes, serving, cpu)")
    ap.add_argument("--stokes-cycles", type=int, default=120, help="V-cycle cap of the Stokes leg")
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--serving-rates", default="10,50,200,500")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path would shard by "
              "macro-cell: each GPU a share of the macro-cells, the interface halos exchanged between them)",
              file=sys.stderr)
        sys.exit(2)

    import hyteg_amd as hg
    ctx = hg.Context(0)
    run = p2_shell if args.workload == "p2_shell" else p1_cube
    L = args.level if args.level is not None else (6 if args.workload == "p2_shell" else 9)
    n, sec, kms, roofline, config, kname, tkey = run(hg, ctx, L, args.steps, args.warmup)
    log(f"{args.steps} applies: {1e3 * sec:.3f} ms/apply, dominant kernel {kms:.3f} ms")
    roofline["frac"] = roofline["achieved"] / roofline["peak"]
    roofline.update({"traffic": load_traffic(tkey), "kernel": kname, "kernel_avg_ms": kms})
    legs = set() if args.only_apply else (
        set(args.only.split(",")) if args.only else
        {"variants", "transfers", "p1", "p2const", "gmg", "cfg1", "cfg4", "stokes", "serving", "cpu"})
    if args.no_gmg:
        legs.discard("gmg")
    if args.no_batched:
        legs.discard("cfg4")
    if args.no_serving:
        legs.discard("serving")
    if args.no_stokes:
        legs.discard("stokes")
    if args.no_cpu:
        legs.discard("cpu")
    extra = {}
    p2 = args.workload == "p2_shell"
    if p2 and "variants" in legs:
        extra["p2_apply_variants_shell_L6"] = p2_apply_variants(hg, ctx)
    if p2 and "transfers" in legs:
        extra["p2_transfers_shell_L6_f64"] = p2_transfers(hg, ctx)
    if p2 and "p1" in legs:
        n1, sec1, kms1, roof1, cfg1, _, _ = p1_cube(hg, ctx, 9, 20, 3)
        extra["p1_apply_cube_L9_f64"] = {"dofs": n1, "ms_per_apply": 1e3 * sec1, "dof_per_s": n1 / sec1,
                                         "kernel_ms": kms1, "kernel_gbs": roof1["achieved"],
                                         "hbm_frac": roof1["achieved"] / HBM_PEAK_GBS,
                                         "traffic": load_traffic("p1_apply_cube_L9_f64"),
                                         "gauss_seidel_sweep": roof1["gauss_seidel_sweep"]}
    if p2 and "p2const" in legs:
        extra["p2_const_apply_cube_L8_f64"] = p2_const_cube(hg, ctx)
    if p2 and "gmg" in legs:
        extra["gmg_p2_shell_f64"] = gmg_p2_shell(hg, ctx, L)
    if "cfg1" in legs:
        extra["gmg_p1_cube_L5_f64"] = gmg_cfg1(hg, ctx)
    if p2 and "cfg4" in legs:
        extra["batched_p1_cube_L6_x64"] = batched_cfg4(hg, ctx)
    if p2 and "stokes" in legs:
        extra["stokes_shell_L5"] = stokes_cfg3(hg, ctx, cycles=args.stokes_cycles)
    if p2 and "serving" in legs:
        # below 200/s a stream of 30 s of arrivals (the device keeps up there, so
        # the latencies are service times, not queueing)
        extra["serving_mixed_stream"] = [serving_cfg5(min(args.serving_requests, max(100, int(30 * float(r)))),
                                                      float(r))
                                         for r in args.serving_rates.split(",")]
    cpu = None
    if "cpu" in legs:
        cpu = cpu_baseline(args.workload, args.cpu_seconds)
        if p2:
            extra["cpu_oracle_solves"] = cpu_extra_baselines()
    out = {
        "metric": METRIC, "value": n / sec, "unit": "DoF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sec, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded uniform[-1,1] input, seeded k)",
        "config": config, "roofline": roofline, "cpu_baseline": cpu, **extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
