#!/usr/bin/env python3
"""Benchmark of the edge-assembly + pressure-solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
                    [--train-steps N] [--epochs E] [--only PART] [--no-cpu]

Headline (one JSON line on stdout): the continuity edge assembly
(ContinuityEdgeSolverAlg, FP64, the pressure-Poisson system) on the
256^3-cell Hex8 box graded toward z = 0 with the z grading of the 128^3 box
(per-cell ratio 1.05^(128/256) = 1.0247, first/last spacing ratio 490),
divergence-free random-Fourier velocity of 8 modes.  A step is one
assembly over all edges with its inputs resident in HBM; value = edges /
step time in Medges/s; `roofline` is that kernel's algorithmic bytes over
its HIP-event launch time on the library's stream.  The line also carries:
the momentum (128^3) and scalar (32^3) assemblies with their own rooflines,
the edge algorithms and device geometry of the 256^3 box, the pressure
solves of that box (GMRES(50) with SA-AMG, Jacobi and Chebyshev
preconditioning, FP64 and FP32-inner / FP64-refined), the learned initial
guess on a 256^3 pressure-projection series, the conforming mixed
Hex8/Tet4/Wed6/Pyr5 mesh, the HBM copy rate measured by a streaming kernel, and the oracle
timed on this host's cores on the same 256^3 continuity assembly
(`cpu_baseline`) and the 128^3 momentum assembly.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nalu-wind_amd"))

HBM_PEAK_GBS = 8000.0
LEARN_RATE = 2e-3  # Adam step of the learned guess's training
METRIC = "Medges/s assembled + %HBM peak; pressure solve ms to 1e-8 rel. residual"


class Hip:
    """hipEvent timing on a given stream through the HIP runtime the
    library already loaded (same process-wide runtime)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.lib.hipEventSynchronize.argtypes = [C.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p,
                                                 C.c_void_p]
        self.lib.hipEventDestroy.argtypes = [C.c_void_p]
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t,
                                            C.c_void_p]
        self.lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t,
                                            C.c_int, C.c_void_p]

    def copy_async(self, dst, src, nbytes, stream):
        assert self.lib.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3,  # D2D
                                       C.c_void_p(stream)) == 0

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p

    def free(self, p):
        assert self.lib.hipFree(p) == 0

    def memset_async(self, p, value, nbytes, stream):
        assert self.lib.hipMemsetAsync(p, value, C.c_size_t(nbytes),
                                       C.c_void_p(stream)) == 0

    def event(self):
        e = C.c_void_p()
        assert self.lib.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.lib.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed(self, a, b):
        assert self.lib.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.lib.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0


# ------------------------------------------------------------ byte models
# Algorithmic bytes per launch: every node field the kernel uses read once
# (FP64), the edge inputs read once (area vector, mdot / Peclet factor when
# used, two int32 node ids), the matrix values and the rhs written once.
# The CSR structure and the row schedule are not counted.

def momentum_bytes(W, nnzb):
    # node fields read once: coords 3, velocity 3, dudx 9, viscosity 1,
    # wall mask 1 (17 f64); edge: area vector 3 f64, mdot, peclet factor,
    # two int32 node ids; block values written once (9 f64 per block), rhs
    # (3 f64 per node)
    return 17 * 8 * W.n + (5 * 8 + 8) * W.ne + 72 * nnzb + 24 * W.n


def continuity_bytes(W, nnz):
    # coords 3, velocity 3, dpdx 3, density, pressure, Udiag (12 f64);
    # edge: area vector + node ids; values + rhs written once
    return 12 * 8 * W.n + (24 + 8) * W.ne + 8 * nnz + 8 * W.n


def scalar_bytes(W, nnz):
    # coords 3, velocity 3, q, dqdx 3, density, diffusion coeff (12 f64);
    # edge: area vector, mdot, node ids
    return 12 * 8 * W.n + (32 + 8) * W.ne + 8 * nnz + 8 * W.n


def time_launches(hip, ctx, fn, steps, warmup):
    """Times `steps` calls of fn with events around each call (kernel
    durations) and around the whole loop; returns (avg_ms_per_call,
    loop_ms)."""
    s = ctx.stream()
    for _ in range(warmup):
        fn()
    hip.sync()
    ev = [hip.event() for _ in range(2 * steps + 2)]
    hip.record(ev[0], s)
    for i in range(steps):
        hip.record(ev[1 + 2 * i], s)
        fn()
        hip.record(ev[2 + 2 * i], s)
    hip.record(ev[-1], s)
    total = hip.elapsed(ev[0], ev[-1])
    per = [hip.elapsed(ev[1 + 2 * i], ev[2 + 2 * i]) for i in range(steps)]
    return sum(per) / steps, total


def stream_peak(hip, stream, nbytes=1 << 30, reps=10):
    """HBM copy rate of this GPU (bytes read + written per second, GB/s) by
    the streaming-copy kernel of libnw_stream.so over 1 GiB buffers, plain
    and non-temporal; plus the runtime's device-to-device copy for
    comparison."""
    lib = C.CDLL(os.path.join(ROOT, "nalu-wind_amd", "lib", "libnw_stream.so"))
    lib.nw_stream_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int,
                                   C.c_int, C.c_void_p, C.POINTER(C.c_double)]
    a, b = hip.malloc(nbytes), hip.malloc(nbytes)
    res = {}
    try:
        hip.memset_async(a, 1, nbytes, stream)
        for name, nt in (("kernel_copy", 0), ("kernel_copy_nontemporal", 1)):
            ms = C.c_double()
            assert lib.nw_stream_copy(b, a, nbytes, reps, nt, C.c_void_p(stream),
                                      C.byref(ms)) == 0
            res[name + "_GBps"] = round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)
        hip.copy_async(b, a, nbytes, stream)
        ev = [hip.event() for _ in range(2)]
        hip.record(ev[0], stream)
        for _ in range(reps):
            hip.copy_async(b, a, nbytes, stream)
        hip.record(ev[1], stream)
        ms = hip.elapsed(ev[0], ev[1]) / reps
        res["hipMemcpy_d2d_GBps"] = round(2 * nbytes / (ms * 1e-3) / 1e9, 1)
    finally:
        hip.sync()
        hip.free(a)
        hip.free(b)
    res["peak_GBps"] = max(res["kernel_copy_GBps"], res["kernel_copy_nontemporal_GBps"])
    res["note"] = ("streaming copy of 1 GiB (16-byte vector loads, 4 in flight per "
                   "lane), best of plain and non-temporal; bytes read + written "
                   "per second")
    return res


def time_cold(hip, ctx, fn, steps, flush_bytes=1 << 29):
    """Like time_launches, but each call starts with cold caches: a 512 MiB
    device write on the same stream before it evicts L2 and the Infinity
    Cache.  Returns the average ms of the calls alone."""
    s = ctx.stream()
    buf = hip.malloc(flush_bytes)
    try:
        ev = [hip.event() for _ in range(2 * steps)]
        for i in range(steps):
            hip.memset_async(buf, i & 0xFF, flush_bytes, s)
            hip.record(ev[2 * i], s)
            fn()
            hip.record(ev[2 * i + 1], s)
        per = [hip.elapsed(ev[2 * i], ev[2 * i + 1]) for i in range(steps)]
    finally:
        hip.sync()
        hip.free(buf)
    return sum(per) / steps


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if one
    exists for this kernel/workload (line 'traffic_bytes_per_launch: X')."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def roofline(kernel, nbytes, kern_ms, peak, profile):
    achieved = nbytes / (kern_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(profile), "kernel": kernel,
            "kernel_avg_ms": round(kern_ms, 4), "algorithmic_bytes_per_launch": nbytes,
            "copy_peak_measured": peak, "frac_of_copy_peak": round(achieved / peak, 4)}


def host_cores():
    cores = len(os.sched_getaffinity(0))
    return min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))


def cpu_rate(run, edges, seconds):
    """Medges/s of run() on 1 thread and on all host cores, each repeated
    for about its share of `seconds` (at least once)."""
    from oracle import orc
    cores = host_cores()

    def rate(threads, secs):
        orc.set_threads(threads)
        reps, t0 = 0, time.perf_counter()
        while True:
            run()
            reps += 1
            el = time.perf_counter() - t0
            if el > secs:
                return reps * edges / el / 1e6, reps, el

    serial, sreps, sel = rate(1, 0.3 * seconds)
    par, reps, el = rate(cores, 0.7 * seconds)
    orc.set_threads(1)
    return par, serial, cores, reps, el, sreps, sel


def cpu_continuity_baseline(W, seconds):
    """The oracle (C restatement of ContinuityEdgeSolverAlg + the CSR
    sum_into, OpenMP over edges with atomic scatter adds, as the
    reference's Kokkos-OpenMP build) on the benchmark's own 256^3 box,
    repeated for about `seconds`, plus a 1-thread run."""
    from oracle import orc
    import naluwind_amd as nw
    f = W.fields
    A = orc.CsrSystem(W.n, W.edges, 1)
    o = orc.ContinuityOpts(W.opts.dt, W.opts.gamma1, W.opts.noc,
                           W.opts.interp_together, W.opts.incompressible)

    def run():
        orc.assemble_continuity(o, W.n, W.edges, W.coords, f[nw.F_VELOCITY],
                                f[nw.F_DPDX], f[nw.F_DENSITY], f[nw.F_PRESSURE],
                                f[nw.F_UDIAG], W.areav, csr=A)

    par, serial, cores, reps, el, sreps, sel = cpu_rate(run, W.ne, seconds)
    return {"value": round(par, 3), "unit": "Medges/s", "cores": cores,
            "kind": "port", "serial_value": round(serial, 3),
            "sample": "oracle continuity assembly (C restatement of "
                      "ContinuityEdgeSolverAlg, CSR sum_into with atomic adds, "
                      "OpenMP over edges on %d threads) on the same 256^3-cell box "
                      "(%d edges) x %d repetitions in %.1f s; serial_value = 1 "
                      "thread, %d repetitions in %.1f s" % (cores, W.ne, reps, el,
                                                            sreps, sel)}


def cpu_momentum_baseline(W, seconds):
    """The oracle Peclet factor + momentum assembly (the GPU step) on the
    benchmark's own 128^3 box, all host cores and 1 thread."""
    from oracle import orc
    import naluwind_amd as nw
    f = W.fields
    pec = orc.Peclet(W.peclet.form, W.peclet.hybrid, W.peclet.c1, W.peclet.c2)
    o = orc.MomentumOpts(W.opts.alpha, W.opts.alpha_upw, W.opts.upw, W.opts.relax,
                         W.opts.include_divU, W.opts.use_limiter)
    A = orc.CsrSystem(W.n, W.edges, 3)

    def run():
        pecfac = orc.momentum_edge_peclet(W.edges, W.coords, f[nw.F_VELOCITY],
                                          f[nw.F_DENSITY], f[nw.F_VISCOSITY], pec)
        orc.assemble_momentum(o, W.n, W.edges, W.coords, f[nw.F_VELOCITY],
                              f[nw.F_DUDX], f[nw.F_VISCOSITY], f[nw.F_DENSITY],
                              W.areav, f[nw.E_MASS_FLOW_RATE], pecfac,
                              f[nw.F_WALL_MASK], csr=A)

    par, serial, cores, reps, el, sreps, sel = cpu_rate(run, W.ne, seconds)
    return {"value": round(par, 3), "unit": "Medges/s", "cores": cores,
            "kind": "port", "serial_value": round(serial, 3),
            "sample": "oracle MomentumEdgePecletAlg + MomentumEdgeSolverAlg + "
                      "CSR sum_into (atomic adds, OpenMP over edges on %d threads) "
                      "on the same 128^3-cell box (%d edges) x %d repetitions in "
                      "%.1f s; serial_value = 1 thread" % (cores, W.ne, reps, el)}


# ------------------------------------------------------------ solves

def solve_record(st):
    return {"ms": round(st.ms, 2), "iters": st.iters, "refinements": st.refinements,
            "final_rel_residual": st.final_rel, "converged": bool(st.converged)}


def pressure_solves_c3(ctx, W, quick, log):
    """The config's pressure solves on the assembled 256^3 system (Dirichlet
    p = 0 on the top face), x0 = 0, tol 1e-8 on ||b - A x||/||b||: GMRES(50)
    right-preconditioned by SA-AMG, by Jacobi and by Chebyshev(8) on D^-1 A,
    each in FP64 and FP32-inner / FP64-refined.  The residual is recomputed
    in FP64 on the device after each solve (nw_linsys_spmv)."""
    import naluwind_amd as nw
    import workloads as wl
    res = {"workload": "the continuity system of the headline box with Dirichlet "
                       "p = 0 on the top face, x0 = 0, tol 1e-8 on ||b-Ax||/||b||, "
                       "GMRES(50), right preconditioning"}
    ctx.assemble_continuity(W.opts)
    wl.apply_pressure_bc(ctx, W)
    cases = [("amg", dict(precond=nw.PREC_AMG, reuse_preconditioner=1), 2000, 1e-4),
             ("jacobi", dict(precond=nw.PREC_JACOBI), 600 if quick else 2000, 1e-5),
             ("chebyshev8", dict(precond=nw.PREC_CHEBYSHEV, cheb_degree=8,
                                 cheb_ratio=30.0), 1500 if quick else 4000, 1e-5)]
    for name, pre, maxit, itol in cases:
        for mode_name, mode in (("fp64", nw.SOLVE_FP64), ("fp32_refined", nw.SOLVE_FP32_REFINE)):
            cfg = nw.solver_cfg(kspace=50, max_iters=maxit, tol=1e-8, mode=mode,
                                inner_tol=itol, x0=nw.X0_ZERO, **pre)
            st = ctx.solve(cfg, raise_on_noconv=False)
            if name == "amg" and mode_name == "fp64":
                res["amg_setup_ms"] = round(st.setup_ms, 1)
                res["amg_levels_rows_nnz"] = ctx.amg_levels()
                st = ctx.solve(cfg, raise_on_noconv=False)
            r = solve_record(st)
            r["max_iters"] = maxit
            if mode == nw.SOLVE_FP32_REFINE:
                r["inner_tol"] = itol
            res[name + "_" + mode_name] = r
            log("C3 solve %s %s: %s" % (name, mode_name, r))
        a, b = res[name + "_fp64"], res[name + "_fp32_refined"]
        if a["converged"] and b["converged"]:
            res[name + "_speedup_fp32_refined"] = round(a["ms"] / b["ms"], 3)
    return res


def learned_guess_series(ctx, W, ntrain, neval, epochs, log, dt_flow=0.02):
    """The learned initial guess on the pressure-projection loop
    (workloads.PressureSeries): steps 1..ntrain solve for the pressure
    increment with the plain FP64 method (from the previous increment) and
    record (features, increment) pairs; the network is trained on them;
    steps ntrain+1.. compare, from the same state, the FP64 method from
    x0 = 0 (the reference's choice), the FP32-inner / FP64-refined method
    from x0 = 0, from the previous-step increment, from the untrained
    network (its prior: quadratic extrapolation in time) and from the
    trained network."""
    import numpy as np
    import naluwind_amd as nw
    import workloads as wl
    res = {}
    t0 = time.perf_counter()
    S = wl.PressureSeries(ctx, W=W, dt_flow=dt_flow)
    res["workload"] = (
        "pressure-projection loop on the headline box: per step the random-Fourier "
        "velocity at t = %g k as the momentum predictor, continuity assembly with "
        "the current p and dp/dx, Dirichlet dp = 0 on z = 1, solve for dp (tol 1e-8, "
        "GMRES(50) + SA-AMG), p += dp, MdotEdgeAlg, nodal velocity projection"
        % dt_flow)
    res["dt_flow"] = dt_flow
    res["nodes"] = W.n
    base = dict(kspace=50, max_iters=2000, tol=1e-8, precond=nw.PREC_AMG,
                reuse_preconditioner=1)
    fp64 = lambda x0: nw.solver_cfg(mode=nw.SOLVE_FP64, x0=x0, **base)
    fp32 = lambda x0: nw.solver_cfg(mode=nw.SOLVE_FP32_REFINE, inner_tol=1e-4,
                                    x0=x0, **base)
    S.step(0)
    st = ctx.solve(fp64(nw.X0_ZERO))
    S.update()
    res["amg_setup_ms"] = round(st.setup_ms, 1)
    log("projection series: setup %.1fs" % (time.perf_counter() - t0))
    net = nw.Mpnn(ctx, 64, 3, seed=7)
    t0 = time.perf_counter()
    gen_ms = []
    for k in range(1, ntrain + 1):
        S.step(k)
        st = ctx.solve(fp64(nw.X0_SOLUTION))
        gen_ms.append(st.ms)
        if k > 2:  # the guess history is complete from step 3 on
            net.record()
        S.update()
        if k % 50 == 0:
            log("pair generation: step %d, %.1fs" % (k, time.perf_counter() - t0))
    t_gen = time.perf_counter() - t0
    hold = max(3, net.samples // 10)
    t0 = time.perf_counter()
    # the head starts at the quadratic extrapolation (output weights zero)
    # and training learns how much of the third and fourth history
    # differences to add, per node; the most recent pairs validate, and the
    # parameters that predict them best (the untrained ones included) are
    # kept
    loss = net.train(nw.mpnn_train_cfg(epochs=epochs, batch=1, lr=LEARN_RATE,
                                       seed=3, holdout=hold))
    t_train = time.perf_counter() - t0
    prior = nw.Mpnn(ctx, 64, 3, seed=7)  # untrained: quadratic extrapolation
    kept_prior = bool(np.array_equal(net.params, prior.params))
    # a fixed rule with the same inputs for comparison: the output head set
    # to (1, 1) by hand, no training (quartic extrapolation in time)
    quartic = nw.Mpnn(ctx, 64, 3, seed=7)
    pq = quartic.params
    pq[-2:] = 1.0
    quartic.params = pq
    b_out = [float(v) for v in net.params[-2:]]
    res["learned_guess"] = {
        "network": "3 message-passing layers, 64 hidden; prediction in bfloat16 "
                   "weights and activations on the matrix cores with FP32 "
                   "accumulation, training in FP32",
        "pairs": net.samples, "validation_pairs": hold, "epochs": epochs,
        "adam_lr": LEARN_RATE, "best_validation_loss": loss,
        "trained_b_out": b_out,
        "validation_kept_untrained_parameters": kept_prior,
        "pair_generation_s": round(t_gen, 1), "pair_solve_ms_avg": round(
            sum(gen_ms) / len(gen_ms), 2), "training_s": round(t_train, 1)}
    log("training: %d pairs in %.1fs, %d epochs in %.1fs, loss %.4g, kept prior %s" %
        (net.samples, t_gen, epochs, t_train, loss, kept_prior))
    runs = {"fp64_x0_zero": [], "fp64_prev_step": [], "fp32_refined_x0_zero": [],
            "fp32_refined_prev_step": [], "fp32_refined_untrained_net": [],
            "fp32_refined_quartic_rule": [], "fp32_refined_learned": []}
    for k in range(ntrain + 1, ntrain + 1 + neval):
        x_prev = ctx.solution()
        S.step(k)
        # learned first: its inputs are x0 = dp(k-1) and xp = dp(k-2), the
        # guess of the previous step's last solve
        sol = {}
        nets = {"fp32_refined_untrained_net": prior,
                "fp32_refined_quartic_rule": quartic}
        for name, cfg in [("fp32_refined_learned", fp32(nw.X0_LEARNED)),
                          ("fp32_refined_untrained_net", fp32(nw.X0_LEARNED)),
                          ("fp32_refined_quartic_rule", fp32(nw.X0_LEARNED)),
                          ("fp32_refined_prev_step", fp32(nw.X0_SOLUTION)),
                          ("fp32_refined_x0_zero", fp32(nw.X0_ZERO)),
                          ("fp64_prev_step", fp64(nw.X0_SOLUTION)),
                          ("fp64_x0_zero", fp64(nw.X0_ZERO))]:
            ctx.set_solution(x_prev)
            ctx.set_learned_guess(nets.get(name, net))
            st = ctx.solve(cfg)
            sol[name] = ctx.solution()
            runs[name].append(st)
        # the state advances with the reference's own solve (the last one)
        S.update()
        ref = sol["fp64_x0_zero"]
        for name in sol:
            runs[name][-1] = (runs[name][-1], float(np.linalg.norm(sol[name] - ref) /
                                                    np.linalg.norm(ref)))
    for name, lst in runs.items():
        sts = [s for s, _ in lst]
        res[name] = {
            "ms": round(sum(s.ms for s in sts) / len(sts), 2),
            "iters": round(sum(s.iters for s in sts) / len(sts), 2),
            "refinements": round(sum(s.refinements for s in sts) / len(sts), 2),
            "final_rel_residual_max": max(s.final_rel for s in sts),
            "x0_rel_residual": sum(s.x0_rel for s in sts) / len(sts),
            "rel_error_vs_fp64": max(e for _, e in lst),
            "converged": all(bool(s.converged) for s in sts)}
        if name in ("fp32_refined_learned", "fp32_refined_untrained_net",
                    "fp32_refined_quartic_rule"):
            res[name]["learned_guess_used"] = sum(s.guess == nw.X0_LEARNED for s in sts)
            res[name]["guess_ms"] = round(sum(s.guess_ms for s in sts) / len(sts), 2)
        log("solve %s: %s" % (name, res[name]))
    res["eval_steps"] = neval
    z = res["fp64_x0_zero"]["ms"]
    res["speedup_vs_fp64_x0_zero"] = {
        n: round(z / res[n]["ms"], 3) for n in runs if n != "fp64_x0_zero"}
    tl, tp, tu = (res[n] for n in ("fp32_refined_learned", "fp32_refined_prev_step",
                                   "fp32_refined_untrained_net"))
    res["trained_vs_prior"] = {
        "iters": [tl["iters"], tu["iters"]], "ms": [tl["ms"], tu["ms"]],
        "x0_rel_residual": [tl["x0_rel_residual"], tu["x0_rel_residual"]],
        # beats the prior only with a different network that needs fewer
        # iterations, or as many from a starting residual lower by > 1 %
        # (equal iterations with a few % in ms is timing noise)
        "trained_beats_prior": bool(
            not res["learned_guess"]["validation_kept_untrained_parameters"]
            and (tl["iters"] < tu["iters"] or
                 (tl["iters"] == tu["iters"] and
                  tl["x0_rel_residual"] < 0.99 * tu["x0_rel_residual"])))}
    res["trained_vs_prev_step"] = {
        "iters": [tl["iters"], tp["iters"]], "ms": [tl["ms"], tp["ms"]],
        "x0_rel_residual": [tl["x0_rel_residual"], tp["x0_rel_residual"]]}
    tq = res["fp32_refined_quartic_rule"]
    res["trained_vs_quartic_rule"] = {
        "iters": [tl["iters"], tq["iters"]], "ms": [tl["ms"], tq["ms"]],
        "x0_rel_residual": [tl["x0_rel_residual"], tq["x0_rel_residual"]]}
    res["note"] = ("fp32_refined_untrained_net starts from the network before "
                   "training, whose output head (zero) gives the quadratic "
                   "extrapolation 3 x(n-1) - 3 x(n-2) + x(n-3) of the increments; "
                   "fp32_refined_learned from the trained network, which adds "
                   "learned per-node multiples of the third and fourth history "
                   "differences; fp32_refined_quartic_rule from the untrained "
                   "network with its head set by hand to (1, 1), the fixed "
                   "quartic extrapolation; pairs entries: [trained, other]")
    net.close()
    prior.close()
    quartic.close()
    return res


def device_geometry(hip, ctx, W, types, steps, warmup):
    """GeometryInteriorAlg on the device for a workload that kept its
    elements: the host setup of the incidence lists (once per mesh), the
    device time of one nw_compute_geometry call, and the largest deviation from the host-computed area vectors
    and dual volumes.  Overwrites the context's area vectors and dual
    volumes with the device values."""
    import numpy as np
    import naluwind_amd as nw
    t0 = time.perf_counter()
    ctx.mesh_elements(types, W.conn)
    setup_s = time.perf_counter() - t0
    kms, _ = time_launches(hip, ctx, ctx.compute_geometry, steps, warmup)
    a = ctx.field_download(nw.E_AREA_VECTOR)
    nel = int(W.conn.shape[0])
    res = {"elements": nel, "incidence_setup_s": round(setup_s, 2),
           "ms": round(kms, 3), "Melems_per_s": round(nel / (kms * 1e-3) / 1e6, 1),
           "max_rel_diff_area_vs_host": float(np.abs(a - W.areav).max() /
                                              np.abs(W.areav).max())}
    if W.dualvol is not None:
        v = ctx.field_download(nw.F_DUAL_VOLUME)
        res["max_rel_diff_dual_volume_vs_host"] = float(
            np.abs(v - W.dualvol).max() / np.abs(W.dualvol).max())
    return res


def mixed_mesh(hip, n, steps, warmup, peak, log):
    """Config of the unstructured mixed mesh: continuity assembly and the
    AMG-preconditioned pressure solve (FP64 and FP32-inner/FP64-refined,
    x0 = 0) on an n^3 box cut into Hex8/Tet4/Wed6 cells (60/30/10 %), with
    transition cells of Pyr5 and Tet4 where faces would not match, jittered,
    randomly permuted and RCM-renumbered."""
    import naluwind_amd as nw
    import workloads as wl
    t0 = time.perf_counter()
    W = wl.continuity_mixed(n, keep_elements=True)
    log("mixed mesh %d^3: %d nodes %d edges (%.1fs)" % (n, W.n, W.ne,
                                                        time.perf_counter() - t0))
    res = {"workload": "continuity on a conforming %d^3 box cut into "
                       "Hex8/Tet4/Wed6 (60/30/10 %% of the cells); a Hex8 or Wed6 "
                       "cell whose whole face meets a cut one becomes Pyr5 + Tet4 "
                       "about a centre node; +-10 %% jitter, nodes randomly "
                       "permuted then RCM-renumbered" % n,
           "nodes": W.n, "edges": W.ne,
           "transition_cells": int(W.n - (n + 1) ** 3),
           "elements_hex_tet_wedge_pyramid": [int(v) for v in W.elements]}
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(1)
        _, nnz = ctx.sizes()
        kms, _ = time_launches(hip, ctx, lambda: ctx.assemble_continuity(W.opts),
                               steps, warmup)
        cb = continuity_bytes(W, nnz)
        res["nnz"] = nnz
        res["assembly"] = {"Medges_per_s": round(W.ne / (kms * 1e-3) / 1e6, 1),
                           "roofline": roofline("k_assemble<ContinuityPhysT<false>>", cb, kms,
                                                peak, "continuity_mixed_160_assembly.txt")}
        log("mixed assembly: %s" % res["assembly"])
        ctx.assemble_continuity(W.opts)
        wl.apply_pressure_bc(ctx, W)
        for name, mode in (("fp64", nw.SOLVE_FP64), ("fp32_refined", nw.SOLVE_FP32_REFINE)):
            cfg = nw.solver_cfg(kspace=50, max_iters=2000, tol=1e-8,
                                precond=nw.PREC_AMG, mode=mode, inner_tol=1e-5,
                                reuse_preconditioner=1)
            st = ctx.solve(cfg, raise_on_noconv=False)
            if name == "fp64":
                res["amg_setup_ms"] = round(st.setup_ms, 1)
                st = ctx.solve(cfg, raise_on_noconv=False)
            res["solve_" + name] = solve_record(st)
            log("mixed solve %s: %s" % (name, res["solve_" + name]))
        res["amg_levels_rows_nnz"] = ctx.amg_levels()
        res["geometry"] = device_geometry(hip, ctx, W, W.types, min(steps, 5), 1)
        log("mixed geometry: %s" % res["geometry"])
        # the edge algorithms on the mixed mesh (dual volumes: the device
        # geometry's)
        ea = {}
        for name, fn in (("mdot", lambda: ctx.compute_mdot(1.0, 1.0)),
                         ("nodal_grad_pressure",
                          lambda: ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX)),
                         ("nodal_grad_velocity",
                          lambda: ctx.nodal_grad(nw.F_VELOCITY, nw.F_DUDX))):
            kms, _ = time_launches(hip, ctx, fn, steps, warmup)
            ea[name] = {"kernel_avg_ms": round(kms, 4),
                        "Medges_per_s": round(W.ne / (kms * 1e-3) / 1e6, 1)}
        res["edge_algorithms"] = ea
        log("mixed edge algorithms: %s" % ea)
    return res


def continuity_part(hip, a, n3, quick, log, out, peak):
    """The headline: continuity assembly on the 256^3 box, its edge
    algorithms and device geometry, the config's solves, the learned-guess
    series and the CPU baseline on the same mesh."""
    import naluwind_amd as nw
    import workloads as wl
    t0 = time.perf_counter()
    W = wl.continuity_rf(n3, keep_elements=True)
    log("continuity workload %d^3: %d nodes %d edges (%.1fs)" %
        (n3, W.n, W.ne, time.perf_counter() - t0))
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(1)
        _, nnz = ctx.sizes()
        s = ctx.stream()
        # the step: ContinuityEdgeSolverAlg over all edges, rows overwritten
        step = lambda: ctx.assemble_continuity(W.opts)
        kern_ms, loop_ms = time_launches(hip, ctx, step, a.steps, a.warmup)
        wall0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        hip.sync()
        wall = time.perf_counter() - wall0
        ms_step = loop_ms / a.steps
        cb = continuity_bytes(W, nnz)
        out["value"] = round(W.ne / (ms_step * 1e-3) / 1e6, 2)
        out["ms_per_step"] = round(ms_step, 4)
        out["config"] = {
            "workload": "continuity edge assembly (ContinuityEdgeSolverAlg), "
                        "%d^3-cell Hex8 box graded toward z=0 at %.5f per cell "
                        "(the z grading of the 128^3 box: first/last spacing ratio "
                        "%.0f), divergence-free random-Fourier velocity (8 modes), "
                        "rho = 1, p = 0, dp/dx = 0, Udiag = 1/dt" %
                        (n3, wl.grading_128(n3), wl.grading_128(n3) ** (n3 - 1)),
            "nodes": W.n, "edges": W.ne, "nnz": nnz, "ndof": 1,
            "step": "one ContinuityEdgeSolverAlg pass over all edges, inputs "
                    "resident in HBM (rows overwritten: fused zeroSystem)",
            "wall_s_untimed_repeat": round(wall, 4)}
        out["roofline"] = roofline("k_assemble<ContinuityPhysT<false>>", cb, kern_ms, peak,
                                   "continuity_rf_256_assembly.txt")
        log("continuity: %.3f ms/step, kernel %.3f ms, %.0f GB/s" %
            (ms_step, kern_ms, cb / (kern_ms * 1e-3) / 1e9))
        # with the inputs handed over as host arrays (the boundary's field
        # uploads) the same step crosses PCIe first
        t0 = time.perf_counter()
        for fid in (nw.F_COORDINATES, nw.F_VELOCITY, nw.F_DPDX, nw.F_DENSITY,
                    nw.F_PRESSURE, nw.F_UDIAG):
            ctx.field_upload(fid, W.coords if fid == nw.F_COORDINATES else W.fields[fid])
        ctx.field_upload(nw.E_AREA_VECTOR, W.areav)
        step()
        ctx.synchronize()
        up = time.perf_counter() - t0
        out["config"]["pcie_inclusive"] = {
            "ms": round(up * 1e3, 2), "Medges_per_s": round(W.ne / up / 1e6, 1),
            "note": "node and edge inputs uploaded from host arrays, then one "
                    "assembly (not the headline value)"}
        # the edge algorithms after the solve on the same mesh
        edge_algs = {}
        for name, fn, nb in (
                ("mdot", lambda: ctx.compute_mdot(1.0, 1.0),
                 12 * 8 * W.n + (24 + 8 + 8) * W.ne),
                ("nodal_grad_pressure", lambda: ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX),
                 (8 + 8 + 24) * W.n + 24 * W.ne),
                ("nodal_grad_velocity", lambda: ctx.nodal_grad(nw.F_VELOCITY, nw.F_DUDX),
                 (24 + 8 + 72) * W.n + 24 * W.ne)):
            kms, _ = time_launches(hip, ctx, fn, a.steps, a.warmup)
            edge_algs[name] = {"kernel_avg_ms": round(kms, 4),
                               "Medges_per_s": round(W.ne / (kms * 1e-3) / 1e6, 1),
                               "GBps_algorithmic": round(nb / (kms * 1e-3) / 1e9, 1),
                               "frac_hbm_peak": round(nb / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            log("%s: %s" % (name, edge_algs[name]))
        out["edge_algorithms_256"] = edge_algs
        # GeometryInteriorAlg on the device, same mesh
        out["geometry_256"] = device_geometry(hip, ctx, W, None, min(a.steps, 5), 1)
        log("geometry: %s" % out["geometry_256"])
        ctx.field_upload(nw.E_AREA_VECTOR, W.areav)
        ctx.field_upload(nw.F_DUAL_VOLUME, W.dualvol)
        W.drop_elements()
        if not a.no_solve:
            out["pressure_solve_256"] = pressure_solves_c3(ctx, W, quick, log)
            out["pressure_solve_256"]["nodes"] = W.n
            out["learned_guess_256"] = learned_guess_series(
                ctx, W, a.train_steps, a.eval_steps, a.epochs, log)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_continuity_baseline(W, 6.0 if quick else 24.0)
        log("cpu baseline: %s" % out["cpu_baseline"])
    else:
        out["cpu_baseline"] = None


def momentum_part(hip, a, n2, quick, log, out, peak):
    import naluwind_amd as nw
    import workloads as wl
    t0 = time.perf_counter()
    W = wl.momentum_abl(n2)
    log("momentum workload %d^3: %d nodes %d edges (%.1fs)" % (n2, W.n, W.ne,
                                                               time.perf_counter() - t0))
    res = {"workload": "momentum edge assembly, %d^3-cell Hex8 box stretched "
                       "1.05 per cell toward z=0, +-10%% jitter, log-law u + 5%% "
                       "noise, nu_t~U(0,1e-2), alpha=0.5, hybrid 1" % n2,
           "nodes": W.n, "edges": W.ne}
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(3)
        _, nnz = ctx.sizes()
        nnzb = nnz // 9
        res["block_nnz"] = nnzb
        # the same step in one pass (the Peclet factor computed in the
        # assembly kernel)
        k1, t1 = time_launches(
            hip, ctx, lambda: ctx.assemble_momentum_with_peclet(W.opts, W.peclet),
            a.steps, a.warmup)
        # the step: MomentumEdgePecletAlg, then MomentumEdgeSolverAlg
        for _ in range(a.warmup):
            ctx.momentum_edge_peclet(W.peclet)
            ctx.assemble_momentum(W.opts)
        hip.sync()
        s = ctx.stream()
        ev = [hip.event() for _ in range(2 * a.steps + 2)]
        hip.record(ev[0], s)
        for i in range(a.steps):
            ctx.momentum_edge_peclet(W.peclet)
            hip.record(ev[1 + 2 * i], s)
            ctx.assemble_momentum(W.opts)
            hip.record(ev[2 + 2 * i], s)
        hip.record(ev[-1], s)
        hip.sync()
        total_ms = hip.elapsed(ev[0], ev[-1])
        kern_ms = sum(hip.elapsed(ev[1 + 2 * i], ev[2 + 2 * i])
                      for i in range(a.steps)) / a.steps
        ms_step = total_ms / a.steps
        res["step"] = ("MomentumEdgePecletAlg + MomentumEdgeSolverAlg (rows "
                       "overwritten: fused zeroSystem)")
        res["Medges_per_s"] = round(W.ne / (ms_step * 1e-3) / 1e6, 2)
        res["ms_per_step"] = round(ms_step, 4)
        res["one_pass_ms_per_step"] = round(t1 / a.steps, 4)
        res["one_pass_note"] = ("nw_assemble_momentum_with_peclet: the same step "
                                "with the Peclet factor computed in the assembly "
                                "kernel")
        res["roofline"] = roofline("k_assemble_persistent<MomentumPhysT<false>>",
                                   momentum_bytes(W, nnzb), kern_ms, peak,
                                   "momentum_abl_128_assembly.txt")
        log("momentum: %.3f ms/step, assembly kernel %.3f ms" % (ms_step, kern_ms))
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_momentum_baseline(W, 5.0 if quick else 12.0)
        log("cpu baseline momentum: %s" % res["cpu_baseline"])
    out["momentum_assembly"] = res


def scalar_part(hip, a, n1, out, peak):
    import naluwind_amd as nw
    import workloads as wl
    W = wl.scalar_tg(n1)
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(1)
        _, nnz = ctx.sizes()
        kms, tot = time_launches(hip, ctx, lambda: ctx.assemble_scalar(W.opts),
                                 a.steps, a.warmup)
        cold = time_cold(hip, ctx, lambda: ctx.assemble_scalar(W.opts), a.steps)
        sb = scalar_bytes(W, nnz)
        out["scalar_assembly"] = {
            "workload": "scalar adv-diff %d^3 uniform, Taylor-Green u, mu=1e-3" % n1,
            "nodes": W.n, "edges": W.ne, "Medges_per_s": round(W.ne / (kms * 1e-3) / 1e6, 1),
            "kernel_avg_ms": round(kms, 5),
            "GBps_algorithmic": round(sb / (kms * 1e-3) / 1e9, 1),
            "cold_kernel_avg_ms": round(cold, 5),
            "cold_Medges_per_s": round(W.ne / (cold * 1e-3) / 1e6, 1),
            "cold_GBps_algorithmic": round(sb / (cold * 1e-3) / 1e9, 1),
            "note": "~10 MB working set: warm = cache-resident (no HBM claim); "
                    "cold = each launch after a 512 MiB device write that "
                    "evicts L2 and the Infinity Cache"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--quick", action="store_true", help="small sizes (testing)")
    ap.add_argument("--train-steps", type=int, default=202,
                    help="pressure-projection steps whose pairs train the network "
                         "(pairs = steps - 2)")
    ap.add_argument("--eval-steps", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--only", choices=["continuity", "momentum", "mixed"],
                    help="run one part (profiling); the JSON line is then partial")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the multi-GPU path (node "
              "partition + shared-row export over RCCL) is a design, not a "
              "build; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import naluwind_amd as nw
    hip = Hip()
    log = lambda *x: print(*x, file=sys.stderr, flush=True)
    n1 = 16 if a.quick else 32
    n2 = 32 if a.quick else 128
    n3 = 48 if a.quick else 256
    if a.quick:
        a.train_steps = min(a.train_steps, 14)

    out = {"metric": METRIC, "unit": "Medges/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    with nw.Context(0) as ctx:
        out["hbm_copy_peak"] = stream_peak(hip, ctx.stream())
    peak = out["hbm_copy_peak"]["peak_GBps"]
    log("copy peak: %s" % out["hbm_copy_peak"])

    if a.only in (None, "momentum"):
        momentum_part(hip, a, n2, a.quick, log, out, peak)
        scalar_part(hip, a, n1, out, peak)
    if a.only in (None, "continuity"):
        continuity_part(hip, a, n3, a.quick, log, out, peak)
    if not a.no_solve and a.only in (None, "mixed"):
        out["mixed_mesh"] = mixed_mesh(hip, 20 if a.quick else 160, a.steps,
                                       a.warmup, peak, log)
    out.setdefault("cpu_baseline", None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
