#!/usr/bin/env python3
"""Benchmark of the edge-assembly + pressure-solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
                    [--solve-size N] [--no-cpu]

Headline (one JSON line on stdout): momentum edge assembly (FP64, 3 dof,
3x3 block CSR) on a 128^3-cell box stretched 1.05 toward z = 0 with +-10 %
node jitter, log-law velocity + 5 % noise, nu_t ~ U(0, 1e-2), alpha = 0.5.
A step is MomentumEdgePecletAlg + MomentumEdgeSolverAlg over all edges
(inputs resident in HBM), value = edges / step time in Medges/s.  The line
also carries the continuity (256^3) and scalar (32^3) assembly rates, the
pressure-Poisson solves, the roofline of the assembly kernel measured with
HIP events on the library's stream, and the CPU oracle timed on this host.
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


def copy_peak(hip, stream, nbytes=1 << 30, reps=5):
    """Device-to-device copy bandwidth (bytes read + written per second,
    GB/s) of a 1 GiB buffer on this GPU: the achievable HBM rate that the
    nominal 8 TB/s overstates."""
    a, b = hip.malloc(nbytes), hip.malloc(nbytes)
    try:
        hip.memset_async(a, 1, nbytes, stream)
        hip.copy_async(b, a, nbytes, stream)
        ev = [hip.event() for _ in range(2)]
        hip.record(ev[0], stream)
        for _ in range(reps):
            hip.copy_async(b, a, nbytes, stream)
        hip.record(ev[1], stream)
        ms = hip.elapsed(ev[0], ev[1]) / reps
    finally:
        hip.sync()
        hip.free(a)
        hip.free(b)
    return 2 * nbytes / (ms * 1e-3) / 1e9


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


def cpu_momentum_baseline(seconds=12.0, n=48):
    """The oracle (C restatement of MomentumEdgePecletAlg +
    MomentumEdgeSolverAlg + sum_into, the same step as the GPU's) on a
    bounded sample: a 48^3 box of the same configuration, repeated for
    about `seconds`, with OpenMP over edges and atomic scatter adds on the
    host cores (the reference's OpenMP execution), plus a 1-thread run."""
    sys.path.insert(0, ROOT)
    from oracle import orc
    import workloads as wl
    import naluwind_amd as nw
    W = wl.momentum_abl(n)
    f = W.fields
    pec = orc.Peclet(W.peclet.form, W.peclet.hybrid, W.peclet.c1, W.peclet.c2)
    o = orc.MomentumOpts(0.5, 1.0, 1.0, 1.0, 0.0, 0)
    A = orc.CsrSystem(W.n, W.edges, 3)
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))

    def rate(threads, secs):
        orc.set_threads(threads)
        reps, t0 = 0, time.perf_counter()
        while True:
            pecfac = orc.momentum_edge_peclet(W.edges, W.coords, f[nw.F_VELOCITY],
                                              f[nw.F_DENSITY], f[nw.F_VISCOSITY],
                                              pec)
            orc.assemble_momentum(o, W.n, W.edges, W.coords, f[nw.F_VELOCITY],
                                  f[nw.F_DUDX], f[nw.F_VISCOSITY], f[nw.F_DENSITY],
                                  W.areav, f[nw.E_MASS_FLOW_RATE], pecfac,
                                  f[nw.F_WALL_MASK], csr=A)
            reps += 1
            el = time.perf_counter() - t0
            if el > secs:
                return reps * W.ne / el / 1e6, reps, el

    serial, _, _ = rate(1, 0.3 * seconds)
    par, reps, el = rate(cores, 0.7 * seconds)
    orc.set_threads(1)
    return {"value": round(par, 3), "unit": "Medges/s", "cores": cores,
            "kind": "port", "serial_value": round(serial, 3),
            "sample": "oracle Peclet factor + momentum assembly (C restatement, "
                      "CSR sum_into with atomic adds, OpenMP over edges on %d "
                      "threads), %d^3-cell box "
                      "of the same configuration (%d edges) x %d repetitions in "
                      "%.1f s; serial_value = the same on 1 thread" %
                      (cores, n, W.ne, reps, el)}


def pressure_solves(ns, ntrain, neval, epochs, log):
    """Pressure-Poisson solves on a time-evolving flow (PressureSeries): the
    plain FP64 method from x0 = 0 (the reference's choice), the
    FP32-inner / FP64-refined method, and the refined method started from
    the learned guess.  Steps 1..ntrain are plain FP64 solves from the
    previous-step solution whose (features, solution) pairs train the
    network; steps ntrain+1.. compare the methods from the same state."""
    import numpy as np
    import naluwind_amd as nw
    import workloads as wl
    res = {}
    with nw.Context(0) as ctx:
        t0 = time.perf_counter()
        S = wl.PressureSeries(ctx, ns, seed=1234, dt_flow=0.02)
        W = S.W
        res["workload"] = (
            "pressure Poisson of the continuity assembly, %d^3-cell box with the "
            "z grading of the 128^3 box (per-cell ratio %.5f), random-Fourier "
            "velocity advancing 0.02 time units per step, Dirichlet p=0 on z=1, "
            "tol 1e-8 on ||b-Ax||/||b||, GMRES(50) right-preconditioned by "
            "SA-AMG (Chebyshev(2) smoothing, drop tol 0.02)" % (ns, wl.grading_128(ns)))
        res["nodes"] = W.n
        log("pressure series %d^3: %d nodes (%.1fs)" % (ns, W.n, time.perf_counter() - t0))
        base = dict(kspace=50, max_iters=2000, tol=1e-8, precond=nw.PREC_AMG,
                    reuse_preconditioner=1)
        fp64 = lambda x0: nw.solver_cfg(mode=nw.SOLVE_FP64, x0=x0, **base)
        # inner tolerance 1e-4: measured best of 1e-3 .. 1e-6 (fewest iterations)
        fp32 = lambda x0: nw.solver_cfg(mode=nw.SOLVE_FP32_REFINE, inner_tol=1e-4,
                                        x0=x0, **base)
        S.step(0)
        st = ctx.solve(fp64(nw.X0_ZERO))
        res["amg_setup_ms"] = round(st.setup_ms, 1)
        res["amg_levels"] = st.precond_levels
        net = nw.Mpnn(ctx, 64, 3, seed=7)
        t0 = time.perf_counter()
        for k in range(1, ntrain + 1):
            S.step(k)
            ctx.solve(fp64(nw.X0_SOLUTION))
            if k > 2:  # the guess history is complete from step 3 on
                net.record()
        t_gen = time.perf_counter() - t0
        t0 = time.perf_counter()
        # small steps: the head starts at the quadratic extrapolation and
        # training adjusts it (large Adam steps on the head destroy the prior)
        # the most recent pairs validate: the parameters that predict them
        # best (the untrained ones included) are kept
        loss = net.train(nw.mpnn_train_cfg(epochs=epochs, batch=1, lr=1e-4, seed=3,
                                           holdout=3))
        t_train = time.perf_counter() - t0
        res["learned_guess"] = {
            "network": "3 message-passing layers, 64 hidden; prediction in bfloat16 weights and activations on the matrix cores with FP32 accumulation, training in FP32",
            "pairs": net.samples, "validation_pairs": 3, "epochs": epochs,
            "adam_lr": 1e-4,
            "best_validation_loss": loss, "pair_generation_s": round(t_gen, 1),
            "training_s": round(t_train, 1)}
        log("training: %d pairs in %.1fs, %d epochs in %.1fs, loss %.4g" %
            (ntrain, t_gen, epochs, t_train, loss))
        prior = nw.Mpnn(ctx, 64, 3, seed=7)  # untrained: quadratic extrapolation
        runs = {"fp64_x0_zero": [], "fp64_prev_step": [], "fp32_refined_x0_zero": [],
                "fp32_refined_prev_step": [], "fp32_refined_untrained_net": [],
                "fp32_refined_learned": []}
        for k in range(ntrain + 1, ntrain + 1 + neval):
            x_prev = ctx.solution()
            S.step(k)
            # learned first: its inputs are x0 = p(k-1) and xp = p(k-2), the
            # guess of the previous step's last solve
            sol = {}
            for name, cfg in [("fp32_refined_learned", fp32(nw.X0_LEARNED)),
                              ("fp32_refined_untrained_net", fp32(nw.X0_LEARNED)),
                              ("fp32_refined_prev_step", fp32(nw.X0_SOLUTION)),
                              ("fp32_refined_x0_zero", fp32(nw.X0_ZERO)),
                              ("fp64_x0_zero", fp64(nw.X0_ZERO)),
                              ("fp64_prev_step", fp64(nw.X0_SOLUTION))]:
                ctx.set_solution(x_prev)
                ctx.set_learned_guess(prior if name == "fp32_refined_untrained_net" else net)
                st = ctx.solve(cfg)
                sol[name] = ctx.solution()
                runs[name].append(st)
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
                "x0_rel_residual": round(sum(s.x0_rel for s in sts) / len(sts), 6),
                "rel_error_vs_fp64": max(e for _, e in lst),
                "converged": all(bool(s.converged) for s in sts)}
            if name in ("fp32_refined_learned", "fp32_refined_untrained_net"):
                res[name]["learned_guess_used"] = sum(s.guess == nw.X0_LEARNED for s in sts)
                res[name]["guess_ms"] = round(sum(s.guess_ms for s in sts) / len(sts), 2)
            log("solve %s: %s" % (name, res[name]))
        res["eval_steps"] = neval
        z = res["fp64_x0_zero"]["ms"]
        res["speedup_vs_fp64_x0_zero"] = {
            n: round(z / res[n]["ms"], 3) for n in runs if n != "fp64_x0_zero"}
        res["note"] = ("fp32_refined_untrained_net starts from the network before "
                       "training, whose output head is the quadratic extrapolation "
                       "3 x(n-1) - 3 x(n-2) + x(n-3); fp32_refined_learned from the "
                       "trained network")
        net.close()
        prior.close()
    return res


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


def polynomial_solves(n, log):
    """The pressure solve with the polynomial preconditioner the config
    names (Chebyshev on D^-1 A, degree 8), FP64 against FP32-inner /
    FP64-refined, x0 = 0.  At 256^3 it needs several thousand iterations, so
    it runs on the 128^3 box of the same geometry."""
    import naluwind_amd as nw
    import workloads as wl
    res = {}
    W = wl.continuity_rf(n)
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(1)
        ctx.assemble_continuity(W.opts)
        wl.apply_pressure_bc(ctx, W)
        res["workload"] = ("pressure Poisson of the continuity assembly, %d^3-cell "
                           "box (z grading of the 128^3 box), x0 = 0, tol 1e-8, "
                           "GMRES(50) right-preconditioned by Chebyshev(8) on "
                           "D^-1 A (eigenvalue ratio 30)" % n)
        res["nodes"] = W.n
        for name, mode in (("fp64", nw.SOLVE_FP64), ("fp32_refined", nw.SOLVE_FP32_REFINE)):
            cfg = nw.solver_cfg(kspace=50, max_iters=5000, tol=1e-8,
                                precond=nw.PREC_CHEBYSHEV, cheb_degree=8,
                                mode=mode, inner_tol=1e-5)
            ctx.solve(cfg, raise_on_noconv=False)  # warm-up
            st = ctx.solve(cfg, raise_on_noconv=False)
            res[name] = {"ms": round(st.ms, 2), "iters": st.iters,
                         "refinements": st.refinements,
                         "final_rel_residual": st.final_rel,
                         "converged": bool(st.converged)}
            log("polynomial %s: %s" % (name, res[name]))
        res["speedup_fp32_refined"] = round(res["fp64"]["ms"] / res["fp32_refined"]["ms"], 3)
    return res


def device_geometry(hip, ctx, W, types, steps, warmup):
    """GeometryInteriorAlg on the device for a workload that kept its
    elements: the host setup of the incidence lists (once per mesh), the
    device time of one nw_compute_geometry call (scratch allocation
    included), and the largest deviation from the host-computed area vectors
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


def mixed_mesh(hip, n, steps, warmup, log):
    """Config of the unstructured mixed mesh: continuity assembly and the
    AMG-preconditioned pressure solve (FP64 and FP32-inner/FP64-refined,
    x0 = 0) on an n^3 box cut into Hex8/Tet4/Wed6 cells (60/30/10 %),
    jittered, randomly permuted and RCM-renumbered."""
    import naluwind_amd as nw
    import workloads as wl
    t0 = time.perf_counter()
    W = wl.continuity_mixed(n, keep_elements=True)
    log("mixed mesh %d^3: %d nodes %d edges (%.1fs)" % (n, W.n, W.ne,
                                                        time.perf_counter() - t0))
    res = {"workload": "continuity on a %d^3 box cut into Hex8/Tet4/Wed6 "
                       "(60/30/10 %% of the cells), +-10 %% jitter, nodes randomly "
                       "permuted then RCM-renumbered" % n,
           "nodes": W.n, "edges": W.ne,
           "elements_hex_tet_wedge": [int(v) for v in W.elements]}
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(1)
        _, nnz = ctx.sizes()
        kms, _ = time_launches(hip, ctx, lambda: ctx.assemble_continuity(W.opts),
                               steps, warmup)
        cb = continuity_bytes(W, nnz)
        res["nnz"] = nnz
        res["assembly"] = {"Medges_per_s": round(W.ne / (kms * 1e-3) / 1e6, 1),
                           "kernel_avg_ms": round(kms, 4),
                           "GBps_algorithmic": round(cb / (kms * 1e-3) / 1e9, 1),
                           "frac_hbm_peak": round(cb / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
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
            res["solve_" + name] = {"ms": round(st.ms, 2), "iters": st.iters,
                                    "refinements": st.refinements,
                                    "final_rel_residual": st.final_rel,
                                    "converged": bool(st.converged)}
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--quick", action="store_true", help="small sizes (testing)")
    ap.add_argument("--solve-size", type=int, default=256)
    ap.add_argument("--train-steps", type=int, default=24)
    ap.add_argument("--eval-steps", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--only", choices=["assembly", "solve", "polynomial", "mixed"],
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
    import workloads as wl
    hip = Hip()
    log = lambda *x: print(*x, file=sys.stderr, flush=True)
    n2 = 32 if a.quick else 128
    n3 = 48 if a.quick else 256
    n1 = 16 if a.quick else 32

    out = {"metric": METRIC, "unit": "Medges/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic"}

    if a.only in (None, "assembly"):
        # ---------------------------------------------------------- momentum
        t0 = time.perf_counter()
        W = wl.momentum_abl(n2)
        log("momentum workload %d^3: %d nodes %d edges (%.1fs)" % (n2, W.n, W.ne,
                                                                   time.perf_counter() - t0))
        with nw.Context(0) as ctx:
            W.upload(ctx)
            ctx.build_graph(3)
            _, nnz = ctx.sizes()
            nnzb = nnz // 9

            # the same step in one pass (nw_assemble_momentum_with_peclet:
            # the Peclet factor computed in the assembly kernel), for
            # comparison; timed first, it also brings the GPU to its working
            # clocks before the timed steps
            k1, t1 = time_launches(
                hip, ctx, lambda: ctx.assemble_momentum_with_peclet(W.opts, W.peclet),
                a.steps, a.warmup)

            # the step: MomentumEdgePecletAlg, then MomentumEdgeSolverAlg
            def step():
                ctx.momentum_edge_peclet(W.peclet)
                ctx.assemble_momentum(W.opts)

            for _ in range(a.warmup):
                step()
            hip.sync()
            s = ctx.stream()
            ev = [hip.event() for _ in range(2 * a.steps + 2)]
            wall0 = time.perf_counter()
            hip.record(ev[0], s)
            for i in range(a.steps):
                ctx.momentum_edge_peclet(W.peclet)
                hip.record(ev[1 + 2 * i], s)
                ctx.assemble_momentum(W.opts)
                hip.record(ev[2 + 2 * i], s)
            hip.record(ev[-1], s)
            hip.sync()
            wall = time.perf_counter() - wall0
            total_ms = hip.elapsed(ev[0], ev[-1])
            kern_ms = sum(hip.elapsed(ev[1 + 2 * i], ev[2 + 2 * i])
                          for i in range(a.steps)) / a.steps
            ms_step = total_ms / a.steps
            mbytes = momentum_bytes(W, nnzb)
            achieved = mbytes / (kern_ms * 1e-3) / 1e9
            out["value"] = round(W.ne / (ms_step * 1e-3) / 1e6, 2)
            out["ms_per_step"] = round(ms_step, 4)
            out["config"] = {
                "workload": "momentum edge assembly, %d^3-cell Hex8 box stretched "
                            "1.05 toward z=0, +-10%% jitter, log-law u + 5%% noise, "
                            "nu_t~U(0,1e-2), alpha=0.5, hybrid 1" % n2,
                "nodes": W.n, "edges": W.ne, "block_nnz": nnzb, "ndof": 3,
                "step": "MomentumEdgePecletAlg + MomentumEdgeSolverAlg (rows "
                        "overwritten: fused zeroSystem)",
                "one_pass_ms_per_step": round(t1 / a.steps, 4),
                "one_pass_note": "nw_assemble_momentum_with_peclet: the same "
                                 "step with the Peclet factor computed in the "
                                 "assembly kernel (fewer launches, a small "
                                 "register spill)",
                "wall_s_timed_region": round(wall, 4)}
            traffic = read_traffic("momentum_abl_128_assembly.txt")
            cpk = copy_peak(hip, s)
            out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1),
                               "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(achieved / HBM_PEAK_GBS, 4),
                               "traffic": traffic,
                               "kernel": "k_assemble_persistent<MomentumPhysT<false>>",
                               "kernel_avg_ms": round(kern_ms, 4),
                               "algorithmic_bytes_per_launch": mbytes,
                               "copy_peak_measured": round(cpk, 1),
                               "frac_of_copy_peak": round(achieved / cpk, 4)}
            log("momentum: %.3f ms/step, assembly kernel %.3f ms, %.0f GB/s" %
                (ms_step, kern_ms, achieved))
        del W

        # ---------------------------------------------------------- scalar
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
        del W

        # ---------------------------------------------------------- continuity
        t0 = time.perf_counter()
        W = wl.continuity_rf(n3, keep_elements=True)
        log("continuity workload %d^3: %d nodes %d edges (%.1fs)" % (n3, W.n, W.ne,
                                                                     time.perf_counter() - t0))
        with nw.Context(0) as ctx:
            W.upload(ctx)
            ctx.build_graph(1)
            _, nnz = ctx.sizes()
            kms, tot = time_launches(hip, ctx, lambda: ctx.assemble_continuity(W.opts),
                                     a.steps, a.warmup)
            cb = continuity_bytes(W, nnz)
            out["continuity_assembly"] = {
                "workload": "continuity %d^3-cell box stretched 1.05, random-Fourier u (8 modes)" % n3,
                "nodes": W.n, "edges": W.ne, "nnz": nnz,
                "Medges_per_s": round(W.ne / (kms * 1e-3) / 1e6, 1),
                "kernel_avg_ms": round(kms, 4),
                "GBps_algorithmic": round(cb / (kms * 1e-3) / 1e9, 1),
                "frac_hbm_peak": round(cb / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "algorithmic_bytes_per_launch": cb}
            log("continuity: %.3f ms, %.0f GB/s" % (kms, cb / (kms * 1e-3) / 1e9))
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
        del W

    # ---------------------------------------------------------- solve
    if not a.no_solve and a.only in (None, "solve"):
        ns = 24 if a.quick else a.solve_size
        out["pressure_solve"] = pressure_solves(ns, a.train_steps, a.eval_steps,
                                                a.epochs, log)
    if not a.no_solve and a.only in (None, "polynomial"):
        out["pressure_solve_polynomial"] = polynomial_solves(
            24 if a.quick else 128, log)

    # ---------------------------------------------------------- mixed mesh
    if not a.no_solve and a.only in (None, "mixed"):
        out["mixed_mesh"] = mixed_mesh(hip, 20 if a.quick else 160, a.steps,
                                       a.warmup, log)

    if not a.no_cpu and a.only is None:
        out["cpu_baseline"] = cpu_momentum_baseline(3.0 if a.quick else 12.0,
                                                    16 if a.quick else 48)
        log("cpu baseline: %s" % out["cpu_baseline"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
