"""Profiling driver: runs one assembly kernel a few times on a benchmark
workload, plus a device-to-device copy of known size that calibrates the
HBM byte counters.  Meant to run under rocprofv3, e.g.

    rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc -- python3 \
        nalu-wind_amd/profile_kernels.py --kernel continuity --n 256

    python3 nalu-wind_amd/profile_kernels.py --summarize OUT_FETCH OUT_WRITE
    python3 nalu-wind_amd/profile_kernels.py --stats OUT/x_kernel_stats.csv
"""
import argparse
import ctypes as C
import glob
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def run(kernel, n, reps, copy_mib):
    import naluwind_amd as nw
    import workloads as wl
    hip = C.CDLL("libamdhip64.so.7")
    if kernel == "momentum":
        W = wl.momentum_abl(n)
        ndof = 3
        fn = lambda ctx: ctx.assemble_momentum(W.opts)
    elif kernel == "continuity":
        W = wl.continuity_rf(n)
        ndof = 1
        fn = lambda ctx: ctx.assemble_continuity(W.opts)
    elif kernel == "mixed":
        # continuity on the conforming mixed Hex8/Tet4/Wed6/Pyr5 box (RCM-numbered)
        W = wl.continuity_mixed(n)
        ndof = 1
        fn = lambda ctx: ctx.assemble_continuity(W.opts)
    elif kernel == "edge_algs":
        # MdotEdgeAlg and NodalGradEdgeAlg (dp/dx, du/dx) on the continuity box
        W = wl.continuity_rf(n)
        ndof = 1

        def fn(ctx):
            ctx.compute_mdot(1.0, 1.0)
            ctx.nodal_grad(nw.F_PRESSURE, nw.F_DPDX)
            ctx.nodal_grad(nw.F_VELOCITY, nw.F_DUDX)
    else:
        W = wl.scalar_tg(n)
        ndof = 1
        fn = lambda ctx: ctx.assemble_scalar(W.opts)
    with nw.Context(0) as ctx:
        W.upload(ctx)
        ctx.build_graph(ndof)
        if kernel == "momentum":
            ctx.momentum_edge_peclet(W.peclet)
        for _ in range(reps):
            fn(ctx)
        ctx.synchronize()
    # calibration copy: copy_mib MiB device to device
    nbytes = copy_mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # D2D
    assert hip.hipDeviceSynchronize() == 0
    print("ran %s n=%d reps=%d copy=%d MiB nodes=%d edges=%d" %
          (kernel, n, reps, copy_mib, W.n, W.ne), flush=True)


def short_name(name):
    """Kernel name without argument list and namespaces, templates kept."""
    n = name.replace("void ", "")
    for ns in ("(anonymous namespace)::", "nw::"):
        n = n.replace(ns, "")
    return n.split("(")[0]


def stats(files, top):
    """Per-kernel call count and average duration from rocprofv3
    kernel_stats.csv files."""
    import csv
    for f in files:
        rows = list(csv.DictReader(open(f)))
        rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
        for r in rows[:top]:
            print("  %-70s calls=%5d avg_us=%10.2f min_us=%10.2f max_us=%10.2f" % (
                short_name(r["Name"])[:70], int(r["Calls"]),
                float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                float(r["MaxNs"]) / 1e3))


def summarize(dirs):
    """Average counter value (KiB) per dispatch for each kernel name."""
    import csv
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"),
                           recursive=True):
            acc = {}
            for r in csv.DictReader(open(f)):
                key = (short_name(r["Kernel_Name"])[:60], r["Counter_Name"])
                acc.setdefault(key, []).append(float(r["Counter_Value"]))
            for (k, cn), v in sorted(acc.items()):
                print("%-62s %-12s dispatches=%4d avg=%.1f" % (k, cn, len(v),
                                                              sum(v) / len(v)))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", default="continuity")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--summarize", nargs="*")
    ap.add_argument("--stats", nargs="*")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args()
    if a.stats:
        stats(a.stats, a.top)
    elif a.summarize:
        summarize(a.summarize)
    else:
        run(a.kernel, a.n, a.reps, a.copy_mib)
