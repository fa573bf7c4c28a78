"""Benchmark: time-to-1e-9 residual and HBM GB/s of the 256^3
variable-coefficient (density ratio 1000) Poisson problem, FP64 vs mixed.

One step = one poisson() solve from a = 0 to max|res| <= 1e-9 on the GPU
(the alpha/lambda restriction of poisson() included), with b and alpha
already resident in HBM.  Both precisions are timed; `value` is the mixed
solver's time-to-tolerance.  A separate pass of each precision records HIP
events around the fine-level relax sweeps for the roofline line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--depth 8]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes moved per fine-level relax launch, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE on this workload (KiB per dispatch) and scaled by
# the counters' calibration on a copy kernel of known size (fetch x2.0,
# write x1.0); the derivation is in profiles/hbm_traffic_poisson256_mixed.md.
# Keyed by (precision, depth, red-black kernel): "stream3t" is the default
# two-sweep launch of the FP32 256^3 level (per launch, two sweeps),
# "stream3" the single-pass sweep, "inplace" the two launches of the colour
# kernel (per sweep).
MEASURED_SWEEP_TRAFFIC = {
    ("mixed", 8, "stream3t"): (2 * 183990.1 + 65536.0) * 1024,
    ("mixed", 8, "stream3"): (2 * 181527.1 + 65536.0) * 1024,
    ("mixed", 8, "inplace"): 2 * (333867.0 + 65536.0) * 1024,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--smoother", default="rbgs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-serial", action="store_true",
                    help="skip the one-thread oracle solve (about a minute at 256^3)")
    ap.add_argument("--no-gs", action="store_true", help="skip the GPU solve with the reference's default GS")
    ap.add_argument("--no-view", action="store_true", help="skip poisson() through the reference-storage view")
    ap.add_argument("--no-adapt", action="store_true", help="tree workloads: skip adapt_wavelet, dump and restore")
    ap.add_argument("--precisions", default="fp64,mixed")
    ap.add_argument("--config", type=int, default=2, choices=(2, 3, 5),
                    help="2: the 256^3 multigrid problem (default); 3: 2D adaptive quadtree; "
                         "5: 3D adaptive octree")
    return ap.parse_args()


def tree_bench(args):
    """The adaptive-tree workloads: one step = one poisson() solve to the
    tolerance on the tree (alpha/lambda restriction included), after the
    tree and its slot tables are built (once per mesh, outside the timed
    region)."""
    import numpy as np
    import torch
    import basilisk_amd as bmg
    from basilisk_amd import configs
    make = configs.config3_tree if args.config == 3 else configs.config5_tree
    dim = 2 if args.config == 3 else 3
    res = {}
    for prec in args.precisions.split(","):
        t = bmg.Tree(dim, 1.0, (0.0, 0.0, 0.0), precision=prec, smoother=args.smoother)
        t0 = time.perf_counter()
        info = make(t)
        t.set_niter(200, 1)
        zeros = torch.zeros(info["leaves"], dtype=torch.float64, device="cuda:0")
        t.set_a(zeros)
        st = t.poisson(tolerance=info["tolerance"])  # first call also compiles the slot tables
        build = time.perf_counter() - t0
        for _ in range(args.warmup):
            t.set_a(zeros)
            t.poisson(tolerance=info["tolerance"])
        t.synchronize()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            t.set_a(zeros)
            st = t.poisson(tolerance=info["tolerance"])
        t.synchronize()
        dt = (time.perf_counter() - t1) / args.steps
        d, nl, nc = t.info()
        # one more solve with HIP events around every relax launch
        t.profile(True)
        t.set_a(zeros)
        t.poisson(tolerance=info["tolerance"])
        relax_ms, relax_n, relax_cells = t.profile_get()
        t.profile(False)
        res[prec] = {"seconds": dt, "stats": st, "a": t.get_a(), "build_s": build, "depth": d, "cells": nc,
                     "leaves": nl, "relax": (relax_ms, relax_n, relax_cells)}
        print("# %s: %.2f ms, %d cycles, resa %.3g, %d leaves, setup %.1f s" % (prec, dt * 1e3, st["i"], st["resa"],
              nl, build), file=sys.stderr, flush=True)
        if prec == args.precisions.split(",")[-1] and not args.no_adapt:
            # one adapt_wavelet (grid/tree-common.h:165-303) on the solution:
            # estimate, host passes, recompilation and field remap
            a = res[prec]["a"]
            t.set_s(a)
            tol = 1e-3 * float(np.abs(a).max())
            t.synchronize()
            t2 = time.perf_counter()
            nf, ncs = t.adapt_wavelet(tol, d)
            t.synchronize()
            adapt = {"ms": round((time.perf_counter() - t2) * 1e3, 1), "max": tol, "maxlevel": d, "nf": nf,
                     "nc": ncs, "leaves_after": t.nleaves()}
            print("# adapt_wavelet: %.1f ms, %d refined, %d coarsened, %d leaves" % (adapt["ms"], nf, ncs,
                  adapt["leaves_after"]), file=sys.stderr, flush=True)
            # dump() and restore() (output.h:1043-1390) of the adapted tree
            # with a, b and s, through a file in a scratch directory
            import tempfile
            with tempfile.TemporaryDirectory() as td:
                path = os.path.join(td, "tree.dump")
                names = {"a": "a", "b": "b", "s": "s"}
                t3 = time.perf_counter()
                t.dump(path, names=names)
                t4 = time.perf_counter()
                nbytes = os.path.getsize(path)
                t.restore(path, names=names)
                t.synchronize()
                t5 = time.perf_counter()
            adapt["dump_ms"] = round((t4 - t3) * 1e3, 1)
            adapt["restore_ms"] = round((t5 - t4) * 1e3, 1)
            adapt["dump_mb"] = round(nbytes / 1e6, 1)
            print("# dump: %.1f ms, restore %.1f ms, %.1f MB" % (adapt["dump_ms"], adapt["restore_ms"],
                  adapt["dump_mb"]), file=sys.stderr, flush=True)
        t.close()
    if args.no_adapt:
        adapt = None
    main_prec = "mixed" if "mixed" in res else args.precisions.split(",")[0]
    r = res[main_prec]
    name = {3: "2D quadtree levels 5-11 around 40 circles", 5: "3D octree levels 5-10 around 200 spheres"}
    line = {
        "metric": "time-to-1e-9 residual, %s (density ratio 1000)" % name[args.config],
        "value": round(r["seconds"] * 1e3, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["seconds"] * 1e3, 3), "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None,
        "dtype": "f32/f64" if main_prec == "mixed" else "f64",
        "data": "synthetic (seeded circles/spheres, density ratio 1000)",
        "config": {"workload": "%s, %d leaves, %d allocated cells, tolerance 1e-9, %s smoother"
                               % (name[args.config], r["leaves"], r["cells"], args.smoother),
                   "leaves": r["leaves"], "parallelism": "single GPU"},
        "cycles": r["stats"]["i"],
        "vcycles_per_s": round(r["stats"]["i"] / r["seconds"], 1),
        "leaves_per_s": r["leaves"] * r["stats"]["i"] / r["seconds"],
        "resa": r["stats"]["resa"],
        "setup_s": round(r["build_s"], 2),
        "adapt_wavelet": adapt,
    }
    # the dominant kernel, k_t_relax: per relaxed cell the algorithmic bytes
    # of the sweep (da, res and the dim low faces read, da written: (dim+3)w)
    # plus the 2*dim 4-byte neighbour indices it gathers through
    w = 4 if main_prec == "mixed" else 8
    per_cell = (dim + 3) * w + 4 * 2 * dim
    rms, rn, rcells = r["relax"]
    if rms > 0 and rn > 0:
        achieved = per_cell * rcells / (rms * 1e-3) / 1e9
        line["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                            "kernel": "k_t_relax (%s), %d launches of more than 1024 cells per solve"
                                      % (main_prec, rn),
                            "launch_us": round(rms / rn * 1e3, 2), "cells_per_launch": round(rcells / rn),
                            "bytes_per_cell": per_cell}
    else:
        line["roofline"] = None
    if line["roofline"] is not None:
        # a device-to-device copy of 256 MB on the same box (its counters
        # also calibrate FETCH_SIZE / WRITE_SIZE in a profiled run)
        copy_gbs = copy_bandwidth(torch, 1 << 28)
        line["roofline"]["copy_gbs"] = round(copy_gbs, 1)
        line["roofline"]["frac_of_copy"] = round(line["roofline"]["achieved"] / copy_gbs, 4)
    if not args.no_cpu_baseline:
        line["cpu_baseline"], oref = tree_cpu_baseline(args, r)
        if oref is not None and len(oref) == len(r["a"]):
            scale = float(np.abs(oref - oref.mean()).max())
            for prec, rr in res.items():
                a = rr["a"] - rr["a"].mean()
                line["%s_vs_oracle_rel_maxnorm" % prec] = float(np.abs(a - (oref - oref.mean())).max() / scale)
    if "fp64" in res and "mixed" in res:
        f, m = res["fp64"], res["mixed"]
        line["fp64_ms"] = round(f["seconds"] * 1e3, 3)
        line["fp64_cycles"] = f["stats"]["i"]
        line["mixed_speedup_vs_fp64"] = round(f["seconds"] / m["seconds"], 3)
        af, am = f["a"] - f["a"].mean(), m["a"] - m["a"].mean()
        line["mixed_vs_fp64_rel_maxnorm"] = float(np.abs(am - af).max() / np.abs(af).max())
    print(json.dumps(line), flush=True)


def tree_cpu_baseline(args, r):
    """The tree oracle (one thread, the reference's default in-place GS in
    foreach_cell's Morton order) on the same generator.  Config 3: the whole
    solve to 1e-9, timed, on the benchmark's own tree.  Config 5: the same
    generator at depth 8 (1.1M leaves; the depth-10 tree takes the oracle
    tens of minutes to build), solved to 1e-9 and scaled by leaf-cycles to
    the GPU run's leaves and cycles.  Tree construction is not timed."""
    from oracle.tree import OracleTree
    from basilisk_amd import configs
    dim = 2 if args.config == 3 else 3
    depth = None if args.config == 3 else 8
    o = OracleTree(dim, 1.0, (0.0, 0.0, 0.0))
    o.set_smoother("gs")
    if args.config == 3:
        info = configs.config3_tree(o)
    else:
        info = configs.config5_tree(o, depth=depth)
    o.set_niter(300, 1)
    t0 = time.perf_counter()
    st = o.poisson(tolerance=info["tolerance"])
    sec = time.perf_counter() - t0
    leaves = o.nleaves()
    base = {"unit": "ms", "cores": 1, "kind": "port", "cpu": cpu_model(), "cycles": st["i"], "resa": st["resa"],
            "leaves": leaves}
    if args.config == 3:
        base["value"] = round(sec * 1e3, 1)
        base["sample"] = "the whole poisson() solve to 1e-9 on the same quadtree (%d leaves), GS, one thread" % leaves
        return base, o.get_a()
    per = sec / (leaves * st["i"])
    base["value"] = round(per * r["leaves"] * r["stats"]["i"] * 1e3, 1)
    base["sample"] = ("poisson() to 1e-9 of the same generator at depth 8 (%d leaves, %d cycles, %.1f s, GS, one "
                      "thread), %.3g s per leaf-cycle scaled to the GPU run's %d leaves x %d cycles"
                      % (leaves, st["i"], sec, per, r["leaves"], r["stats"]["i"]))
    return base, None


def solve_once(g, zeros, tol):
    g.set_a(zeros)
    return g.poisson(tolerance=tol)


def run_precision(bmg, torch, prob, prec, args):
    import numpy as np
    dim, depth = prob["dim"], prob["depth"]
    g = bmg.Grid(dim, depth, precision=prec, smoother=args.smoother)
    dev = torch.device("cuda:0")
    b = torch.from_numpy(prob["b"]).to(dev)
    g.set_b(b)
    g.set_alpha([torch.from_numpy(a).to(dev) for a in prob["alpha"]])
    zeros = torch.zeros(b.shape, dtype=torch.float64, device=dev)
    tol = prob["tolerance"]
    for _ in range(args.warmup):
        solve_once(g, zeros, tol)
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        stats.append(solve_once(g, zeros, tol))
    g.synchronize()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # timing pass: HIP events on the library's stream around each kernel class
    g.profile(True)
    st = solve_once(g, zeros, tol)
    kinds = {}
    for k in ("relax_fine", "residual", "restriction", "prolongation", "correction", "relax_coarse",
              "relax_fine2"):
        ms, n = g.profile_get(k)
        kinds[k] = {"ms": ms, "count": n}
    g.profile(False)
    a = g.get_a()
    proj = None
    if prob.get("u") is not None:
        # project() of the face velocity whose divergence is b (dt = 1, the
        # same tolerance): divergence, poisson() and face update on the
        # device; the velocity is reset from a pristine copy between steps
        u0 = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in prob["u"]]
        u = [x.clone() for x in u0]
        g.set_tolerance(tol)
        ptimes = []
        for it in range(args.warmup + args.steps):
            for x, y in zip(u, u0):
                x.copy_(y)
            g.set_a(zeros)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            pst = g.project(u, dt=1.0)
            g.synchronize()
            if it >= args.warmup:
                ptimes.append(time.perf_counter() - t1)
        proj = {"seconds": sum(ptimes) / len(ptimes), "stats": pst}
        del u, u0
    visc = None
    if prob.get("inside") is not None and prec == "fp64":
        # viscosity() (viscosity.h:163-178) of a two-phase velocity field on
        # the same grid: the second mg_solve client, FP64, the grid's smoother
        from basilisk_amd import configs
        vp = configs.config2_viscous(prob)
        v0 = [torch.from_numpy(x).to(dev) for x in vp["u"]]
        mu = [torch.from_numpy(np.ascontiguousarray(m)).to(dev) for m in vp["mu"]]
        rho = torch.from_numpy(vp["rho"]).to(dev)
        del vp["u"], vp["mu"], vp["rho"]
        uv = [x.clone() for x in v0]
        vtimes = []
        for it in range(args.warmup + args.steps):
            for x, y in zip(uv, v0):
                x.copy_(y)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            vst = g.viscosity(uv, mu=mu, rho=rho, dt=vp["dt"])
            g.synchronize()
            if it >= args.warmup:
                vtimes.append(time.perf_counter() - t1)
        visc = {"seconds": sum(vtimes) / len(vtimes), "stats": vst}
        del uv, v0, mu, rho
    g.close()
    return {"seconds": dt, "stats": stats[-1], "profile": kinds, "profiled_stats": st, "a": a, "project": proj,
            "viscosity": visc}


def run_view(bmg, prob, args, nvar=8):
    """poisson() through the drop-in boundary: the problem laid out in the
    reference's multigrid storage (records of nvar doubles: a, b, res,
    alpha.x/y/z and two unrelated fields, two ghost layers) and solved by
    bmg_poisson_view in mixed precision with red-black relax -- the upload
    and download of the record array included in the time."""
    import numpy as np
    dim, depth = prob["dim"], prob["depth"]
    n = 1 << depth
    levels = bmg.view_records(dim, depth, nvar)
    fin = levels[depth].reshape((n + 4,) * dim + (nvar,))
    fin[(slice(2, n + 2),) * dim + (1,)] = prob["b"]
    for c in range(dim):
        shape = list(prob["b"].shape)
        shape[c] += 1
        fin[tuple(slice(2, 2 + e) for e in shape) + (3 + c,)] = prob["alpha"][c]
    bmg.view_configure("mixed", "rbgs")
    times = []
    try:
        for it in range(args.warmup + args.steps):
            fin[..., 0] = 0.0
            t0 = time.perf_counter()
            st = bmg.poisson_view(levels, dim, depth, nvar, 0, 1, (3, 4, 5)[:dim], 0, prob["tolerance"], res=2)
            if it >= args.warmup:
                times.append(time.perf_counter() - t0)
    finally:
        bmg.view_configure("fp64", "rbgs")
        bmg.library().bmg_view_release()
    rec_bytes = levels[depth].nbytes
    return {"ms": round(sum(times) / len(times) * 1e3, 3), "cycles": st["i"], "resa": st["resa"],
            "record_bytes": rec_bytes, "datasize": nvar * 8}


def copy_bandwidth(torch, nbytes, reps=20):
    """GB/s (bytes read + bytes written) of torch's device-to-device copy of
    an nbytes buffer, timed with HIP events on the current stream."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").uniform_()
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * (nbytes // 4) * 4 / (ms * 1e-3) / 1e9 if ms > 0 else 0.0


def run_adjoint(bmg, torch, prob, args):
    """One differentiable solve and its backward through basilisk_amd.autograd
    (mixed precision, the problem's tolerance): phi = poisson(b, alpha),
    J = sum(w phi) with seeded zero-mean weights w, then dJ/db and dJ/dalpha
    from the adjoint solve and the fused face-gradient kernel."""
    import numpy as np
    from basilisk_amd.autograd import PoissonSolver
    dev = torch.device("cuda:0")
    S = PoissonSolver(prob["dim"], prob["depth"], precision="mixed", smoother=args.smoother,
                      tolerance=prob["tolerance"])
    b = torch.from_numpy(prob["b"]).to(dev).requires_grad_(True)
    alpha = [torch.from_numpy(np.ascontiguousarray(a)).to(dev).requires_grad_(True) for a in prob["alpha"]]
    w = np.random.default_rng(3).standard_normal(prob["b"].shape)
    w = torch.from_numpy(w - w.mean()).to(dev)
    times = []
    for it in range(args.warmup + args.steps):
        b.grad = None
        for a in alpha:
            a.grad = None
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        phi = S(b, alpha)
        (w * phi).sum().backward()
        torch.cuda.synchronize()
        if it >= args.warmup:
            times.append(time.perf_counter() - t1)
    out = {"seconds": sum(times) / len(times), "forward": S.last_stats, "adjoint": S.last_adjoint_stats,
           "finite": bool(torch.isfinite(b.grad).all()) and all(bool(torch.isfinite(a.grad).all()) for a in alpha)}
    S.fwd.close()
    S.adj.close()
    return out


def cpu_model():
    """The host CPU's model name (the "model name" of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prob, args):
    """The oracle's build of the reference's default solver (in-place
    lexicographic GS, nrelax 4, AoS storage, schedule(static) loops) on the
    host: poisson() from a = 0 to the problem's tolerance, timed whole, on
    min(nproc, 16) OpenMP threads (the reference's OpenMP path, whose GS
    races between threads like the reference's) and, unless
    --no-cpu-serial, on one thread (the deterministic serial order: its
    cycle count and residual are what the GPU's GS mode reproduces).  The
    threaded solution is then carried on to 1e-11: the reference solution
    the GPU solutions are measured against."""
    import numpy as np
    from oracle.oracle import OracleGrid, set_threads
    threads = min(os.cpu_count() or 1, 16)
    tol = prob["tolerance"]
    set_threads(threads, prob["dim"])
    o = OracleGrid(prob["dim"], prob["depth"], omp=True)
    o.set_b(prob["b"])
    o.set_alpha(prob["alpha"])
    o.set_a(0 * prob["b"])
    t0 = time.perf_counter()
    st = o.poisson(tolerance=tol)
    dt = time.perf_counter() - t0
    base = {"value": round(dt * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "cycles": st["i"], "resa": st["resa"],
            "sample": "the whole solve: poisson() of the %d^3 problem from a = 0 to %g with the reference's "
                      "default relax (GS, nrelax 4) on %d OpenMP threads, restriction of alpha included"
                      % (1 << prob["depth"], tol, threads)}
    serial = None
    if not args.no_cpu_serial:
        s1 = OracleGrid(prob["dim"], prob["depth"], omp=False)
        s1.set_b(prob["b"])
        s1.set_alpha(prob["alpha"])
        s1.set_a(0 * prob["b"])
        t1 = time.perf_counter()
        ss = s1.poisson(tolerance=tol)
        base["one_thread"] = {"ms": round((time.perf_counter() - t1) * 1e3, 1), "cycles": ss["i"],
                              "resa": ss["resa"]}
        serial = {"stats": ss, "a": s1.get_a()}
        del s1
    t2 = time.perf_counter()
    st2 = o.poisson(tolerance=1e-11)
    ref = o.get_a()
    oracle = {"tolerance": 1e-11, "cycles": st["i"] + st2["i"], "resa": st2["resa"],
              "seconds": round(dt + time.perf_counter() - t2, 2), "a": ref - ref.mean()}
    return base, oracle, serial


def gs_gpu(bmg, torch, prob, args):
    """The GPU's GS mode (the reference's default relax in its serial order,
    FP64) on the same problem: one timed poisson() and its solution."""
    g = bmg.Grid(prob["dim"], prob["depth"], precision="fp64", smoother="gs")
    dev = torch.device("cuda:0")
    b = torch.from_numpy(prob["b"]).to(dev)
    g.set_b(b)
    g.set_alpha([torch.from_numpy(a).to(dev) for a in prob["alpha"]])
    zeros = torch.zeros(b.shape, dtype=torch.float64, device=dev)
    solve_once(g, zeros, prob["tolerance"])  # graphs of the cycle built here
    g.synchronize()
    t0 = time.perf_counter()
    st = solve_once(g, zeros, prob["tolerance"])
    g.synchronize()
    out = {"seconds": time.perf_counter() - t0, "stats": st, "a": g.get_a()}
    g.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.config != 2:
        tree_bench(args)
        return
    import numpy as np
    import torch
    import basilisk_amd as bmg
    from basilisk_amd import configs

    prob = configs.config2(depth=args.depth, seed=1)
    N = prob["b"].size
    res = {}
    for prec in args.precisions.split(","):
        res[prec] = run_precision(bmg, torch, prob, prec, args)
        print("# %s: %.2f ms, %d cycles, resa %.3g" % (prec, res[prec]["seconds"] * 1e3, res[prec]["stats"]["i"],
                                                      res[prec]["stats"]["resa"]), file=sys.stderr, flush=True)

    main_prec = "mixed" if "mixed" in res else args.precisions.split(",")[0]
    r = res[main_prec]
    cycles = r["stats"]["i"]
    w = 4 if main_prec == "mixed" else 8
    # fine-level relax sweep: da read, res read, 3 alpha read, da write = 6w
    # bytes per cell and sweep (the reference's relax, poisson.h:303-320).
    # In FP32 the fine level runs its sweeps two per launch (k_rb_stream3t),
    # the odd one alone: the dominant kernel is then the two-sweep one, and
    # its algorithmic bytes are those of the two sweeps it does
    prof2 = r["profile"].get("relax_fine2", {"ms": 0.0, "count": 0})
    two = prof2["count"] > 0
    prof = prof2 if two else r["profile"]["relax_fine"]
    sweeps = 2 if two else 1
    sweep_ms = prof["ms"] / max(prof["count"], 1)
    sweep_bytes = 6 * w * N * sweeps
    achieved = sweep_bytes / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else 0.0
    # algorithmic bytes of the whole solve (per V-cycle and cell, 3D, nrelax
    # 4: ~312 B FP64, ~192 B mixed) for the whole-solve GB/s
    per_cycle_bytes = {"fp64": 312.0, "mixed": 192.0}[main_prec] * N
    rb_kernel = "stream3t" if two else os.environ.get("BMG_RB_KERNEL", "stream3")
    traffic = MEASURED_SWEEP_TRAFFIC.get((main_prec, args.depth, rb_kernel))
    line = {
        "metric": "time-to-1e-9 residual, 256^3 var-coeff Poisson (mixed precision)",
        "value": round(r["seconds"] * 1e3, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["seconds"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32/f64" if main_prec == "mixed" else "f64",
        "data": "synthetic (seeded spheres, density ratio 1000; seeded N(0,1) face velocity)",
        "config": {"workload": "3D multigrid level %d (%d^3), alpha=1/rho face-averaged, symmetry walls, "
                               "tolerance 1e-9, %s smoother" % (args.depth, 1 << args.depth, args.smoother),
                   "cells": N, "parallelism": "single GPU"},
        "cycles": cycles,
        "vcycles_per_s": round(cycles / r["seconds"], 1),
        "cells_per_s": N * cycles / r["seconds"],
        "solve_gbs": round(per_cycle_bytes * cycles / r["seconds"] / 1e9, 1),
        "resa": r["stats"]["resa"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": ("k_rb_stream3t: two fine-level red-black sweeps per launch (%s)" if two else
                                "fine-level relax sweep (%s, both colours)") % main_prec,
                     "sweeps_per_launch": sweeps, "launch_ms": round(sweep_ms, 5),
                     "algorithmic_bytes": sweep_bytes},
    }
    p1 = r["profile"]["relax_fine"]
    if two and p1["count"]:
        # the odd sweep of a cycle, in the one-sweep kernel
        line["roofline"]["one_sweep_launch_ms"] = round(p1["ms"] / p1["count"], 5)
    copy_gbs = copy_bandwidth(torch, 6 * w * N // 2)
    # the same box's device-to-device copy of half the sweep's bytes (read
    # and written once): what plain streaming reaches here, and the sweep's
    # share of it
    line["roofline"]["copy_gbs"] = round(copy_gbs, 1)
    line["roofline"]["frac_of_copy"] = round(achieved / copy_gbs, 4) if copy_gbs > 0 else None
    if "fp64" in res and "mixed" in res:
        f, m = res["fp64"], res["mixed"]
        line["fp64_ms"] = round(f["seconds"] * 1e3, 3)
        line["fp64_cycles"] = f["stats"]["i"]
        line["mixed_speedup_vs_fp64"] = round(f["seconds"] / m["seconds"], 3)
        af, am = f["a"], m["a"]
        af = af - af.mean()
        am = am - am.mean()
        line["mixed_vs_fp64_rel_maxnorm"] = float(np.abs(am - af).max() / np.abs(af).max())
        pf = f["profile"]["relax_fine"]
        line["fp64_sweep_gbs"] = round(6 * 8 * N / (pf["ms"] / max(pf["count"], 1) * 1e-3) / 1e9, 1)
    line["kernel_ms"] = {k: round(v["ms"], 3) for k, v in r["profile"].items()}
    if r.get("project"):
        # project() (poisson.h:469-509) of the config's face velocity: the
        # solver's caller, divergence and face update included
        line["project_ms"] = round(r["project"]["seconds"] * 1e3, 3)
        line["project_cycles"] = r["project"]["stats"]["i"]
    rv = res.get("fp64", {}).get("viscosity")
    if rv:
        # viscosity() of a two-phase velocity (rho 1000/1, mu 1/0.01, dt 1e-3)
        # to the same tolerance: FP64, red-black relax_viscosity (the grid's
        # smoother)
        line["viscosity_ms"] = round(rv["seconds"] * 1e3, 3)
        line["viscosity_cycles"] = rv["stats"]["i"]
        line["viscosity_resa"] = rv["stats"]["resa"]
    if main_prec == "mixed" and prob.get("alpha") is not None:
        # forward solve + backward (adjoint solve, dJ/db, dJ/dalpha) through
        # torch.autograd, on the same problem
        adj = run_adjoint(bmg, torch, prob, args)
        line["adjoint_ms"] = round(adj["seconds"] * 1e3, 3)
        line["adjoint_cycles"] = {"forward": adj["forward"]["i"], "adjoint": adj["adjoint"]["i"],
                                  "adjoint_resa": adj["adjoint"]["resa"], "finite_gradients": adj["finite"]}
    if not args.no_view:
        v = run_view(bmg, prob, args)
        line["view"] = v
        print("# view (mixed): %.2f ms, %d cycles, %.2f GB of records each way" % (v["ms"], v["cycles"],
              v["record_bytes"] / 1e9), file=sys.stderr, flush=True)
    gs = None
    if not args.no_gs:
        gs = gs_gpu(bmg, torch, prob, args)
        line["gs_gpu"] = {"ms": round(gs["seconds"] * 1e3, 3), "cycles": gs["stats"]["i"],
                          "resa": gs["stats"]["resa"]}
        print("# gs (fp64): %.2f ms, %d cycles, resa %.3g" % (gs["seconds"] * 1e3, gs["stats"]["i"],
              gs["stats"]["resa"]), file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        line["cpu_baseline"], oracle, serial = cpu_baseline(prob, args)
        ref = oracle.pop("a")
        scale = float(np.abs(ref).max())
        for prec, rr in res.items():
            a = rr["a"] - rr["a"].mean()
            oracle["%s_rel_maxnorm" % prec] = float(np.abs(a.reshape(ref.shape) - ref).max() / scale)
        if gs is not None:
            a = gs["a"] - gs["a"].mean()
            oracle["gs_rel_maxnorm"] = float(np.abs(a.reshape(ref.shape) - ref).max() / scale)
        line["error_vs_oracle"] = oracle
        # V-cycles to the tolerance against the reference's default relax in
        # its serial order (the oracle on one thread): the GPU's GS mode
        # must give the same count and bit-identical iterates, the red-black
        # default within a cycle or two
        if serial is not None:
            sc = serial["stats"]["i"]
            cyc = {"serial_gs_oracle": sc}
            for prec, rr in res.items():
                cyc["rbgs_%s" % prec] = rr["stats"]["i"]
                cyc["rbgs_%s_minus_serial_gs" % prec] = rr["stats"]["i"] - sc
            if gs is not None:
                cyc["gs_gpu"] = gs["stats"]["i"]
                cyc["gs_gpu_resa_equals_oracle"] = gs["stats"]["resa"] == serial["stats"]["resa"]
                cyc["gs_gpu_a_equals_oracle"] = bool(np.array_equal(gs["a"].reshape(serial["a"].shape),
                                                                    serial["a"]))
            line["cycles_vs_reference_default"] = cyc
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
