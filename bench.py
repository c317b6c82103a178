"""Benchmark: cell-steps/s of the per-timestep hot path on the 3-D 256^3 vortex tank,
the PCG's achieved HBM bandwidth, and the FP32-refined solver against the
FP64 one at the same 1e-8 relative residual.

A step is one Solver::step (RK3 semi-Lagrangian advection of q, u, v, w,
inputs, divergence, Jacobi-PCG to 1e-8, projection) over the whole 256^3
grid, state resident in HBM.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--n N] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "cell-steps/s at 256^3 + PCG HBM GB/s; FP32-refined vs FP64 at 1e-8 resid"
HBM_PEAK_GBS = 8000.0

# Algorithmic vector words per owned cell per launch of each 3-D PCG kernel
# (8 bytes in FP64, 4 in FP32): dir_spmv reads r and the old direction and
# writes the new one (3).  The update alternates two variants: X_SKIP reads
# p, r and writes r (3); X_PAIR reads x, p, the previous p, r and writes x,
# r (6) -- 4.5 words per launch on average, 7.5 per iteration.  The
# reference-shaped split (SpMV / update / direction) moves 11 words per
# iteration.  "pcg.gbs_11_words" counts 11 words per iteration (the
# contract's convention), the roofline object counts each kernel's own words.
WORDS = {"dir_spmv": 3, "update": 4.5}
# rows of the kernel-trace / counter tables for each timed FP64 kernel
KERNEL64 = {"dir_spmv": "k_pcg_dir_spmv3<double, 1>",
            "update": "k_pcg_update3<double, 1> (X_SKIP and X_PAIR, mean)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 leg (profiling runs only; no JSON line)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-mg", action="store_true",
                    help="skip the multigrid-preconditioned FP64 leg")
    ap.add_argument("--mg-only", action="store_true",
                    help="only the multigrid leg (profiling runs; prints its object)")
    ap.add_argument("--mg-nu", type=int, default=None,
                    help="multigrid leg: sweeps before / after the coarse correction")
    ap.add_argument("--mg-alpha", type=float, default=None,
                    help="multigrid leg: coarse-correction scale")
    ap.add_argument("--mg-omega", type=float, default=None,
                    help="multigrid leg: Jacobi damping")
    ap.add_argument("--mg-coarse", type=int, default=None,
                    help="multigrid leg: sweeps on the coarsest level")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dye", action="store_true",
                    help="skip the 2-D 128^2 dye-tank leg (profiling runs)")
    ap.add_argument("--cpu-iters", type=int, default=400)
    ap.add_argument("--timing-sample", type=int, default=16,
                    help="time the PCG kernels with HIP events in one "
                         "iteration out of N inside the timed region (0: off)")
    ap.add_argument("--adjoint", action="store_true",
                    help="also time the reverse mode: the 2-D 512^2 dye tank (100 "
                         "steps, checkpoint every 10) and the 3-D 384^3 periodic "
                         "box (10 steps, checkpoint every 2) against their forwards")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the reverse-mode leg (profiling runs)")
    ap.add_argument("--summarize-counters", metavar="CSV",
                    help="print per-kernel dispatch count, median and sum of a "
                         "rocprofv3 --pmc counter CSV, then exit (no GPU use)")
    ap.add_argument("--summarize-trace", metavar="CSV",
                    help="print per-kernel, per-grid-size dispatch count and mean / total "
                         "duration of a rocprofv3 --kernel-trace CSV, then exit (no GPU use)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only time plain streaming kernels (torch's add: read 2, write "
                         "1; and copy) over 263^3 FP64 arrays, the padded 256^3 size, "
                         "and print their rates")
    ap.add_argument("--dye-only", action="store_true",
                    help="only the 2-D 128^2 dye tank (setup + 50 steps) on the GPU "
                         "(profiling runs; prints its object)")
    ap.add_argument("--advect-only", action="store_true",
                    help="only the advection kernel: 4 launches each on the 256^3 vortex "
                         "tank and the 384^3 periodic box after setup + one step, timed "
                         "with HIP events (profiling runs; prints the mean ms)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a plain 1 GiB device copy, the reference "
                         "kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def copy_probe():
    """One plain device-to-device copy of 2^27 doubles (1 GiB read, 1 GiB
    written), so that a counter run over this command also holds the
    counter values of a copy of known size."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def stream_probe():
    """Rates of plain streaming kernels over arrays of one padded 256^3 FP64
    vector (263^3 elements): torch's add (c = a + s b: read 2, write 1) and
    copy (read 1, write 1), each the mean of 50 launches that alternate
    between two operand sets so that one launch's inputs are not the
    previous launch's outputs.  The reference rate for the PCG kernels'
    streams."""
    import json
    import torch
    n = 263 ** 3
    a, b, c, d, e = (torch.full((n,), 1.0 + i, dtype=torch.float64, device="cuda")
                     for i in range(5))
    out = {}
    for name, words, fn in (
            ("add_2r1w", 3, lambda i: torch.add(a, b, alpha=0.5, out=c) if i & 1
             else torch.add(d, e, alpha=0.5, out=b)),
            ("copy_1r1w", 2, lambda i: c.copy_(a) if i & 1 else b.copy_(d))):
        fn(0)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for i in range(50):
            fn(i)
        t1.record()
        t1.synchronize()
        us = t0.elapsed_time(t1) * 1e3 / 50
        out[name] = {"us": round(us, 1), "tb_s": round(words * n * 8 / us / 1e6, 2)}
    print(json.dumps({"stream_probe": out, "elements": n}), flush=True)


def make_solver(cfg):
    from cajitafluids_amd import configs as K
    from cajitafluids_amd.solver import Solver
    s = Solver(cfg)
    s.init_vortex_tubes(K.VORTEX_TANK_SEED, 1.0)
    return s


def timed_run(sol, steps, warmup, timing=True):
    sol.setup()
    for _ in range(warmup):
        sol.step()
    sol.sync()
    sol.reset_kernel_time()
    sol.set_kernel_timing(timing)
    iters, outer, res = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = sol.step()
        iters.append(st["iterations"])
        outer.append(st["outer_iterations"])
        res.append(st["final_residual"])
    sol.sync()
    t = time.perf_counter() - t0
    kt = {k: sol.kernel_time(k) for k in range(_lib_tk_count())}
    sol.set_kernel_timing(False)
    return t, iters, outer, res, kt


def _lib_tk_count():
    from cajitafluids_amd import _lib
    return _lib.TK_VCYCLE + 1


def summarize_trace(path):
    """Per (kernel, grid size): dispatches, mean and total duration of a
    rocprofv3 kernel_trace CSV, largest total first, and the means of the
    even- and odd-numbered dispatches in launch order.  The grid size tells
    the multigrid levels apart; where two levels share a grid size and
    alternate, the even / odd means separate them."""
    import csv
    vals = {}
    with open(path) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
        for r in rows:
            gsz = 1
            for ax in ("X", "Y", "Z"):
                gsz *= int(r.get(f"Grid_Size_{ax}", 1) or 1)
            if "Grid_Size" in r:
                gsz = int(r["Grid_Size"])
            key = (r["Kernel_Name"], gsz)
            vals.setdefault(key, []).append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    w = csv.writer(sys.stdout)
    w.writerow(["kernel", "grid_items", "dispatches", "mean_us", "total_ms", "mean_even_us",
                "mean_odd_us"])
    for (k, gsz), v in sorted(vals.items(), key=lambda kv: -sum(kv[1])):
        ev, od = v[0::2], v[1::2]
        w.writerow([k[:80], gsz, len(v), round(sum(v) / len(v), 2), round(sum(v) * 1e-3, 3),
                    round(sum(ev) / len(ev), 2), round(sum(od) / len(od), 2) if od else ""])


def summarize_counters(path):
    """Per-kernel dispatch count, median, max and sum of each counter in a
    rocprofv3 counter_collection CSV, largest sum first."""
    import csv
    import statistics
    vals = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            vals.setdefault((r["Kernel_Name"], r["Counter_Name"]), []).append(
                float(r["Counter_Value"]))
    w = csv.writer(sys.stdout)
    w.writerow(["kernel", "counter", "dispatches", "median", "max", "sum"])
    for (k, c), v in sorted(vals.items(), key=lambda kv: -sum(kv[1])):
        w.writerow([k[:120], c, len(v), statistics.median(v), max(v), sum(v)])


def main():
    args = parse()
    if args.summarize_counters:
        summarize_counters(args.summarize_counters)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs one subdomain on one GPU; "
                          "several GPUs (spatial blocks with halo exchange) "
                          "are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from cajitafluids_amd import _lib, configs as K

    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_legs()}))
        return
    if args.stream_probe:
        stream_probe()
        return
    if args.dye_only:
        print(json.dumps({"dye_tank_128": dye_tank_leg(False)}))
        return
    if args.advect_only:
        print(json.dumps({"advect_ms": advect_only()}))
        return
    if args.copy_probe:
        copy_probe()
    n = args.n
    if args.mg_only:
        m64 = mg_leg(n, args, None, None, None)
        m32 = mg_leg(n, args, None, None, None, f32=True)
        m32["speedup_vs_mg_fp64"] = round(m32["value"] / m64["value"], 3)
        print(json.dumps({"multigrid": m64, "multigrid_fp32": m32}))
        return
    ncell = n ** 3
    if args.no_fp64:
        cfg32 = K.vortex_tank_3d(n, precision=K.FP32_REFINED)
        s32 = make_solver(cfg32)
        timed_run(s32, args.steps, args.warmup)
        s32.close()
        return
    cfg = K.vortex_tank_3d(n)
    sol = make_solver(cfg)
    t64, it64, _, res64, kt = timed_run(sol, args.steps, args.warmup,
                                        timing=args.timing_sample)
    dt_eff = sol.dt
    q64, p64 = owned_q_and_p(sol) if not args.no_fp32 else (None, None)
    sol.close()
    del sol

    value = ncell * args.steps / t64
    names = {_lib.TK_DIR_SPMV: "dir_spmv", _lib.TK_UPDATE: "update"}
    per = {}
    for k, nm in names.items():
        ms, cnt = kt[k]
        if cnt:
            per[nm] = dict(avg_us=1e3 * ms / cnt, launches=cnt, total_ms=ms,
                           bytes=WORDS[nm] * 8 * ncell)
    # Kernel time per iteration from the sampled launches, times the
    # iterations of the timed region.
    t_pcg = sum(p["avg_us"] for p in per.values()) * 1e-6 * sum(it64)
    total_iters = sum(it64)
    pcg_gbs = 88.0 * ncell * total_iters / t_pcg / 1e9 if t_pcg > 0 else None
    roofline = None
    if per:
        # dominant kernel = the one with the most device time per iteration
        dom = max(per, key=lambda k: per[k]["avg_us"])
        d = per[dom]
        achieved = d["bytes"] / (d["avg_us"] * 1e-6) / 1e9
        kname = KERNEL64[dom]
        traffic = measured_traffic(kname) if n == 256 else None
        roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "traffic_note": TRAFFIC_NOTE if traffic else None,
                    "bytes_per_launch": d["bytes"], "avg_us": round(d["avg_us"], 2)}
    adv_ms, adv_n = kt[_lib.TK_ADVECT]

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t64 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded divergence-free vortex tubes (seed 1), sphere inflow",
        "config": {"workload": f"3-D {n}^3 vortex tank: MAC grid, FP64, cubic B-spline fields, "
                               f"RK3, Jacobi-PCG to 1e-8, dt clamped to {dt_eff:.6g}",
                   "grid": [n, n, n], "halo": cfg["halo"], "cg_tol": cfg["cg_tol"]},
        "pcg": {"iterations_per_step": round(total_iters / args.steps, 1),
                "iterations": it64, "gbs_11_words": round(pcg_gbs, 1) if pcg_gbs else None,
                "frac_11_words": round(pcg_gbs / HBM_PEAK_GBS, 4) if pcg_gbs else None,
                "time_share": round(t_pcg / t64, 4),
                "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                                for kk, vv in v.items()} for k, v in per.items()},
                "max_final_residual": max(res64)},
        "advect": advect_figures(cfg, adv_ms / adv_n) if adv_n else None,
        "roofline": roofline,
    }

    if not args.no_fp32:
        cfg32 = K.vortex_tank_3d(n, precision=K.FP32_REFINED)
        s32 = make_solver(cfg32)
        t32, it32, outer32, res32, kt32 = timed_run(s32, args.steps, args.warmup,
                                                    timing=args.timing_sample)
        q32, p32 = owned_q_and_p(s32)
        s32.close()
        del s32
        v32 = ncell * args.steps / t32
        k32 = {}
        for k, nm in ((_lib.TK_DIR_SPMV32, "dir_spmv"), (_lib.TK_UPDATE32, "update")):
            ms, cnt = kt32[k]
            if cnt:
                gbs = WORDS[nm] * 4 * ncell / (ms / cnt * 1e-3) / 1e9
                k32[nm] = {"avg_us": round(1e3 * ms / cnt, 3), "launches": cnt,
                           "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        ms, cnt = kt32[_lib.TK_RESIDUAL]
        if cnt:
            k32["fp64_residual"] = {"avg_us": round(1e3 * ms / cnt, 3), "launches": cnt}
        out["fp32_refined"] = {
            "value": v32, "unit": "cell-steps/s", "speedup_vs_fp64": round(v32 / value, 3),
            "ms_per_step": 1e3 * t32 / args.steps,
            "inner_iterations_per_step": round(sum(it32) / args.steps, 1),
            "outer_passes_per_step": round(sum(outer32) / args.steps, 2),
            "max_final_residual": max(res32),
            "refine_mode": "carry" if cfg32["refine_mode"] == K.REFINE_CARRY else "restart",
            "inner_rtol": cfg32["inner_tol"],
            "dye_rel_l2_vs_fp64": rel(q32, q64),
            "pressure_rel_l2_vs_fp64": rel(p32, p64),
            "kernels": k32,
        }

    if not args.no_mg:
        out["multigrid"] = mg_leg(n, args, value, q64, p64)
        out["multigrid_fp32"] = mg_leg(n, args, value, q64, p64, f32=True)
        out["multigrid_fp32"]["speedup_vs_mg_fp64"] = round(
            out["multigrid_fp32"]["value"] / out["multigrid"]["value"], 3)

    if args.adjoint:
        out["adjoint"] = adjoint_legs()

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, n, total_iters / args.steps, args.cpu_iters)
    if not args.no_dye:
        out["dye_tank_128"] = dye_tank_leg(not args.no_cpu)

    print(json.dumps(out))


# FP64 operations (add, sub and mul 1, fma 2) per advected point of a 3-D
# step with cubic fields and linear velocity, counted from the device code:
# RK3 = 3 velocity evaluations x 3 components x (12 for the weights + 28 for
# the 8-tap product sum) + 39 for the stage positions and the combination =
# 399; the cubic field sample = 66 for the weights + 208 for the 64-tap sum.
ADVECT_FLOPS_PER_POINT_3D_CUBIC = 399 + 274
FP64_PEAK_TFLOPS = 78.6


def advect_figures(cfg, avg_ms):
    """Advection's rates: owned points (cells and faces), unique HBM bytes
    (read each state array once, write each Next array once) and FP64
    operations per launch over its mean duration."""
    from cajitafluids_amd.solver import layout
    pts = 0
    for e in range(cfg["dim"] + 1):
        _, lo, hi = layout(cfg, e)
        pts += (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2])
    out = {"avg_ms": round(avg_ms, 3), "points": pts,
           "gbs_unique": round(16 * pts / (avg_ms * 1e-3) / 1e9, 1)}
    if cfg["dim"] == 3 and cfg["field_order"] == 3 and cfg["velocity_order"] == 1:
        tf = ADVECT_FLOPS_PER_POINT_3D_CUBIC * pts / (avg_ms * 1e-3) / 1e12
        out.update(flops_per_point=ADVECT_FLOPS_PER_POINT_3D_CUBIC,
                   tflops=round(tf, 2), fp64_peak_frac=round(tf / FP64_PEAK_TFLOPS, 4),
                   hbm_frac=round(out["gbs_unique"] / HBM_PEAK_GBS, 4),
                   bound="gathers: 136 taps per point through L1/L2, three RK3 stages in sequence")
    return out


def advect_only():
    """Mean advection kernel time (HIP events, 4 launches) on the 256^3 vortex
    tank and on the 384^3 periodic box (FP64, multigrid solver so that the
    setup is quick), each after setup and one step."""
    from cajitafluids_amd import _lib, configs as K
    from cajitafluids_amd.solver import Solver
    out = {}
    tank = K.vortex_tank_3d(256)
    tank["preconditioner"] = K.PRECOND_MG
    box = K.periodic_box_3d(384)
    box.update(precision=K.FP64, preconditioner=K.PRECOND_MG)
    for name, cfg, init in (
            ("vortex_tank_256", tank, lambda s: s.init_vortex_tubes(K.VORTEX_TANK_SEED, 1.0)),
            ("periodic_box_384", box, lambda s: (s.init_vortex_tubes(K.PERIODIC_BOX_VEL_SEED, 1.0),
                                                 s.init_blobs(K.PERIODIC_BOX_DYE_SEED, 8)))):
        with Solver(cfg) as s:
            init(s)
            s.setup()
            s.step()
            s.sync()
            s.reset_kernel_time()
            s.set_kernel_timing(True)
            for _ in range(4):
                s.time_integrate()
            s.sync()
            ms, n = s.kernel_time(_lib.TK_ADVECT)
            out[name] = round(ms / n, 4)
    return out


def dye_tank_leg(with_cpu):
    """The 2-D 128^2 dye tank, setup + 50 steps (linear fields, Jacobi-PCG
    to 1e-8), as a whole run: on the GPU, and with the oracle on the host
    cores.  At 128^2 the GPU path is bound by the latency of the PCG's two
    reductions per iteration (one cooperative launch per solve)."""
    from cajitafluids_amd import configs as K
    from cajitafluids_amd.solver import Solver
    cfg = K.dye_tank_2d(128)
    nsteps = K.DYE_TANK_STEPS
    cells = 128 * 128
    out = {"workload": "2-D 128^2 dye tank, setup + 50 steps, FP64, linear fields",
           "unit": "cell-steps/s"}
    with Solver(cfg) as s:
        s.setup()          # warm-up run (module load, first launches)
        s.sync()
    with Solver(cfg) as s:
        t0 = time.perf_counter()
        s.setup()
        its = 0
        for _ in range(nsteps):
            its += s.step()["iterations"]
        s.sync()
        t = time.perf_counter() - t0
    out["gpu"] = {"value": cells * nsteps / t, "seconds": round(t, 4),
                  "pcg_iterations": its}
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        orc = O.Oracle(cfg)
        t0 = time.perf_counter()
        rc, _ = orc.setup()
        for _ in range(nsteps):
            rc, _ = orc.step()
        t = time.perf_counter() - t0
        orc.close()
        out["cpu_oracle"] = {"value": cells * nsteps / t, "seconds": round(t, 3),
                             "cores": threads, "kind": "port"}
        out["gpu_over_cpu"] = round(out["gpu"]["value"] / out["cpu_oracle"]["value"], 2)
    return out


def mg_leg(n, args, value64, q64, p64, f32=False):
    """The same workload with the V-cycle preconditioner (same 1e-8
    relative residual): FP64, or (f32) with the V-cycle's level 0 on FP32
    vectors inside the FP64 CG (precision CF_FP32_REFINED)."""
    from cajitafluids_amd import _lib, configs as K
    cfg = K.vortex_tank_3d(n, precision=K.FP32_REFINED if f32 else K.FP64)
    cfg["preconditioner"] = K.PRECOND_MG
    for key, val in (("mg_nu", args.mg_nu), ("mg_alpha", args.mg_alpha),
                     ("mg_omega", args.mg_omega), ("mg_coarse_sweeps", args.mg_coarse)):
        if val is not None:
            cfg[key] = val
    sol = make_solver(cfg)
    # 9 iterations a step: every iteration timed (1-in-N sampling would
    # rarely hit one)
    t, its, _, res, kt = timed_run(sol, args.steps, args.warmup,
                                   timing=1 if args.timing_sample else 0)
    qm, pm = owned_q_and_p(sol) if q64 is not None else (None, None)
    sol.close()
    del sol
    v = n ** 3 * args.steps / t
    kern = {}
    for k, nm in ((_lib.TK_DIR_SPMV, "dir_spmv"), (_lib.TK_UPDATE, "update"),
                  (_lib.TK_VCYCLE, "vcycle")):
        ms, cnt = kt[k]
        if cnt:
            kern[nm] = {"avg_us": round(1e3 * ms / cnt, 3), "launches": cnt}
    return {
        "value": v, "unit": "cell-steps/s",
        "vcycle_level0": "fp32" if f32 else "fp64",
        "speedup_vs_jacobi": round(v / value64, 3) if value64 else None,
        "ms_per_step": 1e3 * t / args.steps,
        "iterations_per_step": round(sum(its) / args.steps, 1),
        "max_final_residual": max(res),
        "mg_nu": cfg["mg_nu"], "mg_omega": cfg["mg_omega"],
        "mg_coarse_sweeps": cfg["mg_coarse_sweeps"], "mg_alpha": cfg["mg_alpha"],
        "dye_rel_l2_vs_jacobi": rel(qm, q64) if qm is not None else None,
        "pressure_rel_l2_vs_jacobi": rel(pm, p64) if pm is not None else None,
        "kernels": kern,
    }


def adjoint_leg(cfg, nsteps, ckpt, init, loss_grad):
    """Wall time of setup + nsteps (forward) and of cf_adjoint_run (which
    runs the forward again with checkpoints, recomputes each segment and
    sweeps back), from the same initial state."""
    import numpy as np
    from cajitafluids_amd.solver import Solver
    with Solver(cfg) as s:
        init(s)
        s.sync()
        t0 = time.perf_counter()
        s.setup()
        for _ in range(nsteps):
            s.step()
        s.sync()
        t_fwd = time.perf_counter() - t0
        dqT = loss_grad(s)
    with Solver(cfg) as s:
        init(s)
        s.sync()
        t0 = time.perf_counter()
        g = s.adjoint_run(nsteps, ckpt, dqT)
        t_adj = time.perf_counter() - t0
    cells = int(np.prod([c for c in cfg["num_cell"][:cfg["dim"]]]))
    print(f"adjoint leg: {cells} cells, {nsteps} steps: forward {t_fwd:.3f} s, "
          f"adjoint {t_adj:.3f} s", file=sys.stderr, flush=True)
    return {"cells": cells, "steps": nsteps, "ckpt": ckpt,
            "forward_s": round(t_fwd, 3), "adjoint_s": round(t_adj, 3),
            "adjoint_over_forward": round(t_adj / t_fwd, 2),
            "grad_u_norm": float(np.linalg.norm(g["du"])),
            "dvel": [float(v) for v in g["dvel"]],
            "note": "adjoint_s includes the host copies of dL/dq_T in and of the "
                    "gradients out (cf_adjoint_run takes host arrays); adjoint_device_s, "
                    "where present, is cf_adjoint_run_device on device arrays",
            "_du": g["du"]}


def adjoint_device_s(cfg, nsteps, ckpt, init):
    """Wall time of cf_adjoint_run_device with every array in device memory
    (no host copies), for the loss 0.5 ||q_T||^2: dL/dq_T = q_T, copied
    device to device from a forward run's final state.  Returns the time and
    dL/d(inflow velocity)."""
    import ctypes as C
    from cajitafluids_amd import _lib
    from cajitafluids_amd.solver import Solver, padded_layout
    _lib.load()
    hip = C.CDLL("libamdhip64.so.7")  # the runtime libcfmi.so links
    nbytes = padded_layout(cfg)[3] * 8
    bufs = []
    try:
        for _ in range(5):
            p = C.c_void_p()
            if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
                raise RuntimeError("hipMalloc failed")
            bufs.append(p.value)
        with Solver(cfg) as s:
            init(s)
            s.setup()
            for _ in range(nsteps):
                s.step()
            s.copy_out_device(0, bufs[0])
            s.sync()
        with Solver(cfg) as s:
            init(s)
            s.sync()
            t0 = time.perf_counter()
            dvel = s.adjoint_run_device(nsteps, ckpt, *bufs)
            t = time.perf_counter() - t0
        print(f"device adjoint: {t:.3f} s", file=sys.stderr, flush=True)
        return t, dvel
    finally:
        for p in bufs:
            hip.hipFree(C.c_void_p(p))


def with_mg(cfg):
    """The same workload with the FP64 multigrid-preconditioned solver."""
    from cajitafluids_amd import configs as K
    c = K.copy_config(cfg)
    c.update(preconditioner=K.PRECOND_MG, precision=K.FP64)
    return c


def compare_mg(out, name):
    """Attach the MG leg's speed-up and gradient difference to the Jacobi
    leg of the same workload."""
    import numpy as np
    a, b = out[name], out[name + "_mg"]
    b["forward_speedup_vs_jacobi"] = round(a["forward_s"] / b["forward_s"], 2)
    b["adjoint_speedup_vs_jacobi"] = round(a["adjoint_s"] / b["adjoint_s"], 2)
    b["grad_u_rel_diff_vs_jacobi"] = float(np.linalg.norm(b["_du"] - a["_du"]) /
                                           np.linalg.norm(a["_du"]))
    for leg in (a, b):
        leg.pop("_du", None)


def adjoint_legs():
    import numpy as np
    from cajitafluids_amd import configs as K
    out = {}
    # 2-D 512^2 dye tank: loss 0.5 ||q_T - blobs||^2, initial velocity 0.2 N(0,1)
    c4 = K.adjoint_tank_2d(512)
    from cajitafluids_amd.solver import Solver
    with Solver(c4) as s:
        s.init_blobs(K.DYE_ADJOINT_TARGET_SEED, 8)
        target = s.get(0)
    rng = np.random.default_rng(K.DYE_ADJOINT_FD_SEED)
    faces = {}
    for e in (1, 2):
        from cajitafluids_amd.solver import layout
        ext, lo, hi = layout(c4, e)
        a = np.zeros((ext[2], ext[1], ext[0]))
        sl = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
        a[sl] = 0.2 * rng.normal(size=a[sl].shape)
        faces[e] = a

    def init4(s):
        for e, a in faces.items():
            s.set(e, a)
    out["dye_tank_512"] = adjoint_leg(c4, K.DYE_ADJOINT_STEPS, K.DYE_ADJOINT_CKPT, init4,
                                      lambda s: s.get(0) - target)
    out["dye_tank_512_mg"] = adjoint_leg(with_mg(c4), K.DYE_ADJOINT_STEPS, K.DYE_ADJOINT_CKPT, init4,
                                         lambda s: s.get(0) - target)
    compare_mg(out, "dye_tank_512")
    # 3-D 384^3 periodic box: vortex tubes (seed 5), dye blobs (seed 55),
    # loss 0.5 ||q_T||^2
    c5 = K.periodic_box_3d(384)

    def init5(s):
        s.init_vortex_tubes(K.PERIODIC_BOX_VEL_SEED, 1.0)
        s.init_blobs(K.PERIODIC_BOX_DYE_SEED, 8)
    out["periodic_box_384"] = adjoint_leg(c5, K.PERIODIC_BOX_STEPS, K.PERIODIC_BOX_CKPT, init5,
                                          lambda s: s.get(0))
    out["periodic_box_384_mg"] = adjoint_leg(with_mg(c5), K.PERIODIC_BOX_STEPS, K.PERIODIC_BOX_CKPT,
                                             init5, lambda s: s.get(0))
    t, _ = adjoint_device_s(with_mg(c5), K.PERIODIC_BOX_STEPS, K.PERIODIC_BOX_CKPT, init5)
    out["periodic_box_384_mg"]["adjoint_device_s"] = round(t, 3)
    # the workload's FP32-refined precision with the multigrid solver: the
    # V-cycle's finest level in FP32
    c5m = with_mg(c5)
    c5m["precision"] = K.FP32_REFINED
    out["periodic_box_384_mg_fp32"] = adjoint_leg(c5m, K.PERIODIC_BOX_STEPS,
                                                  K.PERIODIC_BOX_CKPT, init5, lambda s: s.get(0))
    out["periodic_box_384_mg_fp32"]["grad_u_rel_diff_vs_mg_fp64"] = float(
        np.linalg.norm(out["periodic_box_384_mg_fp32"]["_du"] - out["periodic_box_384_mg"]["_du"])
        / np.linalg.norm(out["periodic_box_384_mg"]["_du"]))
    out["periodic_box_384_mg_fp32"].pop("_du", None)
    compare_mg(out, "periodic_box_384")
    return out


TRAFFIC_NOTE = ("bytes per launch = 2 x FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc "
                "passes over this script at 256^3 (FETCH_SIZE reads half of a 1 GiB "
                "copy on this GPU, WRITE_SIZE reads it whole); "
                "profiles/pcg_tiles_vortex_tank_256.md")


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel, from the counter summary committed
    beside the kernel-trace summary (None when absent)."""
    path = os.path.join(REPO, "profiles", "pcg_tiles_vortex_tank_256.md")
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 4 and cells[0].startswith(kernel + " ") and cells[2].endswith("MB"):
                    return round(float(cells[2].split()[0]) * 1e6)
    except OSError:
        pass
    return None


def owned_q_and_p(sol):
    """Owned dye and last pressure (mean removed) of a solver, on the host."""
    q = sol.owned_view(sol.get(0), 0).copy()
    p = sol.owned_view(sol.get_pressure(0), 0).copy()
    return q, p - p.mean()


def rel(a, b):
    import numpy as np
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / nb if nb else float(np.linalg.norm(a))


def cpu_baseline(cfg, n, iters_per_step, cpu_iters):
    """The oracle (a plain-C restatement with the reference's stored-stencil
    traffic) timed on the host: one advection + inputs + RHS + projection,
    and cpu_iters PCG iterations, on the same 256^3 state; the step time is
    extrapolated with the GPU's PCG iteration count (the two paths take the
    same iterations)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    from cajitafluids_amd import configs as K
    from cajitafluids_amd.solver import Solver

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sol = Solver(cfg)
    sol.init_vortex_tubes(K.VORTEX_TANK_SEED, 1.0)
    sol.setup()
    c2 = dict(cfg)
    c2["cg_max_iter"] = cpu_iters
    orc = O.Oracle(c2)
    for e in range(4):
        orc.field(e)[...] = sol.get(e)
    sol.close()
    t0 = time.perf_counter()
    orc.time_integrate()
    orc.add_inputs()
    orc.build_rhs()
    t1 = time.perf_counter()
    orc.pressure_solve(O.FP64)   # stops at cpu_iters (non-converged)
    t2 = time.perf_counter()
    orc.apply_pressure()
    t3 = time.perf_counter()
    t_iter = (t2 - t1) / cpu_iters
    t_step = (t1 - t0) + (t3 - t2) + iters_per_step * t_iter
    orc.close()
    return {"value": n ** 3 / t_step, "unit": "cell-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"one {n}^3 step's advection+inputs+RHS+projection and {cpu_iters} "
                      f"PCG iterations of the oracle (OpenMP, {threads} threads), "
                      f"step time extrapolated with {iters_per_step:.0f} iterations/step",
            "sample_seconds": round(t3 - t0, 2), "pcg_iter_ms": round(1e3 * t_iter, 2)}


if __name__ == "__main__":
    main()
