"""Benchmark of the implicit-IPC step on one MI355X.

Workload: the single-garment production case - a 1 m closed tube garment (150 x 133 = 19 950
vertices, 39 600 StVK shell triangles, +-2 mm radial noise) with three 200-segment seam rods,
falling onto a static 12-capsule body (25 368 vertices, 50 688 triangles); hinge bending,
log-barrier contact with Rayleigh stiffness, friction 0.3, dhat 1 mm, dt 1/60 s, 4 Newton
iterations, each linear solve to 1e-6 relative residual (FP32 PCG + FP64 refinement), CCD
line search. A step is one `nguo_step_advance(1)`; inputs are resident on the device.

Prints one JSON line: ms per implicit step (lower is better), the roofline of the PCG
operator (3x3-BSR SpMV + matrix-free contact) timed with HIP events on the library stream,
the same operator on a synthetic 2048 x 2048 cloth BSR that exceeds the last-level cache,
the graded-size batch (8 tubes advanced concurrently vs one after another), 20 timed
gradient-descent iterations of the drape adjoint, and the CPU baseline: the oracle
restatement (FP32, one core) on a bounded sample.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def load_module(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps timed for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-warmup", type=int, default=1, help="untimed oracle steps before the CPU baseline's timed ones")
    ap.add_argument("--full-steps", type=int, default=60, help="steps of the whole-config run timed beside the K timed steps (0 = skip)")
    ap.add_argument("--grid", type=int, default=2048, help="side of the synthetic SpMV grid (0 = skip)")
    ap.add_argument("--c0-steps", type=int, default=36, help="steps of the sheet-on-sphere drape timed as a second plain workload (0 = skip)")
    ap.add_argument("--sort-n", type=int, default=1 << 26, help="pairs of the synthetic radix-sort measurement (0 = skip)")
    ap.add_argument("--c3-steps", type=int, default=10, help="timed steps of the 8-size graded batch vs sequential (0 = skip)")
    ap.add_argument("--c3-warmup", type=int, default=10)
    ap.add_argument("--c3-solve", choices=["both", "batched", "concurrent"], default="both",
                    help="batch solve modes to time (a single mode also skips the sequential runs)")
    ap.add_argument("--c4-iters", type=int, default=20, help="timed gradient-descent iterations of the drape adjoint (0 = skip)")
    ap.add_argument("--c4-hessian", choices=["frozen", "exact", "newton"], default="newton",
                    help="adjoint Hessian: the exact Jacobian with the contact parameters held at x* (frozen), exact elastic "
                         "second derivatives only (exact), or the Newton iteration's projected blocks (newton)")
    ap.add_argument("--schwarz-domain", type=int, default=None, help="override the scene's Schwarz domain size")
    ap.add_argument("--schwarz-weight", type=float, default=None, help="override the scene's two-level coarse weight (0 = one level)")
    # 17.94 MB read + 1.22 MB written per launch of the roofline kernel (the Schwarz apply) on
    # this workload in contact, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a
    # 256 MiB add and fill of known traffic
    ap.add_argument("--traffic-bytes", type=float, default=19.15e6,
                    help="HBM bytes per launch of the roofline kernel, measured with rocprofv3 counter passes")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the bench, copy a buffer of this many MiB once (counter calibration for --pmc passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    nh = load_module("nguo_hip", os.path.join(ROOT, "nguo-3d_amd", "nguo_hip.py"))
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except ImportError:
        sync = lambda: None  # noqa: E731

    ctx = nh.Context(0)
    name, cus = ctx.info()
    scene = nh.Scene(1, seed=1)
    if args.schwarz_domain is not None:
        scene.desc.params.schwarz_domain = args.schwarz_domain
    if args.schwarz_weight is not None:
        scene.desc.params.schwarz_weight = args.schwarz_weight
    progress("tube garment: warm-up and timed steps")
    st = nh.Stepper(ctx, scene)
    for _ in range(args.warmup):
        st.advance(1)
    sync()
    x_w, v_w = st.positions(), st.velocities()  # start state of the timed steps (CPU sample starts here too)
    t0 = time.perf_counter()
    iters, newton, stencils, res, mlist = [], [], [], 0.0, []
    for _ in range(args.steps):
        st.advance(1)
        m = st.metrics()
        mlist.append(m)
        iters.append(m.max_cg_iters)
        newton.append(m.newton_iters)
        stencils.append(m.num_stencils)
        res = max(res, m.max_reresid)
        assert m.all_finite == 1, "non-finite step"
    sync()
    elapsed = time.perf_counter() - t0
    ms = 1000.0 * elapsed / args.steps
    x_e, v_e = st.positions(), st.velocities()  # the drape after warmup + timed steps (start of the adjoint loop)
    inv_ok, inv_checks = nh.check_invariants(mlist, scene.desc.params.cg_tol, scene.desc.params.cg_max_iter)
    invariants = {"all_ok": inv_ok, "failed": [c[0] for c in inv_checks if not c[1]]}

    # the dominant kernel (most kernel time per step): the Schwarz launch of every PCG
    # iteration, timed with HIP events on the library stream on the final state's system
    pc_ms, pc_bytes = st.bench_precond(reps=500)
    pc_ach = pc_bytes / (pc_ms * 1e-3) / 1e9
    roofline = {"kernel": "k_schwarz_pcg<false> (FP32 tiled Schwarz inverses, the rows' contact blocks, r/d update, r.z and r.r partials)",
                "bound": "hbm", "achieved": round(pc_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(pc_ach / HBM_PEAK_GBS, 4),
                "traffic": args.traffic_bytes, "bytes_per_launch": pc_bytes, "avg_launch_us": round(pc_ms * 1000, 3)}
    # the other launch of every PCG iteration: the SpMV with the contact stencils (the
    # metric's SpMV), in its iteration form and as the plain operator
    spmv_ms, spmv_bytes = st.bench_spmv(reps=500, iteration_form=True)
    plain_ms, plain_bytes = st.bench_spmv(reps=500, iteration_form=False)
    spmv_ach = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    roofline["spmv_iteration"] = {"kernel": "k_pcg_iter<SPMV_NEXT> (SELL-64 3x3 BSR SpMV + contact stencil blocks, p = z + beta p and the rz/rr reduction fused)",
                                  "avg_launch_us": round(spmv_ms * 1000, 3), "bytes_per_launch": spmv_bytes, "achieved": round(spmv_ach, 1),
                                  "frac": round(spmv_ach / HBM_PEAK_GBS, 4)}
    roofline["plain_spmv"] = {"avg_launch_us": round(plain_ms * 1000, 3), "bytes_per_launch": plain_bytes,
                              "achieved": round(plain_bytes / (plain_ms * 1e-3) / 1e9, 1)}
    full = None
    if args.full_steps > 0:
        progress("tube garment: whole 60-step config")
        # the whole config: steps 1..full_steps of a fresh stepper (the timed window above
        # is steps warmup+1 .. warmup+steps of its own run)
        sf = nh.Stepper(ctx, scene)
        sync()
        tf = time.perf_counter()
        fl = []
        for _ in range(args.full_steps):
            sf.advance(1)
            fl.append(sf.metrics())
        sync()
        f_ok, f_checks = nh.check_invariants(fl, scene.desc.params.cg_tol, scene.desc.params.cg_max_iter)
        full = {"steps": f"1..{args.full_steps}", "ms_per_step": round(1000.0 * (time.perf_counter() - tf) / args.full_steps, 3),
                "mean_total_pcg_iters": round(sum(m.sum_cg_iters for m in fl) / len(fl), 1),
                "invariants": {"all_ok": f_ok, "failed": [c[0] for c in f_checks if not c[1]]}}
        sf.close()
    grid = None
    if args.grid:
        progress("SpMV beyond the LLC")
        g_ms, g_bytes, g_nnzb = nh.bench_spmv_grid(ctx, args.grid, args.grid, reps=20)
        g_ach = g_bytes / (g_ms * 1e-3) / 1e9
        grid = {"workload": f"{args.grid}x{args.grid} cloth-grid BSR, {int(g_nnzb)} blocks", "achieved": round(g_ach, 1),
                "unit": "GB/s", "frac": round(g_ach / HBM_PEAK_GBS, 4), "avg_launch_us": round(g_ms * 1000, 2),
                "bytes_per_launch": g_bytes}

    c0 = None
    if args.c0_steps > 0:
        progress("sheet on sphere")
        # the reference's regression scene: 32 x 32 sheet dropped onto a static icosphere
        sc0 = nh.Scene(0, seed=0)
        st0 = nh.Stepper(ctx, sc0)
        sync()
        t1 = time.perf_counter()
        it0 = []
        for _ in range(args.c0_steps):
            st0.advance(1)
            it0.append(st0.metrics().max_cg_iters)
        sync()
        m0 = st0.metrics()
        c0 = {"workload": f"32x32 sheet (1024 v, 1922 t) dropped onto a static 2562-vertex icosphere, steps 1..{args.c0_steps}",
              "ms_per_step": round(1000.0 * (time.perf_counter() - t1) / args.c0_steps, 3), "mean_pcg_iters": sum(it0) / len(it0),
              "penetration_free": int(m0.penetration_free), "all_finite": int(m0.all_finite)}
        st0.close()

    sort = None
    if args.sort_n:
        progress("radix sort")
        s_ms, s_bytes = nh.bench_sort(ctx, args.sort_n, reps=10)
        s_ach = s_bytes / (s_ms * 1e-3) / 1e9
        sort = {"workload": f"4-pass LSD radix sort of {args.sort_n} random (u32 key, u32 index) pairs", "avg_sort_ms": round(s_ms, 3),
                "bytes_per_sort": s_bytes, "achieved": round(s_ach, 1), "unit": "GB/s", "frac": round(s_ach / HBM_PEAK_GBS, 4)}

    c3 = None
    if args.c3_steps > 0:
        progress("graded batch")
        # graded sizes: circumference factors 0.85 .. 1.20 of the tube garment, each on its
        # scaled body; all 8 advanced concurrently vs the same 8 advanced one after another
        factors = [0.85, 0.90, 0.95, 1.00, 1.05, 1.10, 1.15, 1.20]
        scenes = [nh.Scene(3, seed=3, param=f) for f in factors]
        seq_ms = []
        for sc in (scenes if args.c3_solve == "both" else []):
            s1 = nh.Stepper(ctx, sc)
            s1.advance(args.c3_warmup)
            sync()
            t1 = time.perf_counter()
            s1.advance(args.c3_steps)
            sync()
            seq_ms.append(1000.0 * (time.perf_counter() - t1) / args.c3_steps)
            s1.close()
        b_ms = {}
        modes = {"both": (True, False), "batched": (True,), "concurrent": (False,)}[args.c3_solve]
        for mode in modes:
            batch = nh.Batch(ctx, scenes)
            batch.set_solve(mode)
            batch.advance(args.c3_warmup)
            sync()
            t1 = time.perf_counter()
            batch.advance(args.c3_steps)
            sync()
            b_ms[mode] = 1000.0 * (time.perf_counter() - t1) / args.c3_steps
            batch.close()
        free = [int((sc.inv_mass > 0).sum()) for sc in scenes]
        best = min(b_ms.values())
        c3 = {"workload": f"8 graded tubes, circumference x{factors[0]}..x{factors[-1]}, {min(free)}..{max(free)} free vertices "
                          f"({sum(free)} total), steps {args.c3_warmup + 1}..{args.c3_warmup + args.c3_steps}",
              "batched_ms_per_step": round(best, 3), "sequential_ms_per_step": round(sum(seq_ms), 3) if seq_ms else None,
              "per_size_ms": [round(v, 3) for v in seq_ms], "speedup": round(sum(seq_ms) / best, 3) if seq_ms else None,
              "batched_pcg_ms_per_step": round(b_ms[True], 3) if True in b_ms else None,
              "concurrent_solves_ms_per_step": round(b_ms[False], 3) if False in b_ms else None,
              "form": "one stream and host thread per instance; Newton solves either as one batched PCG over all instances "
                      "(two launches per iteration) or each on its instance stream"}

    c4 = None
    if args.c4_iters > 0:
        progress("drape adjoint")
        # differentiable drape: 64 panel rest scales ~ N(1, 0.02) (seed 4), target = the drape
        # of a second draw (seed 5) from the same start state (the drape after the timed
        # steps); each gradient-descent iteration = one implicit step from that state + the
        # adjoint solve
        import numpy as np
        panel = nh.drape_panels(scene)
        free = scene.inv_mass > 0
        theta = np.random.default_rng(4).normal(1.0, 0.02, 64)
        theta_t = np.random.default_rng(5).normal(1.0, 0.02, 64)

        def forward(th):
            st.set_rest_scale(th[panel].astype(np.float32))
            st.set_state(x_e, v_e)
            st.advance(1)
            return st.positions().astype(np.float64)

        target = forward(theta_t)
        mode = {"frozen": st.ADJ_EXACT, "exact": st.ADJ_EXACT_ELASTIC, "newton": st.ADJ_NEWTON}[args.c4_hessian]

        def loss_of(x):
            d = (x - target)[free]
            return float((d * d).sum(axis=1).mean()), d

        def gradient(d):
            dLdx = np.zeros((len(free), 3))
            dLdx[free] = 2.0 * d / free.sum()
            nonlocal fallbacks
            try:
                g_face, stats = st.adjoint(dLdx, tol=1e-6, hessian=mode)
            except nh.NguoError:
                fallbacks += 1  # the exact form does not solve at this end state: the Newton blocks
                g_face, stats = st.adjoint(dLdx, tol=1e-6, hessian=st.ADJ_NEWTON)
            adj_iters.append(stats.iters)
            return np.bincount(panel, weights=g_face.astype(np.float64), minlength=64)

        # gradient descent with step control: a step that lowers the loss is taken and the
        # next one grows by 1.2x; one that raises it is undone and the step halves (the loss
        # of one implicit step with contact events is only piecewise smooth)
        losses, fwd_ms, adj_ms, adj_iters, fallbacks, accepted = [], [], [], [], 0, 0
        dist0 = float(np.linalg.norm(theta - theta_t))
        step = 0.004
        sync()
        t1 = time.perf_counter()
        ta = time.perf_counter()
        loss, d = loss_of(forward(theta))
        tb = time.perf_counter()
        g = gradient(d)
        fwd_ms.append(1000.0 * (tb - ta))
        adj_ms.append(1000.0 * (time.perf_counter() - tb))
        losses.append(loss)
        for it in range(args.c4_iters):
            gmax = np.abs(g).max()
            if not gmax > 0:
                break
            cand = theta - step * g / gmax  # at most `step` per panel
            ta = time.perf_counter()
            lc, dc = loss_of(forward(cand))
            tb = time.perf_counter()
            fwd_ms.append(1000.0 * (tb - ta))
            if lc < loss:
                theta, loss, d = cand, lc, dc
                g = gradient(d)
                adj_ms.append(1000.0 * (time.perf_counter() - tb))
                accepted += 1
                step *= 1.2
            else:
                step *= 0.5
            losses.append(loss)
        sync()
        total = time.perf_counter() - t1
        c4 = {"workload": "tube garment (config above) with 64 panel rest scales (8 sectors x 8 height bands), theta ~ N(1, 0.02) "
                          "seed 4, target drape from seed 5; loss = mean |x* - target|^2 over free vertices; "
                          f"one implicit step from the drape after {args.warmup + args.steps} steps + one adjoint solve per iteration",
              "iters": args.c4_iters, "accepted": accepted, "ms_per_iter": round(1000.0 * total / args.c4_iters, 3),
              "forward_ms": round(sum(fwd_ms) / len(fwd_ms), 3), "adjoint_ms": round(sum(adj_ms) / len(adj_ms), 3),
              "adjoint_pcg_iters": round(sum(adj_iters) / len(adj_iters), 1), "hessian": args.c4_hessian,
              "newton_fallbacks": fallbacks, "loss_first": losses[0], "loss_last": losses[-1],
              "theta_dist_first": round(dist0, 6), "theta_dist_last": round(float(np.linalg.norm(theta - theta_t)), 6)}
        st.set_rest_scale(np.ones(scene.desc.n_shell_faces, np.float32))

    cpu = None
    if args.cpu_steps > 0:
        progress("CPU baseline")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py  # the CPU restatement of the reference (checker), timed as the baseline
        cscene = nh.Scene(1, seed=1)
        o = oracle_py.OracleSolver(cscene)
        o.set_state(x_w, v_w)
        o.step(args.cpu_warmup)
        tc = time.perf_counter()
        o.step(args.cpu_steps)
        cpu_s = time.perf_counter() - tc
        model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        cpu = {"value": round(1000.0 * cpu_s / args.cpu_steps, 1), "unit": "ms/implicit step", "cores": 1, "kind": "port",
               "sample": f"steps {args.warmup + args.cpu_warmup + 1}..{args.warmup + args.cpu_warmup + args.cpu_steps} of the same tube-garment "
                         f"workload ({args.cpu_warmup} untimed step from the state the timed GPU steps start at, then {args.cpu_steps} timed), "
                         f"oracle FP32 restatement of the reference CPU tier (f64 block-Jacobi CG), 1 thread of {model}"}

    out = {
        "metric": "ms/implicit step @1e-6 PCG residual (plain single garment); SpMV GB/s vs HBM peak",
        "value": round(ms, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (PCG vectors, elements) + f64 (refinement residual, CCD time of impact, PCG scalars)",
        "data": "synthetic",
        "config": {"workload": f"tube garment 19950 v / 39600 t + 3x200-segment rods on a 12-capsule body (50688 t), 60-step drape config, "
                               f"timed steps {args.warmup + 1}..{args.warmup + args.steps}",
                   "dt": 1.0 / 60.0, "newton_iters": 4, "cg_tol": 1e-6, "device": name, "cus": cus,
                   "precond": {"schwarz_domain": int(scene.desc.params.schwarz_domain), "coarse_weight": float(scene.desc.params.schwarz_weight)}},
        # mean_pcg_iters: the largest Newton solve of a step; total: all solves of a step (the
        # step's time over it bounds the cost of one PCG iteration from above)
        "stats": {"mean_pcg_iters": sum(iters) / len(iters),
                  "mean_total_pcg_iters": sum(m.sum_cg_iters for m in mlist) / len(mlist),
                  "us_per_pcg_iter_max": round(1000.0 * ms / max(1.0, sum(m.sum_cg_iters for m in mlist) / len(mlist)), 2),
                  "mean_newton": sum(newton) / len(newton),
                  "mean_stencils": sum(stencils) / len(stencils), "max_reresid": res,
                  "invariants": invariants},
        "roofline": roofline,
        "full_config": full,
        "spmv_hbm": grid,
        "sort_hbm": sort,
        "plain_c0": c0,
        "batched_c3": c3,
        "adjoint_c4": c4,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    st.close()
    ctx.close()
    if args.copy_calib_mib:
        import torch
        # counter calibration: a fill (writes N MiB), then an elementwise add (reads N MiB,
        # writes N MiB), each one kernel of known traffic
        src = torch.ones(args.copy_calib_mib * 262144, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.add(src, 1.0, out=dst)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
