#!/usr/bin/env python3
"""Benchmark: time to 1e-8 relative residual of hecmw1's 3x3-block CG on one MI355X.

Workload (default): 100^3 HEX8 linear-elastic unit cube, 3,090,903 DOF,
27,270,901 3x3 blocks; E=210000, nu=0.3, bottom face fixed, unit pressure on
top; assembled like fistr1 (frontistr_amd/csrc/gen_hex8.cpp).  One step = one
complete FP64 solve x0=0 -> 1e-8 relative residual with the multicolour-SSOR
preconditioned CG (METHOD=1, PRECOND=1, NCOLOR_IN=10), matrix and vectors
resident in HBM, preconditioner set up inside the step (numeric refresh;
the symbolic RCM+multicolour ordering is built once at load).

Also reported: DIAG-CG on the same system, the FP64/FP32 SpMV bandwidth, and
the reference's own CPU solver (oracle/_ref, built from the FrontISTR sources)
timed on the host on a bounded sample of the same solve; the 160^3
heterogeneous cube ("mixed": FP64 SSOR-CG vs the flexible CG with FP32
operators, alone, with the two-level rigid-body-mode correction, and with the
learned correction on top, same GPU); the batched workload ("batch": 48 TET10
systems x 16 right-hand sides); the 120^3 two-phase inclusion cube
("learned": the same modes plus the learned correction alone, with the
network trained on the GPU on solver-harvested residuals first).

    python bench.py [--steps K] [--warmup W] [--n 100] [--gpus 1] [--c3-n 160] [--c4-nsys 48]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the in-solve SpMV (k_spmv<2>) on the 100^3 cube, from
# two rocprofv3 passes (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000, the factors a
# 1 GiB copy of the same run gives); profiles/spmv_hex8_100_traffic.md
SPMV_TRAFFIC_100 = 2199539866


def spmv_bytes(m, w):
    """Algorithmic bytes of one y = A x: the reference's matvec bandwidth
    formula (hecmw_solver_Iterative.f90:212-217) with merged indices:
    9w per block + 4 B column index per block + row pointers + x read + y write."""
    nnzb = m.N + m.NPL + m.NPU
    return 9 * w * nnzb + 4 * nnzb + 4 * (m.N + 1) + 3 * w * m.NP + 3 * w * m.N


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100, help="elements per cube edge")
    ap.add_argument("--precond", type=int, default=1, help="1 SSOR, 3 DIAG for the headline solve")
    ap.add_argument("--cpu-sample-s", type=float, default=20.0, help="budget of the CPU reference sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=100)
    ap.add_argument("--c3-n", type=int, default=160, help="mixed-precision comparison cube (0: skip)")
    ap.add_argument("--c4-nsys", type=int, default=48, help="batched TET10 systems (0: skip)")
    ap.add_argument("--c5-n", type=int, default=120, help="learned-correction cube (0: skip)")
    ap.add_argument("--c5-variants", type=int, default=32, help="training variants for the learned correction")
    ap.add_argument("--only", choices=("c3", "c4", "c5"), default=None,
                    help="run one workload section alone (for profiling); prints that section's JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: one GPU only (--gpus %d requested); sharding over GPUs (one rank per GPU, a halo "
              "exchange per product over RCCL) is not implemented" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import frontistr_amd as fa

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")

    if args.only:
        net, train = train_net(args, dev) if args.only != "c4" and args.c5_variants > 0 else (None, None)
        if args.only == "c3":
            sec = {"mixed": bench_mixed(args, dev, net)}
        elif args.only == "c4":
            sec = {"batch": bench_batch(args)}
        else:
            sec = {"learned": bench_learned(args, dev, net, train)}
        if net is not None:
            net.close()
        print(json.dumps(dict(only=args.only, **sec)), flush=True)
        return

    t0 = time.time()
    m = fa.gen_hex8(args.n)
    log("generated %d^3 HEX8: N=%d nnzb=%d in %.2fs" % (args.n, m.N, m.N + m.NPL + m.NPU, time.time() - t0))
    m.set_iter(20000)
    m.set_precond(args.precond)
    m.set_i(m.I_TIMELOG, 1)

    ctx = fa.DeviceContext()
    t0 = time.time()
    ctx.load(m)
    t_load = time.time() - t0
    info = ctx.info()
    log("loaded in %.2fs (%s)" % (t_load, info))

    b = torch.from_numpy(m.B).to(dev)
    x = torch.zeros(3 * m.NP, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    ia = m.Iarray.copy()
    ra = m.Rarray.copy()

    def step(ia_step):
        x.zero_()
        torch.cuda.synchronize()
        return ctx.solve_device(b.data_ptr(), x.data_ptr(), ia_step, ra)

    # numeric preconditioner refresh every step (Iarray(97)=1), ordering reused
    ia_step = ia.copy()
    ia_step[96] = 1
    ia_step[97] = 0
    for i in range(args.warmup):
        r = step(ia_step)
        log("warmup %d: iter=%d resid=%.3e err=%d Tset=%.3f Tsol=%.3f" % (i, r.iter, r.resid, r.error, r.Tset, r.Tsol))
    torch.cuda.synchronize()
    results = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        results.append(step(ia_step))
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    for i, r in enumerate(results):
        log("step %d: iter=%d resid=%.3e err=%d Tset=%.3f Tsol=%.3f time_Ax=%.3f time_precond=%.3f" % (
            i, r.iter, r.resid, r.error, r.Tset, r.Tsol, r.time_Ax, r.time_precond))
    r0 = results[-1]
    xs = x.cpu().numpy()
    m.X[:] = xs
    true_resid = fa.hecmw_rel_resid_L2(m)
    log("true relative residual %.3e" % true_resid)
    sec_per_solve = t_total / args.steps
    iters = r0.iter

    # dominant kernel: the CG SpMV (q = A p with p.q), timed by events inside the solves
    ax_ms = sum(r.time_Ax for r in results) / sum(r.iter for r in results) * 1e3
    bytes64 = spmv_bytes(m, 8)
    achieved = bytes64 / (ax_ms * 1e-3) / 1e9

    # standalone SpMV bandwidth, FP64 and FP32
    y = torch.zeros(3 * m.NP, dtype=torch.float64, device=dev)
    xr = torch.rand(3 * m.NP, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ms64 = ctx.matvec_device(xr.data_ptr(), y.data_ptr(), args.spmv_reps)
    x32 = xr.float()
    y32 = torch.zeros(3 * m.NP, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    ms32 = ctx.matvec_device(x32.data_ptr(), y32.data_ptr(), args.spmv_reps, fp32=True)
    gbs64 = bytes64 / (ms64 * 1e-3) / 1e9
    gbs32 = spmv_bytes(m, 4) / (ms32 * 1e-3) / 1e9
    # a plain copy of known size (1 GiB read + 1 GiB written): the reference
    # point that scales the L2 traffic counters of a profiled run
    cpy_src = torch.empty(1 << 27, dtype=torch.float64, device=dev).uniform_()
    cpy_dst = torch.empty_like(cpy_src)
    for _ in range(3):
        cpy_dst.copy_(cpy_src)
    torch.cuda.synchronize()
    # what a plain streaming read reaches on this GPU: the practical ceiling
    # under the published 8 TB/s
    part = torch.empty(65536, dtype=torch.float64, device=dev)
    stream_gbs = fa.stream_read(cpy_src.data_ptr(), cpy_src.numel(), part.data_ptr(), 10)
    del cpy_src, cpy_dst, part
    log("SpMV FP64 %.3f ms %.0f GB/s; FP32 %.3f ms %.0f GB/s; in-solve SpMV %.3f ms %.0f GB/s; stream read %.0f GB/s" % (
        ms64, gbs64, ms32, gbs32, ax_ms, achieved, stream_gbs))

    # DIAG-CG on the same system
    ia_d = ia.copy()
    ia_d[2] = 3
    ia_d[96] = 1
    ia_d[97] = 1
    x.zero_()
    torch.cuda.synchronize()
    rd = ctx.solve_device(b.data_ptr(), x.data_ptr(), ia_d, ra)
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rd = ctx.solve_device(b.data_ptr(), x.data_ptr(), ia_d, ra)
    t_diag = time.perf_counter() - t0
    log("DIAG-CG: iter=%d resid=%.3e time %.3fs" % (rd.iter, rd.resid, t_diag))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, iters, args.cpu_sample_s)
    del ctx, b, x, y, xr, x32, y32
    torch.cuda.empty_cache()
    net, train = train_net(args, dev) if (args.c5_n > 0 or args.c3_n > 0) and args.c5_variants > 0 else (None, None)
    torch.cuda.empty_cache()
    mixed = bench_mixed(args, dev, net) if args.c3_n > 0 else None
    torch.cuda.empty_cache()
    batch = bench_batch(args) if args.c4_nsys > 0 else None
    torch.cuda.empty_cache()
    learned = bench_learned(args, dev, net, train) if args.c5_n > 0 else None
    if net is not None:
        net.close()

    out = {
        "metric": "time to 1e-8 rel. residual: mixed+learned FCG vs plain CG; BCSR SpMV GB/s",
        "value": round(sec_per_solve, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_per_solve * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%d^3 HEX8 elastic cube, %d DOF, %d blocks; plain FP64 multicolour-SSOR CG to 1e-8"
                        % (args.n, 3 * m.N, m.N + m.NPL + m.NPU),
            "method": "CG", "precond": {1: "SSOR", 3: "DIAG"}.get(args.precond, args.precond),
            "ncolor": r0.ncolor, "iterations": iters, "rel_resid_true": true_resid,
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_spmv<kEpiDotX> (q = A p with p.q)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": SPMV_TRAFFIC_100 if args.n == 100 else None,
            "bytes_per_launch": bytes64, "avg_launch_ms": round(ax_ms, 4),
            "stream_read_GBs": round(stream_gbs, 1), "frac_of_stream_read": round(achieved / stream_gbs, 4),
        },
        "spmv": {"fp64_ms": round(ms64, 4), "fp64_GBs": round(gbs64, 1), "fp32_ms": round(ms32, 4),
                 "fp32_GBs": round(gbs32, 1)},
        "diag_cg": {"seconds": round(t_diag, 4), "iterations": rd.iter},
        "solve_parts": {"Tset": round(r0.Tset, 4), "Tsol": round(r0.Tsol, 4),
                        "time_Ax": round(r0.time_Ax, 4), "time_precond": round(r0.time_precond, 4),
                        "load_s": round(t_load, 3)},
        "cpu_baseline": cpu,
        "mixed": mixed,
        "batch": batch,
        "learned": learned,
    }
    print(json.dumps(out), flush=True)


# flexible CG variants: (name, Iarray 41, Iarray 43, Iarray 44); the two-level
# correction symmetric (coarse, SSOR, coarse; 44 = 0) or additive (44 = 1); the
# learned solve adds it additively
FCG_MODES = (("fp64_ssor_cg", 0, 0, 0), ("mixed_fcg", 1, 0, 0), ("mixed_coarse_fcg", 1, 1, 0),
             ("mixed_coarse_additive_fcg", 1, 1, 1), ("learned_coarse_fcg", 2, 1, 1))


def _run_modes(tag, fa, ctx, m, b, x, net, modes):
    """Each mode from x0 = 0 to 1e-8 with the system resident and a numeric
    preconditioner refresh (SSOR factors and, with Iarray 43, the coarse
    inverse) inside the timed solve."""
    import torch
    res, xs = {}, {}
    for mode, opt, coarse, additive in modes:
        if opt == 2 and net is None:
            continue
        ia = m.Iarray.copy()
        ia[96] = 1
        ia[97] = 0
        ia[40] = opt
        ia[42] = coarse
        ia[43] = additive
        x.zero_()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        r = ctx.solve_device(b.data_ptr(), x.data_ptr(), ia, m.Rarray)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        xs[mode] = x.cpu().numpy()
        m.X[:] = xs[mode]
        inf = ctx.info()
        res[mode] = {"seconds": round(wall, 4), "iterations": r.iter, "error": r.error, "Tset": round(r.Tset, 4),
                     "Tsol": round(r.Tsol, 4), "ms_per_iter": round(1e3 * r.Tsol / max(1, r.iter), 4),
                     "time_Ax": round(r.time_Ax, 4), "time_precond": round(r.time_precond, 4),
                     "rel_resid_true": fa.hecmw_rel_resid_L2(m)}
        if opt > 0:
            res[mode]["fp64_residual_updates"] = inf["nreliable"]
            res[mode]["spmv_fp64_values"] = bool(inf["spmv64"])
        if coarse:
            res[mode]["coarse_aggregates"] = inf["coarse"]
            res[mode]["coarse_form"] = "additive" if additive or opt == 2 else "symmetric"
        if opt == 2:
            res[mode]["fallback_iterations"] = inf["nfallback"]
            res[mode]["correction_dropped"] = bool(inf["learned_off"])
            res[mode]["probe_ratio"] = inf["probe_ratio"]
        log("%s %s: %s" % (tag, mode, res[mode]))
    ref_x = xs["fp64_ssor_cg"]
    base = res["fp64_ssor_cg"]["seconds"]
    for mode in res:
        res[mode]["rel_err_vs_fp64"] = float(np.linalg.norm(xs[mode] - ref_x) / np.linalg.norm(ref_x))
        if mode != "fp64_ssor_cg":
            res[mode]["speedup_vs_fp64"] = round(base / res[mode]["seconds"], 3)
    return res


def bench_mixed(args, dev, net):
    """Heterogeneous cube (per-element E log-uniform in [1e3, 1e6], nu uniform
    in [0.2, 0.4], seed 0): FP64 multicolour-SSOR CG against the flexible CG
    with FP32 SpMV/SSOR/vectors and FP64 residual replacement, alone, with the
    two-level (rigid-body-mode coarse space) correction, and with the learned
    correction on top."""
    import torch
    import frontistr_amd as fa
    n = args.c3_n
    t0 = time.time()
    m = fa.gen_hex8(n, kind="lognormal", seed=0)
    log("C3: generated %d^3 lognormal HEX8, %d DOF in %.1fs" % (n, 3 * m.N, time.time() - t0))
    m.set_iter(20000)
    m.set_precond(1)
    m.set_i(m.I_TIMELOG, 1)
    m.set_coarse(1)  # the aggregation is symbolic set-up, built at load like the colouring
    ctx = fa.DeviceContext()
    ctx.set_coords(fa.hex8_coords(n))
    ctx.set_net(net)
    t0 = time.time()
    ctx.load(m)
    t_load = time.time() - t0
    b = torch.from_numpy(m.B).to(dev)
    x = torch.zeros(3 * m.NP, dtype=torch.float64, device=dev)
    res = _run_modes("C3", fa, ctx, m, b, x, net, FCG_MODES)
    ctx.set_net(None)
    ctx.close()
    out = {"workload": "%d^3 HEX8, lognormal E in [1e3,1e6], nu in [0.2,0.4], seed 0, %d DOF, %d blocks"
                       % (n, 3 * m.N, m.N + m.NPL + m.NPU), "load_s": round(t_load, 2)}
    out.update(res)
    return out


def _inclusions(fa, n, seed):
    m = fa.gen_hex8(n, kind="inclusions", seed=seed, E=1000.0, nu=0.3, contrast=1e4, rmin=2, rmax=6, vfrac=0.15)
    m.set_iter(30000)
    m.set_precond(1)
    m.set_i(m.I_TIMELOG, 1)
    return m


def train_net(args, dev):
    """The learned correction's network, trained on residuals the solver
    harvests from FP64 SSOR-CG runs on seeded variants of the two-phase cube
    (24^3, the inclusion statistics of the learned-correction workload, 7 load
    cases x 10 iteration counts each), 4 epochs on the GPU."""
    import torch
    import frontistr_amd as fa
    t0 = time.time()
    net = fa.LearnedNet(seed=0)
    keep = []
    rng = np.random.default_rng(123)
    nv = args.c5_variants
    count = 0
    for v in range(nv):
        m = _inclusions(fa, 24, 1000 + v)
        ctx = fa.DeviceContext()
        ctx.load(m)
        top = np.nonzero(m.B)[0]
        for lc in range(7):
            # the pressure load, then random nodal loads on the loaded face
            bvec = m.B.copy()
            if lc > 0:
                bvec[:] = 0.0
                bvec[top] = rng.standard_normal(top.size) * np.abs(m.B[top]).mean()
            b = torch.from_numpy(bvec).to(dev)
            count = fa.harvest_residuals(net, ctx, m, b)
        keep.append(ctx)
    t_harvest = time.time() - t0
    t0 = time.time()
    loss = net.train(epochs=4, lr=2e-3, seed=0)
    t_train = time.time() - t0
    log("trained: %d residuals from %d variants in %.1fs; training %.1fs, energy %s" % (
        count, nv, t_harvest, t_train, np.round(loss, 4)))
    net.clear_samples()
    for c in keep:
        c.close()
    return net, {"samples": count, "variants": nv, "harvest_s": round(t_harvest, 2), "train_s": round(t_train, 2),
                 "energy_plain_ssor": loss[0], "energy_per_epoch": [float(v) for v in loss[1:]]}


def bench_learned(args, dev, net, train):
    """Two-phase inclusion cube (E contrast 1e4, 15% volume, radii 2-6
    elements): FP64 SSOR-CG, the flexible CG with FP32 operators, with the
    learned correction, and both with the two-level correction."""
    import torch
    import frontistr_amd as fa
    m = _inclusions(fa, args.c5_n, 7)
    m.set_coarse(1)
    ctx = fa.DeviceContext()
    ctx.set_coords(fa.hex8_coords(args.c5_n))
    ctx.load(m)
    b = torch.from_numpy(m.B).to(dev)
    ctx.set_net(net)
    x = torch.zeros(3 * m.NP, dtype=torch.float64, device=dev)
    modes = FCG_MODES[:2] + (("learned_fcg", 2, 0, 0),) + FCG_MODES[2:]
    res = _run_modes("C5", fa, ctx, m, b, x, net, modes)
    ctx.set_net(None)
    ctx.close()
    out = {"workload": "%d^3 HEX8, two-phase inclusions (E 1e3 / 1e7, 15%% volume, radii 2-6 elements, seed 7), "
                       "%d DOF; network trained on %s residuals harvested from SSOR-CG on %s seeded 24^3 variants"
                       % (args.c5_n, 3 * m.N, train["samples"] if train else 0, train["variants"] if train else 0),
           "train": train}
    out.update(res)
    return out


def bench_batch(args):
    """48 TET10 boxes of 5k-600k DOF (log-uniform, seeded) x 16 right-hand
    sides, DIAG flexible CG with FP32 SpMM and FP64 reductions, one
    block-diagonal launch per kernel, graph-captured iterations."""
    import frontistr_amd as fa
    t0 = time.time()
    mats, Bs, specs = fa.gen_tet10_batch(nsys=args.c4_nsys, nrhs=16, seed=0)
    dof = sum(sp[3] for sp in specs)
    log("C4: generated %d TET10 systems, %d DOF total in %.1fs" % (len(mats), dof, time.time() - t0))
    m0 = fa.gen_hex8(1)
    m0.set_iter(20000)
    m0.set_precond(3)
    t0 = time.time()
    bs = fa.BatchSolver(mats, Bs)
    t_load = time.time() - t0
    info = bs.info()
    bs.solve(m0.Iarray, m0.Rarray)  # warm-up: graph capture
    t0 = time.perf_counter()
    it, er, rs, tm = bs.solve(m0.Iarray, m0.Rarray)
    wall = time.perf_counter() - t0
    blocks = info["blocks"]
    spmm_bytes = 9 * 4 * blocks + 4 * blocks + 4 * (info["rows"] + 1) + 16 * 3 * 4 * 2 * info["rows"]
    spmm_ms = info["variant_us"][info["variant"]] / 1e3
    log("C4: %.3fs, iterations %d..%d, max resid %.2e, errors %d, SpMM %s %.3f ms, FP64 residual SpMM %s %.3f ms" % (
        wall, it.min(), it.max(), rs.max(), int((er != 0).sum()), info["variant"], spmm_ms, info["variant64"],
        info["variant64_us"][info["variant64"]] / 1e3))
    bs.close()
    # the same instances one at a time through the single-system path (first
    # right-hand side of every system, x 16): FP64 DIAG-CG (the reference's
    # method) and the single-system mixed DIAG flexible CG
    import torch
    seq = {}
    for name, opt in (("fp64_diag_cg", 0), ("mixed_diag_fcg", 1)):
        total, iters = 0.0, 0
        for s_, mat in enumerate(mats):
            mm = mat.copy()
            mm.set_iter(20000)
            mm.set_precond(3)
            mm.set_i(41, opt)
            ctx = fa.DeviceContext()
            ctx.load(mm)
            b = torch.from_numpy(np.ascontiguousarray(Bs[s_][0])).to("cuda:0")
            xv = torch.zeros_like(b)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            r = ctx.solve_device(b.data_ptr(), xv.data_ptr(), mm.Iarray, mm.Rarray)
            torch.cuda.synchronize()
            total += time.perf_counter() - t1
            iters += r.iter
            ctx.close()
        seq[name] = {"seconds_first_rhs": round(total, 3), "iterations_first_rhs": iters,
                     "seconds_all_rhs_est": round(16 * total, 2)}
        log("C4 one at a time, %s: %.2fs for the 48 first right-hand sides (x16: %.1fs)" % (name, total, 16 * total))
    return {"workload": "%d TET10 boxes (5k-600k DOF log-uniform, aspect 0.5-2, seed 0) x 16 RHS, %d DOF, %d blocks"
                        % (len(mats), dof, blocks),
            "seconds": round(wall, 4), "instances": int(it.size), "converged": int((er == 0).sum()),
            "iterations_min": int(it.min()), "iterations_max": int(it.max()), "max_rel_resid": float(rs.max()),
            "one_at_a_time": seq,
            "speedup_vs_one_at_a_time_fp64": round(seq["fp64_diag_cg"]["seconds_all_rhs_est"] / wall, 2),
            "spmm_variant": info["variant"], "spmm_ms": spmm_ms,
            "spmm64_variant": info["variant64"], "spmm64_res_ms": info["variant64_us"][info["variant64"]] / 1e3,
            "spmm_GBs": round(spmm_bytes / (spmm_ms * 1e-3) / 1e9, 1), "load_s": round(t_load, 2)}


def cpu_baseline(m, gpu_iters, budget_s):
    """The reference's CPU solver (oracle/_ref) on the same system, all host
    cores, on a bounded sample: a solve capped at a number of iterations that
    fits the budget; time to 1e-8 = setup + per-iteration time x the number of
    iterations the full solve needs (the GPU's count: same ordering)."""
    try:
        from oracle import ref
    except Exception as e:  # pragma: no cover
        log("cpu baseline unavailable: %s" % e)
        return None
    if not ref.available():
        log("cpu baseline unavailable: oracle/_ref not built")
        return None
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(2, min(cores, 16))  # the box's CPU share for one GPU is 16 threads
    # probe: 5 iterations to size the sample
    ia = m.Iarray.copy()
    ia[0] = 5
    ia[21] = 0
    t0 = time.time()
    _, info = ref.solve_krylov(m, nthreads=cores, iarray=ia)
    probe = time.time() - t0
    per_it = max(info["Tsol"] / max(1, min(info["iter"], 5)), 1e-6)
    n_it = int(max(10, min(gpu_iters, (budget_s - info["Tset"]) / per_it)))
    ia[0] = n_it
    t0 = time.time()
    _, info = ref.solve_krylov(m, nthreads=cores, iarray=ia)
    wall = time.time() - t0
    ran = min(info["iter"], n_it)
    per_it = info["Tsol"] / ran
    est = info["Tset"] + per_it * gpu_iters
    log("CPU reference: %d threads, %d iterations in %.2fs (Tset %.2fs, %.4fs/iter) -> est %.2fs to 1e-8" % (
        cores, ran, wall, info["Tset"], per_it, est))
    return {"value": round(est, 3), "unit": "s", "cores": cores, "kind": "reference",
            "sample": "reference hecmw_solve_CG (flang -O2 -fopenmp, HECMW_SERIAL) on the same 100^3 system, "
                      "%d of %d iterations run (%.1fs wall incl. setup %.2fs); value = Tset + %.4f s/iter x %d"
                      % (ran, gpu_iters, wall, info["Tset"], per_it, gpu_iters),
            "spmv_s_per_call": round(info["time_Ax"] / max(1, ran + 1), 5)}


if __name__ == "__main__":
    main()
