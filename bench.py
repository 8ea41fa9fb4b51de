"""Benchmark of the MI355X QUDOS single-particle solve.

One step = one solve of the truncated-cone In0.8Ga0.2As dot at 19x19x15 plane waves (N_H = 43320;
--config selects another): Hamiltonian assembly from device-resident FP64 Delta-G tables + the 40
eigenpairs of the gap window, with the FP32 two-stage Householder reduction (dense -> band on
MFMA, band -> tridiagonal by bulge chasing) and FP64 refinement.  Beside the timed steps: the same
GPU's plain FP64 solve (speed ratio, eigenvalue error), the Coulomb elements of the 4 highest hole
and 4 lowest electron states, the gradient of (E_e1 - E_h1) + J_eh on the 11x11x9 lens against its
forward pass, and the reference's CPU path on this host.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s/solve (H build+40 eigpairs) FP32+refine vs FP64 @≤1µeV; grad/fwd time"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFS = 157.3   # MI355X dense FP32 matrix-core peak
FP64_PEAK_TFS = 78.6         # MI355X dense FP64 (matrix) peak


def cfg_paths(name):
    d = os.path.join(ROOT, "qudos_amd", "configs", name)
    return os.path.join(d, "parameters.in"), os.path.join(d, "supercell.in")


def _host_cores():
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()


def cpu_baseline(name, solve_name="lens_11x11x9", columns_per_band=24):
    """The reference's CPU path on this host's cores, on a bounded sample of the workload: the C
    restatement of its per-element H fill (hamiltonian_8_band.c loop structure, OpenMP) on the
    same plane-wave column sample of all 8 band columns of the workload's H, scaled by the
    lower-triangle element count; Intel MKL zheevr (RANGE='I', 40 pairs, as
    diagonalise_hamiltonian.c:55-77 calls it) timed on the smaller solve_name problem and scaled
    by N_H^3."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_problem
    from oracle import qd_cpu
    from oracle import qd_oracle as qo
    ncores = _host_cores()
    pp, sp = cfg_paths(name)
    p, sh = oracle_problem(pp, sp)
    t = qo.coefficient_tables(p, sh)
    cp = qd_cpu.CpuProblem(t)
    nh, ng = cp.NH, cp.NH // 8
    stride = max(1, ng // columns_per_band)
    cols = np.array([b * ng + m for b in range(8) for m in range(0, ng, stride)], dtype=np.int32)
    t0 = time.perf_counter()
    cp.hamiltonian_columns(cols, threads=ncores)
    t_fill = time.perf_counter() - t0
    sampled = float(np.sum(nh - cols))
    fill = t_fill * (nh * (nh + 1) / 2) / sampled
    ps, shs = oracle_problem(*cfg_paths(solve_name))
    ts = qo.coefficient_tables(ps, shs)
    cps = qd_cpu.CpuProblem(ts)
    Hs = cps.hamiltonian(threads=ncores)
    il, iu = qo.window(ps)
    t0 = time.perf_counter()
    qd_cpu.zheevr(Hs, il, iu, threads=ncores)
    t_eig = time.perf_counter() - t0
    eig = t_eig * (nh / cps.NH) ** 3
    return {"value": round(fill + eig, 1), "unit": "s/solve", "cores": ncores, "kind": "port",
            "library": f"Intel MKL zheevr ({qd_cpu.mkl_path()}), GNU OpenMP threads",
            "sample": f"C restatement of the reference H fill on {len(cols)} columns of {name} (8 band columns x "
                      f"{len(cols) // 8} plane waves, {sampled:.3g} entries): {t_fill:.2f} s, scaled by the lower-"
                      f"triangle count to {fill:.1f} s; MKL zheevr of the 40-pair window of {solve_name} (N_H="
                      f"{cps.NH}): {t_eig:.2f} s, scaled by N_H^3 to {eig:.1f} s (extrapolated); host threads "
                      f"{ncores}"}


def cpu_full(name):
    """One unsampled run of the reference's CPU path (C H fill + MKL zheevr) on this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_problem
    from oracle import qd_cpu
    from oracle import qd_oracle as qo
    ncores = _host_cores()
    p, sh = oracle_problem(*cfg_paths(name))
    t = qo.coefficient_tables(p, sh)
    cp = qd_cpu.CpuProblem(t)
    print(f"[cpu-full] {name}: N_H={cp.NH}, {ncores} threads", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    H = cp.hamiltonian(threads=ncores)
    t_fill = time.perf_counter() - t0
    print(f"[cpu-full] H fill {t_fill:.2f} s", file=sys.stderr, flush=True)
    il, iu = qo.window(p)
    t0 = time.perf_counter()
    w, _ = qd_cpu.zheevr(H, il, iu, threads=ncores)
    t_eig = time.perf_counter() - t0
    cpu = open("/proc/cpuinfo").read().split("model name")[1].split(":")[1].split("\n")[0].strip() \
        if os.path.exists("/proc/cpuinfo") else "unknown"
    print(json.dumps({"config": name, "n_h": cp.NH, "h_fill_s": round(t_fill, 2), "zheevr_s": round(t_eig, 2),
                      "total_s": round(t_fill + t_eig, 2), "threads": ncores, "cpu": cpu,
                      "library": f"Intel MKL zheevr ({qd_cpu.mkl_path()}), GNU OpenMP threads",
                      "E_h1": float(w[2 * p.N_vb - 1]), "E_e1": float(w[2 * p.N_vb])}))


def copy_calibration():
    """Four device-to-device copies of 2^28 float32 (1 GiB read + 1 GiB written each): the
    known-size kernel that scales FETCH_SIZE / WRITE_SIZE counter readings to bytes."""
    import torch
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(4):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": x.numel() * 4, "copy_calibration_bytes_written": y.numel() * 4,
                      "copies": 4}))


def draw_theta(seed):
    """Seeded lens parameters: height 4-6 nm, base radius 8-12 nm, In fraction 0.6-1.0, field of
    0-20 kV/cm in a uniformly random direction."""
    rng = np.random.default_rng(seed)
    h, R, f = rng.uniform(4, 6), rng.uniform(8, 12), rng.uniform(0.6, 1.0)
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    return np.array([h, R, f, *(rng.uniform(0, 20) * u)])


def grad_timing(name, precision, device_tables=False):
    """Forward (tables + H build + solve + J_eh Coulomb) and gradient (adjoint + dT/dtheta) times
    at one parameter draw; device_tables: both table passes on the device."""
    import qudos_amd as q
    from qudos_amd.autograd import DotProblem
    pp, sp = cfg_paths(name)
    prob = DotProblem(pp, sp, precision=precision, device_tables=device_tables)
    theta = draw_theta(0)
    prob.forward(theta)           # warm-up (allocations, code objects)
    prob.gradient()
    val, info = prob.forward(theta)
    t_fwd = prob.timings["forward_s"] + prob.timings["tables_s"]
    g, ginfo = prob.gradient()
    t_grad = ginfo["adjoint_s"] + ginfo["table_derivatives_s"]
    prob.close()
    prob.solver.close()
    return {"config": name, "n_h": 8 * prob.inp.N[0] * prob.inp.N[1] * prob.inp.N[2],
            "precision": "fp32+refine" if precision == q.QD_FP32_REFINE else "fp64",
            "theta": [round(x, 6) for x in theta.tolist()], "L_eV": val, "J_eh_eV": info["J_eh"],
            "grad": g.tolist(), "forward_s": round(t_fwd, 4), "gradient_s": round(t_grad, 4),
            "grad_over_forward": round(t_grad / t_fwd, 3),
            "forward_stages_s": {"tables": round(prob.timings["tables_s"], 4),
                                 "device": round(prob.timings["forward_s"], 4)},
            "gradient_stages_s": {"adjoint": round(ginfo["adjoint_s"], 4),
                                  "table_derivatives": round(ginfo["table_derivatives_s"], 4)},
            "adjoint_only_over_device_forward": round(ginfo["adjoint_s"] / prob.timings["forward_s"], 3),
            "adjoint_stages_s": {k: round(ginfo[k], 4) for k in ("t_density", "t_sternheimer", "t_bin")},
            "sternheimer_iters": ginfo["sternheimer_iters"],
            "table_derivatives": ginfo.get("table_derivatives", "")}


def grad_check(name, ndraws):
    import qudos_amd as q
    from qudos_amd.autograd import DotProblem
    pp, sp = cfg_paths(name)
    solver = q.Solver(0)
    p64 = DotProblem(pp, sp, precision=q.QD_FP64, solver=solver)
    p32 = DotProblem(pp, sp, precision=q.QD_FP32_REFINE, solver=solver)
    rows = []
    for seed in range(ndraws):
        theta = draw_theta(seed)
        p32.forward(theta)
        g32, i32 = p32.gradient()
        val, _ = p64.forward(theta)
        g64, i64 = p64.gradient()
        fd = np.zeros(6)
        for j in range(6):
            hstep = 1e-4 * abs(theta[j]) if j < 3 else 1e-4 * 20.0
            tp, tm = theta.copy(), theta.copy()
            tp[j] += hstep
            tm[j] -= hstep
            fd[j] = (p64.forward(tp)[0] - p64.forward(tm)[0]) / (2 * hstep)
        scale = np.abs(fd).max()
        rows.append({"seed": seed, "theta": theta.tolist(), "L": val, "grad_fp64": g64.tolist(),
                     "grad_fp32_refine": g32.tolist(), "fd": fd.tolist(),
                     "max_err_fp64_vs_fd_rel": float(np.abs(g64 - fd).max() / scale),
                     "max_err_fp32_vs_fp64_rel": float(np.abs(g32 - g64).max() / np.abs(g64).max())})
        print(f"[grad-check] seed {seed}: fp64 vs fd {rows[-1]['max_err_fp64_vs_fd_rel']:.2e}, "
              f"fp32+refine vs fp64 {rows[-1]['max_err_fp32_vs_fp64_rel']:.2e}", file=sys.stderr, flush=True)
    p64.close()
    p32.close()
    solver.close()
    print(json.dumps({"config": name, "draws": ndraws, "params": ["height_nm", "radius_nm", "in_fraction",
                                                                   "Fx_kV_cm", "Fy_kV_cm", "Fz_kV_cm"],
                      "max_err_fp64_vs_fd_rel": max(r["max_err_fp64_vs_fd_rel"] for r in rows),
                      "max_err_fp32_vs_fp64_rel": max(r["max_err_fp32_vs_fp64_rel"] for r in rows),
                      "rows": rows}))


def device_tables_timing(q, s, inp, host_tables, reps=3):
    """Wall time of qd_tables_build_device (shape transforms, fields, convolutions and tables on the
    device, synchronised; the first call is a warm-up) and its largest difference from the host tables,
    per table relative to that table's largest entry."""
    import numpy as np

    def arr(t):
        nT = t.NT[0] * t.NT[1] * t.NT[2]
        flat = np.ctypeslib.as_array(t.tab, shape=(q.QD_NTAB * nT * 2,)).copy()
        return flat.view(np.complex128).reshape(q.QD_NTAB, nT)

    H = arr(host_tables)
    times, dev = [], None
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        dev = s.build_tables(inp)
        times.append(time.perf_counter() - t0)
    s.upload(host_tables)   # the solves above used the host tables; keep them resident
    err = float((np.abs(arr(dev) - H).max(axis=1) / np.maximum(np.abs(H).max(axis=1), 1e-300)).max())
    print(f"[bench] device tables: {min(times[1:]):.4f} s, max rel diff {err:.2e}", file=sys.stderr, flush=True)
    return {"seconds": round(min(times[1:]), 4), "max_rel_diff_vs_host": err, "reps": reps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cone_19x19x15")
    ap.add_argument("--cpu-solve-config", default="lens_11x11x9",
                    help="problem on which the CPU baseline times MKL zheevr (scaled by N_H^3)")
    ap.add_argument("--cpu-full", default="",
                    help="only run the reference's CPU path (C H fill + MKL zheevr) unsampled on this config")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--reduction", choices=["two-stage", "one-stage"], default="two-stage",
                    help="tridiagonal reduction: dense->band->tridiagonal (default) or one-stage Householder")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="do not bracket the reduction's dominant kernel with HIP events (for hardware-counter runs)")
    ap.add_argument("--coulomb", default="4,4",
                    help="NH,NE: after the timed solves, time qd_coulomb on the NH highest hole and NE lowest "
                         "electron states of the window (e.g. 4,4)")
    ap.add_argument("--grad-config", default="lens_11x11x9",
                    help="configuration of the gradient timing (L = (E_e1 - E_h1) + J_eh, adjoint)")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--device-tables", action="store_true",
                    help="also build the tables on the device (qd_tables_build_device, outside the timed solves): "
                         "its time and its largest difference from the host tables join the line, and the gradient "
                         "is timed a second time with both table passes on the device")
    ap.add_argument("--grad-check", type=int, default=0,
                    help="only run N seeded parameter draws of the gradient check (adjoint FP32+refine and "
                         "FP64 vs central differences of L) and print a JSON summary")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 plain 1 GiB device copies (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs a single GPU (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.cpu_full:
        cpu_full(args.cpu_full)
        return
    if args.grad_check:
        grad_check(args.grad_config, args.grad_check)
        return
    import qudos_amd as q
    pp, sp = cfg_paths(args.config)
    inp = q.read_inputs(pp, sp)
    t0 = time.perf_counter()
    tables = q.build_tables(inp)
    t_tables = time.perf_counter() - t0
    il, iu = q.window(inp)
    k = iu - il + 1
    nh = 8 * inp.N[0] * inp.N[1] * inp.N[2]
    s = q.Solver(0)
    if args.reduction == "one-stage":
        s.set_reduction(q.QD_REDUCE_ONE_STAGE)
    s.upload(tables)
    print(f"[bench] {args.config}: N_H={nh}, window {il}..{iu}, tables {t_tables:.2f} s", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        st = s.solve_device(il, iu, q.QD_FP32_REFINE)
        print(f"[bench] warmup {i}: {st['t_total']:.3f} s", file=sys.stderr, flush=True)
    s.profile(not args.no_kernel_timing)
    times, stats = [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        st = s.solve_device(il, iu, q.QD_FP32_REFINE)   # returns after the device work completes
        times.append(time.perf_counter() - t0)
        stats.append(st)
        print(f"[bench] step {i}: {times[-1]:.3f} s", file=sys.stderr, flush=True)
    nprof, t_kern, w_kern = s.profile_read()
    s.profile(False)
    w32, _ = s.fetch(k, vectors=False)
    coul = None
    if args.coulomb:
        nhc, nec = (int(x) for x in args.coulomb.split(","))
        s.coulomb(inp, tables, nhc, nec)            # warm-up (allocations, code objects)
        runs = [s.coulomb(inp, tables, nhc, nec)[1] for _ in range(3)]
        best = min(runs, key=lambda r: r["t_total"])
        coul = {"states": f"{nhc} holes + {nec} electrons", "pairs": best["n_pairs"], "fft_grid": list(best["fft"]),
                "nq": best["nq"], "t_pairs_s": round(best["t_pairs"], 6), "t_contract_s": round(best["t_contract"], 6),
                "t_total_s": round(best["t_total"], 6)}
        print(f"[bench] coulomb: {coul}", file=sys.stderr, flush=True)
    extra = {}
    if not args.no_fp64:
        t0 = time.perf_counter()
        st64 = s.solve_device(il, iu, q.QD_FP64)
        t64 = time.perf_counter() - t0
        w64, _ = s.fetch(k, vectors=False)
        print(f"[bench] fp64 solve: {t64:.3f} s", file=sys.stderr, flush=True)
        extra = {"fp64_s_per_solve": round(t64, 4), "fp32_refine_speedup_vs_fp64": round(t64 / np.mean(times), 3),
                 "max_abs_dE_vs_fp64_eV": float(np.max(np.abs(w32 - w64))),
                 "fp64_stages_s": {kk: round(v, 4) for kk, v in st64.items() if kk.startswith("t_")}}
    s.close()
    mean = float(np.mean(times))
    last = stats[-1]
    # HBM bytes per k_hemm_f32 launch from rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this bench on the
    # same workload (profiles/k_hemm_f32_traffic_<config>.json); null for workloads without such passes
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", f"k_hemm_f32_traffic_{args.config}.json")
    if args.reduction == "two-stage" and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
        traffic_src = os.path.relpath(tpath, ROOT)
    chase = None
    if args.reduction == "two-stage":
        # k_chase: band (width b = 32) -> tridiagonal, ~n^2 / (2 b) steps of ~6 b^2 complex
        # multiply-adds (two-sided update of a b x b Hermitian block, one-sided updates of the b x b
        # block below it, the next bulge column): 24 b n^2 flops, FP64 vector arithmetic
        t_chase = float(np.mean([st["t_reduce"] - st["t_band"] for st in stats]))
        fl = 24.0 * 32 * nh * nh
        chase = {"kernel": "k_chase (band -> tridiagonal bulge chasing, FP64, one launch per solve)",
                 "bound": "latency (chain of dependent steps)", "launch_s": round(t_chase, 4),
                 "algorithmic_flops": fl, "achieved": round(fl / t_chase / 1e12, 3), "peak": FP64_PEAK_TFS,
                 "unit": "TFLOP/s", "frac": round(fl / t_chase / 1e12 / FP64_PEAK_TFS, 4)}
    if args.reduction == "two-stage":
        # k_hemm: X = A22 V for an m x m trailing matrix and a 32-column panel, 8 m^2 32 flops (the
        # conventional complex count); the kernel forms each complex product from three real MFMA
        # products, so the matrix cores execute 3/4 of that count
        achieved = w_kern / t_kern / 1e12 if t_kern > 0 else None
        roof = {"bound": "mfma", "kernel": "k_hemm_f32 (stage-1 trailing-matrix product A22 V, FP32 32x32x2 MFMA)",
                "achieved": round(achieved, 2) if achieved else None, "peak": FP32_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_MFMA_PEAK_TFS, 4) if achieved else None, "traffic": traffic,
                "traffic_source": traffic_src, "launches_timed": nprof, "mean_launch_s": t_kern,
                "mean_algorithmic_flops": w_kern,
                "mfma_executed_tflops": round(0.75 * achieved, 2) if achieved else None,
                "note": "complex products as three real MFMA products (Gauss): executed MFMA flops = 3/4 of the algorithmic count"}
    else:
        achieved = w_kern / t_kern / 1e9 if t_kern > 0 else None
        roof = {"bound": "hbm", "kernel": "k_hemv<float> (trailing-matrix Hermitian matvec)",
                "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": None,
                "launches_timed": nprof, "mean_launch_s": t_kern, "mean_algorithmic_bytes": w_kern}
    out = {
        "metric": METRIC, "value": round(mean, 4), "unit": "s/solve", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(mean * 1e3, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex64 reduction + float64 refinement", "data": "synthetic",
        "config": {"workload": f"{args.config.split('_')[0]} InAs/GaAs quantum dot, {args.config.split('_')[1]} "
                               f"plane waves (N_H={nh}), {k} eigenpairs of the gap window, FP32 "
                               f"{'two-stage' if args.reduction == 'two-stage' else 'one-stage'} Householder "
                               f"reduction + FP64 refinement",
                   "n_h": nh, "eigenpairs": k, "reduction": args.reduction,
                   "config_dir": f"qudos_amd/configs/{args.config}"},
        "stages_s": {kk: round(v, 4) for kk, v in last.items() if kk.startswith("t_")},
        "refine_iters": last["refine_iters"], "max_residual_rel": last["max_residual"],
        "max_orth": last["max_orth"], "tables_host_s": round(t_tables, 3),
        "roofline": roof,
    }
    if chase:
        out["roofline_chase"] = chase
    if args.device_tables:
        out["tables_device"] = device_tables_timing(q, s, inp, tables)
    out.update(extra)
    if coul:
        out["coulomb"] = coul
    if not args.no_grad:
        out["gradient"] = grad_timing(args.grad_config, q.QD_FP32_REFINE)
        print(f"[bench] gradient: {out['gradient']['grad_over_forward']} x forward", file=sys.stderr, flush=True)
        if args.device_tables:
            out["gradient_device_tables"] = grad_timing(args.grad_config, q.QD_FP32_REFINE, device_tables=True)
            print(f"[bench] gradient, tables on the device: {out['gradient_device_tables']['grad_over_forward']} x "
                  f"forward", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_solve_config)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
