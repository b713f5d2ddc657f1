"""Benchmark: CSD voxels/s on one MI355X (the dwi2fod csd / msmt_csd hot path).

Default workload (--workload single_shell_brain): the whole-brain-size single-shell phantom
(128x128x72, ellipsoidal mask with 646,128 voxels, 8 b=0 + 64 b=3000 volumes, lmax 8, 300
constraint directions, reference defaults), solved plain (cold start) and with the learned warm
start.  A step is one solve of every masked voxel with the signals already resident in HBM.

Other workloads: --workload super_resolved (same mask, 5 b=0 + 45 b=3000, lmax 12: 91
coefficients from 45 measurements), --workload multi_shell (same mask, b=0/1000/2000/3000 with
10/30/60/90 volumes, three tissues, the ICLS solver behind msmt_csd) and --workload peaks
(sh2peaks on the whole-brain FODs the GPU CSD writes: 60 search directions, 3 peaks).

Prints one JSON line with the throughput, the roofline position of the solver kernel (FP64
algorithmic FLOPs of the reference algorithm, counted from per-voxel iteration statistics the
kernel accumulates) and a CPU baseline: the oracle (this project's C restatement of the
reference algorithm) on the host's cores over a bounded sample of the same voxels.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PROTOCOL_NET = os.path.join(ROOT, "tests", "golden", "warm_net_single_shell_brain.npz")
sys.path.insert(0, ROOT)

METRIC = "dwi2fod voxels/s on 1 MI355X: plain vs BF16+NN warm start at ≤1e-4 SH err"
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector)
CPU_SAMPLE = {"single_shell_brain": 131072, "super_resolved": 16384, "multi_shell": 4096, "peaks": 65536,
              "response": 8192, "gt_energy": 4096, "fixel": 8192, "tournier": 16384}


def csd_flops(n_sh, n_init, n_dw, n_hr, stats):
    """Algorithmic FLOPs of the reference CSD for the counted work.
    stats = (voxels, solve iterations, sum of |neg| over solves, converged voxels)."""
    v, solves, ksum, conv = [float(x) for x in stats[:4]]
    n, N, R, nr = float(n_sh), float(n_dw), float(n_hr), float(n_init)
    set_ = v * (2 * nr * N + 2 * n * N)
    gemv = (solves + conv) * 2 * R * n
    syrk = ksum * n * (n + 1)
    llt = solves * (n ** 3 / 3.0 + 2 * n * n)
    return dict(total=set_ + gemv + syrk + llt, set=set_, amplitudes=gemv, normal_matrix=syrk, solve=llt)


def icls_flops(n, nmeas, ncon, stats):
    """Algorithmic FLOPs of the reference ICLS solver for the counted work.  stats = (voxels, outer
    iterations, inner solves, sum k, sum k^2, sum k^3 over inner solves, sum k over constraint
    updates, ...): per voxel 2n(Nmeas + Ncon) + n^2, per inner solve k(k+1)n + k^3/3 + 2k^2, per
    outer iteration 2nk + 2n Ncon."""
    v, outer, inner, s1, s2, s3, sc = [float(x) for x in stats[:7]]
    n, m, c = float(n), float(nmeas), float(ncon)
    setup = v * (2 * n * m + 2 * n * c + n * n)
    solves = n * (s2 + s1) + s3 / 3.0 + 2 * s2
    updates = 2 * n * sc + 2 * n * c * outer
    return dict(total=setup + solves + updates, setup=setup, solves=solves, updates=updates)


def load_traffic(kernel_tag, leg=None):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration), when a profile summary has been recorded for this kernel (per leg —
    "plain" or "warm" — where the summary holds both instantiations)."""
    path = os.path.join(ROOT, "profiles", kernel_tag + "_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        if leg is not None and leg in d:
            d = d[leg]
        return d.get("bytes_per_launch")
    return None


def threads():
    t = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return max(1, min(t, os.cpu_count() or 1))


def host_info():
    """CPU model, logical CPUs, physical cores (distinct (package, core) pairs) and the CPUs this
    process may run on."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "physical_cores": len(cores) or None,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def median_rate(run, n, reps=3):
    """Median voxels/s over reps runs of run() on n voxels; also returns the last run's output."""
    rates, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = run()
        rates.append(n / (time.perf_counter() - t0))
    return float(np.median(rates)), [float(r) for r in rates], out


def cpu_baseline_csd(ph, lmax, gpu_fod, sample):
    """The oracle's CSD on the host cores: the -O3 -march=native build (the value) and the generic
    x86-64 -O3 build beside it, median of 3 runs each, with the reference's x-line scheduling;
    plus single-thread GFLOP/s of its rank-k update and LLT at the workload's n."""
    from oracle import orc
    sample = min(sample, ph.nvox)
    sample -= sample % 128
    sub = np.ascontiguousarray(ph.dw[:sample])
    nt = threads()
    res = {}
    for name, generic in (("native", False), ("generic", True)):
        o = orc.CsdOracle(ph.grad, ph.response, lmax=lmax, generic=generic)
        med, runs, (fod, _) = median_rate(lambda: o.run_volume(sub, (128, sample // 128, 1), nthreads=nt), sample)
        res[name] = (o, med, runs, fod)
    o, med, runs, fod = res["native"]
    ref = fod.astype(np.float64)
    g = gpu_fod[:sample].astype(np.float64)
    rel = np.linalg.norm(g - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    single = [o.time_syrk_llt(k) for k in (60, 144)]
    return dict(value=med, unit="voxels/s", cores=nt, kind="port",
                sample="first %d masked voxels of the same phantom; oracle C restatement of CSD (FP64, "
                       "reference x-line scheduling, %d threads), median of 3 runs" % (sample, nt),
                runs=runs, build="gcc -O3 -march=native",
                generic_build={"build": "gcc -O3 (generic x86-64)", "value": res["generic"][1],
                               "runs": res["generic"][2]},
                single_thread_kernels=single, host=host_info()), rel, ref


def cpu_baseline_icls(ph, gpu_x, sample):
    from oracle import orc
    sample = min(sample, ph.nvox)
    sub = np.ascontiguousarray(ph.dw[:sample])
    nt = threads()
    res = {}
    for name, generic in (("native", False), ("generic", True)):
        o = orc.MsmtOracle(ph.grad, ph.response, generic=generic)
        med, runs, (x, _) = median_rate(lambda: o.run_batch(sub, nthreads=nt), sample)
        res[name] = (med, runs, x)
    med, runs, x = res["native"]
    ref = x.astype(np.float64)
    g = gpu_x[:sample].astype(np.float64)
    rel = np.linalg.norm(g - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    return dict(value=med, unit="voxels/s", cores=nt, kind="port",
                sample="first %d masked voxels of the same phantom; oracle C restatement of "
                       "ICLS::Solver (FP64, %d threads), median of 3 runs" % (sample, nt),
                runs=runs, build="gcc -O3 -march=native",
                generic_build={"build": "gcc -O3 (generic x86-64)", "value": res["generic"][0],
                               "runs": res["generic"][1]},
                host=host_info()), rel


class Timer:
    """Steps timed with HIP events on the launch stream; wall time around them."""

    def __init__(self, torch, stream, steps):
        self.torch, self.stream, self.steps = torch, stream, steps

    def run(self, run_once, label):
        torch = self.torch
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(self.steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(self.steps)]
        print("[bench] timing %d steps (%s)" % (self.steps, label), file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(self.steps):
            starts[i].record(self.stream)
            run_once(i)
            ends[i].record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))


def calibrate_copy(torch, dev, mb):
    if mb:
        src = torch.empty(mb << 18, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst


def write_response(path, rows):
    with open(path, "w") as f:
        for row in np.atleast_2d(np.asarray(rows, dtype=np.float64)):
            f.write(" ".join(repr(float(v)) for v in row) + "\n")


def end_to_end_cli(ph, algo, extra_args=()):
    """dwi2fod run as a user would: the phantom written as .mif images (4-D DWI with its gradient
    table in dw_scheme, a mask, response text files), the command run as a child process, and the
    FOD images written.  Reports the command's wall time and its own stage breakdown (-info)."""
    import shutil
    import subprocess
    import tempfile
    from mrtrix3_amd import mif
    exe = os.path.join(ROOT, "mrtrix3_amd", "bin", "dwi2fod")
    if not os.path.exists(exe):
        raise RuntimeError("dwi2fod is not built: run __graft_entry__.build()")
    work = tempfile.mkdtemp(prefix="mrcsd_e2e_")
    try:
        nx, ny, nz = ph.shape
        vol = np.zeros((nz, ny, nx, ph.dw.shape[1]), dtype=np.float32)
        vol[ph.mask] = ph.dw
        dwi, msk = os.path.join(work, "dwi.mif"), os.path.join(work, "mask.mif")
        mif.write(dwi, vol.transpose(2, 1, 0, 3), vox=[2.0, 2.0, 2.0],
                  keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
        del vol
        mif.write(msk, ph.mask.transpose(2, 1, 0).astype(np.float32), vox=[2.0, 2.0, 2.0])
        cmd = [exe, algo, dwi]
        responses = ph.response if algo == "msmt_csd" else [ph.response]
        for t, r in enumerate(responses):
            rp, op = os.path.join(work, "response%d.txt" % t), os.path.join(work, "fod%d.mif" % t)
            write_response(rp, r)
            cmd += [rp, op]
        cmd += ["-mask", msk, "-force", "-info"] + list(extra_args)
        print("[bench] end to end: dwi2fod %s on %d voxels" % (algo, ph.nvox), file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("dwi2fod failed: " + r.stderr[-2000:])
        line = [l for l in r.stderr.splitlines() if "timing:" in l]
        stages = {}
        if line:   # "timing: read 1.128 s, setup 0.036 s, GPU solve incl. transfers 0.088 s, ..."
            for part in line[-1].split("timing:", 1)[1].split(","):
                name, sec, unit = part.strip().rsplit(" ", 2)
                stages[name] = float(sec)
        solve = stages.get("GPU solve incl. transfers")
        return {"command": "dwi2fod %s dwi.mif %s-mask mask.mif -force -info%s" % (
                    algo, "".join("response%d.txt fod%d.mif " % (t, t) for t in range(len(responses))),
                    "".join(" " + a for a in extra_args)),
                "voxels": ph.nvox, "wall_s": wall, "stages_s": stages,
                "voxels_per_s_whole_command": ph.nvox / wall,
                "voxels_per_s_solve_incl_transfers": ph.nvox / solve if solve else None,
                "note": "child process: .mif read, solver set-up, HIP initialisation, pageable host "
                        "buffers over PCIe, GPU solve, .mif write"}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def csd_workload(workload):
    from mrtrix3_amd import phantom
    if workload == "super_resolved":
        return (phantom.super_resolved, 12, "csd_persistent_kernel<91, 128, %s>",
                "super_resolved 128x128x72 ellipsoid mask, %d voxels, 5 b=0 + 45 b=3000, lmax 12, "
                "300 constraint directions")
    return (phantom.single_shell_brain, 8, "csd_persistent_kernel<45, 64, %s>",
            "single_shell_brain 128x128x72 ellipsoid mask, %d voxels, 8 b=0 + 64 b=3000, lmax 8, "
            "300 constraint directions")


def bench_csd(args, torch, workload, ph):
    from mrtrix3_amd import csd

    dev = torch.device("cuda:0")
    make, lmax, tag, desc = csd_workload(workload)
    shared = csd.CSDShared(ph.grad, ph.response, lmax=lmax)
    solver = csd.CSD(shared)
    dw = torch.from_numpy(ph.dw).to(dev)
    fod = torch.empty((ph.nvox, shared.n_sh), dtype=torch.float32, device=dev)
    stats = torch.zeros(6, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    calibrate_copy(torch, dev, args.copy_calibration_mb)

    print("[bench] warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        solver.run(dw, fod=fod, iters=False, nneg=False)
    torch.cuda.synchronize()
    from mrtrix3_amd import tune
    # the persistent grid of each instantiation chosen by the autotuner on this volume
    print("[bench] autotuning the plain solve's grid", file=sys.stderr, flush=True)
    tuning = {"plain": tune.autotune_csd(solver, dw)}
    stream = torch.cuda.current_stream(dev)
    timer = Timer(torch, stream, args.steps)
    def roofline(st, kern_ms, leg):
        fl = csd_flops(shared.n_sh, shared.n_init, shared.n_dw, shared.n_hr, st)
        achieved = fl["total"] / (kern_ms * 1e-3) / 1e12
        return {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": load_traffic("csd_persistent_kernel_%s" % workload, leg),
                "kernel": tag % ("true" if leg == "warm" else "false"), "kernel_ms": kern_ms, "flop_per_launch": fl["total"],
                "note": "FP64-bound; peak is the MI355X FP64 rate (vector = matrix)"}

    # ---- plain GPU CSD (cold start from rconv b) ----
    wall, kern_ms = timer.run(lambda i: solver.run(dw, fod=fod, iters=False, nneg=False,
                                                   stats=stats if i == 0 else None), "plain")
    st = stats.cpu().numpy()
    plain = {"value": ph.nvox * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
             "solves_per_voxel": float(st[1]) / max(1, st[0]), "mean_neg": float(st[2]) / max(1, st[1]),
             "converged_frac": float(st[3]) / max(1, st[0]), "roofline": roofline(st, kern_ms, "plain")}
    fod_plain = fod.clone()
    fod_w = None
    # the same step captured as a HIP graph and replayed (one launch per step)
    fod_g = torch.empty_like(fod)
    graph_p = tune.SolveGraph(solver, dw, fod_g)
    wall_g, _ = timer.run(lambda i: graph_p.launch(), "plain, graph replays")
    plain["eager_ms_per_step"] = plain["ms_per_step"]
    plain["graph"] = {"ms_per_step": wall_g * 1e3 / args.steps, "identical_to_eager": bool(torch.equal(fod_g, fod_plain))}
    if plain["graph"]["identical_to_eager"] and wall_g < wall:
        plain["value"] = ph.nvox * args.steps / wall_g
        plain["ms_per_step"] = wall_g * 1e3 / args.steps
    del graph_p, fod_g

    out = {
        "metric": METRIC,
        "value": plain["value"],
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc % ph.nvox,
                   "form": "plain GPU CSD: FP64 projection and solve, cold start, persistent kernel"},
        "roofline": plain["roofline"],
        "plain": plain,
        "autotune": tuning,
    }

    # ---- learned warm start: bf16 MLP proposal, fixed-point certificate, plain fallback ----
    if not args.no_warm:
        from mrtrix3_amd import warm
        dwis = torch.from_numpy(shared.dwis.astype(np.int32)).to(dev)
        b0 = torch.from_numpy(np.nonzero(shared.matrix("grad")[:, 3] <= 10.0)[0].astype(np.int32)).to(dev)
        if args.load_net:
            net = warm.WarmNet.load(args.load_net)
            assert (net.n_in, net.nsh, net.nlog) == (shared.n_dw, shared.n_sh, shared.n_hr), "network/protocol mismatch"
            train_info = {"loaded": args.load_net}
            ntrain = None
        else:
            t0 = time.perf_counter()
            print("[bench] generating %d training phantoms" % args.train_phantoms, file=sys.stderr, flush=True)
            train = np.concatenate([make(seed=100 + i).dw for i in range(args.train_phantoms)])
            gen_s = time.perf_counter() - t0
            net = warm.WarmNet(n_in=shared.n_dw, hidden=256, nsh=shared.n_sh, nlog=shared.n_hr, seed=0)
            dtrain = torch.from_numpy(train).to(dev)
            del train
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            hist = warm.train_warm_start(net, solver, dtrain, dwis, b0, epochs=args.epochs, batch=8192, lr=args.lr,
                                         wc=args.wc, wl=args.wl,
                                         log=lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
            torch.cuda.synchronize()
            train_s = time.perf_counter() - t1
            ntrain = dtrain.shape[0]
            del dtrain
            train_info = {"voxels": int(ntrain), "epochs": args.epochs, "lr": args.lr, "wc": args.wc,
                          "wl": args.wl, "seconds": train_s, "phantom_generation_seconds": gen_s,
                          "final_loss": [float(v) for v in hist[-1]]}
            if args.save_net:
                net.save(args.save_net)
        wstats = torch.zeros(6, dtype=torch.int64, device=dev)
        fod_w = torch.empty_like(fod)
        for _ in range(args.warmup):
            warm.warm_solve(net, solver, dw, dwis, b0, budget=args.warm_budget, fod=fod_w)
        print("[bench] autotuning the warm-started solve's grid", file=sys.stderr, flush=True)
        _, neg_t = net.infer(dw, dwis, b0)
        tuning["warm"] = tune.autotune_csd(solver, dw, warm_neg=neg_t, budget=args.warm_budget)
        del neg_t
        # events split each step into the fused network inference and the warm-started solver kernel
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]

        def warm_once(i):
            ev[i][0].record(stream)
            _, neg = net.infer(dw, dwis, b0)
            ev[i][1].record(stream)
            solver.run(dw, fod=fod_w, iters=False, nneg=False, warm_neg=neg, warm_budget=args.warm_budget,
                       stats=wstats if i == 0 else None)
            ev[i][2].record(stream)

        wall_w, _ = timer.run(warm_once, "warm")
        infer_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
        kern_w = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
        fod_we = fod_w.clone()
        fod_wg = torch.empty_like(fod)
        graph_w = tune.SolveGraph(solver, dw, fod_wg, net=net, dwis=dwis, b0=b0, budget=args.warm_budget)
        wall_wg, _ = timer.run(lambda i: graph_w.launch(), "warm, graph replays")
        warm_graph = {"ms_per_step": wall_wg * 1e3 / args.steps,
                      "identical_to_eager": bool(torch.equal(fod_wg, fod_we))}
        del graph_w, fod_wg
        sw = wstats.cpu().numpy()
        mlp_flop = 2.0 * (net.n_in * net.hidden + 2 * net.hidden * net.hidden + net.hidden * net.nout) * ph.nvox
        mlp_tflops = mlp_flop / (infer_ms * 1e-3) / 1e12
        # parity of the warm-started answers against the plain solver on every voxel
        _, _, _, nm_p = solver.run(dw, negmask=True)
        _, _, _, nm_w = warm.warm_solve(net, solver, dw, dwis, b0, budget=args.warm_budget, negmask=True)
        a = fod_plain.double()
        rel = ((fod_w.double() - a).norm(dim=1) / a.norm(dim=1).clamp_min(1e-30)).cpu().numpy()
        same = (nm_p == nm_w).all(dim=1).cpu().numpy()
        ok = float(np.mean((rel <= 1e-4) & same))
        wr = {"value": ph.nvox * args.steps / wall_w, "ms_per_step": wall_w * 1e3 / args.steps,
              "speedup_vs_plain": (ph.nvox * args.steps / wall_w) / plain["value"],
              "solves_per_voxel": float(sw[1]) / max(1, sw[0]),
              "fallback_frac": float(sw[4]) / max(1, sw[0]),
              "parity_vs_plain": {"max_rel_sh_err": float(rel.max()), "p999_rel_sh_err": float(np.quantile(rel, 0.999)),
                                  "identical_neg_set_frac": float(same.mean()), "ok_frac": ok},
              "train": train_info,
              "roofline": roofline(sw, kern_w, "warm"),
              "inference": {"kernel": "infer_propose_kernel (features + 4 bf16 MFMA layers + set proposal)",
                            "ms": infer_ms, "flop": mlp_flop, "achieved": mlp_tflops,
                            "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": mlp_tflops / BF16_PEAK_TFLOPS},
              "note": "time per step = fused bf16 network inference + warm-started FP64 solve (incl. "
                      "fallbacks); training is a one-off per acquisition protocol, excluded"}
        wr["eager_ms_per_step"] = wr["ms_per_step"]
        wr["graph"] = warm_graph
        if warm_graph["identical_to_eager"] and warm_graph["ms_per_step"] < wr["ms_per_step"]:
            wr["ms_per_step"] = warm_graph["ms_per_step"]
            wr["value"] = ph.nvox * 1e3 / warm_graph["ms_per_step"]
        wr["speedup_vs_plain"] = wr["value"] / plain["value"]
        out["warm_start"] = wr
        if ok >= 0.999:
            out["value"] = wr["value"]
            out["ms_per_step"] = wr["ms_per_step"]
            trained = ("trained on %d plain-solver voxels of other phantoms" % ntrain if ntrain is not None
                       else "loaded from %s" % os.path.basename(args.load_net))
            out["config"]["form"] = ("BF16 MLP warm start (%d-256-256-256-%d, %s) + FP64 persistent CSD with "
                                     "fixed-point certificate and plain fallback" % (net.n_in, net.nout, trained))
            out["roofline"] = wr["roofline"]
        fod.copy_(fod_plain)

    # ---- signals handed over in host memory: the PCIe-inclusive rate (not the metric) ----
    print("[bench] host-buffer legs", file=sys.stderr, flush=True)
    host_dw = np.ascontiguousarray(ph.dw)
    med, runs, _ = median_rate(lambda: solver.run(host_dw, iters=False, nneg=False), ph.nvox)
    hostleg = {"plain_host_pointer_entry": {
        "value": med, "runs": runs,
        "note": "mrcsd_run_host: pageable host signals in, device buffers allocated, H2D, plain solve, "
                "D2H of the coefficients (the path dwi2fod takes)"}}
    # the transfer-inclusive solve-stage rate of the metric's definition (H2D of signals -> D2H of
    # FODs), plain and warm measured the same way: pinned host buffers, copies on the launch stream
    pin_in = torch.from_numpy(host_dw).pin_memory()
    pin_out = torch.empty((ph.nvox, shared.n_sh), dtype=torch.float32).pin_memory()
    dplain = torch.empty_like(dw)

    def plain_pinned():
        dplain.copy_(pin_in, non_blocking=True)
        solver.run(dplain, fod=fod, iters=False, nneg=False)
        pin_out.copy_(fod, non_blocking=True)
        torch.cuda.synchronize()

    plain_pinned()
    med, runs, _ = median_rate(plain_pinned, ph.nvox)
    hostleg["plain_pinned"] = {"value": med, "runs": runs,
                               "note": "pinned host signals, H2D on the stream, plain solve, D2H of the coefficients"}
    del dplain
    fod.copy_(fod_plain)
    if fod_w is not None:

        def warm_pinned():
            d = pin_in.to(dev, non_blocking=True)
            _, neg = net.infer(d, dwis, b0)
            solver.run(d, fod=fod_w, iters=False, nneg=False, warm_neg=neg, warm_budget=args.warm_budget)
            pin_out.copy_(fod_w, non_blocking=True)
            torch.cuda.synchronize()

        warm_pinned()
        med, runs, _ = median_rate(warm_pinned, ph.nvox)
        hostleg["warm_pinned"] = {"value": med, "runs": runs,
                                  "note": "pinned host signals, H2D on the stream, fused inference + warm-started "
                                          "solve, D2H of the coefficients"}
    del pin_in, pin_out
    out["host_buffers"] = hostleg

    if not args.no_cpu:
        sample = args.cpu_sample or CPU_SAMPLE[workload]
        print("[bench] CPU baseline on %d voxels" % sample, file=sys.stderr, flush=True)
        base, rel, ref = cpu_baseline_csd(ph, lmax, fod_plain.cpu().numpy(), sample)
        out["cpu_baseline"] = base
        out["parity_on_cpu_sample"] = {"plain": {"max_rel_sh_err": float(rel.max()),
                                                 "p999_rel_sh_err": float(np.quantile(rel, 0.999)),
                                                 "frac_le_1e-4": float(np.mean(rel <= 1e-4))}}
        if fod_w is not None:
            g = fod_w.cpu().numpy()[:ref.shape[0]].astype(np.float64)
            relw = np.linalg.norm(g - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
            out["parity_on_cpu_sample"]["warm_start"] = {"max_rel_sh_err": float(relw.max()),
                                                         "p999_rel_sh_err": float(np.quantile(relw, 0.999)),
                                                         "frac_le_1e-4": float(np.mean(relw <= 1e-4))}
    return out


def bench_msmt(args, torch, ph):
    from mrtrix3_amd import icls

    dev = torch.device("cuda:0")
    shared = icls.MSMTShared(ph.grad, ph.response)
    prob = shared.problem()
    b = torch.from_numpy(ph.dw).to(dev)
    x = torch.empty((ph.nvox, prob.n), dtype=torch.float32, device=dev)
    stats = torch.zeros(8, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    calibrate_copy(torch, dev, args.copy_calibration_mb)
    print("[bench] warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        prob.solve(b, x=x)
    torch.cuda.synchronize()
    timer = Timer(torch, torch.cuda.current_stream(dev), args.steps)
    wall, kern_ms = timer.run(lambda i: prob.solve(b, x=x, stats=stats if i == 0 else None), "msmt")
    st = stats.cpu().numpy()
    fl = icls_flops(prob.n, prob.nmeas, prob.ncon, st)
    achieved = fl["total"] / (kern_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": ph.nvox * args.steps / wall,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "multi_shell 128x128x72 ellipsoid mask, %d voxels, b=0/1000/2000/3000 x "
                               "10/30/60/90, WM lmax 8 + GM + CSF (%d parameters, %d constraints)"
                               % (ph.nvox, prob.n, prob.ncon),
                   "form": "msmt_csd: ICLS::Solver dual active set, FP64, persistent kernel"},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": load_traffic("icls_active_set_kernel_multi_shell"),
                     "kernel": "icls_active_set_kernel<%d, %d>" % (prob.capacity, next(w for w in (4, 5, 8, 16) if 64 * w >= prob.ncon)),
                     "kernel_ms": kern_ms, "flop_per_launch": fl["total"],
                     "note": "FP64-bound; peak is the MI355X FP64 rate (vector = matrix)"},
        "icls": {"outer_per_voxel": float(st[1]) / max(1, st[0]), "inner_per_voxel": float(st[2]) / max(1, st[0]),
                 "mean_active": float(st[3]) / max(1, st[2]), "over_capacity": int(st[7])},
    }
    print("[bench] host-buffer leg", file=sys.stderr, flush=True)
    host_b = np.ascontiguousarray(ph.dw, dtype=np.float32)
    med, runs, _ = median_rate(lambda: prob.solve(host_b), ph.nvox)
    out["host_buffers"] = {"icls_host_pointer_entry": {
        "value": med, "runs": runs,
        "note": "mricls_solve_host: pageable host signals in, device buffers allocated, H2D, ICLS solve, "
                "D2H (the path dwi2fod msmt_csd takes)"}}
    if not args.no_cpu:
        sample = args.cpu_sample or CPU_SAMPLE["multi_shell"]
        print("[bench] CPU baseline on %d voxels" % sample, file=sys.stderr, flush=True)
        base, rel = cpu_baseline_icls(ph, x.cpu().numpy(), sample)
        out["cpu_baseline"] = base
        out["parity_on_cpu_sample"] = {"max_rel_err": float(rel.max()),
                                       "p999_rel_err": float(np.quantile(rel, 0.999)),
                                       "frac_le_1e-4": float(np.mean(rel <= 1e-4))}
    return out


def peaks_flops_per_iteration(L):
    """Floating-point operations of one Newton iteration of SH::get_peak at lmax L (SH.h:514-647),
    counted from the restatement's source: +, -, *, / and square roots as one each, and each sin,
    cos, atan2 or hypot as one (their software cost on the GPU is several times that)."""
    f = 0
    for m in range(L + 1):                      # Plm_sph for every order (legendre.h:129-154)
        f += 1                                  # sin^2
        if m:
            f += 4 * m + 3                      # Plm_sph_helper (m levels) and the seed value
        if L > m:
            f += 2                              # P_{m+1}
        f += 4 * max(0, L - m - 1)              # the recursion
    f += 3                                      # sin, cos of the elevation; c0 * P_00
    f += 11 * (L // 2)                          # the m = 0 terms
    for m in range(1, L + 1):
        f += 6                                  # m az, cos, sin, the two sqrt(2) products
        for l in range(m + (m & 1), L + 1, 2):
            f += 30 + (3 if l > m else 0) + (3 if l > m + 1 else 0)
    f += 4                                      # the divisions by sin(el)
    f += 3 + 30 + 9                             # atan2, hypot, atan2; the step; normalize
    return f


def cpu_baseline_peaks(fod, gpu_out, sample, lmax):
    """The oracle's sh2peaks on the host cores (both builds, median of 3) and its agreement with the
    GPU on the same voxels (direction within 1e-3 rad, amplitude within 1e-4 of the FOD scale)."""
    from oracle import orc
    sub = np.ascontiguousarray(fod[:sample])
    nt = threads()
    res = {}
    for name, generic in (("native", False), ("generic", True)):
        med, runs, (out, _) = median_rate(lambda: orc.sh2peaks(sub, lmax=lmax, nthreads=nt, generic=generic), len(sub))
        res[name] = (med, runs, out)
    med, runs, ref = res["native"]
    g = gpu_out[:sample].reshape(len(sub), -1, 3).astype(np.float64)
    r = ref.reshape(len(sub), -1, 3).astype(np.float64)
    scale = 0.28 * np.linalg.norm(sub.astype(np.float64), axis=1)
    ok = np.ones(len(sub), dtype=bool)
    for v in range(len(sub)):
        fg, fr = np.isfinite(g[v, :, 0]), np.isfinite(r[v, :, 0])
        if not np.array_equal(fg, fr):
            ok[v] = False
            continue
        for i in np.nonzero(fr)[0]:
            ag, ar = np.linalg.norm(g[v, i]), np.linalg.norm(r[v, i])
            ang = np.arccos(min(1.0, abs(g[v, i] @ r[v, i]) / (ag * ar)))
            if ang > 1e-3 or abs(ag - ar) > 1e-4 * max(ar, scale[v]):
                ok[v] = False
    return dict(value=med, unit="voxels/s", cores=nt, kind="port",
                sample="first %d masked voxels of the same FOD volume; oracle C restatement of sh2peaks (single "
                       "precision as the reference, %d threads), median of 3 runs" % (len(sub), nt),
                runs=runs, build="gcc -O3 -march=native",
                generic_build={"build": "gcc -O3 (generic x86-64)", "value": res["generic"][0], "runs": res["generic"][1]},
                host=host_info()), float(ok.mean())


def end_to_end_peaks(ph):
    """The pipeline as a user runs it: dwi2fod csd writes the FOD image, sh2peaks reads it and writes
    the peaks image (both child processes, .mif files on disk); each command's wall time and stages."""
    import shutil
    import subprocess
    import tempfile
    from mrtrix3_amd import mif
    bindir = os.path.join(ROOT, "mrtrix3_amd", "bin")
    work = tempfile.mkdtemp(prefix="mrcsd_e2e_")
    try:
        nx, ny, nz = ph.shape
        vol = np.zeros((nz, ny, nx, ph.dw.shape[1]), dtype=np.float32)
        vol[ph.mask] = ph.dw
        dwi, msk, rp = (os.path.join(work, f) for f in ("dwi.mif", "mask.mif", "response.txt"))
        fod, pk = os.path.join(work, "fod.mif"), os.path.join(work, "peaks.mif")
        mif.write(dwi, vol.transpose(2, 1, 0, 3), vox=[2.0, 2.0, 2.0], keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
        del vol
        mif.write(msk, ph.mask.transpose(2, 1, 0).astype(np.float32), vox=[2.0, 2.0, 2.0])
        write_response(rp, ph.response)
        out = {}
        for name, cmd in (("dwi2fod", [os.path.join(bindir, "dwi2fod"), "csd", dwi, rp, fod, "-mask", msk]),
                          ("sh2peaks", [os.path.join(bindir, "sh2peaks"), fod, pk, "-mask", msk])):
            print("[bench] end to end: %s" % name, file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            r = subprocess.run(cmd + ["-force", "-info"], capture_output=True, text=True, timeout=900)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError("%s failed: %s" % (name, r.stderr[-2000:]))
            stages = {}
            for line in r.stderr.splitlines():
                if "timing:" in line:
                    for part in line.split("timing:", 1)[1].split(","):
                        k, sec, unit = part.strip().rsplit(" ", 2)
                        stages[k] = float(sec)
            out[name] = {"wall_s": wall, "stages_s": stages}
        nvox_img = nx * ny * nz
        out["command"] = "dwi2fod csd dwi.mif response.txt fod.mif -mask mask.mif; sh2peaks fod.mif peaks.mif -mask mask.mif"
        out["voxels"] = {"masked": ph.nvox, "image": nvox_img}
        out["note"] = ("sh2peaks visits every voxel of the image, as the reference does (masked-out voxels are "
                       "written as NaN without a search)")
        return out
    finally:
        shutil.rmtree(work, ignore_errors=True)


def bench_peaks(args, torch, ph):
    """sh2peaks on the whole-brain FODs: the GPU CSD writes the 646,128 FODs into HBM, then a step is
    one peak search of every voxel (60 search directions, 3 peaks)."""
    from mrtrix3_amd import csd, peaks
    dev = torch.device("cuda:0")
    calibrate_copy(torch, dev, args.copy_calibration_mb)
    shared = csd.CSDShared(ph.grad, ph.response)
    solver = csd.CSD(shared)
    dw = torch.from_numpy(ph.dw).to(dev)
    fod, _, _ = solver.run(dw, iters=False, nneg=False)
    del dw
    finder = peaks.SH2Peaks(shared.n_sh)
    out = torch.empty((ph.nvox, 3 * finder.npeaks), dtype=torch.float32, device=dev)
    stats = torch.zeros(4, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        finder.run(fod, out=out)
    timer = Timer(torch, torch.cuda.current_stream(dev), args.steps)
    wall, kern_ms = timer.run(lambda i: finder.run(fod, out=out, stats=stats if i == 0 else None), "peaks")
    st = stats.cpu().numpy()
    fl = float(st[1]) * peaks_flops_per_iteration(shared.lmax)
    achieved = fl / (kern_ms * 1e-3) / 1e12
    res = {
        "metric": "sh2peaks voxels/s on 1 MI355X (peaks of the dwi2fod FODs)",
        "value": ph.nvox * args.steps / wall,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "sh2peaks on the single_shell_brain FODs (646,128 voxels, lmax 8), 60 search "
                               "directions, 3 peaks; FODs from the plain GPU CSD, resident in HBM"},
        "roofline": {"bound": "fp32", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": load_traffic("peaks_kernel_single_shell_brain"),
                     "kernel": "peaks_search_kernel<%d> + peaks_select_kernel" % shared.lmax,
                     "kernel_ms": kern_ms, "flop_per_launch": fl,
                     "note": "FP32 VALU-bound (no matrix work); FLOPs counted per Newton iteration from the "
                             "restatement, transcendental calls as one; one run is the search kernel (all the "
                             "FLOPs) and the per-voxel selection kernel, timed together"},
        "peaks": {"newton_iterations_per_voxel": float(st[1]) / max(1, st[0]),
                  "searches_not_converged_frac": float(st[2]) / max(1.0, 60.0 * st[0]),
                  "flop_per_iteration": peaks_flops_per_iteration(shared.lmax)},
    }
    if not args.no_cpu:
        sample = args.cpu_sample or CPU_SAMPLE["peaks"]
        print("[bench] CPU baseline on %d voxels" % sample, file=sys.stderr, flush=True)
        base, agree = cpu_baseline_peaks(fod.cpu().numpy(), out.cpu().numpy(), sample, shared.lmax)
        res["cpu_baseline"] = base
        res["parity_on_cpu_sample"] = {"voxels_agreeing": agree}
    return res


def bench_fixel(args, torch, ph):
    """fod2fixel on the whole-brain FODs: the GPU CSD writes the 646,128 FODs into HBM, then a step
    is one segmentation of every voxel (1281 directions, default thresholds): amplitudes, lobes,
    peak revision, offsets and fixels in voxel order."""
    from mrtrix3_amd import csd, fixel
    from oracle import orc
    dev = torch.device("cuda:0")
    shared = csd.CSDShared(ph.grad, ph.response)
    solver = csd.CSD(shared)
    dw = torch.from_numpy(ph.dw).to(dev)
    fod, _, _ = solver.run(dw, iters=False, nneg=False)
    del dw
    seg = fixel.Segmenter(shared.n_sh)
    first = seg.segment(fod)
    nfix = int(first.fixels.shape[0])
    lib = fixel._lib.load()
    count = torch.empty(ph.nvox, dtype=torch.int32, device=dev)
    offset = torch.empty(ph.nvox, dtype=torch.int32, device=dev)
    fix = torch.empty((max(nfix, 1), 5), dtype=torch.float32, device=dev)
    total = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step(i):
        fixel._lib.check(lib.mrfixel_segment(seg._h, fod.data_ptr(), fod.shape[1], None, ph.nvox, count.data_ptr(),
                                             offset.data_ptr(), fix.data_ptr(), fix.shape[0], None, total.data_ptr(),
                                             stream))

    for _ in range(args.warmup):
        step(0)
    torch.cuda.synchronize()
    timer = Timer(torch, torch.cuda.current_stream(dev), args.steps)
    wall, kern_ms = timer.run(step, "fixel")
    assert int(total.item()) == nfix and torch.equal(count, first.count) and torch.equal(fix, first.fixels)
    # algorithmic traffic: the float32 FODs in, count and offset out, five floats per fixel out
    hbm = ph.nvox * (4.0 * shared.n_sh + 8.0) + 20.0 * nfix
    res = {
        "metric": "fod2fixel voxels/s on 1 MI355X (FMLS segmentation of the dwi2fod FODs)",
        "value": ph.nvox * args.steps / wall,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fod2fixel on the single_shell_brain FODs (646,128 voxels, lmax 8), 1281 "
                               "directions, default thresholds; FODs from the plain GPU CSD, resident in HBM"},
        "roofline": {"bound": "hbm", "achieved": hbm / (kern_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hbm / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "fmls_amp_kernel + fmls_sort_kernel + fmls_grow_kernel + fmls_revise_kernel<%d> "
                               "+ item, finalize, offset and emit kernels" % shared.lmax,
                     "kernel_ms": kern_ms, "algorithmic_bytes": hbm,
                     "note": "sequential per voxel (lobe growth in visiting order): a wavefront per voxel "
                             "sorts, a lane per voxel grows lobes (labels in LDS), a lane per lobe revises "
                             "peaks; the amplitudes and per-voxel work arrays in HBM are extra traffic; the "
                             "amplitude GEMM is 2 x 1281 x n_sh FP64 FLOPs per voxel"},
        "fixels": {"total": nfix, "per_voxel": nfix / ph.nvox},
    }
    if not args.no_cpu:
        sample = args.cpu_sample or CPU_SAMPLE["fixel"]
        print("[bench] CPU baseline on %d voxels" % sample, file=sys.stderr, flush=True)
        o = orc.FmlsOracle(shared.lmax)
        f_host = fod[:sample].cpu().numpy().astype(np.float64)
        t0 = time.perf_counter()
        lobes = [o.segment(f_host[v], max_out=2048)[0] for v in range(sample)]
        cpu_s = time.perf_counter() - t0
        cnt = first.count[:sample].cpu().numpy()
        off = first.offset[:sample].cpu().numpy()
        fx = first.fixels.cpu().numpy()
        same, worst = 0, 0.0
        for v in range(sample):
            want = np.concatenate([lobes[v][:, 0:3], lobes[v][:, 6:8]], axis=1).astype(np.float32)
            got = fx[off[v]:off[v] + cnt[v]]
            if cnt[v] == len(want):
                same += 1
                if len(want):
                    worst = max(worst, float(np.max(np.abs(got - want) / (np.abs(want) + 1e-6))))
        res["cpu_baseline"] = {"value": sample / cpu_s, "unit": "voxels/s", "cores": 1, "kind": "port",
                               "sample": "first %d voxels; oracle FP64 restatement of Segmenter (oracle/fmls_oracle.c), "
                                         "one thread" % sample}
        res["parity_on_cpu_sample"] = {"voxels_same_fixel_count": same, "voxels": sample,
                                       "max_rel_err_fixels": worst}
    return res


def bench_tournier(args, ph):
    """dwi2response tournier on the whole-brain phantom, run as a user would: the DWI (72 volumes,
    646,128 masked voxels of 1,179,648) and the mask written as .mif images, the command run as a
    child process (-number 300 -iter_voxels 3000 -dilate 1 -max_iters 10, the script's defaults).
    A step is one whole estimation; the command reports each iteration's GPU CSD and FMLS times
    (HIP events) and wall time."""
    import re
    import shutil
    import subprocess
    import tempfile
    from mrtrix3_amd import mif
    from oracle import dwi2response_oracle as T
    exe = os.path.join(ROOT, "mrtrix3_amd", "bin", "dwi2response")
    if not os.path.exists(exe):
        raise RuntimeError("dwi2response is not built: run __graft_entry__.build()")
    work = tempfile.mkdtemp(prefix="mrcsd_tournier_")
    try:
        nx, ny, nz = ph.shape
        vol = np.zeros((nz, ny, nx, ph.dw.shape[1]), dtype=np.float32)
        vol[ph.mask] = ph.dw
        dwi, msk = os.path.join(work, "dwi.mif"), os.path.join(work, "mask.mif")
        mif.write(dwi, vol.transpose(2, 1, 0, 3), vox=[2.0, 2.0, 2.0], keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
        del vol
        mif.write(msk, ph.mask.transpose(2, 1, 0).astype(np.float32), vox=[2.0, 2.0, 2.0])
        runs = []
        for i in range(args.warmup + args.steps):
            out = os.path.join(work, "resp%d.txt" % i)
            print("[bench] dwi2response tournier run %d" % i, file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            r = subprocess.run([exe, "tournier", dwi, out, "-mask", msk, "-info"], capture_output=True, text=True,
                               timeout=900)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError("dwi2response failed: " + r.stderr[-2000:])
            its = []
            for m in re.finditer(r"iteration (\d+): (\d+) voxels in the mask; CSD ([\d.]+) ms, FMLS ([\d.]+) ms \(GPU\); "
                                 r"(\d+) voxels selected; iteration ([\d.]+) s", r.stderr):
                its.append({"iteration": int(m.group(1)), "voxels": int(m.group(2)), "csd_ms": float(m.group(3)),
                            "fmls_ms": float(m.group(4)), "selected": int(m.group(5)), "wall_s": float(m.group(6)),
                            "voxels_per_s": int(m.group(2)) / float(m.group(6))})
            runs.append({"wall_s": wall, "iterations": its, "response": open(out).read().split(),
                         "converged": "Convergence of SF voxel selection" in r.stderr})
        timed = runs[args.warmup:]
        loop_s = float(np.mean([sum(x["wall_s"] for x in t["iterations"]) for t in timed]))
        wall = float(np.mean([t["wall_s"] for t in timed]))
        it0 = timed[-1]["iterations"][0]
        res = {
            "metric": "dwi2response tournier: brain-mask voxels per second of the whole estimation on 1 MI355X",
            "value": ph.nvox / loop_s,
            "unit": "voxels/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": loop_s * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "dwi2response tournier on the single_shell_brain phantom (646,128 masked voxels, "
                                   "8 b=0 + 64 b=3000), -number 300 -iter_voxels 3000 -dilate 1 -max_iters 10; "
                                   "the DWI read once, resident on the GPU across iterations"},
            "roofline": {"bound": "latency", "achieved": None, "peak": None, "unit": None, "frac": None,
                         "traffic": None, "kernel": "FMLS segmentation kernels (iteration 0, the whole mask)",
                         "kernel_ms": it0["fmls_ms"],
                         "note": "iteration 0 segments every masked voxel's lmax-4 FOD and dominates; its roofline is "
                                 "the fod2fixel workload's (latency-bound, sequential lobe growth per voxel); the "
                                 "later iterations see the dilated top 3000 voxels only"},
            "iterations": timed[-1]["iterations"],
            "converged": timed[-1]["converged"],
            "response": [float(x) for x in timed[-1]["response"]],
            "command_wall_s": wall,
            "note": "value = masked voxels / the iterations' wall time (GPU CSD, GPU FMLS, host selection, GPU "
                    "amp2response accumulation and solve, host transfers of the per-voxel summary); "
                    "command_wall_s adds reading the 340 MB .mif and HIP initialisation",
        }
        if not args.no_cpu:
            sample = args.cpu_sample or CPU_SAMPLE["tournier"]
            # a centred block of the brain with the same protocol, the loop's defaults scaled to it
            side = int(round(sample ** (1.0 / 3.0)))
            bx, by, bz = side, side, max(1, sample // (side * side))
            x0, y0, z0 = (nx - bx) // 2, (ny - by) // 2, (nz - bz) // 2
            full = np.zeros((nz, ny, nx, ph.dw.shape[1]), dtype=np.float32)
            full[ph.mask] = ph.dw
            blk = full[z0:z0 + bz, y0:y0 + by, x0:x0 + bx]
            del full
            bm = ph.mask[z0:z0 + bz, y0:y0 + by, x0:x0 + bx].reshape(-1)
            keep = ph.grad[:, 3] > 10
            dwb = np.ascontiguousarray(blk.reshape(-1, ph.dw.shape[1])[:, keep])
            nb = int(bm.sum())
            number = max(1, 300 * nb // ph.nvox + 1)
            print("[bench] CPU baseline: oracle loop on a %dx%dx%d block (%d voxels)" % (bx, by, bz, nb),
                  file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            _, _, n_it, conv, _ = T.run(dwb, ph.grad[keep], (bx, by, bz), bm, number=number, iter_voxels=10 * number,
                                        nthreads=1)
            cpu_s = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": nb / cpu_s, "unit": "voxels/s", "cores": 1, "kind": "port",
                                   "sample": "centred %dx%dx%d block of the phantom (%d voxels), -number %d "
                                             "-iter_voxels %d, %d iterations; oracle loop (oracle/dwi2response_oracle.py: "
                                             "C CSD and FMLS, numpy amp2response), one thread" % (
                                                 bx, by, bz, nb, number, 10 * number, n_it)}
        return res
    finally:
        shutil.rmtree(work, ignore_errors=True)


def resp_flops_per_sample(lmax):
    """FP64 FLOPs of one (voxel, direction) step of the amp2response accumulation
    (cmd/amp2response.cpp:152-191 as the kernel restates it): the dot product with the fibre
    direction (5), the direction's norm with its square root (6) and the division (1), the zonal
    Legendre recursion to lmax (2 for degree 1, then 4 per degree: product, division, difference,
    scaling), and the rank-1 updates of M (2 per lower-triangle entry) and b (2 per coefficient)."""
    nz = lmax // 2 + 1
    leg = 2 + 4 * (lmax - 1) if lmax >= 1 else 0
    return 12 + leg + 2 * nz * (nz + 1) // 2 + 2 * nz


def bench_response(args, torch, ph):
    """amp2response's accumulation over the whole-brain phantom: fibre directions from the GPU CSD
    and sh2peaks (first peak), the b=3000 shell (64 directions) at lmax 10; a step is one
    accumulation over all 646,128 voxels (the per-shell constrained solve is timed separately)."""
    from mrtrix3_amd import csd, peaks, resp
    from oracle import resp_oracle as R
    dev = torch.device("cuda:0")
    shared = csd.CSDShared(ph.grad, ph.response)
    dw = torch.from_numpy(ph.dw).to(dev)
    fod, _, _ = csd.CSD(shared).run(dw, iters=False, nneg=False)
    fib = peaks.SH2Peaks(shared.n_sh).run(fod)[:, :3].contiguous()
    ok = torch.isfinite(fib).all(dim=1).to(torch.uint8)
    del fod
    vols = np.nonzero(ph.grad[:, 3] > 100)[0]
    dirs = ph.grad[vols, :3] / np.linalg.norm(ph.grad[vols, :3], axis=1, keepdims=True)
    lmax = 10
    acc = resp.Amp2Response(dirs, vols, lmax)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        acc.accumulate(dw, fib, ok)
    timer = Timer(torch, torch.cuda.current_stream(dev), args.steps)
    holder = {}
    wall, kern_ms = timer.run(lambda i: holder.__setitem__("r", acc.accumulate(dw, fib, ok)), "response")
    M, b, count = holder["r"]
    t0 = time.perf_counter()
    rf, iters = resp.solve(lmax, M, b, True)
    solve_s = time.perf_counter() - t0
    samples = float(count) * len(vols)
    fl = samples * resp_flops_per_sample(lmax)
    achieved = fl / (kern_ms * 1e-3) / 1e12
    res = {
        "metric": "amp2response voxels/s on 1 MI355X (normal-equation accumulation over a mask)",
        "value": ph.nvox * args.steps / wall,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "amp2response on the single_shell_brain phantom (646,128 voxels), b=3000 shell "
                               "(64 directions), lmax 10, fibre directions from GPU CSD + sh2peaks, resident in HBM"},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "resp_accumulate_kernel<10> + resp_sum_kernel (+ the 36-value copy back)",
                     "kernel_ms": kern_ms, "flop_per_launch": fl,
                     "note": "FP64; a step is short enough that the launch, the final in-order sum and the "
                             "copy of the sums to the host are a visible part of it"},
        "response": {"voxels_in_mask": int(count), "solve_ms": solve_s * 1e3, "icls_iterations": iters,
                     "rf": [float(x) for x in rf]},
    }
    if not args.no_cpu:
        sample = args.cpu_sample or CPU_SAMPLE["response"]
        print("[bench] CPU baseline on %d voxels" % sample, file=sys.stderr, flush=True)
        amp_h = ph.dw[:sample]
        fib_h = fib[:sample].cpu().numpy()
        ok_h = ok[:sample].cpu().numpy()
        t0 = time.perf_counter()
        Mo, bo, co = R.accumulate(amp_h, fib_h, ok_h, dirs, vols, lmax)
        cpu_s = time.perf_counter() - t0
        sub = torch.zeros_like(ok)
        sub[:sample] = ok[:sample]
        Mg, bg, cg = acc.accumulate(dw, fib, sub)
        rel = float(max(np.abs(Mg - Mo).max() / np.abs(Mo).max(), np.abs(bg - bo).max() / np.abs(bo).max()))
        res["cpu_baseline"] = {"value": sample / cpu_s, "unit": "voxels/s", "cores": 1, "kind": "port",
                               "sample": "first %d voxels of the same phantom; oracle float64 numpy restatement "
                                         "(oracle/resp_oracle.py, one voxel at a time, the reference's rotation "
                                         "route), one thread" % sample}
        res["parity_on_cpu_sample"] = {"max_rel_err_sums": rel, "voxels_in_mask": int(co)}
    return res


def gt_flops_per_voxel(nrows, ncols, nf):
    """Algorithmic FP64 FLOPs of calcEnergy per voxel (externalenergy.cpp:245-252), the NNLS counted
    as the ICLS solver's unconstrained part (2 nrows (nf+1)) plus one inner solve per constraint."""
    n = nf + 1
    return 2.0 * nrows * ncols + 2.0 * nrows * n + n * n * (n + 1) + 2.0 * nrows * nf + 3.0 * nrows


def bench_gt(args, torch, ph):
    """Global tractography's external energy over the whole-brain phantom: the DW signal of 646,128
    voxels (8 b=0 + 64 b=3000), a TOD per voxel (lmax 8, random coefficients around the phantom's
    fibres' scale), the WM response of the phantom and one isotropic tissue (ADC 3e-3 mm^2/s); a step
    is calcEnergy over every voxel and the sum (ExternalEnergyComputer::resetEnergy)."""
    from mrtrix3_amd import gt, phantom as P
    from oracle import gt_oracle
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(21)
    bvals = sorted(set(np.round(ph.grad[:, 3], -2)))
    wm = np.zeros((len(bvals), np.size(ph.response)))
    wm[-1] = np.ravel(ph.response)
    wm[0, 0] = np.mean(ph.dw[:, ph.grad[:, 3] < 10]) * np.sqrt(4 * np.pi)
    iso = [[P.isotropic_response(b, 3.0e-3)[0] for b in bvals]]
    E = gt.ExternalEnergy(ph.grad, wm, iso, lmax=8, weight=1.0, ppot=0.01)
    tod = (rng.standard_normal((ph.nvox, E.ncols)) * 0.01).astype(np.float32)
    tod[:, 0] = np.abs(tod[:, 0]) + 0.05
    dw = torch.from_numpy(ph.dw).to(dev)
    td = torch.from_numpy(tod).to(dev)
    for _ in range(args.warmup):
        E.energy(dw, td)
    torch.cuda.synchronize()
    timer = Timer(torch, torch.cuda.current_stream(dev), args.steps)
    holder = {}
    wall, kern_ms = timer.run(lambda i: holder.__setitem__("r", E.energy(dw, td)), "gt_energy")
    e, fi, total = holder["r"]
    fl = ph.nvox * gt_flops_per_voxel(E.nrows, E.ncols, E.nf)
    achieved = fl / (kern_ms * 1e-3) / 1e12
    hbm = ph.nvox * 4.0 * (E.nrows + E.ncols + 1 + E.nf)
    res = {
        "metric": "global tractography external energy voxels/s on 1 MI355X (calcEnergy over a volume)",
        "value": ph.nvox * args.steps / wall,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "external energy on the single_shell_brain phantom's signal (646,128 voxels, 72 "
                               "volumes), lmax-8 TOD, WM + one isotropic tissue, resident in HBM"},
        "roofline": {"bound": "hbm", "achieved": hbm / (kern_ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": hbm / (kern_ms * 1e-3) / 1e9 / 8000.0, "traffic": None,
                     "kernel": ("gt_fused_kernel<%d> + gt_sum_kernel" % (E.nf + 1)) if E.nf + 1 <= 4 else
                               "gt_residual_kernel + icls_active_set_kernel + gt_energy_kernel + gt_sum_kernel",
                     "kernel_ms": kern_ms, "algorithmic_bytes": hbm, "fp64_flop": fl,
                     "fp64_achieved_tflops": achieved,
                     "note": "algorithmic bytes = float32 signal and TOD in, energy and fractions out; with up to 4 "
                             "NNLS unknowns one fused kernel (a lane per voxel, LDS-staged TOD and K blocks, the "
                             "NNLS by active-set enumeration) makes no other traffic; its occupancy is bounded by "
                             "the LDS tiles, so it is latency-bound below both the HBM and the FP64 rooflines"},
    }
    if not args.no_cpu:
        sample = args.cpu_sample or CPU_SAMPLE["gt_energy"]
        print("[bench] CPU baseline on %d voxels" % sample, file=sys.stderr, flush=True)
        O = gt_oracle.Energy(E.matrix("K"), E.matrix("Ak"), mu=0.01 * np.sqrt(4 * np.pi))
        t0 = time.perf_counter()
        eo = np.array([O.voxel(ph.dw[v].astype(np.float64), tod[v].astype(np.float64))[0] for v in range(sample)])
        cpu_s = time.perf_counter() - t0
        eg = e[:sample].cpu().numpy()
        res["cpu_baseline"] = {"value": sample / cpu_s, "unit": "voxels/s", "cores": 1, "kind": "port",
                               "sample": "first %d voxels; oracle float64 restatement (oracle/gt_oracle.py: numpy "
                                         "residual, the oracle's C ICLS for the NNLS), one thread" % sample}
        res["parity_on_cpu_sample"] = {"max_rel_err_energy": float(np.max(np.abs(eg - eo) / np.abs(eo)))}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="single_shell_brain", choices=sorted(CPU_SAMPLE))
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="voxels in the CPU baseline sample (0: the workload's default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-warm", action="store_true", help="skip the learned warm start")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the end-to-end dwi2fod command on the phantom written as .mif files")
    ap.add_argument("--train-phantoms", type=int, default=4,
                    help="phantoms (646k voxels each, further seeds of the protocol) to train on")
    ap.add_argument("--epochs", type=int, default=64)
    ap.add_argument("--lr", type=float, default=8e-3, help="peak learning rate of the warm-start training")
    ap.add_argument("--wc", type=float, default=1.0, help="weight of the coefficient loss")
    ap.add_argument("--wl", type=float, default=0.3, help="weight of the negative-set loss")
    ap.add_argument("--warm-budget", type=int, default=6)
    ap.add_argument("--save-net", default="", help="write the trained warm-start network to this .npz")
    ap.add_argument("--load-net", default="",
                    help="use a warm-start network saved by --save-net instead of training one")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="launch one device copy of this many MiB first (a known-traffic kernel for "
                         "calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the CSD voxel solve shards by voxel slabs with no "
              "exchange, and this build runs and reports a single MI355X (--gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    # The phantom is generated, and the end-to-end command run as a child process, before this
    # process touches the GPU.
    from mrtrix3_amd import phantom
    print("[bench] generating %s phantom" % args.workload, file=sys.stderr, flush=True)
    e2e_warm = None
    if args.workload == "multi_shell":
        ph = phantom.multi_shell()
        e2e = None if args.no_e2e else end_to_end_cli(ph, "msmt_csd")
    elif args.workload == "peaks":
        ph = phantom.single_shell_brain()
        e2e = None if args.no_e2e else end_to_end_peaks(ph)
    elif args.workload in ("response", "gt_energy", "fixel", "tournier"):
        ph = phantom.single_shell_brain()
        e2e = None
    else:
        make, lmax, _, _ = csd_workload(args.workload)
        ph = make()
        e2e = None if args.no_e2e else end_to_end_cli(ph, "csd", ["-lmax", str(lmax)])
        if e2e is not None and args.workload == "single_shell_brain" and not args.no_warm:
            # the command with the network trained for this protocol by this script's defaults
            # (tests/golden/warm_net_single_shell_brain.npz, written with --save-net)
            e2e_warm = end_to_end_cli(ph, "csd", ["-lmax", str(lmax), "-warm_net", PROTOCOL_NET])

    if args.workload == "tournier":
        # the command is the whole workload, run as child processes: this process never uses the GPU
        print(json.dumps(bench_tournier(args, ph)))
        return
    import torch
    if args.workload == "multi_shell":
        out = bench_msmt(args, torch, ph)
    elif args.workload == "peaks":
        out = bench_peaks(args, torch, ph)
    elif args.workload == "response":
        out = bench_response(args, torch, ph)
    elif args.workload == "gt_energy":
        out = bench_gt(args, torch, ph)
    elif args.workload == "fixel":
        out = bench_fixel(args, torch, ph)
    else:
        out = bench_csd(args, torch, args.workload, ph)
    if e2e is not None:
        out["end_to_end"] = e2e
    if e2e_warm is not None:
        train_s = out.get("warm_start", {}).get("train", {}).get("seconds")
        saving = e2e["wall_s"] - e2e_warm["wall_s"]
        e2e_warm["amortisation"] = {
            "training_s": train_s, "saving_per_command_s": saving,
            "commands_to_repay_training": (train_s / saving) if (train_s and saving > 0) else None,
            "note": "the network is trained once per acquisition protocol (training_s: this run's training "
                    "on the GPU) and then reused by every dwi2fod -warm_net run of that protocol"}
        out["end_to_end_warm"] = e2e_warm
    print(json.dumps(out))


if __name__ == "__main__":
    main()
