"""Benchmark: batched RBF-FD operator assembly on one MI355X.

Workload: 1,048,576 nodes on a jittered 1024x1024 grid of the unit square
(jitter +-0.25h, seeded), Hilbert-renumbered, n = 35 nearest neighbours,
PHS r^5 + degree-3 polynomials (45x45 systems), operators Laplacian, d/dx,
d/dy into CSR (FP64 weights, int32 column indices shared with ja).

One step = one assembly pass over all 1M stencils, inputs resident in HBM.
Both forms are timed: the plain FP64 kernel and the mixed form (FP32 LU +
FP64 refinement, <= 3 steps, FP64 fallback). The headline value is the
stencils/s of the faster form whose weights stay within 1e-8 (max relative,
per stencil and operator) of the CPU FP64 oracle over the whole workload;
both forms, their ratio, the weight errors, the fallback fraction and the
analytic consistency error are reported beside it. The oracle run doubles
as the CPU baseline (all host cores, same 1M stencils).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-assembly_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix alike)
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_errors(w_gpu, w_ref, n):
    """Per stencil and operator: max_k |w - w_ref| / max_k |w_ref|."""
    out = []
    for g, r in zip(w_gpu, w_ref):
        g = g.reshape(-1, n)
        r = r.reshape(-1, n)
        out.append(np.abs(g - r).max(1) / np.maximum(np.abs(r).max(1), 1e-300))
    return np.stack(out, 1)


def consistency(w, ja, x, y, n):
    """Max relative error of the assembled Laplacian, d/dx, d/dy applied to
    f = sin(2 pi x) cos(2 pi y), on nodes at least 4 h from the boundary."""
    J = ja.reshape(-1, n)
    f = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    exact = [-8 * np.pi ** 2 * f,
             2 * np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y),
             -2 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)]
    inner = (x > 0.004) & (x < 0.996) & (y > 0.004) & (y < 0.996)
    fJ = f[J]
    errs = []
    for wi, ex in zip(w, exact):
        v = np.einsum("ij,ij->i", wi.reshape(-1, n), fJ)
        errs.append(float(np.abs(v - ex)[inner].max() / np.abs(ex).max()))
    return errs


def hip_timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    return t0.elapsed_time(t1) / steps, wall


def read_traffic(kernel_tag):
    """HBM bytes per launch of the kernel, from the committed rocprofv3 --pmc
    summary (profiles/<kernel_tag>.md, line "traffic_bytes_per_launch: N")."""
    path = os.path.join(ROOT, "profiles", kernel_tag + ".md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20,
                    help="stencils in the CPU-baseline / oracle-parity sample (default: all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity")
    ap.add_argument("--form", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--no-tune", action="store_true", help="skip the variant autotuner")
    ap.add_argument("--variant", type=int, default=None,
                    help="use this kernel variant (no autotuning)")
    ap.add_argument("--profile-copy", action="store_true",
                    help="also run 3 plain 256 MiB device copies (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the assembly "
              "shards by node range with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import rbf_amd as R
    from rbf_amd import workloads

    dev = torch.device("cuda:0")
    log("building 1M-node workload (host: jittered grid, Hilbert order, kNN n=35)")
    W = workloads.build("square_1m_n35", seed=args.seed)
    n, P, Q, nst = W["n"], W["P"], W["Q"], W["nstencils"]
    log(f"host prep: {W['host_s']}")
    x = torch.from_numpy(W["x"]).to(dev)
    y = torch.from_numpy(W["y"]).to(dev)
    ja = torch.from_numpy(W["ja"]).to(dev)
    nnz = nst * n
    info = torch.empty(nst, dtype=torch.int32, device=dev)
    status = torch.empty(nst, dtype=torch.uint8, device=dev)

    forms = ["fp64", "mixed"] if args.form == "both" else [args.form]
    res = {}
    for form in forms:
        plan = R.Plan(2, n, P, Q, "diff", precision=form, max_refine=3)
        nt, r = plan.nt, plan.nops
        w = [torch.empty(nnz, dtype=torch.float64, device=dev) for _ in range(r)]

        def step():
            plan.assemble(nst, x, y, ja, w, info=info, status=status)

        tune_ms = None
        if args.variant is not None:
            plan.variant = args.variant
        elif not args.no_tune:
            tune_ms = plan.autotune(nst, x, y, ja, w, info=info, status=status, reps=2)
            log(f"[{form}] autotune: " + ", ".join(
                f"{plan.variant_name(v)}={t:.3f}ms" for v, t in enumerate(tune_ms)))
        log(f"[{form}] variant {plan.variant_name()}; warmup {args.warmup}, "
            f"timing {args.steps} steps")
        ms, wall = hip_timed(step, args.steps, args.warmup)
        bad = int((info != 0).sum().item())
        fb = int(((status & R.STAT_FALLBACK) != 0).sum().item()) if form == "mixed" else 0
        steps_hist = None
        if form == "mixed":
            st = (status & R.STAT_STEPS_MASK).to(torch.int64)
            steps_hist = [int((st == k).sum().item()) for k in range(4)]
        res[form] = {"ms": ms, "wall_s": wall, "singular": bad, "fallback": fb,
                     "variant": plan.variant_name(),
                     "autotune_ms": None if tune_ms is None else
                     {plan.variant_name(v): round(t, 4) for v, t in enumerate(tune_ms)},
                     "refine_steps_hist": steps_hist, "nt": nt, "r": r,
                     "w": [t.cpu().numpy() for t in w]}
        log(f"[{form}] {ms:.3f} ms/step, {nst / (ms * 1e-3) / 1e6:.1f} Mstencils/s, "
            f"singular {bad}, fallback {fb}")
        del w
        torch.cuda.empty_cache()

    # headline: the faster form whose weights meet the 1e-8 bar (checked
    # against the oracle below; without the oracle run, the faster form)
    head = min(res, key=lambda f: res[f]["ms"])
    nt, r = res[head]["nt"], res[head]["r"]
    flop = workloads.flops_per_stencil(nt, r)
    byt = workloads.bytes_per_stencil(n, r)
    ms = res[head]["ms"]
    value = nst / (ms * 1e-3)
    achieved_tflops = flop * nst / (ms * 1e-3) / 1e12

    extra = {}
    for form, d in res.items():
        sps = nst / (d["ms"] * 1e-3)
        extra[form] = {"ms_per_step": round(d["ms"], 4), "stencils_per_s": sps,
                       "gflops": flop * sps / 1e9,
                       "frac_fp64_peak": flop * sps / 1e12 / PEAK_FP64_TFLOPS,
                       "scatter_gbs": byt * sps / 1e9, "singular": d["singular"],
                       "fallback_fraction": d["fallback"] / nst, "variant": d["variant"],
                       "autotune_ms": d["autotune_ms"]}
        if d["refine_steps_hist"] is not None:
            extra[form]["refine_steps_hist"] = d["refine_steps_hist"]
    if "fp64" in res and "mixed" in res:
        extra["mixed_over_fp64"] = res["fp64"]["ms"] / res["mixed"]["ms"]

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = len(os.sched_getaffinity(0))
        cnt = min(args.cpu_sample, nst)
        s0 = (nst - cnt) // 2
        log(f"CPU oracle: {cnt} stencils on {cores} threads")
        t0 = time.perf_counter()
        wref, iref = O.assemble(2, n, P, Q, W["x"], W["y"], W["ja"], s0=s0, count=cnt,
                                nthreads=cores)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": cnt / cpu_s, "unit": "stencils/s", "cores": cores, "kind": "port",
               "sample": f"{cnt} contiguous stencils (from {s0}) of the same 1M workload, "
                         f"FP64 oracle (gather, fill, partial-pivot LU, solve, scatter), "
                         f"{cpu_s:.2f} s"}
        for form, d in res.items():
            wg = [a[s0 * n:(s0 + cnt) * n] for a in d["w"]]
            e = rel_errors(wg, wref, n)
            extra[form]["weight_rel_err_max"] = float(e.max())
            extra[form]["weight_rel_err_median"] = float(np.median(e))
            extra[form]["oracle_sample"] = cnt
            extra[form]["consistency_rel_err"] = consistency(d["w"], W["ja"], W["x"], W["y"], n)
            log(f"[{form}] weight rel err max {e.max():.3e} median {np.median(e):.3e}")

    ok = [f for f in res if extra[f].get("weight_rel_err_max", 0.0) <= 1e-8]
    if ok and head not in ok:
        head = min(ok, key=lambda f: res[f]["ms"])
        nt, r = res[head]["nt"], res[head]["r"]
        ms = res[head]["ms"]
        value = nst / (ms * 1e-3)
        achieved_tflops = flop * nst / (ms * 1e-3) / 1e12
    if args.profile_copy:
        # a plain device copy of known size, for scaling the HBM counters
        src = torch.empty(32 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 256 MiB device copies")
    kernel_tag = "assemble_%s_n35_1m" % head
    line = {
        "metric": "stencils/s assembled, 1M nodes n=35; mixed vs FP64 at <=1e-8 weight rel. err",
        "value": value,
        "unit": "stencils/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if head == "mixed" else "f64",
        "form": head,
        "data": "synthetic (seeded jittered grid, Hilbert order, host kNN)",
        "config": {"workload": "square_1m_n35: 1024x1024 jittered unit square, Hilbert, "
                               "n=35, PHS r^5 + P3 (45x45), Laplacian/dx/dy, form=" + head,
                   "nodes": nst, "n": n, "nt": nt, "ops": r},
        "roofline": {"bound": "mfma", "peak_kind": "fp64 (vector = matrix on MI355X)",
                     "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS,
                     "flop_per_stencil": flop, "traffic": read_traffic(kernel_tag)},
        "cpu_baseline": cpu,
        "forms": extra,
        "host_prep_s": W["host_s"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
