"""Benchmark: pixellated-source log-evidence evaluations per second on one MI355X.

Workload: the 200x200-pixel / 60x60-source numerical-profile configuration
(qlens_amd.workloads.c2): 0.04"/pix with 2x2 supersampling, elliptical NFW (Gauss-Patterson
quadrature, tol 1e-6) + SIE + shear, three Sersic source components, Moffat 21x21 PSF, sigma =
0.01, curvature lambda = 1e3.  The data are synthetic, made on the GPU by the library's own mock
generator from a fixed seed.  A step is one qlx_eval call over one batch of lens-model
proposals drawn from a seeded 1% Gaussian ball around the truth; inputs (image, PSF, settings)
are resident in HBM before the timed region.

Paths measured on the same proposals:
  plain         FP64 everywhere (quadrature ray shooting, FP64 MFMA F and Cholesky), one proposal
                per call: the reference's one-likelihood-per-call pattern
  fp64_batched  the same FP64 arithmetic, a batch of proposals per call (reference precision)
  fast          `value`: mixed precision (F from an FP16 pool on MFMA with FP32 accumulation, FP32
                Cholesky, FP64 iterative refinement and log-determinants), a batch per call
|delta ln E| of the fast path against FP64 on the same proposals is reported per run (contract
< 0.1).  Beside them: the learned-deflection leg, the batched uneven-grid workload, the rooflines
of the Lb build and the SYRK, and the reference compiled from its sources timed on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--plain-steps P]
                    [--no-cpu-baseline] [--no-surrogate] [--no-batched] [--no-fp64-batched]
                    [--no-plain] [--cpu-reps N]
    python bench.py --copy-calibration MIB   # one device copy of MIB MiB (HBM counter scale)
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

# MI355X published peaks (dense)
PEAK_F16_TFLOPS = 2500.0
PEAK_FP64_TFLOPS = 78.6
# flops of one integrand evaluation in the reference's arithmetic (j_integrand_prime +
# NFW::kappa_rsq, profile.cpp:2778-2783, models.cpp:978-984; divisions and transcendentals count 1)
C_INT = 23
PEAK_HBM_GBS = 8000.0
METRIC = "log-evidence evals/s (200² img, 60² src): surrogate vs quad, |ΔlogE|<0.1"  # the benchmark's own label


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE) and scaled by
    a copy-kernel calibration; stored as 'key: value' lines in profiles/<name>.txt."""
    path = os.path.join(REPO, "profiles", name + ".txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, data, nrep):
    """The QLens reference compiled from its sources (oracle/_ref), or the oracle's C restatement
    when the reference build is absent, timed on this host: one warm-up evaluation of the workload,
    then the median of nrep - 1 more (wall time of the inversion only; mock data excluded)."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    model = cpu_model()
    try:
        from oracle import refrun
        if refrun.have_reference():
            with tempfile.TemporaryDirectory() as d:
                t0 = time.time()
                sc, _ = refrun.run_invert(case, data, d, nrep=nrep, threads=threads)
                wall = time.time() - t0
            sec = sc["median_seconds"]
            return {"value": round(1.0 / sec, 5), "unit": "evals/s", "cores": int(sc.get("omp_threads", threads)),
                    "kind": "reference", "cpu_model": model,
                    "sample": "the same 200^2/60^2 workload (truth model) by QLens built from its sources with -O3 "
                              "-fopenmp (default dense path): median of %d evaluations after 1 warm-up, %.2f s each "
                              "(min %.2f, max %.2f); %.1f s wall with setup" % (
                                  nrep - 1, sec, sc.get("min_seconds", sec), sc.get("max_seconds", sec), wall),
                    "minus2lnE": sc["minus2lnE"]}
    except Exception as e:  # noqa: BLE001 - any harness failure falls back to the restatement
        sys.stderr.write("reference baseline unavailable: %s\n" % e)
    try:
        from oracle import oracle as O
        O.invert(case, data, want_arrays=False)
        ts = []
        for _ in range(max(1, nrep - 1)):
            t0 = time.time()
            r, _ = O.invert(case, data, want_arrays=False)
            ts.append(time.time() - t0)
        sec = float(np.median(ts))
        return {"value": round(1.0 / sec, 5), "unit": "evals/s", "cores": threads, "kind": "port", "cpu_model": model,
                "sample": "median of %d evaluations after 1 warm-up by the oracle's C/OpenMP restatement of the "
                          "reference path" % len(ts),
                "minus2lnE": r["minus2lnE"]}
    except Exception as e:  # noqa: BLE001
        sys.stderr.write("oracle baseline unavailable: %s\n" % e)
        return None


KERNEL_TIMERS = ("syrk", "lbkernel", "raytrace", "cholesky", "solve", "lbuild", "total")


def timed_run(inv, props, batch, warmup, steps):
    """Run warmup + steps calls of `batch` proposals; returns (evals/s, ms/step, results of timed
    calls, {stage: [per-call HIP-event times (ms)]})."""
    k = 0
    for _ in range(warmup):
        inv.invert_image_surface_brightness_map(props[k:k + batch])
        k += batch
    # qlx_eval returns after its stream has drained (it reads the results back), so the clock
    # around the calls brackets completed device work; torch is not loaded in this process (its
    # own HIP runtime would be a second one next to the library's)
    results, kt = [], {n: [] for n in KERNEL_TIMERS}
    t0 = time.perf_counter()
    for _ in range(steps):
        results += inv.invert_image_surface_brightness_map(props[k:k + batch])
        k += batch
        for n in KERNEL_TIMERS:
            ms, _n = inv.kernel_time(n)
            if ms:
                kt[n].append(ms)
    el = time.perf_counter() - t0
    return steps * batch / el, el * 1e3 / steps, results, kt


def mean_s(v):
    return float(np.mean(v)) * 1e-3 if v else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=16)
    ap.add_argument("--cpu-reps", type=int, default=6, help="CPU baseline: 1 warm-up + the median of the rest")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-deflection leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-proposal uneven-grid leg")
    ap.add_argument("--no-fp64-batched", action="store_true", help="skip the 256-proposal FP64 leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-proposal FP64 leg")
    ap.add_argument("--ray-fp32", action="store_true", help="FP32 ray shooting in the fast path")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (known HBM bytes: it reads and "
                         "writes that amount once) to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X only: the batched path would shard proposals across GPUs "
              "(each GPU its own batches, only scalars returned), but no multi-GPU figure is produced "
              "by this build; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from qlens_amd import qlx
    from qlens_amd import workloads as W

    case = W.c2()
    inv = qlx.PixelInversion(0)
    nx = len(case["xvals"]) - 1
    inv.load_image(case["xvals"], case["yvals"], np.zeros((nx, nx)), sigma=case["sigma"])
    inv.set_psf(case["psf"])
    inv.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"])
    data = inv.mock_image(case["lenses"], case["sources"], case["sigma"], case["seed"])

    B = args.batch
    n_fast = (args.warmup + args.steps) * B
    props = W.proposals(case, max(n_fast, args.plain_steps + 2))

    # plain path: FP64, one proposal per call (the reference's one-likelihood-per-call pattern)
    plain = None
    if not args.no_plain:
        inv.load_workload(case, data, precision=qlx.QLX_PREC_FP64)
        plain_rate, plain_ms, _res, pk = timed_run(inv, props, 1, 2, args.plain_steps)
        ps = qlx.get_stats(inv)
        syrk_s = mean_s(pk["syrk"])
        plain = {"value": round(plain_rate, 2), "unit": "evals/s", "batch": 1, "ms_per_eval": round(plain_ms, 3),
                 "syrk_f64_ms": round(syrk_s * 1e3, 4) if syrk_s else None,
                 "syrk_f64_executed_tflops": round(ps["syrk_exec_flops"] / syrk_s / 1e12, 2) if syrk_s else None,
                 "syrk_f64_executed_frac_of_fp64_peak": round(ps["syrk_exec_flops"] / syrk_s / 1e12 / PEAK_FP64_TFLOPS, 4)
                 if syrk_s else None,
                 "note": "executed = MFMA flops of the nonzero 64x64 blocks; the dense-equivalent count "
                         "N_img*N_src*(N_src+1) is not a rate of this kernel"}

    # fast path: mixed precision, B proposals per call.  FP64 ray shooting: the FP32 quadrature
    # (--ray-fp32) needs more Patterson levels at tol 1e-6 and is not faster on this GPU
    ray = qlx.QLX_RAY_FP32 if args.ray_fp32 else qlx.QLX_RAY_FP64
    inv.load_workload(case, data, precision=qlx.QLX_PREC_MIXED, ray_precision=ray)
    rate, ms_step, fast_res, kt = timed_run(inv, props, B, args.warmup, args.steps)
    st = qlx.get_stats(inv)
    stage_ms = {k: round(float(np.mean(kt[k])), 3) if kt[k] else None for k in KERNEL_TIMERS}
    bad = sorted(set(r["status"] for r in fast_res if r["status"] != 0))

    # accuracy: the fast path's first timed batch against FP64 on the same proposals; the FP64 call
    # of B proposals is itself timed (one warm-up call first): the reference-precision batched rate
    first = props[args.warmup * B:(args.warmup + 1) * B]
    inv.load_workload(case, data, precision=qlx.QLX_PREC_FP64)
    fp64_batched = None
    if not args.no_fp64_batched:
        inv.invert_image_surface_brightness_map(props[:B])
        t0 = time.perf_counter()
        ref = inv.invert_image_surface_brightness_map(first)
        el = time.perf_counter() - t0
        fp64_batched = {"value": round(B / el, 2), "unit": "evals/s", "batch": B, "ms_per_call": round(el * 1e3, 2),
                        "stage_ms": {k: round(inv.kernel_time(k)[0] or 0.0, 3) for k in KERNEL_TIMERS},
                        "status_errors": sorted(set(r["status"] for r in ref if r["status"] != 0))}
    else:
        ref = inv.invert_image_surface_brightness_map(first)
    dlnE = max(abs(f["minus2lnE"] - r["minus2lnE"]) / 2 for f, r in zip(fast_res[:B], ref))

    # learned-deflection leg: the same batched mixed-precision path with the NFW deflection from
    # the surrogate (trained here on 2 M GPU quadrature samples), against FP64 quadrature
    sur = None
    if not args.no_surrogate:
        t0 = time.time()
        sur_loss = inv.surrogate_train(n_samples=2000000, steps=12000, batch=8192, lr=2e-3, seed=7)
        train_s = time.time() - t0
        inv.load_workload(case, data, precision=qlx.QLX_PREC_MIXED, ray_precision=ray)
        inv.invert_image_surface_brightness_map(first, def_mode=qlx.QLX_DEF_SURROGATE)
        t0 = time.perf_counter()
        sres = inv.invert_image_surface_brightness_map(first, def_mode=qlx.QLX_DEF_SURROGATE)
        s_el = time.perf_counter() - t0
        s_dlnE = max(abs(f["minus2lnE"] - r["minus2lnE"]) / 2 for f, r in zip(sres, ref))
        sur = {"value": round(B / s_el, 2), "unit": "evals/s", "max_abs_dlnE_vs_fp64": float("%.4g" % s_dlnE),
               "meets_dlnE_contract": bool(s_dlnE < 0.1),
               "deflection_stage_ms": round((inv.kernel_time("surrogate")[0] or 0) + inv.kernel_time("raytrace")[0], 3),
               "quadrature_raytrace_ms": stage_ms["raytrace"], "train_seconds": round(train_s, 2),
               "train_samples": 2000000, "train_final_loss": float("%.3g" % sur_loss),
               "note": "BF16 4x128 MLP: ~1e-3 relative deflection error; not used for `value` unless it meets |dlnE| < 0.1"}

    # batched evidence over uneven source grids: 256 proposals around config 1's truth on the 200^2
    # image, each with its own 20^2..60^2 grid, in one call, FP64 and mixed (one warm-up call each)
    batched = None
    if not args.no_batched:
        bcase, bmodels, bgrids = W.c4()
        inv.load_image(bcase["xvals"], bcase["yvals"], np.zeros((200, 200)), sigma=bcase["sigma"])
        inv.set_psf(bcase["psf"])
        inv.set_settings(bcase["nsplit"], bcase["srcgrid"], bcase["src_n"], bcase["regparam"], bcase["integral_tol"])
        bdata = inv.mock_image(bcase["lenses"], bcase["sources"], bcase["sigma"], bcase["seed"])
        bt = {}
        for tag, prec in (("fp64", qlx.QLX_PREC_FP64), ("mixed", qlx.QLX_PREC_MIXED)):
            inv.load_workload(bcase, bdata, precision=prec)
            inv.invert_image_surface_brightness_map(bmodels, src_n=bgrids)
            t0 = time.perf_counter()
            bres = inv.invert_image_surface_brightness_map(bmodels, src_n=bgrids)
            bt[tag] = (time.perf_counter() - t0, bres)
        bd = max(abs(m["minus2lnE"] - f["minus2lnE"]) / 2 for m, f in zip(bt["mixed"][1], bt["fp64"][1]))
        batched = {"workload": "256 proposals (1% ball) around config 1's SIE+shear truth, 200x200 img 0.05\"/pix "
                               "nsplit 2, Gaussian PSF, per-item square source grids",
                   "evals": len(bmodels), "grid_sides": [20, 30, 40, 50, 60],
                   "fp64_evals_per_s": round(len(bmodels) / bt["fp64"][0], 2),
                   "mixed_evals_per_s": round(len(bmodels) / bt["mixed"][0], 2),
                   "max_abs_dlnE_mixed_vs_fp64": float("%.3g" % bd),
                   "status_errors": sorted(set(r["status"] for t in bt.values() for r in t[1] if r["status"] != 0))}

    # rooflines, per launch (= one wave of B items), from HIP-event times on the library's stream.
    # Lb build: the largest kernel; HBM-bound by algorithm: it writes the FP16 pool of blurred rows
    # (bytes counted from the support masks) and reads only the CSR L rows (under 3 % of that).
    # SYRK: block-sparse MFMA work, 2 * 64 * 128^2 flops per pair of nonzero 64 x 128 pool blocks of
    # one row chunk (the dense-equivalent count would exceed the peak and is not a rate).
    lb_s, syrk_s, ray_s = mean_s(kt["lbkernel"]), mean_s(kt["syrk"]), mean_s(kt["raytrace"])
    roof, roof_syrk, roof_ray = None, None, None
    # ray shooting: VALU-bound; algorithmic flops = counted integrand evaluations x C_INT, the
    # reference's arithmetic per evaluation of j_integrand_prime + NFW::kappa_rsq
    if ray_s and st["integrand_evals"] > 0:
        achieved = st["integrand_evals"] * C_INT / ray_s / 1e12
        roof_ray = {"bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                    "kernel": "k_raytrace_map_b (ray shooting with the NFW Gauss-Patterson quadrature, FP64)",
                    "items_per_launch": st["n_items"], "integrand_evals_per_launch": st["integrand_evals"],
                    "c_int_flops_per_eval": C_INT, "avg_launch_ms": round(ray_s * 1e3, 3)}
    if lb_s:
        ach = st["lb_bytes"] / lb_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": read_traffic("lb_build_c2_traffic"),
                "kernel": "k_build_lb_b (PSF (x) L into the compact FP16 pool of blurred rows)",
                "items_per_launch": st["n_items"], "pool_bytes_per_launch": st["lb_bytes"],
                "avg_launch_ms": round(lb_s * 1e3, 3)}
    if syrk_s:
        achieved = st["syrk_exec_flops"] / syrk_s / 1e12
        roof_syrk = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_F16_TFLOPS, 4), "traffic": read_traffic("syrk_f16_c2_traffic"),
                     "kernel": "k_syrk_f16_b (F = Lb^T C Lb: FP16 pool, FP32 accumulation, 128x128 tiles over "
                               "nonzero blocks)",
                     "items_per_launch": st["n_items"], "flops_block_sparse_per_launch": st["syrk_exec_flops"],
                     "pool_bytes_per_launch": st["lb_bytes"], "avg_launch_ms": round(syrk_s * 1e3, 3)}

    out = {
        "metric": METRIC,
        "value": round(rate, 2),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 F / fp32 Cholesky / fp64 refinement, log-dets and rays",
        "data": "synthetic (GPU mock of the workload from a fixed seed; proposals from a seeded 1% ball)",
        "config": {"workload": "200x200 img 0.04\"/pix nsplit 2, 60x60 source, NFW(quadrature tol 1e-6)+SIE+shear, "
                               "Moffat 21x21 PSF, lambda 1e3",
                   "batch": B, "path": "mixed precision + batched (FP64 quadrature deflections; the surrogate "
                                       "leg is reported beside it)",
                   "n_img_active": st["n_img"], "n_src_active": st["n_src"]},
        "fp64_batched": fp64_batched,
        "plain_fp64": plain,
        "fast_stage_ms_per_step": stage_ms,
        "speedup_vs_plain": round(rate / plain["value"], 2) if plain else None,
        "max_abs_dlnE_vs_fp64": float("%.3g" % dlnE),
        "item_status_errors": bad,
        "surrogate_leg": sur,
        "batched_uneven_grids": batched,
        "roofline": roof,
        "roofline_syrk": roof_syrk,
        "roofline_raytrace": roof_ray,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(case, data, args.cpu_reps)
        if cb:
            cb.pop("minus2lnE", None)
            out["cpu_baseline"] = cb
            if plain:
                out["plain_vs_cpu"] = round(plain["value"] / cb["value"], 1)
            if fp64_batched:
                out["fp64_batched_vs_cpu"] = round(fp64_batched["value"] / cb["value"], 1)
    print(json.dumps(out), flush=True)
    inv.close()


if __name__ == "__main__":
    main()
