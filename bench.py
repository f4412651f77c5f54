#!/usr/bin/env python3
"""E-step throughput on one MI355X: the 3D-classification workload (20,000 particles of 256^2,
K=4 references of 256^3 at 1.1 A/px, HEALPix order 3, offsets 6/1.5 px): the plain FP32 path
and the split-precision path (FP16 hi+lo cross term on MFMA) timed in the same run.

A step is one full E-step iteration over all particles, which are synthesised once on the GPU
and stay resident in HBM: reference upload + coarse projections, preprocessing, both sampling
passes, significance, weighted sums and backprojection. The coarse-diff2 GEMM (the dominant
kernel) is timed with HIP events on the library's own stream inside the step.

The CPU baseline times the FP64 oracle (the restatement of relion_refine's CPU E-step) on a
bounded sample of the same workload with one particle per thread.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "t2-relion_amd", "python"))
import t2x_py as T  # noqa: E402
from t2x_py import workloads as W  # noqa: E402

METRIC = "E-step particles/s, 1 MI355X: plain FP32 vs mixed+learned at equal FSC-0.143"
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2500.0, "fp16x3": 2500.0}
PRECISIONS = {"fp32": 0, "bf16x3": 1, "fp16x3": 2}
DTYPES = {"fp32": "fp32", "bf16x3": "bf16x3 (bf16 hi+lo MFMA, fp32 accumulate, fp64 normalisation)",
          "fp16x3": "fp16x3 (fp16 hi+lo MFMA, fp32 accumulate, fp64 normalisation)"}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_dataset(lib, dev, w, n, cfg, sz, seed):
    rng = np.random.default_rng(seed)
    K, nb = cfg.nr_classes, w["n_blobs"]
    blob_sets = [W.gen_blobs(rng, nb, w["blob_radius"], w["blob_sigma"], w["blob_amp"]) for _ in range(K)]
    allb = (T.Blob * (nb * K))()
    for c in range(K):
        for k in range(nb):
            allb[c * nb + k] = blob_sets[c][k]
    boff = (C.c_int * (K + 1))(*[c * nb for c in range(K + 1)])
    pin, tr = W.gen_particles(rng, n, K, w["max_shift"])
    N = cfg.ori_size
    dimg = C.c_void_p()
    T.check(lib.t2x_device_alloc(dev, n * N * N * 4, C.byref(dimg)), "alloc images")
    snr = np.full(n, w["snr"])
    noise_var = np.zeros(n)
    T.check(lib.t2x_synth_particles_device(dev, N, cfg.pixel_size, n, allb, boff, pin,
                                           T.dptr(tr["rot"]), T.dptr(tr["tilt"]), T.dptr(tr["psi"]),
                                           T.dptr(tr["x"]), T.dptr(tr["y"]),
                                           tr["cls"].ctypes.data_as(C.POINTER(C.c_int)),
                                           T.dptr(snr), seed + 7, dimg, T.dptr(noise_var)),
            "synth particles")
    pad, rmax = sz["pad_size"], sz["r_max"]
    nvox = pad * pad * (pad // 2 + 1)
    refs = []
    for c in range(K):
        dre, dim = C.c_void_p(), C.c_void_p()
        T.check(lib.t2x_device_alloc(dev, nvox * 4, C.byref(dre)), "alloc ref")
        T.check(lib.t2x_device_alloc(dev, nvox * 4, C.byref(dim)), "alloc ref")
        one = (T.Blob * nb)(*[blob_sets[c][k] for k in range(nb)])
        T.check(lib.t2x_synth_reference_device(dev, N, cfg.pixel_size, pad, rmax, 2.0, one, nb,
                                               w["resolution"], dre, dim), "synth reference")
        refs.append((dre, dim))
    info = T.SamplingInfo()
    T.check(lib.t2x_sampling_info_get(C.byref(cfg), C.byref(info)), "sampling")
    sigma2, pdf_class, pdf_dir = W.model_arrays(cfg, info.nr_dir, noise_var)
    return dict(dimg=dimg, pin=pin, truth=tr, refs=refs, sigma2=sigma2, pdf_class=pdf_class,
                pdf_dir=pdf_dir, info=info, allb=allb, boff=boff, blob_sets=blob_sets,
                noise_var=noise_var)


def run_step(lib, dev, cfg, sz, ds, n, outs, ws):
    pad, rmax = sz["pad_size"], sz["r_max"]
    for c, (dre, dim) in enumerate(ds["refs"]):
        # AccProjector::initMdl of every iteration (device-resident model -> projector volume)
        T.check(lib.t2x_projector_set(dev, c, 3, pad, rmax, 2.0, C.cast(dre, C.POINTER(C.c_float)),
                                      C.cast(dim, C.POINTER(C.c_float))), "projector")
        T.check(lib.t2x_backprojector_clear(dev, c), "bp clear")
    model = T.Model(T.dptr(ds["sigma2"]), T.dptr(ds["pdf_class"]), T.dptr(ds["pdf_dir"]), None)
    T.check(lib.t2x_estep(dev, C.byref(cfg), C.byref(model), n, ds["dimg"], 1, ds["pin"], outs,
                          C.byref(ws), None), "estep")


def cpu_baseline(w, cfg, sz, seconds_target=20.0):
    """FP64 oracle on a bounded sample (same workload shape), one particle per thread."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    sys.path.insert(0, os.path.join(REPO, "tests"))
    try:
        import estep_harness as H
    except Exception as e:  # oracle not built
        return {"value": None, "unit": "particles/s", "cores": 0, "kind": "port",
                "sample": f"unavailable: {e}"}
    ncpu = len(os.sched_getaffinity(0))
    threads = max(1, min(16, ncpu))
    # one particle per thread, scaled down when a single particle would exceed the target
    wl = H.Workload(w_name(w), n=threads, seed_offset=99, current_size=cfg.current_size,
                    coarse_size=cfg.coarse_size)
    t0 = time.time()
    wl.run_oracle(nthreads=threads)
    dt = time.time() - t0
    return {"value": threads / dt, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": f"{threads} particles of the same workload (FP64 oracle, {threads} OpenMP "
                      f"threads, one particle each), {dt:.1f} s"}


def w_name(w):
    for k, v in W.WORKLOADS.items():
        if v is w or v == w:
            return k
    return "class256"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="class256")
    ap.add_argument("--particles", type=int, default=0, help="default: the workload's count")
    ap.add_argument("--warmup-particles", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy kernel whose FETCH_SIZE/WRITE_SIZE "
                         "counters calibrate the per-dispatch HBM traffic of the GEMM in --pmc runs")
    ap.add_argument("--precision", default="fp32,fp16x3",
                    help="comma list of cross-term variants to time (fp32, bf16x3, fp16x3); the "
                         "last one is the headline value, the others are reported beside it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the E-step "
              "shards over particles with one process per GPU, which is not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    w = W.WORKLOADS[args.workload]
    n = args.particles or w["n_particles"]
    cfg, sz = W.make_config(w)
    lib = T.load()
    dev = C.c_void_p()
    T.check(lib.t2x_device_open(0, C.byref(dev)), "open")
    log(f"workload {args.workload}: n={n} N={cfg.ori_size} K={cfg.nr_classes} current={cfg.current_size} "
        f"coarse={cfg.coarse_size} pad={sz['pad_size']}")
    t0 = time.time()
    ds = build_dataset(lib, dev, w, n, cfg, sz, w["seed"])
    log(f"dataset synthesised on the GPU in {time.time() - t0:.1f} s")
    for c in range(cfg.nr_classes):
        T.check(lib.t2x_backprojector_init(dev, c, 3, sz["pad_size"], sz["r_max"], 2.0), "bp init")
    outs = (T.ParticleOut * n)()
    ws, wa = T.make_wsum(cfg.nr_classes, ds["info"].nr_dir, cfg.ori_size)
    if args.copy_calibration:
        src, dst = C.c_void_p(), C.c_void_p()
        T.check(lib.t2x_device_alloc(dev, 1 << 30, C.byref(src)), "alloc")
        T.check(lib.t2x_device_alloc(dev, 1 << 30, C.byref(dst)), "alloc")
        T.check(lib.t2x_device_copy(dev, dst, src, 1 << 30), "copy")
        T.check(lib.t2x_device_synchronize(dev))
        T.check(lib.t2x_device_free(dev, src))
        T.check(lib.t2x_device_free(dev, dst))
        log("copy calibration: 1 GiB device-to-device copy (t2x::k_copy16) done")
    nw = min(n, args.warmup_particles)
    runs = {}
    for pname in args.precision.split(","):
        cfg.precision = PRECISIONS[pname]
        for i in range(args.warmup):
            t = time.time()
            run_step(lib, dev, cfg, sz, ds, nw, outs, ws)
            T.check(lib.t2x_device_synchronize(dev))
            log(f"[{pname}] warmup {i}: {nw} particles in {time.time() - t:.2f} s")
        T.check(lib.t2x_stats_reset(dev))
        T.check(lib.t2x_device_synchronize(dev))
        t0 = time.time()
        for i in range(args.steps):
            run_step(lib, dev, cfg, sz, ds, n, outs, ws)
            log(f"[{pname}] step {i} done at {time.time() - t0:.2f} s")
        T.check(lib.t2x_device_synchronize(dev))
        dt = time.time() - t0
        st = T.Stats()
        T.check(lib.t2x_stats_get(dev, C.byref(st)))
        best = np.array([o.best_hidden for o in outs])
        runs[pname] = dict(dt=dt, st=st, value=n * args.steps / dt, best=best)
        log(f"[{pname}] {runs[pname]['value']:.1f} particles/s")
    head = args.precision.split(",")[-1]
    r = runs[head]
    st, dt, value = r["st"], r["dt"], r["value"]
    # dominant kernel: coarse diff2 GEMM; algorithmic 4*T*O*P flops per particle (x3 products
    # on 16-bit MFMA for the split-precision variants)
    launches = max(1, st.n_coarse_diff2_launches)
    nprod = 1 if head == "fp32" else 3
    flop_per_launch = nprod * st.coarse_diff2_flop / launches
    avg_ms = st.ms_coarse_gemm / launches
    peak = MFMA_PEAK_TFLOPS[head]
    achieved = flop_per_launch / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    traffic = os.environ.get("T2X_GEMM_TRAFFIC_BYTES")
    kname = {"fp32": "k_gemm_tn<EPI_DIFF2> (FP32 MFMA 16x16x4)",
             "bf16x3": "k_gemm_x3<bf16, EPI_DIFF2> (BF16 MFMA 16x16x32, 3 products)",
             "fp16x3": "k_gemm_x3<f16, EPI_DIFF2> (FP16 MFMA 16x16x32, 3 products)"}[head]
    roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": float(traffic) if traffic else None,
            "kernel": "coarse diff2 cross term, " + kname,
            "avg_launch_ms": avg_ms, "launches": launches, "flop_per_launch": flop_per_launch}
    stages = {k: getattr(st, k) for k in ("ms_preprocess", "ms_coarse_project", "ms_coarse_diff2",
                                           "ms_coarse_weights", "ms_fine_diff2", "ms_fine_weights",
                                           "ms_store")}
    nsig = st.sig_coarse / max(1, st.n_particles)
    cpu = None if args.no_cpu else cpu_baseline(w, cfg, sz)
    variants = {}
    for pname, rr in runs.items():
        variants[pname] = {"particles_per_s": rr["value"], "ms_per_step": rr["dt"] * 1e3 / args.steps,
                           "coarse_gemm_ms": rr["st"].ms_coarse_gemm,
                           "pose_agreement_vs_fp32": (float((rr["best"] == runs["fp32"]["best"]).mean())
                                                      if "fp32" in runs else None)}
    line = {
        "metric": METRIC, "value": value, "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": DTYPES[head],
        "data": "synthetic (GPU-generated blob volumes and CTF-modulated noisy projections)",
        "config": {"workload": f"{args.workload}: {n} particles {cfg.ori_size}^2 @ {cfg.pixel_size} A/px, "
                               f"K={cfg.nr_classes}, HEALPix order {cfg.healpix_order}, offsets "
                               f"{w['offset_range_px']}/{w['offset_step_px']} px, current_size "
                               f"{cfg.current_size}, coarse_size {cfg.coarse_size}, {head} cross term",
                   "global_batch": n, "parallelism": "single GPU",
                   "P_coarse": st.P_coarse, "P_fine": st.P_fine,
                   "mean_significant_coarse": nsig,
                   "fine_samples_per_particle": st.fine_samples / max(1, st.n_particles),
                   "stage_ms": stages},
        "variants": variants,
        "speedup_vs_plain_fp32": (value / runs["fp32"]["value"]) if "fp32" in runs else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    lib.t2x_device_close(dev)


if __name__ == "__main__":
    main()
