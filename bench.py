"""Benchmark of the PPM-PA hot path on one MI355X.

Workload "caustic": a Cornell box of five 1 x 1 m quad walls with a diffuse, a
mirror and a glass sphere, 1024x768 pixels, 2M photons per pass, max 16 bounces,
r_1 = 0.005 m, alpha = 2/3, seed 1.  One step = one PPM-PA pass (photon trace
+ map build + eye rays + gather + FP64 accumulation) on device-resident data.

  python bench.py [--gpus 1] [--steps 512] [--warmup 8]

Prints one JSON line.  value = photons emitted and gathered per second over
the timed passes (Mphotons/s).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=512)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-passes", type=int, default=8)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="after the timed region, fill one buffer of this many MiB and copy it device-to-device "
                         "(kernels of known traffic to scale profiler byte counters against; with HSA_ENABLE_SDMA=0 "
                         "the copy runs as a blit kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the PPM-PA passes would shard "
              "across GPUs by pass with one final all-reduce of the FP64 image, which is not built here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import ppmpa_amd as P
    from ppmpa_amd import _abi as A

    cfgd = P.CONFIGS["caustic"]
    sc = cfgd["scene"]()
    ctx = P.Context(0)
    ds = P.DeviceScene(ctx, sc)

    def cfg(flags=0, first=1):
        return P.pass_cfg(seed=cfgd["seed"], n_photon=cfgd["n_photon"], radius0=cfgd["radius0"],
                          alpha=cfgd["alpha"], max_depth=cfgd["max_depth"], first_pass=first, flags=flags)

    # warmup: a separate session (also builds the pass graph path once)
    if args.warmup > 0:
        w = P.Session(ctx, ds, cfg())
        w.run(args.warmup)
        w.close()

    # timed: passes 1..K of the schedule, hipGraph replay, device-resident
    s = P.Session(ctx, ds, cfg())
    ctx.sync()
    t0 = time.perf_counter()
    s.run(args.steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    t = s.timings()
    mean, var, npass = s.image()
    s.close()
    dev_s = t["ms_total"] / 1e3
    T = max(wall, dev_s)
    n_photon = cfgd["n_photon"]
    pix = sc.camera.xreso * sc.camera.yreso
    value = n_photon * args.steps / T / 1e6
    stored_per_pass = t["photons_stored"] / max(1, args.steps)

    # per-stage device time over a few passes spread over the schedule
    stage = P.Session(ctx, ds, cfg(flags=A.SESSION_STAGE_TIMING))
    stage.run(4)
    st = stage.timings()
    stage.close()
    ms = {k: st[k] / 4 for k in ("ms_photon_trace", "ms_map_build", "ms_eye_trace", "ms_gather", "ms_accumulate")}

    # roofline of the shade kernel (k_shade: gather + direct light, timed
    # alone between HIP events on the context stream): algorithmic bytes per pass = the
    # photon map read once (28 B/photon) + per query its record (48 B), its
    # sorted index (4 B) and its result (16 B).
    S = stored_per_pass
    Q = t["gather_queries"] / max(1, args.steps)
    gather_bytes = 28.0 * S + 68.0 * Q
    achieved = gather_bytes / (ms["ms_gather"] / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_shade", "kernel_ms": round(ms["ms_gather"], 4),
            "algorithmic_bytes": round(gather_bytes), "queries_per_pass": round(Q)}
    traffic = profiled_traffic("k_shade")
    if traffic is not None:
        roof["traffic"] = traffic
        roof["traffic_source"] = "profiles/pmc_traffic_caustic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled)"

    out = {
        "metric": "Mphotons/s traced+gathered",
        "value": round(value, 3),
        "unit": "Mphotons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(T * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "caustic: Cornell box (five 1x1 m quad walls) + diffuse/mirror/glass spheres, 1024x768, 2M photons/pass, "
                               "max 16 bounces, r1=0.005 m, alpha=2/3, seed 1",
                   "photons_per_pass": n_photon, "pixels": pix, "passes": args.steps,
                   "accumulation": "f64"},
        "gathered_px_per_s": round(pix * args.steps / T, 1),
        "stored_photons_per_pass": round(stored_per_pass, 1),
        "stage_ms_per_pass": {k[3:]: round(v, 4) for k, v in ms.items()},
        "image_mean": [round(float(x), 6) for x in mean.reshape(-1, 3).mean(0)],
        "roofline": roof,
    }

    if args.copy_probe_mb > 0:
        out["copy_probe_bytes"] = copy_probe(args.copy_probe_mb << 20)

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, cfgd, args.cpu_sample_passes)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


def copy_probe(nbytes):
    """hipMemset of nbytes (a fill kernel writing nbytes) then a
    device-to-device hipMemcpy of them (reads and writes nbytes), through the
    HIP runtime directly."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(C.byref(buf), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed in the copy probe")
    try:
        if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed in the copy probe")
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed in the copy probe")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed in the copy probe")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return nbytes


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (two rocprofv3 --pmc runs of this same workload), scaled by the factors
    recorded there; None when absent."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_caustic.md")
    if not os.path.exists(path):
        return None
    sf, sw = 1.0, 1.0
    for line in open(path):
        if line.startswith("scale_fetch:"):
            sf = float(line.split(":")[1])
        if line.startswith("scale_write:"):
            sw = float(line.split(":")[1])
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:  # kernel | dispatches | FETCH KiB | WRITE KiB | MB
            return round((float(cells[2]) * sf + float(cells[3]) * sw) * 1024)
    return None


def cpu_baseline(sc, cfgd, n_sample):
    """The FP64 oracle (a restatement of the reference's CPU path) on this
    host's cores, on a bounded sample: passes 1..n_sample (the widest radii,
    the most gather work) of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    import ppmpa_amd as P
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    osc = O.OScene(sc)
    c = P.pass_cfg(seed=cfgd["seed"], n_photon=cfgd["n_photon"], radius0=cfgd["radius0"], alpha=cfgd["alpha"],
                   max_depth=cfgd["max_depth"], first_pass=1)
    t0 = time.perf_counter()
    _, _, stored, ms_trace, ms_eye = osc.run_passes(c, n_sample, threads=threads)
    T = time.perf_counter() - t0
    return {"value": round(cfgd["n_photon"] * n_sample / T / 1e6, 4), "unit": "Mphotons/s", "cores": threads,
            "kind": "port", "sample": f"passes 1-{n_sample} of the workload, FP64 oracle, "
                                      f"{T:.1f} s; trace+kd {ms_trace / 1e3:.1f} s, eye+gather {ms_eye / 1e3:.1f} s"}


if __name__ == "__main__":
    main()
