"""Benchmark of the PPM-PA hot path on one MI355X.

Default workload "caustic": a Cornell box of five 1 x 1 m quad walls with a
diffuse, a mirror and a glass sphere, 1024x768 pixels, 2M photons per pass,
max 16 bounces, r_1 = 0.005 m, alpha = 2/3, seed 1.  One step = one PPM-PA pass
(photon trace + map build + eye rays + gather + FP64 accumulation) on
device-resident data.  Other workloads: "mesh" (the box with a 327,680-triangle
icosphere and 64 spheres, 1920x1080, 8M photons per pass, through the scene
BVH) and "two_spheres" (256x256, 200k photons per pass).

  python bench.py [--gpus 1] [--steps K] [--warmup 8] [--workload caustic]

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

DESCRIPTION = {
    "caustic": "caustic: Cornell box (five 1x1 m quad walls) + diffuse/mirror/glass spheres, 1024x768, "
               "2M photons/pass, max 16 bounces, r1=0.005 m, alpha=2/3, seed 1",
    "mesh": "mesh: Cornell box + icosphere of 327,680 triangles (radial noise 2%) + 64 spheres "
            "(diffuse/mirror/glass), 1920x1080, 8M photons/pass, max 16 bounces, r1=0.005 m, alpha=2/3, seed 1, "
            "scene BVH",
    "caustic_plane": "caustic_plane: the caustic workload with the five walls as infinite planes (the reference's "
                     "Plain shape) instead of 1x1 m quads, 1024x768, 2M photons/pass, max 16 bounces, r1=0.005 m, "
                     "alpha=2/3, seed 1",
    "two_spheres": "two_spheres: Cornell box + diffuse and mirror spheres, 256x256, 200k photons/pass, "
                   "max 8 bounces, r1=0.01 m, alpha=2/3, seed 1",
    "inverse": "inverse: caustic box with 32 Lambertian materials (3 wall + 29 spheres r=0.05) as unknowns, 512x512, "
               "1M photons/pass, 64 passes per render, target rendered from hidden albedos (seed 1001, pass seed 7); "
               "Adam (lr 0.02) on the 96 albedo channels and the light scale from 0.5 / 0.8, forward + adjoint",
    "batch": "batch: 64 Cornell-box rooms (random wall albedos, 2-512 spheres log-uniform), images 128^2-512^2, "
             "50k-2M photons/pass log-uniform, 32-256 passes, max 8 bounces, seed 1; all rooms in one launch per "
             "stage after autotuning, against the same rooms run one after another as plain sessions",
}

# stage -> (kernel, algorithmic bytes per pass from S stored photons, Q
# queries and P pixels)
STAGE_KERNEL = {
    "ms_photon_trace": ("k_trace_photons", lambda S, Q, P: 32.0 * S),
    "ms_eye_trace": ("k_eye", lambda S, Q, P: 48.0 * Q + 16.0 * P),
    "ms_gather": ("k_shade", lambda S, Q, P: 28.0 * S + 68.0 * Q),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed passes (default: the workload's schedule)")
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", choices=sorted(DESCRIPTION), default="caustic")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-3sigma", action="store_true", help="skip the time-to-3-sigma comparison")
    ap.add_argument("--cpu-sample-passes", type=int, default=32,
                    help="passes of the oracle's sample on all allowed cores (also the agreement's pass count)")
    ap.add_argument("--no-plane-walls", action="store_true",
                    help="caustic workload: skip the 64-pass run of the same box with infinite plane walls")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch workload: skip the one-room-after-another comparison")
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

    import ppmpa_amd as P

    if args.workload == "batch":
        print(json.dumps(run_batch(P, args)), flush=True)
        return
    if args.workload == "inverse":
        print(json.dumps(run_inverse(P, args)), flush=True)
        return
    cfgd = P.CONFIGS[args.workload]
    steps = args.steps if args.steps is not None else cfgd["passes"]
    sc = cfgd["scene"]()
    ctx = P.Context(0)
    ds = P.DeviceScene(ctx, sc)
    r = run_workload(P, ctx, ds, cfgd, steps, args.warmup)
    pix = sc.camera.xreso * sc.camera.yreso
    T = r["T"]
    value = cfgd["n_photon"] * steps / T / 1e6
    ms = r["ms"]
    S, Q = r["stored_per_pass"], r["queries_per_pass"]

    # roofline of the dominant kernel, timed alone between HIP events on the
    # context stream; its algorithmic bytes per pass are those of STAGE_KERNEL
    # (k_shade: the photon map read once, 28 B per stored photon, plus per
    # query its record 48 B, sorted index 4 B and result 16 B).
    stage = max(STAGE_KERNEL, key=lambda k: ms[k])
    kern, fbytes = STAGE_KERNEL[stage]
    nbytes = fbytes(S, Q, pix)
    achieved = nbytes / (ms[stage] / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": kern, "kernel_ms": round(ms[stage], 4), "algorithmic_bytes": round(nbytes),
            "queries_per_pass": round(Q)}
    traffic, src = profiled_traffic(kern, args.workload)
    if traffic is not None:
        roof["traffic"] = traffic
        roof["traffic_source"] = src + " (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a copy probe)"
    issue, isrc = profiled_issue(kern, args.workload)
    if issue is not None:
        roof["valu_busy"], roof["salu_busy"] = issue
        roof["issue_source"] = isrc + " (rocprofv3 SQ_INSTS_VALU / SQ_INSTS_SALU per dispatch over GRBM_GUI_ACTIVE)"

    out = {
        "metric": "Mphotons/s traced+gathered",
        "value": round(value, 3),
        "unit": "Mphotons/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(T * 1e3 / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": DESCRIPTION[args.workload], "photons_per_pass": cfgd["n_photon"], "pixels": pix,
                   "passes": steps, "accumulation": "f64"},
        "gathered_px_per_s": round(pix * steps / T, 1),
        "stored_photons_per_pass": round(S, 1),
        "stage_ms_per_pass": {k[3:]: round(v, 4) for k, v in ms.items()},
        "image_mean": r["image_mean"],
        "roofline": roof,
    }

    if args.copy_probe_mb > 0:
        out["copy_probe_bytes"] = copy_probe(args.copy_probe_mb << 20)

    if args.workload == "caustic" and not args.no_plane_walls:
        out["walls_plane"] = plane_walls(P, ctx, args.warmup)

    if not args.no_cpu_baseline:
        orc = {}
        out["cpu_baseline"] = cpu_baseline(sc, cfgd, args.workload, args.cpu_sample_passes, keep=orc)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
        if orc:
            out["oracle_agreement"] = oracle_agreement(P, ctx, ds, cfgd, orc)
    if args.workload == "caustic" and not args.no_3sigma and not args.no_cpu_baseline:
        out["time_to_3sigma"] = time_to_3sigma(P, ctx)
    print(json.dumps(out), flush=True)


def run_batch(P, args):
    """The batched workload: tune, run every room's schedule in the batch,
    then the same rooms one after another as plain sessions (same passes and
    seeds, so the same images and the same error)."""
    ctx = P.Context(0)
    items = P.batch_scenes(n=64, seed=1)
    if args.steps is not None:  # cap the passes per room
        items = [(sc, c, min(p, args.steps)) for sc, c, p in items]
    dss = [P.DeviceScene(ctx, sc) for sc, _, _ in items]

    def cfg(c):
        return P.pass_cfg(seed=c["seed"], n_photon=c["n_photon"], radius0=c["radius0"], alpha=c["alpha"],
                          max_depth=c["max_depth"])

    cfgs = [cfg(c) for _, c, _ in items]
    passes = [p for _, _, p in items]
    b = P.Batch(ctx, dss, cfgs, passes)
    t0 = time.perf_counter()
    launch = b.tune(rounds=2)
    tune_s = time.perf_counter() - t0
    ctx.sync()
    t0 = time.perf_counter()
    b.run()
    ctx.sync()
    T = time.perf_counter() - t0
    tb = b.timings()
    b.close()
    photons = sum(c["n_photon"] * p for (_, c, _), p in zip(items, passes))
    pixels = sum(sc.camera.xreso * sc.camera.yreso * p for (sc, _, _), p in zip(items, passes))
    # plain sessions, one room after another (graph replay per room); only
    # the runs are timed, session set-up is not
    T_seq = 0.0
    for ds, cf, p in ([] if args.no_sequential else zip(dss, cfgs, passes)):
        s = P.Session(ctx, ds, cf)
        s.run(1)  # captures the session's pass graph
        ctx.sync()
        t0 = time.perf_counter()
        s.run(p - 1)
        ctx.sync()
        T_seq += time.perf_counter() - t0
        s.close()
    T_seq *= sum(passes) / max(1, sum(passes) - len(passes))
    if args.no_sequential:
        T_seq = float("nan")
    value = photons / T / 1e6
    return {
        "metric": "Mphotons/s traced+gathered", "value": round(value, 3), "unit": "Mphotons/s", "n_gpus": 1,
        "steps": max(passes), "warmup": 0, "ms_per_step": round(T * 1e3 / max(passes), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": DESCRIPTION["batch"], "scenes": len(items), "photons_total": photons,
                   "pixel_passes_total": pixels, "rooms_with_bvh": sum(1 for d in dss if d.stats()[1] > 0),
                   "accumulation": "f64"},
        "gathered_px_per_s": round(pixels / T, 1),
        "batch_s": round(T, 3), "batch_device_s": round(tb["ms_total"] / 1e3, 3),
        "sequential_sessions_s": round(T_seq, 3), "speedup_vs_sequential": round(T_seq / T, 2),
        "autotuned_launch": launch, "tune_s": round(tune_s, 2),
    }


def run_inverse(P, args):
    """Inverse material recovery: Adam steps through the autograd pair
    (forward passes + adjoint replay) from a wrong guess towards the albedos
    that rendered the target."""
    import numpy as np
    import torch

    from ppmpa_amd.autograd import Renderer
    cfgd = P.CONFIGS["inverse"]
    steps = args.steps if args.steps is not None else 200
    hidden = P.inverse_albedos(1001)
    sc = P.inverse_scene(hidden, 512, 512, seed=1)
    ctx = P.Context(0)
    ds = P.DeviceScene(ctx, sc)

    def cfg(seed):
        return P.pass_cfg(seed=seed, n_photon=cfgd["n_photon"], radius0=cfgd["radius0"], alpha=cfgd["alpha"],
                          max_depth=cfgd["max_depth"])

    rt = Renderer(ctx, ds, cfg(7), cfgd["passes"])
    with torch.no_grad():
        target = rt(torch.tensor(hidden), torch.tensor(1.0, dtype=torch.float64))
    rt.close()
    r = Renderer(ctx, ds, cfg(1), cfgd["passes"])
    x = torch.full((32, 3), 0.5, dtype=torch.float64, requires_grad=True)
    light = torch.tensor(0.8, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([x, light], lr=0.02)
    err0 = float(np.sqrt(((x.detach().numpy() - hidden) ** 2).mean()))
    losses = []
    t_fwd = t_bwd = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad()
        ta = time.perf_counter()
        img = r(x, light)
        loss = ((img - target) ** 2).mean()
        tb = time.perf_counter()
        loss.backward()
        t_fwd += tb - ta
        t_bwd += time.perf_counter() - tb
        opt.step()
        with torch.no_grad():
            x.clamp_(0.02, 0.98)
            light.clamp_(0.1, 10.0)
        losses.append(loss.item())
    T = time.perf_counter() - t0
    r.close()
    err = float(np.sqrt(((x.detach().numpy() - hidden) ** 2).mean()))
    photons = cfgd["n_photon"] * cfgd["passes"] * steps
    return {
        "metric": "Mphotons/s traced+gathered", "value": round(photons / t_fwd / 1e6, 3), "unit": "Mphotons/s",
        "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": round(T * 1e3 / steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": DESCRIPTION["inverse"], "photons_per_pass": cfgd["n_photon"],
                   "passes_per_render": cfgd["passes"], "pixels": 512 * 512, "accumulation": "f64"},
        "forward_ms_per_step": round(t_fwd * 1e3 / steps, 2), "backward_ms_per_step": round(t_bwd * 1e3 / steps, 2),
        "loss_first": losses[0], "loss_last": losses[-1],
        "albedo_rms_error_initial": round(err0, 4), "albedo_rms_error_final": round(err, 4),
        "albedo_max_error_final": round(float(np.abs(x.detach().numpy() - hidden).max()), 4),
        "light_scale_final": round(light.item(), 4), "light_scale_true": 1.0,
    }


def run_workload(P, ctx, ds, cfgd, steps, warmup):
    """Warmup session, then `steps` passes timed on the device (hipGraph
    replay) and by the wall clock, then 4 passes with per-stage HIP events."""
    from ppmpa_amd import _abi as A

    def cfg(flags=0):
        return P.pass_cfg(seed=cfgd["seed"], n_photon=cfgd["n_photon"], radius0=cfgd["radius0"],
                          alpha=cfgd["alpha"], max_depth=cfgd["max_depth"], first_pass=1, flags=flags)

    if warmup > 0:
        w = P.Session(ctx, ds, cfg())
        w.run(warmup)
        w.close()
    s = P.Session(ctx, ds, cfg())
    ctx.sync()
    t0 = time.perf_counter()
    s.run(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    t = s.timings()
    mean, _, _ = s.image(False)
    s.close()
    T = max(wall, t["ms_total"] / 1e3)
    stage = P.Session(ctx, ds, cfg(flags=A.SESSION_STAGE_TIMING))
    stage.run(4)
    st = stage.timings()
    stage.close()
    ms = {k: st[k] / 4 for k in ("ms_photon_trace", "ms_map_build", "ms_eye_trace", "ms_gather", "ms_accumulate")}
    return {"T": T, "ms": ms, "stored_per_pass": t["photons_stored"] / max(1, steps),
            "queries_per_pass": t["gather_queries"] / max(1, steps),
            "image_mean": [round(float(x), 6) for x in mean.reshape(-1, 3).mean(0)]}


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


def profiled_traffic(kernel, workload):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    of this workload (two rocprofv3 --pmc runs), scaled by the factors
    recorded there; (None, None) when absent."""
    rel = os.path.join("profiles", f"pmc_traffic_{workload}.md")
    path = os.path.join(ROOT, rel)
    if not os.path.exists(path):
        return None, None
    sf, sw = 1.0, 1.0
    for line in open(path):
        if line.startswith("scale_fetch:"):
            sf = float(line.split(":")[1])
        if line.startswith("scale_write:"):
            sw = float(line.split(":")[1])
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:  # kernel | dispatches | FETCH KiB | WRITE KiB | MB
            return round((float(cells[2]) * sf + float(cells[3]) * sw) * 1024), rel
    return None, None


def profiled_issue(kernel, workload):
    """(valu_busy, salu_busy) of `kernel` from the committed issue-rate
    summary of this workload; (None, None) when absent."""
    rel = os.path.join("profiles", f"pmc_issue_{workload}.md")
    path = os.path.join(ROOT, rel)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 9 and cells[0] == kernel:
            return (float(cells[6]), float(cells[7])), rel
    return None, None


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return O


def _threads():
    """Every core this process may run on (its affinity mask)."""
    return max(1, len(os.sched_getaffinity(0)))


def host_info():
    return {"nproc": os.cpu_count(), "affinity_cores": _threads(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu": cpu_model()}


def plane_walls(P, ctx, warmup):
    """64 passes of the caustic box with infinite Plain walls (the
    reference's wall shape) timed like the main run: the photons that leave
    through the open front land on the planes' far parts, so a pass stores
    more of them."""
    cfgd = P.CONFIGS["caustic_plane"]
    sc = cfgd["scene"]()
    ds = P.DeviceScene(ctx, sc)
    r = run_workload(P, ctx, ds, cfgd, 64, warmup)
    return {"workload": DESCRIPTION["caustic_plane"], "passes": 64,
            "value": round(cfgd["n_photon"] * 64 / r["T"] / 1e6, 3), "unit": "Mphotons/s",
            "ms_per_step": round(r["T"] * 1e3 / 64, 4), "stored_photons_per_pass": round(r["stored_per_pass"], 1),
            "stage_ms_per_pass": {k[3:]: round(v, 4) for k, v in r["ms"].items()}}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, cfgd, workload, n_sample, keep=None):
    """The FP64 oracle (a restatement of the reference's CPU path, which
    tests every shape for every ray) on this host's cores, on a bounded
    sample of the same workload."""
    O = _oracle()
    import ppmpa_amd as P
    threads = _threads()
    osc = O.OScene(sc)
    c = P.pass_cfg(seed=cfgd["seed"], n_photon=cfgd["n_photon"], radius0=cfgd["radius0"], alpha=cfgd["alpha"],
                   max_depth=cfgd["max_depth"], first_pass=1)
    if workload == "mesh":
        # one pass over 2M pixels with 25 shadow rays per hit against 327,750
        # shapes is out of reach; the sample is photon tracing alone
        n = 400
        t0 = time.perf_counter()
        osc.trace_photons(cfgd["seed"], 1, n, True, cfgd["max_depth"])
        T = time.perf_counter() - t0
        return {"value": round(n / T / 1e6, 6), "unit": "Mphotons/s", "cores": threads, "kind": "port",
                "sample": f"{n} photons of pass 1 traced (no map, no eye pass), FP64 oracle testing all "
                          f"{sc.n_shapes} shapes per ray, {T:.1f} s; CPU {cpu_model()}"}
    n_sample = min(n_sample, cfgd["passes"])
    t0 = time.perf_counter()
    mean, var, _, ms_trace, ms_eye = osc.run_passes(c, n_sample, threads=threads)
    T = time.perf_counter() - t0
    if keep is not None:
        keep.update(mean=mean, var=var, passes=n_sample)
    # the same code on one core, over pass 1 alone
    t0 = time.perf_counter()
    osc.run_passes(c, 1, threads=1)
    T1 = time.perf_counter() - t0
    h = host_info()
    return {"value": round(cfgd["n_photon"] * n_sample / T / 1e6, 4), "unit": "Mphotons/s", "cores": threads,
            "kind": "port", "sample": f"passes 1-{n_sample} of the workload, FP64 oracle on {threads} threads "
                                      f"(all cores the process may use), {T:.1f} s; trace+kd {ms_trace / 1e3:.1f} s, "
                                      f"eye+gather {ms_eye / 1e3:.1f} s; CPU {h['cpu']}",
            "one_core": {"value": round(cfgd["n_photon"] / T1 / 1e6, 4), "unit": "Mphotons/s",
                         "sample": f"pass 1 on 1 thread, {T1:.1f} s"},
            "host": h}


def oracle_agreement(P, ctx, ds, cfgd, orc):
    """The GPU's image of the same passes as the oracle's sample, on an
    independent random stream (seed + 1000): the fraction of pixels whose
    means agree within 3 sigma of the per-pass spread of both sides (equal
    error at equal pass count, at this workload's own size)."""
    import numpy as np
    K = orc["passes"]
    cfg = P.pass_cfg(seed=cfgd["seed"] + 1000, n_photon=cfgd["n_photon"], radius0=cfgd["radius0"],
                     alpha=cfgd["alpha"], max_depth=cfgd["max_depth"], first_pass=1)
    s = P.Session(ctx, ds, cfg)
    s.run(K)
    mg, vg, n = s.image()
    s.close()
    mo, vo = orc["mean"], orc["var"]
    frac, expected, sigma, ncomp = band_fraction(mg, vg, mo, vo, K, K)
    rel = float(np.abs(mg.mean() - mo.mean()) / mo.mean())
    return {"passes": K, "frac_pixels_within_3sigma": round(frac, 4),
            "expected_frac_student_t": round(expected, 4), "binomial_sigma": round(sigma, 5),
            "z": round((frac - expected) / sigma, 2), "passes_3sigma_rule": bool(frac >= expected - 3 * sigma),
            "pixels_compared": ncomp, "image_mean_rel_diff": round(rel, 6)}


def band_fraction(ma, va, mb, vb, ka, kb):
    """Pixel channels whose means agree within 3 standard errors,
    |ma - mb| <= 3 sqrt(va/ka + vb/kb), with the variances estimated from
    ka and kb passes; and the fraction expected if both sides estimate the
    same image: with estimated variances the ratio follows Student's t
    whose degrees of freedom are Welch-Satterthwaite's,
    nu = (a + b)^2 / (a^2/(ka-1) + b^2/(kb-1)), a = va/ka, b = vb/kb, so each
    channel lies within the band with probability P(|T_nu| <= 3) (0.9904 at
    ka = kb = 8, 0.9961 at 32; 0.9973 as nu grows).  Returns (observed,
    expected, binomial sigma of the observed fraction for independent
    channels, channels compared).  The rule is one-sided: a difference
    between the images lowers the fraction, so it passes when observed >=
    expected - 3 sigma.  Neighbouring pixels share photons and a pixel's
    channels its paths, so the binomial sigma understates the spread both
    ways."""
    import numpy as np
    from scipy import stats
    a, b = va / ka, vb / kb
    se = np.sqrt(a + b)
    ok = se > 0
    nu = (a + b)[ok] ** 2 / (a[ok] ** 2 / (ka - 1) + b[ok] ** 2 / (kb - 1))
    frac = float((np.abs(ma - mb)[ok] / se[ok] <= 3).mean())
    expected = float((1.0 - 2.0 * stats.t.sf(3.0, nu)).mean())
    n = int(ok.sum())
    sigma = float(np.sqrt(expected * (1.0 - expected) / max(n, 1)))
    return frac, expected, sigma, n


def _blocks(img, b=4):
    h, w, _ = img.shape
    return img[:h - h % b, :w - w % b].reshape(h // b, b, w // b, b, 3).mean(axis=(1, 3))


def time_to_3sigma(P, ctx):
    """Equal error at equal pass count: the two_spheres workload at full size
    (256x256, 200k photons/pass, 16 passes) on the GPU and in the FP64
    oracle, on independent random streams (seeds 1 and 1001).  Their images
    agree within the 3-sigma band of the per-pass spread (z on 4x4-pixel
    block means, and per pixel), so both reach the same error after the same
    passes; the times to get there are compared."""
    import numpy as np
    O = _oracle()
    cfgd = P.CONFIGS["two_spheres"]
    K = cfgd["passes"]
    sc = cfgd["scene"]()
    ds = P.DeviceScene(ctx, sc)

    def cfg(seed, first=1):
        return P.pass_cfg(seed=seed, n_photon=cfgd["n_photon"], radius0=cfgd["radius0"], alpha=cfgd["alpha"],
                          max_depth=cfgd["max_depth"], first_pass=first)

    s = P.Session(ctx, ds, cfg(1))
    ctx.sync()
    t0 = time.perf_counter()
    s.run(K)
    ctx.sync()
    gpu_s = time.perf_counter() - t0
    s.close()
    # per-pass GPU images (untimed) for the band
    s = P.Session(ctx, ds, cfg(1))
    imgs_g, prev = [], 0.0
    for _ in range(K):
        s.run(1)
        m, _, n = s.image(False)
        imgs_g.append(m * n - prev)
        prev = m * n
    s.close()
    osc = O.OScene(sc)
    threads = _threads()
    imgs_o = []
    t0 = time.perf_counter()
    for k in range(K):
        m1, _, _, _, _ = osc.run_passes(cfg(1001, k + 1), 1, threads=threads)
        imgs_o.append(m1)
    cpu_s = time.perf_counter() - t0
    G = np.stack([_blocks(x) for x in imgs_g])
    Ob = np.stack([_blocks(x) for x in imgs_o])
    se = np.sqrt(G.var(0, ddof=1) / K + Ob.var(0, ddof=1) / K)
    ok = se > 0
    fb = float((np.abs(G.mean(0) - Ob.mean(0))[ok] / se[ok] <= 3).mean())
    pg, po = np.stack(imgs_g), np.stack(imgs_o)
    fp, ep, sgp, _ = band_fraction(pg.mean(0), pg.var(0, ddof=1), po.mean(0), po.var(0, ddof=1), K, K)
    return {"workload": DESCRIPTION["two_spheres"], "passes": K, "frac_blocks_within_3sigma": round(fb, 4),
            "frac_pixels_within_3sigma": round(fp, 4), "expected_frac_student_t": round(ep, 4),
            "passes_3sigma_rule": bool(fp >= ep - 3 * sgp),
            "gpu_s": round(gpu_s, 4), "cpu_s": round(cpu_s, 3), "cpu_cores": threads,
            "speedup": round(cpu_s / gpu_s, 1)}


if __name__ == "__main__":
    main()
