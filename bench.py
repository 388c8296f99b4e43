#!/usr/bin/env python3
"""Throughput of the D3Q19 TGV time step on one MI355X, as one JSON line.

Workload: uniform periodic D3Q19 BGK Taylor-Green vortex, 512^3, FP32 populations, Re 1600, u0 0.05
(two population buffers, 20.4 GB), the lattice's own closed-form initial state. A step is the pull
stream-collide over the whole lattice; the library runs two fused steps per launch (k_dense_step2: each
workgroup streams a column of 128 x 4, 64 x 8 or 32 x 16 cells over 256 or 128 z-planes and keeps the
intermediate step in LDS), the launch form picked by the library's autotuner (jp_autotune) before the
timed steps. W untimed steps, then K steps timed with HIP events on the context's stream (inside the
library, jp_bench). value = GLUPS (1e9 cell updates / s).

roofline: a launch must read the lattice once and write it once, whatever number of steps it fuses:
19 x 4 B read + 19 x 4 B written = 152 B per cell per launch (the 152 B per cell update of a single step).
achieved = 152 B x N^3 / (average launch duration), against MI355X's 8 TB/s HBM3E peak. The rate a
single-step kernel would need for the same GLUPS (152 B per cell update) is reported separately as
extra.effective_gbps_152B_per_update. traffic (measured HBM bytes per launch) comes from the committed
rocprofv3 counter summary for the launch form the autotuner kept (profiles/), scaled by a copy kernel's
counters; extra.hbm_gbps_actual = traffic / launch duration.
cpu_baseline: the CPU oracle (the reference step restated in OpenMP C++) on the host cores, on a
bounded sample of the same lattice update (128^3, a few hundred steps); extra.cpu_baseline_amr times the
oracle's AMR coarse step (static gold at coarse 32, the sensor-driven grid at coarse 48) with the
reference's work formula.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jp-solver_amd"))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
METRIC = "GLUPS + % HBM roofline, uniform & AMR TGV, FP32 vs FP16-store at matched err"
# measured HBM bytes per launch of the 512^3 FP32 lattice, per launch form (x, y, z-segment, lanes of the
# fused column; zeros: single steps)
TRAFFIC_FILES = {(128, 4, 256, 832): os.path.join(ROOT, "profiles", "uniform512_fp32_traffic.json"),
                 (64, 8, 256, 704): os.path.join(ROOT, "profiles", "uniform512_fp32_traffic_64x8.json"),
                 (32, 16, 256, 640): os.path.join(ROOT, "profiles", "uniform512_fp32_traffic_32x16.json"),
                 (128, 4, 128, 832): os.path.join(ROOT, "profiles", "uniform512_fp32_traffic_128x4x128.json"),
                 (64, 8, 128, 704): os.path.join(ROOT, "profiles", "uniform512_fp32_traffic_64x8x128.json"),
                 (32, 16, 128, 640): os.path.join(ROOT, "profiles", "uniform512_fp32_traffic_32x16x128.json"),
                 (0, 0, 0, 0): os.path.join(ROOT, "profiles", "uniform512_fp32_single_step_traffic.json")}


def cpu_baseline(target_s=15.0):
    import oracle

    n = 128
    o = oracle.Uniform(n, re=1600.0, u0=0.05)
    o.init()
    t1 = o.time_steps(2)
    steps = int(max(4, min(400, target_s / max(t1 / 2, 1e-6))))
    t = o.time_steps(steps)
    glups = n ** 3 * steps / t / 1e9
    return {"value": glups, "unit": "GLUPS", "cores": oracle.lib().orc_threads(), "kind": "port",
            "sample": f"oracle uniform {n}^3 FP32 Re 1600, {steps} steps in {t:.1f} s "
                      f"(OpenMP, {oracle.lib().orc_threads()} threads)",
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_amr(target_s=8.0):
    """The oracle's AMR coarse step on the host cores: static gold (coarse 32, --l1 all --l2 all) and the
    sensor-driven grid of the dynamic case at coarse 48, the smallest coarse edge where its three levels all
    exist (frac 0.3 / 0.1, regrid every 10 coarse steps).
    Work per coarse step is (cnb0 + 2 nf1 + 4 nf2) x 512 cell updates (src/tgv3d_amr.cu:843)."""
    import oracle

    out = {"cores": oracle.lib().orc_threads()}
    cases = {"amr_gold_32_128": dict(n=32, re=800.0, l1=oracle.MODE_ALL, l2=oracle.MODE_ALL, adapt=0),
             "amr_dynamic_48_192": dict(n=48, re=1600.0, l1=oracle.MODE_SENSOR, l2=oracle.MODE_SENSOR, frac1=0.3,
                                        frac2=0.1, adapt=10)}
    for name, kw in cases.items():
        adapt = kw.pop("adapt")
        a = oracle.Amr(u0=0.05, **kw)
        a.init()
        work, steps, t = 0.0, 0, 0.0
        while t < target_s and steps < 200:
            c = a.counts()
            t0 = time.time()
            a.step(1)
            if adapt and (steps + 1) % adapt == 0:
                a.regrid()
            t += time.time() - t0
            work += (c["cnb0"] + 2 * c["nf1"] + 4 * c["nf2"]) * 512
            steps += 1
        out[name] = {"value": work / t / 1e9, "unit": "GLUPS", "coarse_steps": steps, "seconds": t,
                     "counts": a.counts()}
    return out


def bench_uniform(n, prec, steps, warmup=20, pairs=True, pair_shape=-1):
    import jpabi

    g = jpabi.Uniform(n, re=1600.0, u0=0.05, prec=jpabi.PREC_NAMES[prec])
    g.init()
    g.set_step_pairs(pairs)
    if pairs and pair_shape >= 0:
        g.set_step_pair_shape(pair_shape)
    r = g.bench(steps, warmup=warmup)
    r["pair_shape"] = list(g.step_pair_shape())
    per_launch = g.steps_per_launch()
    g.close()
    # a fused launch reads and writes the lattice once for per_launch steps: frac_hbm is priced on those
    # compulsory bytes; gbps stays the rate at bytes_per_update per cell update
    r["steps_per_launch"] = per_launch
    r["compulsory_gbps"] = r["gbps"] / per_launch
    r["frac_hbm"] = r["compulsory_gbps"] / PEAK_HBM_GBPS
    return r


def bench_amr(n0, re, l1, l2, frac1, frac2, adapt, steps, prec="fp32", graph=False):
    """AMR coarse steps timed as the reference's --bench (src/tgv3d_amr.cu:839-864): 20 warm-up steps,
    regrids every `adapt` steps inside the timed loop for a sensor-driven run; work weighted 1:2:4."""
    import jpabi

    a = jpabi.Amr(n0, re=re, u0=0.05, l1=jpabi.MODE_NAMES[l1], l2=jpabi.MODE_NAMES[l2], frac1=frac1, frac2=frac2,
                  adapt=adapt, prec=jpabi.PREC_NAMES[prec])
    a.init()
    r = a.bench(steps, warmup=20, graph=graph)
    r["counts"] = a.counts()
    a.close()
    r["frac_hbm"] = r["gbps"] / PEAK_HBM_GBPS
    return r


def bench_ensemble(steps, graph=True, tune=True):
    import jpabi

    with open(os.path.join(ROOT, "jp-solver_amd", "ensemble32.json")) as f:
        members = json.load(f)["members"]
    e = jpabi.Ensemble(members)
    e.init()
    groups = e.autotune() if tune else None
    r = e.bench(steps, warmup=4, graph=graph)
    e.close()
    r["groups"] = groups
    r["frac_hbm"] = r["gbps"] / PEAK_HBM_GBPS
    return r


def bench_field(n=512, reps=3):
    """Frame output of the uniform lattice: macroscopic field, Q-criterion, speed and velocity on the device,
    then the arrays to the host (what one .vti frame needs, without the file write)."""
    import jpabi

    g = jpabi.Uniform(n, re=1600.0, u0=0.05)
    g.init()
    g.step(2)
    g.field(velocity=True)  # allocates the output buffers
    t0 = time.time()
    for _ in range(reps):
        g.field(velocity=True)
    dt = (time.time() - t0) / reps
    g.close()
    out_bytes = 5 * 4 * n ** 3
    return {"n": n, "ms_per_frame": dt * 1e3, "bytes_to_host": out_bytes, "host_gbps": out_bytes / dt / 1e9}


def bench_diag(n=512, reps=5):
    """Energy, enstrophy and mass of the uniform lattice: macroscopic field and FP64 reduction on the device,
    three doubles to the host (jp_diag), the per-sample cost of the driver's diagnostics."""
    import jpabi

    g = jpabi.Uniform(n, re=1600.0, u0=0.05)
    g.init()
    g.diag()  # allocates the field and partial buffers
    t0 = time.time()
    for _ in range(reps):
        g.diag()
    dt = (time.time() - t0) / reps
    g.close()
    return {"n": n, "ms_per_sample": dt * 1e3, "field_bytes": 16 * n ** 3}


def bench_amr_diag(reps=5):
    """Energy, enstrophy and mass of the dynamic AMR case (coarse 96, finest 384): composite field on the finest
    grid from the three levels and the FP64 reduction on the device, three doubles to the host (jp_diag)."""
    import jpabi

    a = jpabi.Amr(96, re=1600.0, u0=0.05, l1=jpabi.MODE_NAMES["sensor"], l2=jpabi.MODE_NAMES["sensor"], frac1=0.3,
                  frac2=0.1, adapt=10)
    a.init()
    a.diag()  # allocates the composite field
    t0 = time.time()
    for _ in range(reps):
        a.diag()
    dt = (time.time() - t0) / reps
    a.close()
    return {"n_finest": 384, "ms_per_sample": dt * 1e3, "field_bytes": 16 * 384 ** 3}


def extras(quick=False):
    """Secondary workloads of the same metric (reported beside the headline, never as `value`)."""
    out = {}
    jobs = {
        "uniform512_fp32_single_steps": lambda: bench_uniform(512, "fp32", 300 if quick else 500, pairs=False),
        "uniform512_fp16": lambda: bench_uniform(512, "fp16", 300 if quick else 1000),
        "uniform512_bf16": lambda: bench_uniform(512, "bf16", 300 if quick else 1000),
        # regrids every 10 coarse steps inside the timed loop (the _graph line: graphs, regrids on
        # device-resident counts)
        "amr_dynamic_96_384": lambda: bench_amr(96, 1600.0, "sensor", "sensor", 0.3, 0.1, 10, 100 if quick else 300),
        "amr_dynamic_96_384_graph": lambda: bench_amr(96, 1600.0, "sensor", "sensor", 0.3, 0.1, 10,
                                                      100 if quick else 300, graph=True),
        "amr_dynamic_96_384_fp16": lambda: bench_amr(96, 1600.0, "sensor", "sensor", 0.3, 0.1, 10,
                                                     100 if quick else 300, prec="fp16"),
        "amr_gold_64_256": lambda: bench_amr(64, 800.0, "all", "all", 0.35, 0.35, 100, 50 if quick else 200,
                                             graph=True),
        # the precision configuration's sizes: uniform 768^3 and dynamic AMR coarse 128 / finest 512
        "uniform768_fp32": lambda: bench_uniform(768, "fp32", 100 if quick else 200),
        "uniform768_fp16": lambda: bench_uniform(768, "fp16", 100 if quick else 200),
        "uniform768_bf16": lambda: bench_uniform(768, "bf16", 100 if quick else 200),
        "amr_dynamic_128_512": lambda: bench_amr(128, 1600.0, "sensor", "sensor", 0.3, 0.1, 10, 50 if quick else 100),
        "amr_dynamic_128_512_fp16": lambda: bench_amr(128, 1600.0, "sensor", "sensor", 0.3, 0.1, 10,
                                                      50 if quick else 100, prec="fp16"),
        "ensemble32": lambda: bench_ensemble(20 if quick else 60),
        "field_output_512": lambda: bench_field(512, 2 if quick else 3),
        "diag_512": lambda: bench_diag(512, 3 if quick else 5),
        "amr_diag_96_384": lambda: bench_amr_diag(3 if quick else 5),
    }
    for k, fn in jobs.items():
        t0 = time.time()
        try:
            out[k] = fn()
        except Exception as ex:  # an extra must not hide the headline measurement
            out[k] = {"error": str(ex)}
        out[k]["wall_s"] = time.time() - t0
        v = out[k].get("glups", out[k].get("ms_per_frame", out[k].get("ms_per_sample", float("nan"))))
        print(f"# extra {k}: {v:.3f}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--prec", default="fp32", choices=["fp32", "fp16", "bf16", "fp64"])
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the secondary workloads")
    ap.add_argument("--quick-extras", action="store_true", help="shorter secondary workloads")
    ap.add_argument("--pair-shape", type=int, default=-1,
                    help="--only uniform: the fused pair's column shape (0: 128 x 4, 1: 64 x 8, 2: 32 x 16, each "
                         "over 256 z-planes; 3-5: the same over 128 z-planes; -1: default)")
    ap.add_argument("--only", default="", help="run one secondary workload and print it (profiling): amr-dynamic, "
                    "amr-dynamic-graph, amr-gold (these take --prec, and --n as the coarse edge), ensemble, uniform, "
                    "uniform-single, field, diag, amr-diag")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import jpabi

    if args.only:
        # a plain copy of known size in the same process calibrates the HBM byte counters
        jpabi.probe_copy(1 << 30, iters=2)
        n0 = args.n if args.n != 512 else None
        fn = {"amr-dynamic": lambda: bench_amr(n0 or 96, 1600.0, "sensor", "sensor", 0.3, 0.1, 10, args.steps,
                                               prec=args.prec),
              "amr-dynamic-graph": lambda: bench_amr(n0 or 96, 1600.0, "sensor", "sensor", 0.3, 0.1, 10,
                                                     args.steps, prec=args.prec, graph=True),
              "amr-gold": lambda: bench_amr(n0 or 64, 800.0, "all", "all", 0.35, 0.35, 100, args.steps,
                                            prec=args.prec, graph=True),
              "ensemble": lambda: bench_ensemble(args.steps),
              "uniform": lambda: bench_uniform(args.n, args.prec, args.steps, pair_shape=args.pair_shape),
              "uniform-single": lambda: bench_uniform(args.n, args.prec, args.steps, pairs=False),
              "field": lambda: bench_field(args.n, 3),
              "diag": lambda: bench_diag(args.n, 5),
              "amr-diag": lambda: bench_amr_diag(5)}[args.only]
        print(json.dumps({"workload": args.only, **fn()}), flush=True)
        return
    prec = jpabi.PREC_NAMES[args.prec]
    n = args.n
    copy_gbps, _ = jpabi.probe_copy(4 << 30, iters=10)
    attainable = jpabi.probe_bandwidth(4 << 30, iters=10)
    g = jpabi.Uniform(n, re=1600.0, u0=0.05, prec=prec)
    g.init()
    # the library's autotuner times single steps and the fused column shapes that divide N, keeps the fastest
    shape = g.autotune()  # (x, y, z, lanes) of the fused column; zeros for single steps
    tuned = g.tuned
    print(f"# uniform {n}^3 {args.prec}: launch form {shape}, warmup {args.warmup}, timing {args.steps} steps",
          file=sys.stderr, flush=True)
    per_launch = tuned["steps_per_launch"]
    r = g.bench(args.steps, warmup=args.warmup, graph=bool(args.graph))
    E, ens, mass = g.diag()
    g.close()
    bpu = r["bytes_per_update"]
    ms_launch = r["ms_per_step"] * per_launch
    # compulsory bytes per launch: one read and one write of the lattice (152 B per cell in FP32)
    achieved = bpu * n ** 3 / (ms_launch * 1e-3) / 1e9
    # the rate a one-step-per-launch kernel would need for the same cell updates (152 B per update)
    effective = bpu * n ** 3 * per_launch / (ms_launch * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tfile = TRAFFIC_FILES.get(tuple(shape))
    if args.prec == "fp32" and n == 512:
        if tfile and os.path.exists(tfile):
            with open(tfile) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
            traffic_src = os.path.relpath(tfile, ROOT)
        else:
            traffic_src = f"no committed counter summary for launch form {tuple(shape)}"
    line = {
        "metric": METRIC,
        "value": r["glups"],
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp16": "f16-store/f32", "bf16": "bf16-store/f32", "fp64": "f64"}[args.prec],
        "data": "synthetic (closed-form TGV initial state, no dataset)",
        "config": {"workload": f"uniform D3Q19 BGK TGV {n}^3 {args.prec.upper()} populations, Re 1600, u0 0.05, "
                               f"{args.steps} timed steps", "n": n, "re": 1600.0, "u0": 0.05,
                   "precision": args.prec, "graph": bool(args.graph),
                   # the library's default device layout for the form (jp_uniform_params.plane_pad = 0)
                   "layout": "direction planes" if args.prec == "fp64" else "z-layers",
                   "launch": (f"two fused steps per launch, columns {shape[0]} x {shape[1]} x {shape[2]}, "
                              f"{shape[3]} lanes" if shape[0] else f"one step per launch, variant {tuned['variant']}")
                             + f", {tuned['waves_per_cu']} waves per CU (picked by jp_autotune)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBPS, "traffic": traffic},
        "extra": {"copy_kernel_gbps": copy_gbps, "attainable_hbm_gbps": attainable,
                  "steps_per_launch": per_launch, "ms_per_launch": ms_launch,
                  "compulsory_bytes_per_launch": bpu * n ** 3,
                  "frac_of_attainable_copy": achieved / attainable["copy"],
                  "effective_gbps_152B_per_update": effective,
                  "frac_effective_152B_per_update": effective / PEAK_HBM_GBPS,
                  "traffic_source": traffic_src,
                  "hbm_gbps_actual": traffic / (ms_launch * 1e-3) / 1e9 if traffic else None,
                  "frac_hbm_actual": traffic / (ms_launch * 1e-3) / 1e9 / PEAK_HBM_GBPS if traffic else None,
                  "frac_of_attainable_copy_actual": (traffic / (ms_launch * 1e-3) / 1e9 / attainable["copy"]
                                                     if traffic else None),
                  "bytes_per_update": bpu, "pool_gb": r["pool_gb"], "diag_after": [E, ens, mass]},
    }
    if not args.no_extras:
        line["extra"]["workloads"] = extras(quick=args.quick_extras)
    if not args.no_cpu:
        t0 = time.time()
        line["cpu_baseline"] = cpu_baseline()
        line["extra"]["cpu_baseline_amr"] = cpu_baseline_amr()
        line["extra"]["cpu_baseline_wall_s"] = time.time() - t0
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
