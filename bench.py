"""Benchmark of the propagation and conjunction hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--legs fp64,split,pairs,layouts,adjoint,cpu,refpoint,sources,perturb,copy]

Prints ONE JSON line. The headline `value` is FP64 object-steps/s of the mixed-regime workload:
100,000 objects (80% LEO at 30 s, 15% MEO and 5% GEO at 300 s, spans U(6, 24) h), JGM3
degree/order 70, inputs resident in HBM (ogp_propagate_device), startup included. One step = one
complete propagation of that population. Side legs (same line):
  split     50,000 LEO, 60 s, 7 days, degree 70: FP64 and split-precision object-steps/s and
            the split path's final-position error against the FP64 GPU path for all objects
  pairs     20,000 objects with 1,000 forced encounters, set 0 = first 2,000, 3 days at 60 s,
            degree 20: propagation, screen (10 km) and refinement; pair-steps/s of the screen
  layouts   the same objects as one object against the other 19,999 and as three sets (8,
            1,992, 18,000): the screen with square tiles, narrow tiles and the tuned choice
  adjoint   10,000 LEO, 60 s, 24 h, degree 20: the 6x6 d(final)/d(initial) per object by the
            discrete adjoint and by the variational equations; agreement and cost vs the forward
            propagation
  cpu       the FP64 host oracle on a stratified 2,500-object sample of the headline workload,
            timed on this host's cores (cpu_baseline; plus a single-core rate on 125 objects).
            Run last, it is also the checker: the GPU-vs-oracle difference on that sample, the
            full pairs workload propagated, screened and refined by the oracle against the GPU's
            candidate set, 32 objects of the split workload, and the adjoint against central
            differences of the oracle for two objects
  refpoint  10,000 LEO, 30 s, 720 steps, degree 70, end to end from host arrays (the workload
            of the reference's published H100 figure), and the same through the command-line
            tool as a whole process (JSON in, JSON out)
  sources   the headline workload with the Sun, the Moon and drag added to gravity (the
            acceleration-source extension point): object-steps/s and cost vs gravity alone
  perturb   the Monte-Carlo population generator: 10^8 satellites written to HBM (satellites/s
            and the write bandwidth against the HBM peak)
  copy      a 1 GiB device copy, the calibration dispatch of the HBM byte counters (not default)
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "object-steps/s at deg-70 gravity: FP64 vs split-prec (≤1 m @7 d); pairs/s"
CPU_SAMPLE_STRIDE = 40  # cpu leg: every 40th object of the headline workload (~12 s on 16 threads)
CPU_SINGLE_STRIDE = 800  # single-core rate: every 800th object (~8 s on one thread)
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (equal to its FP64 matrix figure)
PEAK_FP32_TFLOPS = 157.3  # packed FP32 (v_pk_fma_f32), which these builds never emit
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # non-packed FP32 issues at the FP64 rate
PEAK_HBM_GBS = 8000.0


def terms(L):
    return (L + 1) * (L + 2) // 2 - 3


@contextlib.contextmanager
def quiet_gc():
    """A timed region without Python garbage-collection pauses: collected before, paused
    during (a full collection with torch loaded takes tens of ms; the timed work is the C
    library's)."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def flops_eval(L):
    """Algorithmic FLOPs of one force evaluation: 25 per harmonic term (2 <= n <= L) plus the
    per-column and per-evaluation work (FMA = 2)."""
    return 25 * terms(L) + 15 * L + 60


def flops_step(L, k=7):
    """Algorithmic FLOPs per GJ8 object-step: k distinct evaluations + predictor/corrector."""
    return k * flops_eval(L) + 228 + 33 * k


def flops_chunks(L, gj, evals):
    """Algorithmic FLOPs of the GJ8 chunk kernels over gj object-steps in which the kernels ran
    `evals` object-evaluations (ogp_timing.evaluations): a step's evaluations end at the
    corrector's bitwise fixed point, so k varies per step; each evaluation costs F_E + 33 (its
    corrector share) and each object-step 228 (predictor and history sums)."""
    return evals * (flops_eval(L) + 33) + 228 * gj


def profile_value(name, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/, or None."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic_per_object_step():
    """HBM bytes per GJ8 object-step from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated against a copy of known size), or None when it is absent."""
    return profile_value("gj8_chunk_mixed100k_deg70_hbm_bytes.md", "traffic_bytes_per_object_step")


def hw_rate(profile, key, units, seconds, peak):
    """Issued FLOP rate from the hardware counters: the committed summary's FLOPs per unit
    (SQ_INSTS_VALU_FLOPS_* x 64 lanes over a profiled run of the same workload) times the units
    this run processed, over this run's time."""
    per = profile_value(profile, key)
    if per is None or seconds <= 0:
        return None
    rate = per * units / seconds / 1e12
    return dict(flop_per_unit=per, tflops=rate, frac=rate / peak, source="profiles/" + profile)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--legs",
                   default="fp64,split,pairs,layouts,adjoint,cpu,refpoint,sources,perturb")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="0 = every CPU in the affinity mask, at most 16")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def leg_fp64(ogp, scenarios, torch, args):
    c1 = scenarios.config1()
    n = len(c1["r"])
    steps = c1["steps"]
    ctx = ogp.Context(degree=70)
    dev = torch.device("cuda:0")
    r0 = torch.tensor(np.ascontiguousarray(c1["r"].T), device=dev)
    v0 = torch.tensor(np.ascontiguousarray(c1["v"].T), device=dev)
    r = torch.empty_like(r0)
    v = torch.empty_like(v0)
    torch.cuda.synchronize()

    def run():
        ctx.propagate_device(r0, v0, n, c1["step"], steps, r, v)

    for i in range(args.warmup):
        run()
        log(f"fp64 warmup {i + 1}/{args.warmup}: {ctx.timing()['total_ms']:.1f} ms")
    torch.cuda.synchronize()
    tm = []
    with quiet_gc():
        t0 = time.perf_counter()
        for i in range(args.steps):
            run()
            tm.append(ctx.timing())
            log(f"fp64 step {i + 1}/{args.steps}: {tm[-1]['total_ms']:.1f} ms (device)")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    obj_steps = float(steps.sum())
    ms_per_step = wall * 1e3 / args.steps
    gj = tm[-1]["gj_object_steps"]
    steps_ms = sum(t["steps_ms"] for t in tm) / len(tm)
    launches = tm[-1]["step_launches"]
    # each launch's own duration (HIP events on the stream it runs on); the object slices run
    # concurrently, so the chip's rate is the per-launch rate times the mean number of launches
    # in flight (= algorithmic FLOPs of the GJ8 phase / the phase's elapsed time)
    avg_launch_ms = sum(t["launch_ms_sum"] for t in tm) / len(tm) / launches
    concurrency = avg_launch_ms * launches / steps_ms
    evals = tm[-1]["evaluations"]
    flops_launch = flops_chunks(70, gj, evals) / launches
    per_launch = flops_launch / (avg_launch_ms * 1e-3) / 1e12
    achieved = per_launch * concurrency
    dev_ms = sum(t["total_ms"] for t in tm) / len(tm)
    bps = measured_traffic_per_object_step()
    traffic = None if bps is None else bps * gj / launches
    out = dict(
        value=obj_steps / (ms_per_step * 1e-3),
        ms_per_step=ms_per_step,
        device=dict(total_ms=dev_ms, startup_ms=sum(t["startup_ms"] for t in tm) / len(tm),
                    gj8_ms=steps_ms, object_steps_per_s=obj_steps / (dev_ms * 1e-3)),
        roofline=dict(bound="valu", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                      frac=achieved / PEAK_FP64_TFLOPS, traffic=traffic,
                      kernel="gj8_chunk_kernel<kGravity>",
                      flops_per_object_step=flops_chunks(70, gj, evals) / gj,
                      evaluations_per_object_step=evals / gj,
                      flops_basis="F_E = 64,935 per executed evaluation + 33 corrector share, "
                                  "228 per object-step; evaluations counted on the device",
                      avg_launch_ms=avg_launch_ms, launches_per_step=launches,
                      launches_in_flight=concurrency, per_launch_tflops=per_launch,
                      achieved_basis="per-launch FLOP rate x mean launches in flight "
                                     "(= GJ8-phase FLOPs / GJ8-phase elapsed time)",
                      object_steps_per_launch=gj / launches,
                      traffic_unit="bytes per launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
                      algorithmic_bytes_per_launch=9.75 * gj / launches,
                      hw_fp64=hw_rate("gj8_chunk_mixed100k_deg70_counters.md",
                                      "hw_fp64_flop_per_object_step", gj, steps_ms * 1e-3,
                                      PEAK_FP64_TFLOPS)),
        object_steps=obj_steps,
        tuning={str(b): ctx.get_tuning(64 * 2 ** b)[0] for b in range(0, 11)},
    )
    # the same propagation with every step running all 7 corrector evaluations (the reference's
    # count): the fixed-point exit must leave every final state bit-identical
    r_fp, v_fp = r.clone(), v.clone()
    ctx.set_fixed_point_exit(False)
    run()
    t7 = ctx.timing()
    ctx.set_fixed_point_exit(True)
    out["all_corrections"] = dict(
        device_ms=t7["total_ms"], object_steps_per_s=obj_steps / (t7["total_ms"] * 1e-3),
        evaluations_per_object_step=t7["evaluations"] / t7["gj_object_steps"],
        bit_identical=bool(torch.equal(r, r_fp) and torch.equal(v, v_fp)))
    log(f"fp64 all 7 corrections: {t7['total_ms']:.1f} ms, bit-identical "
        f"{out['all_corrections']['bit_identical']}")
    # keep the final states of the stratified sample for the cpu leg's comparison
    idx = np.arange(0, n, CPU_SAMPLE_STRIDE)
    out["_sample"] = dict(idx=idx, r=r.cpu().numpy().T[idx].copy(), v=v.cpu().numpy().T[idx].copy(),
                          c1=c1)
    ctx.close()
    return out


def leg_split(ogp, scenarios, args):
    c2 = scenarios.config2()
    n = len(c2["r"])
    res = {}
    for prec in ("fp64", "split"):
        with ogp.Context(degree=70, precision=prec) as ctx:
            # tuning pass on a 1-day span, then the timed 7-day run
            ctx.propagate(c2["r"], c2["v"], c2["step"], np.minimum(c2["steps"], 1440))
            t0 = time.perf_counter()
            r, v = ctx.propagate(c2["r"], c2["v"], c2["step"], c2["steps"])
            wall = time.perf_counter() - t0
            t = ctx.timing()
            res[prec] = dict(r=r, v=v, dev_ms=t["total_ms"], wall=wall,
                             rate=float(c2["steps"].sum()) / (t["total_ms"] * 1e-3),
                             gj8_ms=t["steps_ms"], gj=t["gj_object_steps"],
                             evals=t["evaluations"],
                             launches=t["step_launches"], launch_ms_sum=t["launch_ms_sum"])
            log(f"split leg {prec}: {t['total_ms']:.0f} ms device, {res[prec]['rate']:.3e} obj-steps/s")
    err = np.linalg.norm(res["split"]["r"] - res["fp64"]["r"], axis=1)
    out = dict(workload="50,000 LEO, 60 s, 10,080 steps (7 d), JGM3 70x70",
               fp64_object_steps_per_s=res["fp64"]["rate"],
               split_object_steps_per_s=res["split"]["rate"],
               speedup=res["split"]["rate"] / res["fp64"]["rate"],
               split_vs_fp64_gpu_m=dict(max=float(err.max()), p99=float(np.percentile(err, 99)),
                                        median=float(np.median(err))))
    # split-path roofline: FP32 terms n >= 3 against the FP32 peak, the rest against FP64
    L, ns = 70, 2
    s = res["split"]
    k = s["evals"] / s["gj"]  # executed evaluations per object-step
    f32 = k * 25 * (terms(L) - terms(ns))
    f64 = flops_chunks(L, s["gj"], s["evals"]) / s["gj"] - f32
    t_launch = s["gj8_ms"] / s["launches"] * 1e-3
    per_launch = s["gj"] / s["launches"]
    bound_s = (f32 / (PEAK_FP32_TFLOPS * 1e12) + f64 / (PEAK_FP64_TFLOPS * 1e12)) * per_launch
    # chip-level: the GJ8 phase's bound time over its elapsed time (slices run concurrently)
    prof = "gj8_chunk_split_leo50k_7d_deg70_counters.md"
    hw32 = hw_rate(prof, "hw_fp32_flop_per_object_step_split", s["gj"], s["gj8_ms"] * 1e-3,
                   PEAK_FP32_TFLOPS)
    hw64 = hw_rate(prof, "hw_fp64_flop_per_object_step_split", s["gj"], s["gj8_ms"] * 1e-3,
                   PEAK_FP64_TFLOPS)
    # non-packed FP32 (the only FP32 form these builds emit) issues at the FP64 rate
    nonpacked = None
    if hw32 and hw64:
        nonpacked = (hw32["tflops"] + hw64["tflops"]) / PEAK_FP32_NONPACKED_TFLOPS
    out["split_roofline"] = dict(frac=bound_s / t_launch, fp32_flops_per_object_step=f32,
                                 fp64_flops_per_object_step=f64,
                                 evaluations_per_object_step=k,
                                 fp64_path_evaluations_per_object_step=(
                                     res["fp64"]["evals"] / res["fp64"]["gj"]),
                                 elapsed_ms_per_launch=t_launch * 1e3,
                                 avg_launch_ms=s["launch_ms_sum"] / s["launches"],
                                 frac_basis="algorithmic FP32 at the packed 157.3 TFLOP/s peak, "
                                            "FP64 at 78.6",
                                 algorithmic_frac_nonpacked_fp32=(f32 + f64) / (
                                     PEAK_FP32_NONPACKED_TFLOPS * 1e12) * per_launch / t_launch,
                                 hw_fp32=hw32, hw_fp64=hw64,
                                 hw_issue_frac_nonpacked=nonpacked,
                                 ceiling=("non-packed FP32 issues at the FP64 rate and the split "
                                          "path issues as many FLOPs per object-step as the "
                                          "FP64 path (9% more instructions), so its ceiling is "
                                          "its own issue time without waits: with the FP64 "
                                          "path's VALU about 68% busy, 1.35-1.4x the FP64 "
                                          "path, not 3x"))
    # 32 objects for the cpu leg's check against the oracle
    k = np.arange(0, n, n // 32)[:32]
    out["_oracle_sample"] = dict(r0=c2["r"][k], v0=c2["v"][k], step=c2["step"][k],
                                 steps=c2["steps"][k], split_r=res["split"]["r"][k],
                                 fp64_r=res["fp64"]["r"][k])
    return out


def leg_pairs(ogp, scenarios):
    c4 = scenarios.config4()
    with ogp.Context(degree=20) as ctx:
        # warm-up on a short window (node tables, tuning, allocations)
        ca = ogp.ClosestApproach(ctx, c4["step"])
        ca.initialize(c4["r"], c4["v"], c4["sets"])
        ca.propagate(64)
        ca.find_closest(threshold=10e3)
        ca = ogp.ClosestApproach(ctx, c4["step"])
        ca.initialize(c4["r"], c4["v"], c4["sets"])
        with quiet_gc():
            t0 = time.perf_counter()
            ca.propagate(c4["steps"])
            t1 = time.perf_counter()
            prop_evals = ctx.timing()["evaluations"]
            res = ca.find_closest(threshold=10e3)
            t2 = time.perf_counter()
        t = ca.timing()
        # the same search screening every tile
        ctx.set_radial_filter(0)
        res_all = ca.find_closest(threshold=10e3)
        t_all = ca.timing()
        ctx.set_radial_filter(1)
    unfiltered = dict(screen_ms=t_all["screen_ms"],
                      same_results=all(np.array_equal(getattr(res, f), getattr(res_all, f))
                                       for f in ("a", "b", "step", "distance", "time", "pos_a",
                                                 "pos_b")))
    pair_steps = t["pair_steps"]
    # the pair-steps the screen kernels evaluated: the radial filter skips the tiles of
    # radially separated objects, whose pairs cannot come within the threshold
    screened = t["screened_pair_steps"]
    prop_gj = float(len(c4["r"])) * (c4["steps"] - 4)
    gpu = dict(a=np.asarray(res.a), b=np.asarray(res.b), step=np.asarray(res.step),
               distance=np.asarray(res.distance), time=np.asarray(res.time))

    return dict(_oracle_sample=dict(c4=c4, gpu=gpu),
        workload="20,000 objects (1,000 forced pairs), set 0 = first 2,000, 3 d at 60 s, "
                 "JGM3 20x20, 10 km screen",
        pairs=int(t["pairs"]), pair_steps=pair_steps, candidates=int(t["candidates"]),
        results=len(res),
        screen_ms=t["screen_ms"], refine_ms=t["refine_ms"], propagate_ms=t["propagate_ms"],
        screened_pair_steps=screened, radial_filter_skipped=1.0 - screened / pair_steps,
        every_tile=unfiltered,
        # evaluated pair-steps per second of screen time (sorting and gathering included)
        pair_steps_per_s=screened / (t["screen_ms"] * 1e-3),
        # every pair-step resolved (evaluated or skipped) per second of screen time
        pair_steps_resolved_per_s=pair_steps / (t["screen_ms"] * 1e-3),
        pairs_per_s_end_to_end=t["pairs"] / (t2 - t0),
        propagate_wall_s=t1 - t0, find_wall_s=t2 - t1,
        screen_roofline=dict(bound="valu",
                             achieved=9 * screened / (t["screen_ms"] * 1e-3) / 1e12,
                             peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                             frac=9 * screened / (t["screen_ms"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                             basis="9 FLOPs per evaluated pair-step",
                             hw_fp64=hw_rate("screen_pairs20k_3d_counters.md",
                                             "hw_fp64_flop_per_pair_step", screened,
                                             t["screen_ms"] * 1e-3, PEAK_FP64_TFLOPS)),
        propagate_roofline=dict(
            bound="valu", unit="TFLOP/s", peak=PEAK_FP64_TFLOPS,
            flops_per_object_step=flops_chunks(20, prop_gj, prop_evals) / prop_gj,
            evaluations_per_object_step=prop_evals / prop_gj,
            achieved=flops_chunks(20, prop_gj, prop_evals) / (t["propagate_ms"] * 1e-3) / 1e12,
            frac=flops_chunks(20, prop_gj, prop_evals) / (t["propagate_ms"] * 1e-3) / 1e12
            / PEAK_FP64_TFLOPS,
            hw_fp64=hw_rate("gj8_chunk_conj20k_3d_deg20_counters.md",
                            "hw_fp64_flop_per_object_step", prop_gj, t["propagate_ms"] * 1e-3,
                            PEAK_FP64_TFLOPS)))


def leg_pairs_shapes(ogp, c4):
    """The same 20,000 objects and 3-day window with other set layouts: one object against the
    other 19,999, and three sets (8, 1,992 and 18,000 objects). The screen with square tiles,
    with narrow tiles and with the tuner's choice per set rectangle; the candidate sets must be
    equal."""
    out = {}
    n = len(c4["r"])
    for name, sizes in (("1_vs_19999", (1, n - 1)), ("sets_8_1992_18000", (8, 1992, n - 2000))):
        sets = np.repeat(np.arange(len(sizes), dtype=np.uint32), sizes)
        with ogp.Context(degree=20) as ctx:
            ca = ogp.ClosestApproach(ctx, c4["step"])
            ca.initialize(c4["r"], c4["v"], sets)
            ca.propagate(c4["steps"])
            rec = {}
            cands = []
            for label, shape in (("square", 1), ("narrow", 2), ("tuned", 0)):
                ctx.set_screen_shape(shape)
                ca.find_closest(threshold=10e3)  # the tuner's trial, allocations
                res = ca.find_closest(threshold=10e3)
                t = ca.timing()
                rec[label] = dict(screen_ms=t["screen_ms"],
                                  pair_steps_per_s=t["screened_pair_steps"] / (t["screen_ms"] * 1e-3),
                                  radial_filter_skipped=1.0 - t["screened_pair_steps"]
                                  / t["pair_steps"])
                cands.append((np.asarray(res.a).tobytes(), np.asarray(res.b).tobytes(),
                              np.asarray(res.step).tobytes()))
            rec["tuned_shape_first_set"] = "narrow" if ctx.get_screen_shape(sizes[0]) == 2 else "square"
            rec["pairs"] = int(t["pairs"])
            rec["candidates"] = int(t["candidates"])
            rec["same_candidates"] = all(c == cands[0] for c in cands)
        out[name] = rec
        log(f"pairs leg, {name}: square {rec['square']['screen_ms']:.2f} ms, narrow "
            f"{rec['narrow']['screen_ms']:.2f} ms, tuned {rec['tuned']['screen_ms']:.2f} ms")
    return out


def leg_cpu(sample, threads):
    """FP64 oracle on every 40th object of the headline workload (stratified over regimes), the
    cpu_baseline, on `threads` cores; then every 800th object on one core (the single-core rate).
    Returns it with the GPU-vs-oracle difference on the first sample."""
    from oracle import pyoracle as po
    c1, idx = sample["c1"], sample["idx"]
    o = po.Oracle(degree=70)
    t0 = time.perf_counter()
    r, v = o.propagate(po.GRAVITY, c1["r"][idx], c1["v"][idx], c1["step"][idx], c1["steps"][idx],
                       threads=threads)
    dt = time.perf_counter() - t0
    work = float(c1["steps"][idx].sum())
    dr = np.linalg.norm(r - sample["r"], axis=1)
    dv = np.linalg.norm(v - sample["v"], axis=1)
    one = np.arange(0, len(c1["r"]), CPU_SINGLE_STRIDE)
    t0 = time.perf_counter()
    o.propagate(po.GRAVITY, c1["r"][one], c1["v"][one], c1["step"][one], c1["steps"][one],
                threads=1)
    dt1 = time.perf_counter() - t0
    work1 = float(c1["steps"][one].sum())
    return dict(value=work / dt, unit="object-steps/s", cores=threads, kind="port",
                sample=f"{len(idx)} objects (every {CPU_SAMPLE_STRIDE}th of the 100,000-object mixed "
                       f"workload), "
                       f"{int(work)} object-steps, FP64 oracle, OpenMP over objects, {dt:.1f} s",
                single_core=dict(value=work1 / dt1, unit="object-steps/s",
                                 sample=f"{len(one)} objects (every {CPU_SINGLE_STRIDE}th), "
                                        f"{int(work1)} object-steps, 1 thread, {dt1:.1f} s"),
                cpu_model=_cpu_model(), nproc=os.cpu_count(),
                affinity_cpus=len(os.sched_getaffinity(0)),
                threads_rule="every CPU in this process's affinity mask, at most 16 (the CPU "
                             "share of one GPU on the GPU box, whose mask holds 16)"), \
        dict(objects=len(idx), max_dr_m=float(dr.max()), max_dv_mps=float(dv.max()))


def pairs_oracle_check(sample, threads):
    """The conjunction search of the pairs leg at full size against the oracle: the oracle
    propagates all 20,000 objects over the 3-day window (4,320 steps, degree 20), screens the
    36 M cross-set pairs at 10 km with the per-pair rule of the reference (in a blocked loop
    order) and refines its candidates. The candidate pair indices and their sampled minimum steps
    must be equal; TCA and miss distance are compared with the oracle's refinement."""
    from oracle import pyoracle as po
    c4, gpu = sample["c4"], sample["gpu"]
    o = po.Oracle(degree=20)
    t0 = time.perf_counter()
    _, _, traj = o.propagate(po.GRAVITY, c4["r"], c4["v"], c4["step"], c4["steps"],
                             trajectory=True, threads=threads)
    t1 = time.perf_counter()
    log(f"pairs oracle: propagation {t1 - t0:.1f} s")
    mind, mink = o.screen_blocked(traj, c4["sets"], 10e3, threads=threads)
    t2 = time.perf_counter()
    log(f"pairs oracle: screen {t2 - t1:.1f} s")
    cand = np.nonzero(mind < 10e3)[0]
    na = int((c4["sets"] == 0).sum())
    nb = len(c4["sets"]) - na
    oa, ob = cand // nb, na + cand % nb  # two sets: pair index = a * |B| + (b - |A|)
    same_set = (len(cand) == len(gpu["a"]) and np.array_equal(oa, gpu["a"])
                and np.array_equal(ob, gpu["b"]))
    out = dict(candidates_oracle=int(len(cand)), candidates_gpu=int(len(gpu["a"])),
               candidate_set_equal=bool(same_set),
               min_step_equal=bool(same_set and np.array_equal(mink[cand], gpu["step"])),
               nearest_to_threshold_m=float(np.abs(mind - 10e3).min()),
               oracle_propagate_s=t1 - t0, oracle_screen_s=t2 - t1, threads=threads)
    if same_set:
        dt = dd = 0.0
        for q in range(len(cand)):
            ref = o.refine(traj, c4["step"], int(oa[q]), int(ob[q]), int(mink[cand[q]]),
                           float(mind[cand[q]]))
            dt = max(dt, abs(ref["time"] - gpu["time"][q]))
            dd = max(dd, abs(ref["distance"] - gpu["distance"][q]))
        out.update(max_tca_diff_s=dt, max_miss_diff_m=dd)
    del traj
    return out


def oracle_checks(split_sample, adjoint_sample, threads):
    """Checks of the split and adjoint legs against the oracle, run within the cpu leg: the
    final positions of 32 objects of the split workload, and central differences of the
    oracle's propagation for two objects of the adjoint workload."""
    from oracle import pyoracle as po
    out = {}
    if split_sample is not None:
        s = split_sample
        t0 = time.perf_counter()
        ro, _ = po.Oracle(degree=70).propagate(po.GRAVITY, s["r0"], s["v0"], s["step"], s["steps"],
                                               threads=threads)
        log(f"split oracle sample: {time.perf_counter() - t0:.1f} s")
        out["split"] = dict(objects=len(ro),
                            split_max=float(np.linalg.norm(s["split_r"] - ro, axis=1).max()),
                            fp64_max=float(np.linalg.norm(s["fp64_r"] - ro, axis=1).max()))
    if adjoint_sample is not None:
        a = adjoint_sample
        o = po.Oracle(degree=20)
        n = len(a["r0"])
        fd = np.zeros((n, 6, 6))
        x0 = np.hstack([a["r0"], a["v0"]])
        for col in range(6):
            d = np.zeros((n, 6))
            d[:, col] = 1e-3 if col < 3 else 1e-6
            rp, vp = o.propagate(po.GRAVITY, (x0 + d)[:, :3], (x0 + d)[:, 3:], a["h"], a["steps"])
            rm, vm = o.propagate(po.GRAVITY, (x0 - d)[:, :3], (x0 - d)[:, 3:], a["h"], a["steps"])
            fd[:, :, col] = np.hstack([rp - rm, vp - vm]) / (2 * d[0, col])
        e = np.linalg.norm(a["J"] - fd, axis=(1, 2)) / np.linalg.norm(fd, axis=(1, 2))
        out["adjoint"] = dict(objects=n, max=float(e.max()))
    return out


def leg_refpoint(ogp, scenarios):
    r0, v0 = scenarios.leo(10_000, seed=1)
    t0 = time.perf_counter()
    with ogp.Context(degree=70) as ctx:
        ctx.propagate(r0, v0, 30.0, 720)
        t = ctx.timing()
    wall = time.perf_counter() - t0
    out = dict(workload="10,000 LEO, 30 s, 720 steps (6 h), JGM3 70x70, host arrays, "
                        "context creation included (first use of the wave tuner)",
               wall_s=wall, object_steps_per_s_end_to_end=7.2e6 / wall,
               device_ms=t["total_ms"], published_h100_wall_s=18.0,
               published_h100_object_steps_per_s=3.99e5)
    # the same workload through the command-line tool, whole process (JSON in, JSON out), as the
    # published figure is measured
    import subprocess
    import tempfile
    cli = os.path.join(ROOT, "opengpop_amd", "opengpop")
    out["cli_wall_s"] = None
    if os.path.exists(cli):
        with tempfile.TemporaryDirectory() as td:
            inp, res = os.path.join(td, "in.json"), os.path.join(td, "out.json")
            with open(inp, "w") as f:
                json.dump({"propagationStart": {"Year": 2025, "Month": 1, "Day": 1, "Hour": 0,
                                                "Min": 0, "Sec": 0.0},
                           "propagationSeconds": 21600, "positions": r0.tolist(),
                           "velocities": v0.tolist()}, f)
            t0 = time.perf_counter()
            subprocess.run([cli, "propagate", "-i", inp, "-o", res, "--data-dir",
                            os.path.join(ROOT, "data")], check=True, capture_output=True,
                           timeout=300)
            out["cli_wall_s"] = time.perf_counter() - t0
    return out


def leg_adjoint(ogp, scenarios):
    """Differentiation workload: 10,000 LEO, 60 s, 1,440 steps (24 h), JGM3 20x20: the 6x6
    d(final)/d(initial) per object by the adjoint (the forward pass recording the gravity
    Jacobians of the whole span, 4.8 GB; and with checkpoints every 64 steps) and by the
    variational equations, their agreement, and their cost against the plain propagation."""
    c3 = scenarios.config3()
    r0, v0, h, steps = c3["r"], c3["v"], 60.0, c3["steps"]
    with ogp.Context(degree=20) as ctx:
        ctx.propagate(r0[:640], v0[:640], h, steps[:640])  # node tables, wave tuning
        ctx.propagate(r0, v0, h, steps)
        fwd_ms = ctx.timing()["total_ms"]
        fwd_evals = ctx.timing()["evaluations"]
        _, _, J = ctx.adjoint_jacobian(r0, v0, h, steps)
        ta = ctx.adjoint_timing()
        _, _, J64 = ctx.adjoint_jacobian(r0, v0, h, steps, checkpoint_steps=64)
        tc = ctx.adjoint_timing()
        rng = np.random.default_rng(3)
        ctx.adjoint_vjp(r0, v0, h, steps, rng.normal(size=(len(r0), 3)),
                        rng.normal(size=(len(r0), 3)))
        tv = ctx.adjoint_timing()
        _, _, Phi = ctx.stm(r0, v0, h, steps)
        stm_ms = ctx.adjoint_timing()["stm_ms"]
    rel = np.linalg.norm(J - Phi, axis=(1, 2)) / np.linalg.norm(Phi, axis=(1, 2))
    gj = float(np.maximum(steps - 4, 0).sum())
    out = dict(workload="10,000 LEO, 60 s, 1,440 steps (24 h), JGM3 20x20, up to 7 corrector "
                        "evaluations per step (the propagation and the record pass stop at the "
                        "corrector's fixed point); Jacobians recorded in the forward pass",
               forward_ms=fwd_ms,
               jacobian_ms=ta["forward_ms"] + ta["adjoint_ms"],
               jacobian_forward_ms=ta["forward_ms"], jacobian_adjoint_ms=ta["adjoint_ms"],
               jacobian_checkpointed_ms=tc["forward_ms"] + tc["adjoint_ms"],
               checkpointed_equal=bool(np.array_equal(J, J64)),
               vjp_ms=tv["forward_ms"] + tv["adjoint_ms"], stm_ms=stm_ms,
               jacobian_cost_vs_forward=(ta["forward_ms"] + ta["adjoint_ms"]) / fwd_ms,
               vjp_cost_vs_forward=(tv["forward_ms"] + tv["adjoint_ms"]) / fwd_ms,
               stm_cost_vs_forward=stm_ms / fwd_ms,
               adjoint_vs_stm_rel=dict(max=float(rel.max()), median=float(np.median(rel))),
               forward_roofline=dict(
                   bound="valu", unit="TFLOP/s", peak=PEAK_FP64_TFLOPS,
                   evaluations_per_object_step=fwd_evals / gj,
                   achieved=flops_chunks(20, gj, fwd_evals) / (fwd_ms * 1e-3) / 1e12,
                   frac=flops_chunks(20, gj, fwd_evals) / (fwd_ms * 1e-3) / 1e12
                   / PEAK_FP64_TFLOPS,
                   basis="algorithmic FLOPs of the GJ8 steps (executed evaluations) over the "
                         "whole propagation (startup included)"),
               # the Jacobian's forward pass is the record pass over the whole span and the
               # gradient evaluations at its points (the recording startup included in the time)
               record_hw_fp64=hw_rate("gj8_record_adjoint10k_24h_deg20_counters.md",
                                      "hw_fp64_flop_per_record_object_step", gj,
                                      ta["forward_ms"] * 1e-3, PEAK_FP64_TFLOPS))
    # two objects for the cpu leg's central differences of the oracle's propagation
    k = [0, 5000]
    out["_oracle_sample"] = dict(r0=r0[k], v0=v0[k], h=h, steps=steps[k], J=J[k])
    return out


def leg_sources(ogp, scenarios, torch, gravity_only_ms):
    """The headline workload with the Sun, the Moon and drag (C_D A / m = 0.01 m^2/kg) added to
    the degree-70 gravity inside the integrator kernels: object-steps/s and the cost relative to
    gravity alone."""
    c1 = scenarios.config1()
    n = len(c1["r"])
    dev = torch.device("cuda:0")
    r0 = torch.tensor(np.ascontiguousarray(c1["r"].T), device=dev)
    v0 = torch.tensor(np.ascontiguousarray(c1["v"].T), device=dev)
    r = torch.empty_like(r0)
    v = torch.empty_like(v0)
    with ogp.Context(degree=70) as ctx:
        ctx.set_sources(ogp.SOURCE_SUN | ogp.SOURCE_MOON | ogp.SOURCE_DRAG, 0.01)
        ctx.propagate_device(r0, v0, n, c1["step"], c1["steps"], r, v)  # tuning, tables
        best = 1e30
        for _ in range(2):
            ctx.propagate_device(r0, v0, n, c1["step"], c1["steps"], r, v)
            best = min(best, ctx.timing()["total_ms"])
        evals = ctx.timing()["evaluations"]
    rate = float(c1["steps"].sum()) / (best * 1e-3)
    gj = float(np.maximum(c1["steps"] - 4, 0).sum())
    out = dict(workload="the headline workload with Sun + Moon + drag (C_D A/m = 0.01 m^2/kg)",
               object_steps_per_s=rate, device_ms=best,
               # drag reads the velocity, so a step ends when position and velocity repeat
               evaluations_per_object_step=evals / gj)
    if gravity_only_ms:
        out["cost_vs_gravity_only"] = best / gravity_only_ms
    return out


def leg_perturb(ogp, torch):
    """Monte-Carlo population in HBM: 10^8 satellites around one LEO state (SoA, 48 B written per
    satellite, nothing read), timed around the synchronous device entry point."""
    n = 100_000_000
    r0 = [-3884872.0, -5749111.0, 14723.0]
    v0 = [-795.304, 573.772, 7518.179]
    rd = torch.empty((3, n), dtype=torch.float64, device="cuda:0")
    vd = torch.empty_like(rd)
    with ogp.Context(force=ogp.FORCE_NONE) as ctx:
        ctx.perturbed_satellites_device(r0, v0, n, 1.0, rd, vd)  # warm-up
        best = 1e9
        for _ in range(5):
            t0 = time.perf_counter()
            ctx.perturbed_satellites_device(r0, v0, n, 1.0, rd, vd)
            best = min(best, time.perf_counter() - t0)
    gbs = 48.0 * n / best / 1e9
    return dict(workload="10^8 satellites, Philox4x32-10 + Box-Muller normals, SoA output in HBM",
                satellites_per_s=n / best, ms=best * 1e3,
                roofline=dict(bound="hbm", achieved=gbs, peak=PEAK_HBM_GBS, unit="GB/s",
                              frac=gbs / PEAK_HBM_GBS, bytes_per_satellite=48))


def leg_copy(torch):
    """A plain device copy of known size (1 GiB read + 1 GiB written): the calibration dispatch
    for the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of the profiler runs."""
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dict(bytes_read=a.numel() * 8, bytes_written=b.numel() * 8,
                gbs=2 * a.numel() * 8 / dt / 1e9)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported. Propagation "
              "shards by objects with no exchange between GPUs, so N GPUs would run N times "
              "this line's per-GPU work.", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("no HIP device visible")
    import opengpop_amd as ogp
    from opengpop_amd import scenarios
    ogp.lib()  # fails loudly when libogp.so is missing
    threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))

    line = dict(metric=METRIC, value=None, unit="object-steps/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=None, higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload="100,000 mixed objects (80% LEO at 30 s, 15% MEO and 5% GEO "
                                     "at 300 s), spans U(6,24) h, JGM3 70x70, FP64, inputs in HBM",
                            objects=100_000, degree=70, object_steps=None))
    sample = None
    if "fp64" in legs:
        f = leg_fp64(ogp, scenarios, torch, args)
        sample = f.pop("_sample")
        line.update(value=f["value"], ms_per_step=f["ms_per_step"], roofline=f["roofline"])
        line["config"]["object_steps"] = f["object_steps"]
        line["device"] = f["device"]
        line["all_corrections"] = f["all_corrections"]
        line["wave_tuning"] = f["tuning"]
    split_sample = adjoint_sample = pairs_sample = None
    if "split" in legs:
        line["split"] = leg_split(ogp, scenarios, args)
        split_sample = line["split"].pop("_oracle_sample")
    if "pairs" in legs:
        line["pairs"] = leg_pairs(ogp, scenarios)
        pairs_sample = line["pairs"].pop("_oracle_sample")
        log(f"pairs leg: {line['pairs']['pair_steps_per_s']:.3e} pair-steps/s")
    if "layouts" in legs:
        line.setdefault("pairs", {})["set_layouts"] = leg_pairs_shapes(ogp, scenarios.config4())
    if "adjoint" in legs:
        line["adjoint"] = leg_adjoint(ogp, scenarios)
        adjoint_sample = line["adjoint"].pop("_oracle_sample")
        log(f"adjoint leg: jacobian {line['adjoint']['jacobian_cost_vs_forward']:.1f}x forward")
    if "refpoint" in legs:
        line["reference_point"] = leg_refpoint(ogp, scenarios)
    if "sources" in legs:
        line["sources"] = leg_sources(ogp, scenarios, torch,
                                      line["device"]["total_ms"] if "device" in line else None)
        log(f"sources leg: {line['sources']['object_steps_per_s']:.3e} object-steps/s")
    if "perturb" in legs:
        line["perturb"] = leg_perturb(ogp, torch)
        log(f"perturb leg: {line['perturb']['satellites_per_s']:.3e} satellites/s")
    if "copy" in legs:
        line["copy_calibration"] = leg_copy(torch)
    # the cpu leg, last: the oracle timed on the host (cpu_baseline) and, as the checker, the
    # GPU results of the other legs against it
    if "cpu" in legs:
        if sample is not None:
            line["cpu_baseline"], line["parity_sample"] = leg_cpu(sample, threads)
            log(f"cpu leg: {line['cpu_baseline']['value']:.3e} object-steps/s on {threads} threads")
        if pairs_sample is not None:
            line["pairs"]["vs_oracle"] = pairs_oracle_check(pairs_sample, threads)
            pairs_sample = None
        checks = oracle_checks(split_sample, adjoint_sample, threads)
        if "split" in checks:
            line["split"]["vs_oracle_sample_m"] = checks["split"]
        if "adjoint" in checks:
            line["adjoint"]["adjoint_vs_oracle_central_differences_rel"] = checks["adjoint"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
