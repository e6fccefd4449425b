"""Benchmark: converged glint walks/s of unbiased glint SMS on one MI355X.

Workload (one step = one whole frame: the camera footprint queries and their SMS estimate):
  1x1 mirror plane with a seed-generated 2048^2 RGB8 normal map (4096 line scratches of
  width 1-3 texels + Gaussian bumps, slope std 0.3) under a point light, 512x512 pixels at
  4 spp (1,048,576 footprint queries), pixel-footprint UV seeding, unbiased Bernoulli
  estimator with at most 1e5 trials, Newton max 20 iterations, threshold 4e-5, uniqueness
  2e-4 (the glint settings of results/Figure_19_GlintsComparison/shoes_sms.xml:9-11 and
  render_shoes.py:84).

Prints ONE JSON line.  `value` = converged walks/s (walks that returned success) over the
whole timed region, in which each frame is one replay of the frame's captured HIP graph
(sms_glint_frame_graph_create) with that frame's seed: the camera queries are generated in
the kernels that read them, so the frame includes query generation and reads no query
records.  The same frames are first run as two-step frames (sms_glint_make_queries +
sms_glint_estimate_batch) and as eager whole-frame calls, whose HIP events time each kernel;
the replays must reproduce their statistics.  `roofline` refers to the dominant kernel
(k_glint_first: camera query, emitter sample, visibility, footprint target and the first
Newton walk of every query) timed with HIP events on the stream it runs on; `cpu_baseline`
times the CPU oracle on whole frames of the same workload on this host's cores.
`query_buffer_frame` is the frame over camera queries made once and kept in HBM (140-byte
records), as a caller holding its own queries runs it.

Beside it: `dome_ss_biased` (the plain single-scatter caustic config: glass dome, 256^2 x 16
spp, biased M = 64) and `pool_ss` (the water pool, 1024^2 x 8 spp pass, unbiased), each with an
FP32 roofline of its walk kernel (FLOP per Newton iteration frozen by the op-counting oracle,
tests/golden/flop_counts.json) and a CPU-oracle baseline on 16 threads and on one core.  The
glint CPU baseline runs whole frames with the GPU frames' seeds and must reproduce their
queries, statistics and the first frames' RGB, bit for bit, or the benchmark fails.

`equal_time_rmse` is the second half of the metric, on the water-pool config (see
equal_time_rmse): RMSE of uniform-seed and learned-proposal renders given the same wall-clock
budget, against a uniform-seed reference; `ratio` charges the training to the same frame (the
network is trained online between its passes, every pass counts in the image),
`ratio_amortized` compares with the trained network (training done before the frame).
`ms_slab_equal_time_rmse` is the same experiment on the multi-scatter slab, and
`glint_equal_time_rmse` on this glint frame with the learned proposal over the pixel footprint.
`image_parity` renders the dome and the pool with SMS (path_sms_ss) and with the filtered path
tracer (path_filtered_ss) and reports their relative RMSE, caustic term and whole image.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "specular-manifold-sampling_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 vector
BYTES_PER_ITER = 12            # algorithmic: 4 RGB8 texels per Newton iteration (the map's source format)
# HBM bytes per launch of k_glint_first on this workload (default sizes and schedule), from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs scaled by a plain 256 MiB copy's counters
# (FETCH_SIZE reads half the copied bytes: factor 2; WRITE_SIZE factor 1): the whole frame, whose
# camera queries are generated in the kernel, and the frame over a buffer of 140-byte query
# records; see profiles/glint_whole_frame_512x512x4spp_pmc.md.
TRAFFIC_FIRST_PASS = {"frame": {"fetch_kib": 16013.5, "write_kib": 13485.8},
                      "query_buffer": {"fetch_kib": 87663.0, "write_kib": 13504.8}}
FETCH_FACTOR, WRITE_FACTOR = 2.0, 1.0
# FP32 FLOP per Newton iteration of each estimator (its whole work divided by its iterations),
# frozen by the op-counting build of the CPU oracle (tests/golden/make_flop_counts.py)
FLOP_COUNTS = json.load(open(os.path.join(ROOT, "tests", "golden", "flop_counts.json")))
FLOPS_PER_ITER = FLOP_COUNTS["glint"]["flop32_per_iteration"]
CPU_THREADS = 16               # the host share of one GPU on the benchmark machine


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_cpu(run, seconds, max_reps=1000):
    """Calls run(rep) -> (walks, converged) until `seconds` have passed; returns the rates."""
    walks = conv = reps = 0
    t = time.perf_counter()
    while True:
        w, c = run(reps)
        walks += w
        conv += c
        reps += 1
        dt = time.perf_counter() - t
        if dt >= seconds or reps >= max_reps:
            return {"walks_per_s": walks / dt, "converged_per_s": conv / dt, "reps": reps, "walks": walks,
                    "seconds": round(dt, 2)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=512)
    ap.add_argument("--height", type=int, default=512)
    ap.add_argument("--spp", type=int, default=4)
    ap.add_argument("--nm-res", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample length")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library default schedule)")
    ap.add_argument("--lane-trials", type=int, default=None, help="override the tuned schedule")
    ap.add_argument("--blocks-per-cu", type=int, default=None, help="override the tuned schedule")
    ap.add_argument("--trial-group", type=int, default=None, help="override the tuned schedule")
    ap.add_argument("--tail-rounds", type=int, default=None, help="override the tuned schedule")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device-to-device copy (HBM counter calibration)")
    ap.add_argument("--no-rmse", action="store_true", help="skip the equal-time RMSE experiment")
    ap.add_argument("--no-pool", action="store_true", help="skip the pool single-scatter throughput line")
    ap.add_argument("--pool-only", action="store_true", help="run only the pool single-scatter throughput")
    ap.add_argument("--dome-only", action="store_true", help="run only the glass-dome biased single-scatter line")
    ap.add_argument("--no-dome", action="store_true", help="skip the glass-dome biased single-scatter line")
    ap.add_argument("--ss-first-pass", type=int, default=None, help="pool: 0 walker, 1 lane per query (default: tuned)")
    ap.add_argument("--rmse-only", action="store_true", help="run only the equal-time RMSE experiment")
    ap.add_argument("--ms-only", action="store_true", help="run only the multi-scatter slab experiment")
    ap.add_argument("--no-ms", action="store_true", help="skip the multi-scatter slab experiment")
    ap.add_argument("--ms-spp", type=int, default=32, help="uniform-seed samples per pixel of the slab budget")
    ap.add_argument("--rmse-reps", type=int, default=16, help="frames per method of the pool experiment")
    ap.add_argument("--ms-reps", type=int, default=8, help="frames per method of the slab experiment")
    ap.add_argument("--ms-ref-spp", type=int, default=128, help="slab reference samples per pixel")
    ap.add_argument("--ms-pass-spp", type=int, default=2, help="samples per pixel of a slab pass")
    ap.add_argument("--glint-rmse-only", action="store_true",
                    help="run only the glint equal-time RMSE experiment (learned footprint proposal)")
    ap.add_argument("--no-glint-rmse", action="store_true", help="skip the glint equal-time RMSE experiment")
    ap.add_argument("--glint-rmse-spp", type=int, default=512, help="uniform-seed samples per pixel of the glint budget")
    ap.add_argument("--glint-ref-spp", type=int, default=8192, help="glint reference samples per pixel")
    ap.add_argument("--glint-reps", type=int, default=4, help="frames per method of the glint experiment")
    ap.add_argument("--whole-frame-only", action="store_true",
                    help="run only the C2 whole-frame measurement (camera queries generated in the kernels)")
    ap.add_argument("--parity-only", action="store_true",
                    help="run only the image parity of SMS against the filtered path tracer")
    ap.add_argument("--no-parity", action="store_true", help="skip the image parity experiment")
    ap.add_argument("--parity-res", type=int, default=128, help="image width = height of the parity renders")
    ap.add_argument("--parity-sms-seconds", type=float, default=5.0,
                    help="time per scene and term of the SMS renders of the parity experiment")
    ap.add_argument("--parity-ref-seconds", type=float, default=15.0,
                    help="time per scene and term of the filtered path tracer's reference")
    ap.add_argument("--mlp-only", action="store_true", help="run only the proposal network's training / inference timing")
    ap.add_argument("--no-mlp", action="store_true", help="skip the proposal network's timing")
    ap.add_argument("--mlp-steps", type=int, default=2000, help="timed training steps of the network leg")
    ap.add_argument("--rmse-res", type=int, default=1024, help="pool image width = height")
    ap.add_argument("--rmse-spp", type=int, default=64, help="uniform-seed samples per pixel (sets the time budget)")
    ap.add_argument("--rmse-ref-spp", type=int, default=512, help="reference samples per pixel (uniform seeds)")
    ap.add_argument("--rmse-pass-spp", type=int, default=8)
    ap.add_argument("--first-pass-spp", type=int, default=2,
                    help="samples per pixel of the charged frame's uniform first pass")
    ap.add_argument("--train-samples", type=int, default=2_000_000, help="records the online training uses")
    ap.add_argument("--train-epochs", type=int, default=8, help="Adam steps per 8192 captured records")
    ap.add_argument("--defensive", type=float, default=0.1,
                    help="pool: weight of the uniform seed in the anchored mixture")
    ap.add_argument("--defensive-other", type=float, default=0.2,
                    help="slab and glints: weight of the uniform seed in the mixture")
    ap.add_argument("--anchor-radius", type=float, default=0.125,
                    help="pool: half side of the anchored mixtures' uv box (0: mixtures over the whole uv square)")
    return ap.parse_args()


def copy_calibration():
    """A plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE of the same rocprofv3 --pmc
    run, divided by the bytes below, give the factor that turns those counters into bytes."""
    import torch
    n = 256 << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.rmse_only:
        print(json.dumps(equal_time_rmse(args)), flush=True)
        return
    if args.pool_only:
        print(json.dumps(pool_throughput(args)), flush=True)
        return
    if args.dome_only:
        print(json.dumps(dome_throughput(args)), flush=True)
        return
    if args.ms_only:
        print(json.dumps(equal_time_rmse(args, "slab")), flush=True)
        return
    if args.mlp_only:
        print(json.dumps(mlp_throughput(args)), flush=True)
        return
    if args.glint_rmse_only:
        print(json.dumps(equal_time_rmse(args, "glint")), flush=True)
        return
    if args.parity_only:
        print(json.dumps(image_parity(args)), flush=True)
        return
    import numpy as np
    from smsamd import abi, gpu, scenes

    t0 = time.time()
    sc = scenes.glint_plane(args.width, args.height, nm_res=args.nm_res, n_scratches=4096, seed=1)
    cfg = abi.SmsConfig.default(max_trials=100000, solver_threshold=4e-5, uniqueness_threshold=2e-4)
    ctx = gpu.Context(sc, cfg)
    dev, ncu = ctx.device_info()
    qbuf, n = ctx.glint_queries_device(args.spp, seed=1)
    out = gpu.DeviceBuffer(ctx, n * 12)
    # the autotuner runs before the timed region, on the same queries with other seeds
    t_tune = time.time()
    sched = ctx.get_schedule() if args.no_tune else ctx.glint_autotune_device(qbuf, n, 7777, out)
    t_tune = time.time() - t_tune
    for key in ("lane_trials", "blocks_per_cu", "trial_group", "tail_rounds"):
        v = getattr(args, key)
        if v is not None:
            sched[key] = v
    ctx.set_schedule(**sched)
    print(f"# setup {time.time() - t0:.1f}s (autotune {t_tune:.2f}s) device={dev} cus={ncu} queries={n} "
          f"schedule={sched}", file=sys.stderr, flush=True)
    if args.whole_frame_only:
        print(json.dumps(glint_whole_frame(args, ctx, qbuf, n, out)[0]), flush=True)
        out.free()
        qbuf.free()
        ctx.close()
        return

    for w in range(args.warmup):
        ctx.glint_estimate_device(qbuf, n, 1000 + w, out)
    ctx.sync()
    tot = {}
    k1_ms = k2_ms = 0.0
    iters_pass1 = 0
    t_start = time.perf_counter()
    for s in range(args.steps):
        st = ctx.glint_estimate_device(qbuf, n, 1 + s, out)
        a, b = ctx.last_kernel_ms()
        k1_ms += a
        k2_ms += b
        iters_pass1 += st["newton_iterations"] - st["tail_iterations"]
        for k, v in st.items():
            tot[k] = tot.get(k, 0) + v
    ctx.sync()
    eager_s = time.perf_counter() - t_start
    # the same frames (same seeds) replayed from a captured graph: one launch per frame, no host
    # synchronisation between frames, statistics folded on the device
    graph = ctx.glint_graph(qbuf, n, out)
    graph.launch(1000)
    graph.stats(reset=True)
    ctx.sync()
    t_start = time.perf_counter()
    for s in range(args.steps):
        graph.launch(1 + s)
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    gtot = graph.stats(reset=True)
    graph.close()
    for k in ("walks", "walks_converged", "newton_iterations"):
        if gtot[k] != tot[k]:
            raise RuntimeError(f"graph replays differ from the eager frames in {k}: {gtot[k]} != {tot[k]}")
    ms_per_step = 1e3 * elapsed / args.steps
    conv_per_s = tot["walks_converged"] / elapsed
    walks_per_s = tot["walks"] / elapsed
    print(f"# steps={args.steps} graph {ms_per_step:.3f} ms/step (eager {1e3 * eager_s / args.steps:.3f}) "
          f"walks/s={walks_per_s:.4g} conv/s={conv_per_s:.4g} first={k1_ms / args.steps:.3f}ms "
          f"rest={k2_ms / args.steps:.3f}ms stats={tot}", file=sys.stderr, flush=True)

    default_workload = (args.width, args.height, args.spp, args.nm_res) == (512, 512, 4, 2048)
    qb_roofline = first_pass_roofline(iters_pass1 / args.steps, k1_ms / args.steps,
                                      "query_buffer" if default_workload else None)
    query_buffer_frame = {
        "ms_per_frame": round(ms_per_step, 4), "converged_walks_per_s": conv_per_s, "walks_per_s": walks_per_s,
        "eager_ms_per_frame": round(1e3 * eager_s / args.steps, 4),
        "trials_tail_finalize_ms": round(k2_ms / args.steps, 3), "roofline": qb_roofline,
        "note": "sms_glint_make_queries once (seed 1, outside the timed region), then graph replays of "
                "sms_glint_estimate_batch over the 140-byte query records with estimator seeds 1..steps"}

    whole_frame, wf = glint_whole_frame(args, ctx, qbuf, n, out)
    roofline = first_pass_roofline(wf["iters_per_launch"], wf["first_pass_ms"],
                                   "frame" if default_workload else None)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc, cfg, ctx, qbuf, n, out, wf["step_stats"], args)
    out.free()
    qbuf.free()
    ctx.close()
    dome = None if args.no_dome else dome_throughput(args)
    pool = None if args.no_pool else pool_throughput(args)
    mlp = None if args.no_mlp else mlp_throughput(args)
    rmse = None if args.no_rmse else equal_time_rmse(args)
    ms_rmse = None if args.no_ms else equal_time_rmse(args, "slab")
    glint_rmse = None if args.no_glint_rmse else equal_time_rmse(args, "glint")
    parity = None if args.no_parity else image_parity(args)
    result = {
        "metric": "converged walks/s at 1 GPU; equal-time RMSE learned vs uniform seeds",
        "value": wf["converged_per_s"],
        "unit": "walks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wf["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"glint SMS, {args.width}x{args.height}x{args.spp}spp camera footprint queries "
                               f"generated per frame, {args.nm_res}^2 RGB8 scratch+bump normal map, unbiased "
                               f"(max_trials 1e5)",
                   "queries_per_step": n, "walks_per_step": wf["walks"] / args.steps,
                   "walks_per_s": whole_frame["walks_per_s"], "device": dev, "schedule": sched,
                   "launch": "hipGraph replay per frame (sms_glint_frame_graph_create), camera seed = estimator seed",
                   "eager_ms_per_step": whole_frame["eager_ms_per_frame"]},
        "roofline": roofline,
        "whole_frame": whole_frame,
        "query_buffer_frame": query_buffer_frame,
        "cpu_baseline": cpu,
        "dome_ss_biased": dome,
        "pool_ss": pool,
        "mlp": mlp,
        "equal_time_rmse": rmse,
        "ms_slab_equal_time_rmse": ms_rmse,
        "glint_equal_time_rmse": glint_rmse,
        "image_parity": parity,
    }
    if cpu and cpu.get("value"):
        result["gpu_over_cpu"] = wf["converged_per_s"] / cpu["value"]
    print(json.dumps(result), flush=True)


def pool_throughput(args):
    """Unbiased single-scatter SMS on the water pool (uniform seeds): one 1024^2 x 8 spp pass of
    camera queries resident in HBM, `steps` timed estimates after `warmup` (per-step seeds);
    schedule from sms_ss_autotune unless --no-tune / --ss-first-pass / --lane-trials are given."""
    from smsamd import abi, gpu, scenes
    res, spp = args.rmse_res, args.rmse_pass_spp
    sc = scenes.pool(res, res)
    cfg = abi.SmsConfig.default(max_trials=10_000_000, uniqueness_threshold=1e-5, solver_threshold=1e-5)
    ctx = gpu.Context(sc, cfg)
    qb, wb, n = ctx.ss_queries_device(spp, 1)
    out = gpu.DeviceBuffer(ctx, n * 12)
    if args.no_tune or args.ss_first_pass is not None or args.lane_trials is not None:
        kw = {}
        if args.ss_first_pass is not None:
            kw["ss_first_pass"] = args.ss_first_pass
        if args.lane_trials is not None:
            kw["lane_trials"] = args.lane_trials
        if kw:
            ctx.set_schedule(**kw)
        sched = ctx.get_schedule()
    else:
        sched = ctx.ss_autotune_device(qb, n, 7777, out)
    for w in range(args.warmup):
        ctx.ss_estimate_device(qb, n, 2000 + w, out)
    ctx.sync()
    tot = {}
    k1 = k2 = 0.0
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        st = ctx.ss_estimate_device(qb, n, 1 + s_, out)
        a, b = ctx.last_kernel_ms()
        k1 += a
        k2 += b
        for k, v in st.items():
            tot[k] = tot.get(k, 0) + v
    ctx.sync()
    dt = time.perf_counter() - t0
    res_ = {"workload": f"water pool caustics, {res}x{res}x{spp}spp camera queries ({n}), unbiased, uniform seeds",
            "converged_walks_per_s": tot["walks_converged"] / dt, "walks_per_s": tot["walks"] / dt,
            "newton_iterations_per_s": tot["newton_iterations"] / dt, "ms_per_step": 1e3 * dt / args.steps,
            "first_pass_ms": k1 / args.steps, "tail_ms": k2 / args.steps, "steps": args.steps,
            "walks_per_step": tot["walks"] / args.steps, "schedule": sched}
    kernel = "k_ss_walker" if sched.get("ss_first_pass", 0) == 0 else "k_ss_estimate"
    res_["roofline"] = valu_roofline(kernel, "pool", (tot["newton_iterations"] - tot["tail_iterations"]) / args.steps,
                                     k1 / args.steps)
    print(f"# pool: {json.dumps(res_)}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        import numpy as np
        from smsamd import abi
        q = np.zeros(n, abi.SS_QUERY)
        qb.download(q)
        res_["cpu_baseline"] = query_sample_baseline(sc, cfg, q, "ss", args.cpu_seconds / 1.5, 256)
        res_["gpu_over_cpu"] = res_["converged_walks_per_s"] / res_["cpu_baseline"]["value"]
        print(f"# pool cpu: {json.dumps(res_['cpu_baseline'])}", file=sys.stderr, flush=True)
    out.free()
    qb.free()
    wb.free()
    ctx.close()
    return res_


def valu_roofline(kernel, flop_key, iters_per_launch, kernel_ms):
    """FP32 roofline of a walk kernel: frozen FLOP per Newton iteration x the iterations one
    launch ran / its HIP-event time."""
    fpi = FLOP_COUNTS[flop_key]["flop32_per_iteration"]
    tflops = fpi * iters_per_launch / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(tflops, 4), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / FP32_PEAK_TFLOPS, 6), "traffic": None, "kernel": kernel,
            "kernel_ms": round(kernel_ms, 3), "flop32_per_iteration": round(fpi, 1),
            "iters_per_launch": round(iters_per_launch)}


def dome_throughput(args):
    """The plain single-scatter caustic config: glass sphere (r 1, IOR 1.5) centred on a diffuse
    floor (a dome, as the reference's sphere scene places it), sphere light r 0.1 radiance 100 at
    (0,5,0), pinhole 45 deg; 256^2 x 16 spp camera queries resident in HBM, biased SMS with trial
    sets of 64 walks, Newton max 20, threshold 1e-5, uniform seeds.  `steps` timed estimates after
    `warmup`; the CPU oracle runs strided samples of the same queries."""
    from smsamd import abi, gpu, scenes
    sc = scenes.sphere_dome(256, 256)
    cfg = abi.SmsConfig.default(biased=1, max_trials=64)
    ctx = gpu.Context(sc, cfg)
    qb, wb, n = ctx.ss_queries_device(16, 1)
    out = gpu.DeviceBuffer(ctx, n * 12)
    for w in range(args.warmup):
        ctx.ss_estimate_device(qb, n, 3000 + w, out)
    ctx.sync()
    tot, k1 = {}, 0.0
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        st = ctx.ss_estimate_device(qb, n, 1 + s_, out)
        k1 += ctx.last_kernel_ms()[0]
        for k, v in st.items():
            tot[k] = tot.get(k, 0) + v
    ctx.sync()
    dt = time.perf_counter() - t0
    res_ = {"workload": f"glass dome caustics, 256x256x16spp camera queries ({n}), biased M=64, uniform seeds",
            "converged_walks_per_s": tot["walks_converged"] / dt, "walks_per_s": tot["walks"] / dt,
            "newton_iterations_per_s": tot["newton_iterations"] / dt, "ms_per_step": 1e3 * dt / args.steps,
            "steps": args.steps, "walks_per_step": tot["walks"] / args.steps}
    res_["roofline"] = valu_roofline("k_ss_biased", "dome_biased", tot["newton_iterations"] / args.steps,
                                     k1 / args.steps)
    print(f"# dome: {json.dumps(res_)}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        import numpy as np
        q = np.zeros(n, abi.SS_QUERY)
        qb.download(q)
        res_["cpu_baseline"] = query_sample_baseline(sc, cfg, q, "ss", args.cpu_seconds / 1.5, 256)
        cb = res_["cpu_baseline"]
        res_["gpu_over_cpu"] = {"walks": res_["walks_per_s"] / cb["walks_per_s"],
                                "walks_one_core": res_["walks_per_s"] / cb["one_core"]["walks_per_s"]}
        print(f"# dome cpu: {json.dumps(cb)}", file=sys.stderr, flush=True)
    out.free()
    qb.free()
    wb.free()
    ctx.close()
    return res_


BF16_PEAK_TFLOPS = 2500.0      # MI355X dense BF16 matrix peak
# algorithmic multiply-adds per sample: forward 153x64 + 3x64x64 + 64x48 (features unpadded);
# training adds the weight gradients (same count) and the input gradients of layers 2-5
MLP_FWD_MAC = 153 * 64 + 3 * 64 * 64 + 64 * 48
MLP_TRAIN_MAC = 2 * MLP_FWD_MAC + 3 * 64 * 64 + 48 * 64


def mlp_throughput(args):
    """The learned proposal's network on synthetic records resident in HBM: `mlp_steps` training
    steps (sms_mlp_train: gradient of the weighted mixture NLL over a batch of 8192, fused
    reduction + Adam) and one inference over 1024^2 x 8 queries (sms_mlp_infer).  Wall-clock
    per step / per inference (the calls are queued on the context stream and timed to the
    synchronisation that ends them)."""
    import numpy as np
    from smsamd import abi, gpu, scenes
    ctx = gpu.Context(scenes.sphere_dome(), abi.SmsConfig.default())
    rng = np.random.default_rng(0)
    n_rec, batch = 1 << 21, 8192
    x = rng.uniform(-1.0, 1.0, (n_rec, abi.MLP_INPUTS)).astype(np.float32)
    uv = rng.uniform(0.0, 1.0, (n_rec, 2)).astype(np.float32)
    w = rng.lognormal(0.0, 1.0, n_rec).astype(np.float32)
    bufs = [gpu.DeviceBuffer(ctx, a.nbytes) for a in (x, uv, w)]
    for b, a in zip(bufs, (x, uv, w)):
        b.upload(a)
    mlp = gpu.Mlp(ctx, seed=1)
    offs = rng.integers(0, n_rec - batch + 1, args.mlp_steps)
    mlp.train_device(bufs[0], bufs[1], n_rec, batch, offs[:50], 1e-3, wbuf=bufs[2], want_losses=True)
    t0 = time.perf_counter()
    losses = mlp.train_device(bufs[0], bufs[1], n_rec, batch, offs, 1e-3, wbuf=bufs[2], want_losses=True)
    train_s = (time.perf_counter() - t0) / args.mlp_steps
    n_inf = 1024 * 1024 * 8
    xi = gpu.DeviceBuffer(ctx, n_inf * abi.MLP_INPUTS * 4)
    oi = gpu.DeviceBuffer(ctx, n_inf * abi.MLP_OUTPUTS * 4)
    for k in range(n_inf // n_rec):                         # the records, repeated
        gpu.lib().sms_copy_h2d(ctx.h, xi.ptr + k * x.nbytes, x.ctypes.data, x.nbytes)
    mlp.infer_device(xi, n_inf, oi)
    ctx.sync()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        mlp.infer_device(xi, n_inf, oi)
    ctx.sync()
    infer_s = (time.perf_counter() - t0) / reps
    train_tflops = 2.0 * MLP_TRAIN_MAC * batch / train_s / 1e12
    infer_tflops = 2.0 * MLP_FWD_MAC * n_inf / infer_s / 1e12
    out = {"workload": "proposal network 9 -> 153 Fourier features -> 4x64 ReLU -> 48 (8-component mixture), "
                       "BF16 MFMA with FP32 accumulation",
           "train_batch": batch, "train_steps": args.mlp_steps, "train_us_per_step": round(1e6 * train_s, 2),
           "train_samples_per_s": batch / train_s, "train_loss_first_last": [float(losses[0]), float(losses[-1])],
           "infer_queries": n_inf, "infer_ms": round(1e3 * infer_s, 3), "infer_queries_per_s": n_inf / infer_s,
           "roofline": {"bound": "mfma", "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "train_achieved": round(train_tflops, 2), "train_frac": round(train_tflops / BF16_PEAK_TFLOPS, 5),
                        "infer_achieved": round(infer_tflops, 2), "infer_frac": round(infer_tflops / BF16_PEAK_TFLOPS, 5),
                        "flops_per_sample_train": 2 * MLP_TRAIN_MAC, "flops_per_query_infer": 2 * MLP_FWD_MAC}}
    print(f"# mlp: {json.dumps(out)}", file=sys.stderr, flush=True)
    for b in bufs + [xi, oi]:
        b.free()
    mlp.close()
    ctx.close()
    return out


def equal_time_rmse(args, case="pool"):
    """case "pool": water pool (512^2-cell Gerstner mesh, IOR 1.33, floor 1 below, sphere light
    r 0.05), single-scatter caustic term of the camera image at rmse_res^2, unbiased (cap 1e7,
    thresholds 1e-5 as pool_sms.xml); budget rmse_spp, reference rmse_ref_spp.  case "slab":
    glass slab 0.2 thick (IOR 1.5) with value-noise faces over a diffuse floor, two-vertex
    multi-scatter chains (bounces 2, unbiased, cap 1e7, thresholds 1e-5 as slab_sms.xml), the
    proposal on the chain's seed point on the top face; budget ms_spp, reference ms_ref_spp.
    case "glint": the glint frame of the headline line (512^2, 2048^2 scratch/bump map, passes
    of 4 spp, unbiased, cap 1e5, thresholds 4e-5 / 2e-4) with the learned proposal over the
    pixel footprint, trained on converged first walks; budget glint_rmse_spp, reference
    glint_ref_spp.  Uniform seeding renders the budget's samples per pixel; its wall-clock time T is the budget.
    The charged run gets the same T for everything, training included: its first pass (of
    first_pass_spp samples per pixel) uses uniform seeds, every later pass the
    defensive mixture with the network as trained so far; after each pass the network trains on
    the estimate-weighted solutions that pass captured (batch 8192, Adam lr 1e-3, train_epochs
    steps per 8192 records) until train_samples records were used; every pass counts in the
    image.  The amortized run renders T with the trained network (training done before the
    frame).  RMSE against a uniform-seed reference, with the reference's own variance removed.
    Each method renders `frames` independent frames (the charged ones each training a fresh
    network); a ratio is sqrt(mean MSE uniform / mean MSE learned) over them, the per-frame
    ratios and their median are reported beside it -- the errors are heavy-tailed: a frame
    that draws one of the rare samples whose solution the mixture proposes with little mass
    (at worst the defensive weight x the uniform density: up to 10x the uniform inverse-PDF on
    the pool, 5x on the slab and glints) can hold most
    of its squared error in a few pixels."""
    import numpy as np
    from smsamd import abi, gpu, render, scenes
    defensive = args.defensive if case == "pool" else args.defensive_other
    res, pass_spp = args.rmse_res, {"pool": args.rmse_pass_spp, "slab": args.ms_pass_spp, "glint": args.spp}[case]
    if case == "glint":
        res = args.width
        sc = scenes.glint_plane(args.width, args.height, nm_res=args.nm_res, n_scratches=4096, seed=1)
        cfg = abi.SmsConfig.default(max_trials=100000, solver_threshold=4e-5, uniqueness_threshold=2e-4)
        budget_spp, ref_spp, estimator = args.glint_rmse_spp, args.glint_ref_spp, "glint"
    elif case == "pool":
        sc = scenes.pool(res, res)
        cfg = abi.SmsConfig.default(max_trials=10_000_000, uniqueness_threshold=1e-5, solver_threshold=1e-5)
        budget_spp, ref_spp, estimator = args.rmse_spp, args.rmse_ref_spp, "ss"
    else:
        sc = scenes.slab(res, res, n=256)
        cfg = abi.SmsConfig.default(bounces=2, max_trials=10_000_000, uniqueness_threshold=1e-5,
                                    solver_threshold=1e-5)
        budget_spp, ref_spp, estimator = args.ms_spp, args.ms_ref_spp, "ms"
    ctx = gpu.Context(sc, cfg)
    if case == "pool":
        ctx.set_proposal_anchor(args.anchor_radius)
    runner = render.PassRunner(ctx, pass_spp, estimator)
    first_spp = args.first_pass_spp or pass_spp
    first_runner = render.PassRunner(ctx, first_spp, estimator) if first_spp != pass_spp else None
    film = render.Film(ctx, runner.pixels)
    runner.run(999_999, film)                                   # first launches
    if first_runner is not None:
        first_runner.run(999_997, film)
    ctx.sync()
    if estimator == "glint":
        ctx.glint_queries_device(pass_spp, 999_998, runner.q)
        sched = ctx.glint_autotune_device(runner.q, runner.n, 999_998, runner.rgb)
    elif estimator == "ss":
        ctx.ss_queries_device(pass_spp, 999_998, runner.q, runner.w)
        sched = ctx.ss_autotune_device(runner.q, runner.n, 999_998, runner.rgb)
    else:
        ctx.ss_queries_device(pass_spp, 999_998, runner.q, runner.w)
        sched = ctx.ms_autotune_device(runner.q, runner.n, 999_998, runner.rgb)
    print(f"# rmse {case}: schedule {sched}", file=sys.stderr, flush=True)

    def frame(seed0, n_passes, st=None):
        film.clear()
        for p in range(n_passes):
            runner.run(seed0 + p, film, st)
        ctx.sync()
        return film.download()

    (ref, ref_var), t_ref = render.timed(lambda: frame(10_000_000, ref_spp // pass_spp))
    ref_var_mean = ref_var / max(film.spp, 1)
    dropped_ref = film.dropped()
    print(f"# rmse {case}: reference {film.spp} spp in {t_ref:.1f}s", file=sys.stderr, flush=True)
    lum = ref.mean(1)
    mask = lum > 2.0 * lum.mean()
    reps = {"pool": args.rmse_reps, "slab": args.ms_reps, "glint": args.glint_reps}[case]

    def errors(img):
        return (render.rmse(img, ref, ref_var_mean) ** 2, render.rmse(img, ref, ref_var_mean, mask) ** 2,
                render.rmse_trimmed(img, ref, ref_var_mean) ** 2)

    def until_budget(seed0, st):
        """Passes until the next would end after T; then, while time is left, passes of the
        smaller runner (every method closes its frame the same way)."""
        t0 = time.perf_counter()
        times, p = [], 0
        while True:
            est = sum(times) / len(times) if times else 0.0
            if time.perf_counter() - t0 + est > T:
                break
            t = time.perf_counter()
            runner.run(seed0 + p, film, st)
            ctx.sync()
            times.append(time.perf_counter() - t)
            p += 1
        if first_runner is not None and times:
            est_small = min(times) * first_runner.pass_spp / runner.pass_spp
            while time.perf_counter() - t0 + est_small <= T:
                t = time.perf_counter()
                first_runner.run(seed0 + p, film, st)
                ctx.sync()
                est_small = time.perf_counter() - t
                p += 1
        return time.perf_counter() - t0

    # uniform seeds: the first frame's time is the budget T
    st_u, err_u, dropped = {}, [], {"reference": dropped_ref, "uniform": 0, "learned": 0, "amortized": 0}
    (img, _), T = render.timed(lambda: frame(20_000_000, budget_spp // pass_spp, st_u))
    spp_u = film.spp
    max_trials_u = [st_u.get("bernoulli_max", 0)]
    err_u.append(errors(img))
    dropped["uniform"] += film.dropped()
    for r in range(1, reps):
        film.clear()
        st_f = {}
        until_budget(20_000_000 + 1000 * r, st_f)
        max_trials_u.append(st_f.get("bernoulli_max", 0))
        err_u.append(errors(film.download()[0]))
        dropped["uniform"] += film.dropped()
    # learned, training charged to each frame (a fresh network per frame)
    err_l, spp_l, tr0, mlp, frame_stats_l = [], [], None, None, []
    for r in range(reps):
        film.clear()
        m = gpu.Mlp(ctx, seed=1 + r)
        st_l = {}
        frame_stats_l.append(st_l)
        tr = render.render_online(ctx, runner, film, m, 0, T, samples=args.train_samples, epochs=args.train_epochs,
                                  defensive=defensive, seed0=30_000_000 + 1000 * r, rng_seed=r, stats=st_l,
                                  uniform_runner=first_runner)
        err_l.append(errors(film.download()[0]))
        spp_l.append(film.spp)
        dropped["learned"] += film.dropped()
        if r == 0:
            tr0, mlp = tr, m
        else:
            m.close()
    steps, losses, train_s = tr0["steps"], tr0["losses"], tr0["train_s"]
    if args.train_samples - tr0["used"] >= 8192:
        # training of the first frame's network goes on (outside any budget) to the full record
        # count for the amortized frames
        film.clear()
        tr2 = render.render_online(ctx, runner, film, mlp, 0, float("inf"), samples=args.train_samples - tr0["used"],
                                   epochs=args.train_epochs, defensive=defensive, seed0=35_000_000,
                                   uniform_passes=0, rng_seed=99, stop_when_trained=True)
        steps += tr2["steps"]
        losses += tr2["losses"]
        train_s += tr2["train_s"]
    # learned, network already trained
    ctx.set_proposal(mlp, 0, defensive)
    err_a, spp_a, st_a, max_trials_a = [], [], {}, []
    for r in range(reps):
        film.clear()
        st_f = {}
        t_a = until_budget(50_000_000 + 1000 * r, st_f)
        max_trials_a.append(st_f.get("bernoulli_max", 0))
        for k, v in st_f.items():
            st_a[k] = max(st_a.get(k, 0), v) if k == "bernoulli_max" else st_a.get(k, 0) + v
        err_a.append(errors(film.download()[0]))
        spp_a.append(film.spp)
        dropped["amortized"] += film.dropped()
    ctx.set_proposal(None, -1)

    def ratio(eu, el, j):
        return float(np.sqrt(np.mean([e[j] for e in eu]) / np.mean([e[j] for e in el])))

    def bootstrap(eu, el, j=0, n=4000):
        """95 % percentile interval of the ratio over frames resampled with replacement (the
        uniform and the learned frames independently: they share no samples)."""
        rng = np.random.default_rng(12345)
        u = np.array([e[j] for e in eu])
        l_ = np.array([e[j] for e in el])
        r = np.sqrt(u[rng.integers(0, len(u), (n, len(u)))].mean(1) / l_[rng.integers(0, len(l_), (n, len(l_)))].mean(1))
        return [round(float(np.percentile(r, 2.5)), 3), round(float(np.percentile(r, 97.5)), 3)]

    def per_frame(el):
        return [round(float(np.sqrt(err_u[i][0] / el[i][0])), 3) for i in range(len(el))]
    mean_ref = float(ref.mean())

    def conv_rate(st):
        return st.get("walks_converged", 0) / max(st.get("walks", 1), 1)

    def trials(st):
        return st.get("bernoulli_trials", 0) / max(st.get("bernoulli_calls", 1), 1)
    workload = {"pool": f"water pool caustics (single scatter), {res}x{res} camera image, unbiased (max_trials 1e7, "
                        f"uniqueness 1e-5)",
                "slab": f"glass slab double refraction (multi scatter, 2 vertices), {res}x{res} camera image, "
                        f"unbiased (max_trials 1e7, uniqueness 1e-5)",
                "glint": f"glints, {res}x{res} frame, {args.nm_res}^2 scratch+bump map, unbiased (max_trials 1e5, "
                         f"threshold 4e-5, uniqueness 2e-4), learned proposal over the pixel footprint"}[case]
    out = {
        "workload": f"{workload}, passes of {pass_spp} spp",
        "frames": reps, "budget_s": round(T, 3), "uniform_spp": spp_u,
        "learned_spp": spp_l, "learned_passes_uniform": 1, "first_pass_spp": first_spp,
        "train_s_within_budget": round(tr0["train_s"], 3), "train_steps_within_budget": tr0["steps"],
        "train_records_within_budget": tr0["used"],
        "train_s": round(train_s, 3), "train_samples": args.train_samples, "train_steps": steps,
        "train_epochs": args.train_epochs,
        "train_loss_first_last": [round(losses[0], 4), round(losses[-1], 4)] if losses else None,
        "amortized_spp": spp_a,
        "reference_spp": ref_spp, "reference_s": round(t_ref, 2), "mean_pixel": mean_ref,
        "rmse_uniform": float(np.sqrt(np.mean([e[0] for e in err_u]))),
        "rmse_learned": float(np.sqrt(np.mean([e[0] for e in err_l]))),
        "rmse_amortized": float(np.sqrt(np.mean([e[0] for e in err_a]))),
        "ratio": ratio(err_u, err_l, 0), "ratio_amortized": ratio(err_u, err_a, 0),
        "ratio_ci95": bootstrap(err_u, err_l), "ratio_amortized_ci95": bootstrap(err_u, err_a),
        "ci95": "bootstrap over frames (4000 resamples, uniform and learned frames resampled independently)",
        "ratio_per_frame": per_frame(err_l), "ratio_amortized_per_frame": per_frame(err_a),
        "max_trials_per_frame": {"uniform": max_trials_u,
                                 "learned": [s.get("bernoulli_max", 0) for s in frame_stats_l],
                                 "amortized": max_trials_a},
        "ratio_median_frame": float(np.median(per_frame(err_l))),
        "ratio_amortized_median_frame": float(np.median(per_frame(err_a))),
        "caustic_mask_pixels": int(mask.sum()),
        "ratio_caustic_mask": ratio(err_u, err_l, 1), "ratio_caustic_mask_amortized": ratio(err_u, err_a, 1),
        "ratio_trimmed": ratio(err_u, err_l, 2), "ratio_trimmed_amortized": ratio(err_u, err_a, 2),
        "trimmed": "RMSE without the 0.1 % of pixels with the largest squared error",
        "converged_per_walk": {"uniform": round(conv_rate(st_u), 4), "learned": round(conv_rate(st_a), 4)},
        "trials_per_estimate": {"uniform": round(trials(st_u), 3), "learned": round(trials(st_a), 3)},
        "invalid_samples_dropped": dropped,
        "uniform_frame": {"walks_per_s": st_u.get("walks", 0) / T, "converged_walks_per_s": st_u.get("walks_converged", 0) / T,
                          "newton_iterations_per_s": st_u.get("newton_iterations", 0) / T,
                          "note": "whole frames: camera passes, estimates and film"},
        "defensive": defensive, "schedule": sched,
    }
    print(f"# rmse {case}: {json.dumps(out)}", file=sys.stderr, flush=True)
    mlp.close()
    runner.free()
    if first_runner is not None:
        first_runner.free()
    film.free()
    ctx.close()
    return out


def first_pass_roofline(iters_per_launch, kernel_ms, traffic_key):
    """Roofline of k_glint_first: algorithmic bytes (BYTES_PER_ITER per first-pass Newton
    iteration) and FP32 FLOP per launch over its average launch time (HIP events on the library
    stream); `traffic` = the calibrated PMC bytes of that launch on the default workload."""
    k1_s = kernel_ms / 1e3
    achieved_gbs = BYTES_PER_ITER * iters_per_launch / k1_s / 1e9
    achieved_tflops = FLOPS_PER_ITER * iters_per_launch / k1_s / 1e12
    traffic = None
    if traffic_key is not None:
        t = TRAFFIC_FIRST_PASS[traffic_key]
        traffic = round(1024.0 * (t["fetch_kib"] * FETCH_FACTOR + t["write_kib"] * WRITE_FACTOR))
    return {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 5), "traffic": traffic,
            "algorithmic_bytes": round(BYTES_PER_ITER * iters_per_launch),
            "bytes_per_iter": BYTES_PER_ITER, "iters_per_launch": round(iters_per_launch),
            "kernel": "k_glint_first", "kernel_ms": round(kernel_ms, 4),
            "valu": {"achieved": round(achieved_tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 5), "flops_per_iter": FLOPS_PER_ITER}}


def glint_whole_frame(args, ctx, qbuf, n, out):
    """The C2 frame with its camera queries generated by the kernels that read them
    (sms_glint_estimate_frame / sms_glint_frame_graph_create): first the two-step frames of seeds
    1..steps (sms_glint_make_queries + sms_glint_estimate_batch, camera seed = estimator seed),
    whose statistics the whole frames, eager and as graph replays, must reproduce; then the
    query-buffer line's queries (seed 1) are restored in qbuf.  Returns the report and the raw
    figures of the replays and of the eager frames' first passes."""
    for w in range(args.warmup):
        ctx.glint_estimate_frame_device(args.spp, 1000 + w, out)
    ttot = {}
    ctx.sync()
    t_start = time.perf_counter()
    for s in range(args.steps):
        ctx.glint_queries_device(args.spp, 1 + s, qbuf)
        st = ctx.glint_estimate_device(qbuf, n, 1 + s, out)
        for k, v in st.items():
            ttot[k] = ttot.get(k, 0) + v
    ctx.sync()
    two_step_s = time.perf_counter() - t_start
    wtot, k1f_ms, iters_pass1, step_stats = {}, 0.0, 0, []
    t_start = time.perf_counter()
    for s in range(args.steps):
        st = ctx.glint_estimate_frame_device(args.spp, 1 + s, out)
        k1f_ms += ctx.last_kernel_ms()[0]
        iters_pass1 += st["newton_iterations"] - st["tail_iterations"]
        step_stats.append(st)
        for k, v in st.items():
            wtot[k] = wtot.get(k, 0) + v
    ctx.sync()
    eager_s = time.perf_counter() - t_start
    fgraph = ctx.glint_frame_graph(args.spp, out)
    fgraph.launch(1000)
    fgraph.stats(reset=True)
    ctx.sync()
    t_start = time.perf_counter()
    for s in range(args.steps):
        fgraph.launch(1 + s)
    ctx.sync()
    f_elapsed = time.perf_counter() - t_start
    ftot = fgraph.stats(reset=True)
    fgraph.close()
    for k in ("walks", "walks_converged", "newton_iterations"):
        if ftot[k] != ttot[k] or wtot[k] != ttot[k]:
            raise RuntimeError(f"whole frames differ from the two-step frames in {k}: {ftot[k]} / {wtot[k]} != {ttot[k]}")
    ctx.glint_queries_device(args.spp, 1, qbuf)       # the query-buffer line's queries again
    ctx.sync()
    raw = {"converged_per_s": ftot["walks_converged"] / f_elapsed, "walks": ftot["walks"],
           "iters_per_launch": iters_pass1 / args.steps, "first_pass_ms": k1f_ms / args.steps,
           "ms": 1e3 * f_elapsed / args.steps, "step_stats": step_stats}
    res = {"ms_per_frame": round(raw["ms"], 4),
           "converged_walks_per_s": raw["converged_per_s"],
           "walks_per_s": ftot["walks"] / f_elapsed,
           "first_pass_ms": round(raw["first_pass_ms"], 4),
           "eager_ms_per_frame": round(1e3 * eager_s / args.steps, 4),
           "two_step_eager_ms_per_frame": round(1e3 * two_step_s / args.steps, 4),
           "note": "sms_glint_frame_graph_create replays of frames with seeds 1..steps (camera seed = "
                   "estimator seed): camera queries generated in the kernels that read them, no "
                   "140-byte query records; statistics equal to sms_glint_make_queries + "
                   "sms_glint_estimate_batch of the same seeds (two_step, eager calls)"}
    print(f"# whole frame: {json.dumps(res)}", file=sys.stderr, flush=True)
    return res, raw


def image_parity(args):
    """SMS against the filtered path tracer (path_sms_ss vs path_filtered_ss, the reference's own
    comparison pair): the caustic term -- the single-scatter caustic paths through the first
    non-specular vertex -- and the whole image of the glass dome (C1 scene: light r 0.1) and the
    water pool (C3 mesh: light r 0.05), parity_res^2 with the Gaussian film (stddev 0.5),
    max_depth 6, unbiased SMS (cap 1e7, thresholds 1e-5).  SMS renders passes of 64 spp for
    parity_sms_seconds, the path tracer passes of 256 for parity_ref_seconds.  Relative RMSE =
    sqrt(mean over pixels of (SMS - PT)^2 minus both images' own variance of the mean, from
    their pass-to-pass spread) / mean PT pixel, per pixel and over 4 x 4-pixel blocks; the raw
    RMSE and both noise floors are reported beside it."""
    import numpy as np
    from smsamd import abi, gpu, scenes
    res = args.parity_res
    out = {"resolution": res, "film": "gaussian stddev 0.5", "max_depth": 6, "sms_seconds_per_term": args.parity_sms_seconds,
           "reference_seconds_per_term": args.parity_ref_seconds,
           "rel_rmse": "sqrt(mean((SMS - PT)^2) - Var[SMS mean] - Var[PT mean]) / mean(PT)"}
    for name, sc in (("dome", scenes.sphere_dome(res, res)), ("pool", scenes.pool(res, res))):
        cfg = abi.SmsConfig.default(max_trials=10_000_000, uniqueness_threshold=1e-5, solver_threshold=1e-5)
        ctx = gpu.Context(sc, cfg)
        px = res * res
        n_max = px * 256
        rgb, pos = gpu.DeviceBuffer(ctx, n_max * 12), gpu.DeviceBuffer(ctx, n_max * 8)
        img, wt = gpu.DeviceBuffer(ctx, px * 12), gpu.DeviceBuffer(ctx, px * 4)
        zero_i, zero_w = np.zeros((px, 3), np.float32), np.zeros(px, np.float32)
        leg = {}

        def render(pc, pass_spp, passes=None, seconds=None, seed0=1):
            """per-pass developed images: mean image, variance of the mean, passes, seconds, stats"""
            s1, s2, k, st_sum = np.zeros((px, 3)), np.zeros((px, 3)), 0, {}
            t0 = time.perf_counter()
            while (passes is not None and k < passes) or (seconds is not None and time.perf_counter() - t0 < seconds):
                img.upload(zero_i)
                wt.upload(zero_w)
                st = ctx.path_samples_device(pc, pass_spp, seed0 + k, rgb, pos)
                ctx.film_splat_device(rgb, pos, pass_spp, 0.5, img, wt)
                a, b = np.zeros((px, 3), np.float32), np.zeros(px, np.float32)
                img.download(a)
                wt.download(b)
                d = a.astype(np.float64) / np.maximum(b.astype(np.float64), 1e-30)[:, None]
                s1 += d
                s2 += d * d
                k += 1
                for key, v in st.items():
                    st_sum[key] = max(st_sum.get(key, 0), v) if key == "bernoulli_max" else st_sum.get(key, 0) + v
            secs = time.perf_counter() - t0
            mean = s1 / k
            var_mean = np.maximum(s2 / k - mean * mean, 0.0) / max(k - 1, 1)
            return mean, var_mean, k, secs, st_sum

        def blocks(a, b=4):
            return a.reshape(res // b, b, res // b, b, 3).mean((1, 3)).reshape(-1, 3)

        for term, only in (("caustic", 1), ("image", 0)):
            sms = abi.SmsPathConfig.default(max_depth=6, caustics_only=only)
            flt = abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_SS, max_depth=6, caustics_only=only)
            render(sms, 64, passes=1, seed0=10**6)                  # first launches
            ms, vs, ks, ts, st = render(sms, 64, seconds=args.parity_sms_seconds, seed0=1)
            mf, vf, kf, tf, _ = render(flt, 256, seconds=args.parity_ref_seconds, seed0=10**7)
            ref_mean = float(mf.mean())
            rec = {"mean_sms": float(ms.mean()), "mean_pt": ref_mean, "mean_ratio": float(ms.mean()) / ref_mean, "sms_spp": ks * 64, "sms_s": round(ts, 3),
                   "pt_spp": kf * 256, "pt_s": round(tf, 2), "sms_samples_per_s": ks * 64 * px / ts,
                   "pt_samples_per_s": kf * 256 * px / tf, "sms_walks": st.get("walks", 0),
                   "sms_walks_converged": st.get("walks_converged", 0)}
            for tag, a, b, va, vb in (("pixel", ms, mf, vs, vf),
                                      ("block4", blocks(ms), blocks(mf), blocks(vs) / 16, blocks(vf) / 16)):
                d2 = ((a - b) ** 2).mean()
                noise = (va + vb).mean()
                rec[tag] = {"rel_rmse": float(np.sqrt(max(d2 - noise, 0.0)) / ref_mean),
                            "rel_rmse_raw": float(np.sqrt(d2) / ref_mean),
                            "rel_noise_sms": float(np.sqrt(va.mean()) / ref_mean),
                            "rel_noise_pt": float(np.sqrt(vb.mean()) / ref_mean)}
            leg[term] = rec
            print(f"# parity {name} {term}: {json.dumps(rec)}", file=sys.stderr, flush=True)
            if name == "pool" and only:
                # the estimator's uniqueness test merges solutions whose directions differ by less
                # than the threshold (near caustic folds): a tighter threshold against the same PT
                ctx.set_config(abi.SmsConfig.default(max_trials=10_000_000, uniqueness_threshold=1e-6,
                                                     solver_threshold=1e-5))
                mu, vu, ku, tu, _ = render(sms, 64, seconds=args.parity_sms_seconds, seed0=1)
                ctx.set_config(cfg)
                d2 = ((blocks(mu) - blocks(mf)) ** 2).mean()
                noise = (blocks(vu) / 16 + blocks(vf) / 16).mean()
                leg["caustic_uniqueness_1e-6"] = {
                    "mean_ratio": float(mu.mean()) / ref_mean, "sms_spp": ku * 64,
                    "block4_rel_rmse": float(np.sqrt(max(d2 - noise, 0.0)) / ref_mean)}
                print(f"# parity pool caustic, uniqueness 1e-6: {json.dumps(leg['caustic_uniqueness_1e-6'])}",
                      file=sys.stderr, flush=True)
        out[name] = leg
        for b in (rgb, pos, img, wt):
            b.free()
        ctx.close()
    return out


def cpu_baseline(sc, cfg, ctx, qbuf, n, out, step_stats, args):
    """CPU oracle on whole frames of the same workload as the GPU's: camera queries of seed
    1 + rep (oracle_glint_queries) estimated with seed 1 + rep, on CPU_THREADS host threads for
    ~cpu_seconds, then on one thread.  Every frame both ran must agree: the first frame's queries
    equal the GPU's byte for byte, the oracle's statistics equal the GPU whole frame's, and for
    the first frames the oracle's RGB equals the GPU's bit for bit (raises otherwise)."""
    import numpy as np
    from smsamd import abi
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    q1 = np.zeros(n, abi.GLINT_QUERY)
    qbuf.download(q1)                                 # holds the camera queries of seed 1
    checked_rgb = 0
    orc = oracle_lib.Oracle(sc, cfg, threads=CPU_THREADS)
    keys = ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max")

    def frame(rep):
        nonlocal checked_rgb
        q = orc.glint_queries(args.spp, 1 + rep)
        if rep == 0 and q.tobytes() != q1.tobytes():
            raise RuntimeError("CPU oracle and GPU camera queries of seed 1 differ")
        rgb, st = orc.glint_estimate(q, 1 + rep)
        if rep < len(step_stats):
            for k in keys:
                if st[k] != step_stats[rep][k]:
                    raise RuntimeError(f"CPU oracle and GPU differ on frame seed {1 + rep} in {k}: "
                                       f"{st[k]} != {step_stats[rep][k]}")
        if rep < 3:
            ctx.glint_estimate_frame_device(args.spp, 1 + rep, out)
            g = out.download(np.zeros((n, 3), np.float32))
            if not np.array_equal(g, rgb, equal_nan=True):
                raise RuntimeError(f"CPU oracle and GPU RGB differ on frame seed {1 + rep}: "
                                   f"{int((g != rgb).any(1).sum())} queries")
            checked_rgb += 1
        return st["walks"], st["walks_converged"]

    r16 = timed_cpu(frame, args.cpu_seconds)
    orc.close()
    orc1 = oracle_lib.Oracle(sc, cfg, threads=1)

    def frame1(rep):
        st = orc1.glint_estimate(orc1.glint_queries(args.spp, 1 + rep), 1 + rep)[1]
        return st["walks"], st["walks_converged"]
    r1 = timed_cpu(frame1, args.cpu_seconds / 3)
    orc1.close()
    return {"value": r16["converged_per_s"], "unit": "walks/s", "cores": CPU_THREADS, "kind": "port",
            "sample": f"{r16['reps']} whole frames of the same workload (camera queries generated, {n} each, "
                      f"{r16['walks']} walks, {r16['seconds']}s on {CPU_THREADS} threads); frames checked equal "
                      f"to the GPU's: {min(r16['reps'], len(step_stats))} by statistics, {checked_rgb} bit for "
                      f"bit in RGB",
            "walks_per_s": r16["walks_per_s"], "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "one_core": {"value": r1["converged_per_s"], "walks_per_s": r1["walks_per_s"], "cores": 1,
                         "sample": f"{r1['reps']} whole frames, {r1['seconds']}s"}}


def query_sample_baseline(sc, cfg, q_all, kind, seconds, stride):
    """CPU oracle on a strided sample of the given queries (every stride-th; its own random
    streams), CPU_THREADS threads for `seconds`, then one thread for a third of that."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    res = {}
    for threads, secs in ((CPU_THREADS, seconds), (1, seconds / 3)):
        orc = oracle_lib.Oracle(sc, cfg, threads=threads)
        fn = orc.ss_estimate if kind == "ss" else orc.ms_estimate

        def run(rep):
            qs = np.ascontiguousarray(q_all[(rep * 7919) % stride::stride])
            _, st = fn(qs, 1 + rep)
            return st["walks"], st["walks_converged"]
        res[threads] = timed_cpu(run, secs)
        orc.close()
    r16, r1 = res[CPU_THREADS], res[1]
    return {"value": r16["converged_per_s"], "unit": "walks/s", "cores": CPU_THREADS, "kind": "port",
            "sample": f"{r16['reps']} strided samples (1/{stride}) of the same queries, {r16['walks']} walks in "
                      f"{r16['seconds']}s on {CPU_THREADS} threads",
            "walks_per_s": r16["walks_per_s"], "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "one_core": {"value": r1["converged_per_s"], "walks_per_s": r1["walks_per_s"], "cores": 1,
                         "sample": f"{r1['reps']} samples, {r1['walks']} walks, {r1['seconds']}s"}}


if __name__ == "__main__":
    main()
