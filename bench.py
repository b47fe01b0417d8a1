"""Benchmark: Mpaths/s of the book-2 final scene at 1080x1080, 1024 spp (32x32 strata), depth 40, FP32, on one GPU.

A step is one whole render of that image (1024 passes of 1080x1080 paths, 1.19 G paths) with the scene and the
image's double sums resident in device memory. `value` is paths per second over the K timed steps (wall clock around
them; every launch is waited for). Also reported:
  - roofline: the dominant kernel's algorithmic bytes per launch (bytes per path, counted by the instrumented kernel
    over 16 passes of the same image, times the paths of one launch) over its average launch time (HIP events on the
    session's own stream, over the timed region), against 8 TB/s of HBM; `traffic` is the HBM bytes per launch that
    rocprofv3's FETCH_SIZE and WRITE_SIZE counters measured for the same workload, when profiles/hbm_traffic.txt
    holds a line for it (scaled by the counters' reading for a copy kernel of known size);
    `hbm_frac` is that measured traffic over the live launch time against 8 TB/s, `l2` the L2's bytes per launch and
    GB/s from rocprofv3's TCC request counters (profiles/l2_traffic.txt), and `issue` the kernel's active
    lanes per VALU instruction, VALU busy share and waves per SIMD from rocprofv3's SQ counters
    (profiles/valu_issue.txt) for the same workload;
  - cpu_baseline: the CPU restatement of the reference renderer (double, mt19937_64 per row, the reference's
    algorithms, a thread per processor of this job's share of the host, at most 16) timed on a bounded sample: the
    same scene at 1080x1080 with 64 spp; `host` gives nproc, the affinity, the cgroup quota and the CPU model;
  - the plain megakernel (the reference's mapping: a thread per pixel of a 16x8 block, all of a launch's passes in
    its loop), the wavefront and the persistent queue megakernel on the same workload, and the measured integrator's
    ratio to the plain one. All trace the same paths with the same random numbers (their images are the same bytes),
    so at equal spp they are at equal RMSE, and the ratio of their Mpaths/s is the ratio at equal RMSE;
  - config3: the same comparison on the triangle scene (1920x1080, 256 spp, -p mis): the autotuned integrator against
    the plain megakernel, one timed render each (--no-config3 skips it).

Other workloads (--scene): weekend (400x225, 64 spp, depth 50), glass (600x600, 1024 spp, depth 50, -p mixture),
glass-mis (the same with -p mis), triangles (1920x1080, 256 spp, depth 16, -p mis).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--integrator auto|megakernel|wavefront|plain] [--scene NAME]
                [--spp N] [--no-cpu] [--no-megakernel]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# name: (scene, width, spp, sampling, description)
WORKLOADS = {
    "final": ("final", 1080, 1024, None, "final 1080x1080, 1024 spp (32x32 strata), depth 40, FP32, -p mixture, "
                                         "seeded 2048x1024 noise texture"),
    "weekend": ("weekend", 400, 64, None, "weekend 400x225, 64 spp, depth 50, FP32"),
    "glass": ("glass", 600, 1024, None, "glass 600x600, 1024 spp, depth 50, FP32, -p mixture"),
    "glass-mis": ("glass", 600, 1024, "mis", "glass 600x600, 1024 spp, depth 50, FP32, -p mis"),
    "triangles": ("triangles", 1920, 256, None, "triangles 1920x1080, 256 spp, depth 16, FP32, -p mis, "
                                                "2.62 M triangles, 1024x512 HDR sky with a sun"),
}


def measure(rt, scene, integrator, steps, warmup, seed=0, **tuning):
    """(Mpaths/s, ms per step, average launch seconds, launches per step, session) over `steps` timed renders"""
    passes = scene.passes
    paths_per_step = scene.width * scene.height * passes
    s = rt.Session(scene, seed=seed, integrator=integrator, launch_seconds=0.1, **tuning)
    for _ in range(warmup):
        s.set_camera(scene.desc.camera)
        s.render_to(passes)
    launches = 0
    gpu_seconds = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        s.set_camera(scene.desc.camera)
        r = s.render_to(passes)
        launches += r.launches
        gpu_seconds += r.gpu_seconds
    elapsed = time.perf_counter() - t0
    value = paths_per_step * steps / elapsed / 1e6
    return value, 1000 * elapsed / steps, gpu_seconds / max(launches, 1), launches / steps, s


def host_cpus():
    """The host's processors as this process sees them: nproc (online), the ones it may run on (affinity), the cgroup's
    CPU quota (cpu.max, None when unlimited), the model name, and the threads the CPU baseline uses: one per
    processor of this job's share (the smallest of those), at most 16, the CPU share of one GPU's job on the GPU host"""
    online = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = online
    quota = None
    try:
        parts = open("/sys/fs/cgroup/cpu.max").read().split()
        if parts and parts[0] != "max":
            quota = max(1, int(int(parts[0]) / int(parts[1])))
    except (OSError, ValueError, IndexError):
        pass
    model = ""
    try:
        model = next((line.split(":", 1)[1].strip() for line in open("/proc/cpuinfo") if line.startswith("model name")),
                     "")
    except OSError:
        pass
    threads = max(1, min(16, affinity, quota or affinity))
    return {"nproc": online, "affinity": affinity, "cgroup_quota": quota, "model": model, "threads": threads}


def cpu_baseline(name: str, width: int, spp: int):
    """The oracle's C++ restatement of the reference CPU renderer (src/camera.ixx:204-272: a thread per hardware thread,
    rows handed out dynamically), on the job's share of the host's processors"""
    import oracle

    cpus = host_cpus()
    threads = cpus["threads"]
    t0 = time.perf_counter()
    img, rays = oracle.render(name, width=width, spp=spp, seed=0, threads=threads, earth_width=2048)
    seconds = time.perf_counter() - t0
    paths = img.shape[0] * img.shape[1] * spp
    return {"value": round(paths / seconds / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{name} {img.shape[1]}x{img.shape[0]}, {spp} spp, double, {seconds:.1f} s, "
                      f"{rays / paths:.2f} rays/path",
            "host": {k: cpus[k] for k in ("nproc", "affinity", "cgroup_quota", "model")}}


def traffic_key(kernel: str) -> str:
    """The kernel's key in profiles/hbm_traffic.txt: "wavefront", or "megakernel-w<waves>" (the register budget sets
    the spills, which are HBM writes; the block size doesn't change the traffic)"""
    if kernel == "wavefront":
        return kernel
    waves = next((w[1:] for w in kernel.split() if w.startswith("w") and w[1:].isdigit()), "0")
    return f"megakernel-w{waves}"


def measured_issue(workload: str, key: str):
    """Lanes per VALU instruction, VALU busy share and waves per SIMD of the dominant kernel, from the rocprofv3 SQ
    counter summary (profiles/valu_issue.txt), or None"""
    path = os.path.join(ROOT, "profiles", "valu_issue.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 5 and not line.startswith("#") and parts[0] == workload and parts[1] == key:
            return {"lanes_per_valu": float(parts[2]), "valu_busy": float(parts[3]), "waves_per_simd": float(parts[4])}
    return None


def measured_traffic(workload: str, key: str):
    """HBM bytes per launch for this workload and kernel (traffic_key), from the rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE runs of the same command, scaled by the copy kernel's calibration), or None"""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 4 and not line.startswith("#") and parts[0] == workload and parts[1] == key:
            return float(parts[2]) + float(parts[3])
    return None


def measured_l2(workload: str, key: str):
    """The L2's bytes per launch (read requests x 128 B + write requests x 64 B, the copy kernel's sizes) and hit rate
    for this workload and kernel, from the rocprofv3 TCC counter summary (profiles/l2_traffic.txt), or None"""
    path = os.path.join(ROOT, "profiles", "l2_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 5 and not line.startswith("#") and parts[0] == workload and parts[1] == key:
            return 128.0 * float(parts[2]) + 64.0 * float(parts[3]), float(parts[4])
    return None


SERVING_SCENES = ["weekend", "glass", "final", "triangles"]
SERVING_RESOLUTIONS = [(256, 256), (512, 512), (1080, 1080), (1920, 1080)]


def serving(rt, rates, per_rate, seed, tau, paths_per_step, max_active=8):
    """Config 4: a seeded Poisson stream of requests at each rate; each picks a scene of configs 0-3, a resolution and
    a log-uniform spp in [16, 1024], and is rendered progressively (power-of-two milestones) by the server, with the
    four scenes resident. Reports p50/p99 of the time from submission to the first milestone whose estimated relative
    RMSE is below tau (the request's last milestone when its spp never gets there), p50/p99 time to completion, and
    the aggregate Mpaths/s. The relative RMSE of m passes is sigma/sqrt(m), sigma per scene from a pilot: two
    independent 16-pass renders at 128 pixels wide, mean((A - B)^2) / 2 = sigma^2 / 16 per value, over the mean^2."""
    import numpy as np

    views = {}
    scenes = {}
    for name in SERVING_SCENES:
        for w, h in SERVING_RESOLUTIONS:
            built = rt.build_scene(name, width=w, height=h, samples=1)
            views[(name, w, h)] = rt.CameraDesc.from_buffer_copy(built.desc.camera)  # a copy: the scene is freed
    sigma = {}
    for name in SERVING_SCENES:
        pilot = rt.build_scene(name, width=128, samples=1)
        imgs = []
        for s in (101, 202):
            with rt.Session(pilot, seed=s) as sess:
                sess.render_to(16)
                imgs.append(sess.pixels())
        var = 16 * np.mean((imgs[0] - imgs[1]) ** 2) / 2
        sigma[name] = float(np.sqrt(var) / max(np.mean(imgs[0] + imgs[1]) / 2, 1e-9))
        print(f"serving pilot {name}: relative sigma per pass {sigma[name]:.3f}", file=sys.stderr, flush=True)
    server = rt.Server(memory_budget_bytes=64 << 30, max_active=max_active, passes_per_step=8,
                       paths_per_step=paths_per_step)
    for name in SERVING_SCENES:
        scenes[name] = server.add_scene(rt.build_scene(name, width=256, height=256, samples=1))
    results = {}
    rng = np.random.default_rng(seed)
    for rate in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, per_rate))
        picks = [(SERVING_SCENES[rng.integers(4)], SERVING_RESOLUTIONS[rng.integers(4)],
                  int(round(np.exp(rng.uniform(np.log(16), np.log(1024)))))) for _ in range(per_rate)]
        pending = list(range(per_rate))
        live = {}
        paths = 0
        t0 = time.perf_counter()
        milestones = {}
        while pending or live:
            now = time.perf_counter() - t0
            while pending and arrivals[pending[0]] <= now:
                k = pending.pop(0)
                name, (w, h), spp = picks[k]
                rid = server.submit_view(scenes[name], views[(name, w, h)], 1000 + k, spp)
                live[rid] = (k, now)
                milestones[rid] = []
                paths += w * h * spp
            if not live:
                time.sleep(max(0.0, arrivals[pending[0]] - (time.perf_counter() - t0)))
                continue
            server.step()
            now = time.perf_counter() - t0
            for rid in list(live):
                st = server.status(rid)
                if not milestones[rid] or milestones[rid][-1][0] != st.milestone:
                    milestones[rid].append((st.milestone, now))
                if st.state >= 2:
                    live[rid] = live[rid] + (now, st.state)
                    k = live[rid][0]
                    name = picks[k][0]
                    need = (sigma[name] / tau) ** 2
                    hit = next((t for m, t in milestones[rid] if m >= need), milestones[rid][-1][1])
                    results.setdefault(rate, []).append((hit - live[rid][1], now - live[rid][1], st.state))
                    server.release(rid)
                    del live[rid]
        wall = time.perf_counter() - t0
        r = results[rate]
        target = np.array([x[0] for x in r])
        done = np.array([x[1] for x in r])
        results[rate] = {"rate_req_s": rate, "requests": per_rate, "failed": sum(1 for x in r if x[2] != 2),
                         "p50_time_to_target_s": round(float(np.percentile(target, 50)), 3),
                         "p99_time_to_target_s": round(float(np.percentile(target, 99)), 3),
                         "p50_time_to_done_s": round(float(np.percentile(done, 50)), 3),
                         "p99_time_to_done_s": round(float(np.percentile(done, 99)), 3),
                         "aggregate_mpaths_s": round(paths / wall / 1e6, 1), "wall_s": round(wall, 2)}
        print(json.dumps(results[rate]), file=sys.stderr, flush=True)
    server.close()
    return [results[r] for r in rates], sigma


def rmse_protocol(rt, workload, width, oracle_spp):
    """RMSE of a workload's FP32 render against an FP64 render of `oracle_spp` passes (the oracle: the FP64 megakernel,
    RT_FLOAT64_ALL, which renders the quality integrator's scenes in float64 too), at a reduced width (the same scene,
    camera and depth; fewer pixels so that the oracle's 64k passes take a minute), for the megakernel and the
    wavefront at the workload's spp, and the spp each needs to reach the megakernel's RMSE. The two integrators trace
    the same paths, so their RMSEs are equal at equal spp."""
    import numpy as np

    name, _, spp, sampling, description = WORKLOADS[workload]
    oracle_scene = rt.build_scene(name, width=width, samples=oracle_spp, strata=1, sampling=sampling)
    t0 = time.perf_counter()
    with rt.Session(oracle_scene, seed=12345, precision="float64_all", integrator="megakernel",
                    launch_seconds=0.2) as s:
        s.render_to(oracle_spp)
        oracle = s.pixels()
    oracle_seconds = time.perf_counter() - t0
    out = {"workload": description, "width": oracle.shape[1], "height": oracle.shape[0],
           "oracle": f"FP64 megakernel, {oracle_spp} passes, {oracle_seconds:.1f} s",
           "oracle_mean": float(oracle.mean())}
    scene = rt.build_scene(name, width=width, samples=spp, sampling=sampling)
    for integrator in ("megakernel", "wavefront"):
        with rt.Session(scene, seed=7, integrator=integrator) as s:
            t0 = time.perf_counter()
            s.render_to(scene.passes)
            seconds = time.perf_counter() - t0
            img = s.pixels()
        err = float(np.sqrt(np.mean((img - oracle) ** 2)))
        out[integrator] = {"spp": scene.passes, "rmse": err, "relative_rmse": err / max(float(oracle.mean()), 1e-12),
                           "seconds": round(seconds, 3),
                           "mpaths_s": round(img.shape[0] * img.shape[1] * scene.passes / seconds / 1e6, 1)}
    # the oracle's own noise: its RMSE against the truth is about the FP32 render's times sqrt(spp / oracle_spp)
    out["oracle_noise_share"] = float(np.sqrt(scene.passes / oracle_spp))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--integrator", default="auto", choices=["auto", "megakernel", "wavefront", "plain"])
    ap.add_argument("--scene", default="final", choices=sorted(WORKLOADS))
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-megakernel", action="store_true",
                    help="skip the comparison renders (the plain megakernel, the wavefront, the queue megakernel)")
    ap.add_argument("--no-config3", action="store_true",
                    help="skip config 3's comparison (the triangle scene: the autotuned integrator and the plain "
                         "megakernel) that the final scene's line carries")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the plain copy kernel over this many bytes (the HBM counters' calibration)")
    ap.add_argument("--wavefront-graph", type=int, default=0, help="bounces per hipGraph (-1: none, every kernel launched on its own)")
    ap.add_argument("--wavefront-drain", type=int, default=0, help="active slots below which the pool drains")
    ap.add_argument("--wavefront-slots", type=int, default=0, help="the wavefront pool's size")
    ap.add_argument("--megakernel-waves", type=int, default=0, choices=[0, 5, 6, 7],
                    help="the megakernel's register budget in waves per SIMD (0: the compiler's choice), with "
                         "--integrator megakernel")
    ap.add_argument("--block-size", type=int, default=0, choices=[0, 64, 128, 192, 256],
                    help="the megakernel's threads per block (0: the tuned default), with --integrator megakernel")
    ap.add_argument("--tree", default="default", choices=["default", "median", "sah", "two_level"],
                    help="the BVH's splits: the library's default (SAH, one level), the reference's object median, "
                         "the SAH, or the SAH in two levels (meshes placed several times walked through their "
                         "instances)")
    ap.add_argument("--nodes", default="auto", choices=["auto", "float", "packed"],
                    help="the BVH's inner nodes: the library's choice (packed for the quality integrator's trees "
                         "beyond 16 MiB of float nodes), 128-byte float nodes, or 64-byte packed nodes")
    ap.add_argument("--serving", action="store_true", help="config 4: the serving loop under Poisson request streams")
    ap.add_argument("--rates", default="1,4,16,64")
    ap.add_argument("--requests", type=int, default=16, help="requests per rate (serving)")
    ap.add_argument("--tau", type=float, default=0.05, help="target relative RMSE (serving)")
    ap.add_argument("--max-active", type=int, default=8, help="serving: requests rendering at once")
    ap.add_argument("--paths-per-step", type=int, default=1 << 24,
                    help="serving: a small view's steps raised to this many paths (0: 8 passes a step alone)")
    ap.add_argument("--rmse", default="", help="the RMSE protocol for a workload (final, weekend, glass, glass-mis, "
                                               "triangles)")
    ap.add_argument("--rmse-width", type=int, default=270)
    ap.add_argument("--rmse-oracle-spp", type=int, default=65536)
    args = ap.parse_args()
    if args.serving:
        import raytracing_amd as rt

        rows, sigma = serving(rt, [float(r) for r in args.rates.split(",")], args.requests, 4, args.tau,
                              args.paths_per_step, args.max_active)
        print(json.dumps({"metric": "serving p99 time to target noise (config 4)", "value": rows[-1]["p99_time_to_target_s"],
                          "unit": "s", "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
                          "config": {"workload": f"Poisson streams, {args.requests} requests per rate, scenes "
                                                 f"{SERVING_SCENES}, resolutions {SERVING_RESOLUTIONS}, spp "
                                                 f"log-uniform [16, 1024], target relative RMSE {args.tau}",
                                     "passes_per_step": 8, "paths_per_step": args.paths_per_step,
                                     "max_active": args.max_active,
                                     "sigma_per_pass": sigma},
                          "rates": rows}), flush=True)
        return
    if args.rmse:
        import raytracing_amd as rt

        print(json.dumps(rmse_protocol(rt, args.rmse, args.rmse_width, args.rmse_oracle_spp)), flush=True)
        return
    if args.copy_probe:
        import raytracing_amd as rt

        seconds = rt.copy_probe(args.copy_probe, 5)
        print(json.dumps({"copy_bytes": args.copy_probe, "seconds": seconds,
                          "GB/s": 2 * args.copy_probe / seconds / 1e9}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (--gpus {args.gpus} asked)",
              file=sys.stderr)
        sys.exit(2)

    import raytracing_amd as rt

    name, width, spp, sampling, description = WORKLOADS[args.scene]
    width = args.width or width
    spp = args.spp or spp
    scene = rt.build_scene(name, width=width, samples=spp, sampling=sampling)
    device = rt.query_device()
    print(f"{args.scene}: {scene.width}x{scene.height}, {scene.passes} spp; device {device.name.decode()} "
          f"({device.compute_units} CUs)", file=sys.stderr, flush=True)
    tuning = {k: v for k, v in (("wavefront_graph", args.wavefront_graph), ("wavefront_drain", args.wavefront_drain),
                                ("wavefront_slots", args.wavefront_slots),
                                ("megakernel_waves", args.megakernel_waves), ("block_size", args.block_size)) if v}
    tuning["tree"] = args.tree
    tuning["nodes"] = args.nodes
    value, ms, avg_launch, launches_per_step, session = measure(rt, scene, args.integrator, args.steps, args.warmup,
                                                                **tuning)
    kernel = session.kernel().name.decode()
    print(f"{args.integrator}: {value:.1f} Mpaths/s, {ms:.1f} ms/step, {launches_per_step:.1f} launches/step, "
          f"kernel {kernel}", file=sys.stderr, flush=True)

    # algorithmic bytes per path, counted over 16 passes of the same image by the instrumented megakernel
    traffic = session.traffic(16)
    bytes_per_path = traffic.bytes_per_path
    paths_per_launch = scene.width * scene.height * scene.passes / launches_per_step
    achieved = bytes_per_path * paths_per_launch / avg_launch / 1e9
    events = {n: int(traffic.events[i]) for i, n in enumerate(rt.TRAFFIC_EVENTS) if traffic.events[i]}
    phase_cycles = sum(traffic.phase_lane_cycles)
    phases = {n: round(traffic.phase_lane_cycles[i] / phase_cycles, 4) for i, n in enumerate(rt.TRAFFIC_PHASES)} \
        if phase_cycles else None
    print(f"traffic: {bytes_per_path:.0f} B/path, {traffic.rays / traffic.paths:.2f} rays/path, "
          f"{traffic.shadow_rays / traffic.paths:.2f} shadow rays/path, {events}; lane-time by phase {phases}",
          file=sys.stderr, flush=True)
    session.close()

    # the other integrators on the same workload: the plain megakernel (the reference's mapping, the baseline), the
    # wavefront and the persistent queue megakernel. All trace the same paths and sum them in the same order (the same
    # image bytes), so at equal spp they are at equal RMSE and the ratio of their Mpaths/s is the ratio at equal RMSE.
    others = {}
    if not args.no_megakernel:
        for other in ("plain", "wavefront", "megakernel"):
            if other == args.integrator:
                others[other] = round(value, 2)
                continue
            ov, _, _, _, os_ = measure(rt, scene, other, 1, 1, tree=args.tree, nodes=args.nodes)
            os_.close()
            others[other] = round(ov, 2)
            print(f"{other}: {ov:.1f} Mpaths/s", file=sys.stderr, flush=True)

    # config 3 beside the headline: the batched integrator the session's autotuner picks against the plain megakernel
    # on the triangle scene at its full size (1920x1080, 256 spp), one timed render each
    config3 = None
    if args.scene == "final" and not args.no_config3:
        c3_name, c3_width, c3_spp, c3_sampling, c3_description = WORKLOADS["triangles"]
        c3 = rt.build_scene(c3_name, width=c3_width, samples=c3_spp, sampling=c3_sampling)
        config3 = {"workload": c3_description}
        for c3_integrator in ("auto", "plain"):
            v3, ms3, _, _, s3 = measure(rt, c3, c3_integrator, 1, 1)
            k3 = s3.kernel().name.decode()
            s3.close()
            config3[c3_integrator] = {"mpaths_s": round(v3, 2), "ms_per_step": round(ms3, 2), "kernel": k3}
            print(f"config 3 {c3_integrator}: {v3:.1f} Mpaths/s, kernel {k3}", file=sys.stderr, flush=True)
        config3["vs_plain_megakernel"] = round(config3["auto"]["mpaths_s"] / config3["plain"]["mpaths_s"], 3)

    cpu = None
    if not args.no_cpu and name != "triangles" and sampling is None:
        cpu = cpu_baseline(name, width, 64 if name == "final" else 16)
        print(f"cpu baseline: {cpu}", file=sys.stderr, flush=True)

    # (the profiles' lines name the two-level tree's workload apart: its kernels and traffic are its own)
    profiled = args.scene + ("-two-level" if args.tree == "two_level" else "")
    hbm = measured_traffic(profiled, traffic_key(kernel))
    issue = measured_issue(profiled, traffic_key(kernel))
    l2 = measured_l2(profiled, traffic_key(kernel))
    plain = others.get("plain")
    line = {
        "metric": "Mpaths/s (1080^2 final scene)" if args.scene == "final" else f"Mpaths/s ({args.scene})",
        "value": round(value, 2),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": description, "integrator": args.integrator, "tree": args.tree, "nodes": args.nodes,
                   "kernel": kernel,
                   "paths_per_step": scene.width * scene.height * scene.passes,
                   "rays_per_path": round(traffic.rays / traffic.paths, 3),
                   "shadow_rays_per_path": round(traffic.shadow_rays / traffic.paths, 3),
                   "plain_megakernel_mpaths_s": plain,
                   "wavefront_mpaths_s": others.get("wavefront"),
                   "megakernel_mpaths_s": others.get("megakernel"),
                   "vs_plain_megakernel": round(value / plain, 3) if plain else None,
                   "wavefront_vs_plain_megakernel": round(others["wavefront"] / plain, 3)
                   if plain and others.get("wavefront") else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if hbm is None else round(hbm),
                     # measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE) over the live launch time
                     "hbm_frac": None if hbm is None else round(hbm / avg_launch / 1e9 / HBM_PEAK_GBS, 4),
                     "issue": issue,
                     # the L2's measured bytes per launch (requests x their sizes) over the live launch time
                     "l2": None if l2 is None else {"bytes_per_launch": round(l2[0]),
                                                    "GB/s": round(l2[0] / avg_launch / 1e9, 1), "hit_rate": l2[1]},
                     # the counting instance's lanes' time by phase (queue, walk, media, shade)
                     "lane_time_by_phase": phases,
                     "bytes_per_path": round(bytes_per_path, 1), "avg_launch_ms": round(1000 * avg_launch, 3),
                     "paths_per_launch": round(paths_per_launch)},
        "cpu_baseline": cpu,
    }
    if config3 is not None:
        line["config3"] = config3
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
