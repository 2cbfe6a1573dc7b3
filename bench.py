"""Benchmark: Mrays/s of the GPU path tracer on one MI355X.

Workload (the configuration the headline metric is quoted on): a displaced
icosphere of 1,003,520 triangles (frequency-224 geodesic, 3-octave value
noise amplitude 0.05, area-weighted smooth normals, 70% matte / 30% metal by
face hash), two r = 0.1 lights of radiance 50, 1920x1080, 64 spp, 10
bounces, Russian roulette from bounce 3, NEE on, FP32. One step = one full
frame. Rays are counted as the reference counts them: one per closest-hit
query and one per light sample that reaches the occlusion test
(renderer.hpp:19,72,110).

Besides the timed steps it reports:
  roofline      the traversal stage (the dominant kernel) against HBM:
                its algorithmic bytes are what a launch must move through
                HBM at least: per query the ray in and the hit out (32 + 16
                bytes), and the scene's node and leaf arrays once (the
                per-visit node, leaf and primitive reads are served by L2 and
                the 256 MB Infinity Cache, which holds the scene; they are
                reported apart as bytes served), over its HIP-event time per
                launch. "traffic" is the HBM bytes per launch that rocprofv3's
                FETCH_SIZE and WRITE_SIZE counters read for the same launches,
                measured in this run by child rocprofv3 passes over a
                one-frame probe (scaled by what they read for a 1 GiB copy),
                and "issue" the VALU busy fraction and lane use from an SQ
                counter pass: the limit the kernel actually meets;
  cpu_baseline  the reference's own renderer (oracle/_ref, its headers
                compiled with its flags) on the host cores, on a bounded
                sample of the same frame, best of three runs; the CPU model
                and sockets are reported.
  configs       the sphere field (640x360, 16 spp) and the instancing scene
                (4096 instances of 50k/200k/500k-triangle meshes, 1080p,
                32 spp), each against the reference renderer.
  batch         the batched workload: 64 seeded jobs (sphere fields of
                50-2000 spheres or displaced meshes of 10k-1M triangles;
                256^2 to 1024^2; 4, 16 or 64 spp) in one pt_render_batch call
                after the autotuner, against the same jobs rendered one
                pt_render call at a time.
  guided        the hard-lighting room (a 1M-triangle displaced mesh, lamps
                behind a partition with a 0.5 x 0.5 opening) at 1920x1080:
                relMSE against a 4096-spp plain render, plain NEE against the
                learned guide at equal device time (a quarter of it spent
                training), and the time the guided tracer needs to reach the
                plain tracer's relMSE.

python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s (1 GPU); equal-time relMSE guided vs plain NEE, 1080p 1M tris"
HBM_PEAK_GBS = 8000.0
# The traversal stage's kernels, as rocprofv3 names their dispatches: bounce
# 0's camera rays walked as 64-ray packets, then one k_extend per bounce.
EXTEND_KERNEL = ("ptd::k_camera_packet<", "ptd::k_extend<")
COPY_KERNEL = "copyBuffer"
SIMDS_PER_CU = 4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traversal_scene_bytes(ptgp):
    """Bytes of the scene arrays the traversal reads: the 4-wide nodes, the
    leaf records (48 bytes per primitive reference) and the instances."""
    nodes, order, inst = (int(ptgp[3 + 2 * k]) for k in (0, 1, 4))
    return nodes + 48 * (order // 4) + inst


def counter_passes(args, seconds=150):
    """Live counters for the traversal stage: three child rocprofv3 runs of
    this script's one-frame probe (the default frame, then a 1 GiB device
    copy of known traffic), one counter set each (FETCH_SIZE; WRITE_SIZE; SQ
    VALU counters with GRBM_GUI_ACTIVE), each under its own time limit.
    Returns per-kernel sums {counter: {kernel: (value, dispatches)}}, or None
    where rocprofv3 is missing or a pass fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    passes = (("fetch", ["FETCH_SIZE"]), ("write", ["WRITE_SIZE"]),
              ("valu", ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "GRBM_GUI_ACTIVE"]))
    env = dict(os.environ, TMPDIR="/tmp")
    out = {}
    for name, counters in passes:
        d = tempfile.mkdtemp(prefix=f"pt_counters_{name}_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", str(seconds), exe, "--pmc", *counters, "--output-format", "csv", "-d", d,
               "-o", name, "--", sys.executable, os.path.abspath(__file__), "--only", "probe",
               "--width", str(args.width), "--height", str(args.height), "--spp", str(args.spp),
               "--scale", str(args.scale), "--stack", str(args.stack), "--copy-probe", "1024"]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        log(f"counters {name}: rocprofv3 exit {r.returncode} in {time.perf_counter() - t0:.1f} s, {len(files)} files")
        if r.returncode != 0 or not files:
            log(r.stdout[-2000:])
            shutil.rmtree(d, ignore_errors=True)
            return None
        for f in files:
            with open(f, newline="") as fh:
                for row in csv.DictReader(fh):
                    kern, cname = row.get("Kernel_Name", ""), row.get("Counter_Name", "")
                    val = float(row.get("Counter_Value", 0) or 0)
                    acc = out.setdefault(cname, {}).setdefault(kern, [0.0, 0])
                    acc[0] += val
                    acc[1] += 1
        shutil.rmtree(d, ignore_errors=True)
    return out


def kernel_sum(counts, counter, pattern):
    """(sum, dispatches) of a counter over the kernels whose name holds pattern (or one of a tuple of them)."""
    pats = (pattern,) if isinstance(pattern, str) else pattern
    total, n = 0.0, 0
    for kern, (v, k) in counts.get(counter, {}).items():
        if any(p in kern for p in pats):
            total += v
            n += k
    return total, n


def traversal_counters(counts, cus):
    """HBM bytes per traversal-stage launch (FETCH_SIZE and WRITE_SIZE, each
    scaled by what it reads for the 1 GiB copy), and the stage's VALU busy
    fraction and lane use: busy = 4 cycles per wave64 VALU instruction on a
    16-lane SIMD over the SIMDs' cycles in the launches (GRBM_GUI_ACTIVE is
    summed over the 8 XCDs); lane use = VALU thread-cycles over 64 x the
    waves' VALU cycles."""
    copy = float(1 << 30)
    f_copy, _ = kernel_sum(counts, "FETCH_SIZE", COPY_KERNEL)
    w_copy, _ = kernel_sum(counts, "WRITE_SIZE", COPY_KERNEL)
    f_ext, n_ext = kernel_sum(counts, "FETCH_SIZE", EXTEND_KERNEL)
    w_ext, _ = kernel_sum(counts, "WRITE_SIZE", EXTEND_KERNEL)
    if not (f_copy and w_copy and n_ext):
        return None
    f_scale, w_scale = copy / (f_copy * 1024), copy / (w_copy * 1024)
    traffic = (f_ext * 1024 * f_scale + w_ext * 1024 * w_scale) / n_ext
    out = {"traffic": int(traffic), "fetch_scale": round(f_scale, 3), "write_scale": round(w_scale, 3),
           "launches_profiled": n_ext}
    insts, _ = kernel_sum(counts, "SQ_INSTS_VALU", EXTEND_KERNEL)
    active, _ = kernel_sum(counts, "SQ_ACTIVE_INST_VALU", EXTEND_KERNEL)
    thread_cycles, _ = kernel_sum(counts, "SQ_THREAD_CYCLES_VALU", EXTEND_KERNEL)
    gui, _ = kernel_sum(counts, "GRBM_GUI_ACTIVE", EXTEND_KERNEL)
    if insts and active and gui:
        cycles = gui / 8
        out["valu_busy"] = round(4 * insts / (SIMDS_PER_CU * cus * cycles), 3)
        out["lane_use"] = round(thread_cycles / (64 * active), 3)
        out["valu_instructions_per_launch"] = int(insts / n_ext)
    return out


def host_cpu():
    """The host's CPU model, sockets and logical CPUs (from lscpu)."""
    import subprocess

    info = {"model": "unknown", "sockets": None, "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = int(v.strip()) if v.strip().isdigit() else v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(ptsd, width, height, spp_full, seconds, repeats=3):
    """The reference renderer on the host cores over a bounded sample of the
    frame: every row at a reduced spp, sized to about `seconds` of work a
    run; the best of `repeats` runs with the same seed (bench/native.sh's
    best of three)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    threads = max(1, min(16, os.cpu_count() or 1))
    if ob.ref_available():
        world = ob.RefWorld(ptsd)
        kind = "reference"

        def run(spp):
            _, _, rays, secs = world.render(width, height, spp, seed=1, threads=threads)
            return rays, secs
    else:
        import photon_tracer_amd as pt

        ptgp = pt.pack(ptsd)
        world = None
        kind = "port"

        def run(spp):
            t0 = time.perf_counter()
            _, _, rays = ob.or_render(ptgp, width, height, spp, seed=1, threads=threads)
            return rays, time.perf_counter() - t0
    rays, secs = run(1)
    spp = max(1, min(spp_full, int(seconds / max(secs, 1e-3))))
    runs = [run(spp) for _ in range(repeats)]
    rays, secs = min(runs, key=lambda r: r[1])
    if world is not None:
        world.close()
    cpu = host_cpu()
    return {
        "value": round(rays / secs / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": kind,
        "sample": f"full {width}x{height} frame at {spp} spp (of {spp_full}), seed 1, {rays} rays in {secs:.2f} s "
                  f"(best of {repeats}: {', '.join(f'{r[1]:.2f}' for r in runs)} s), rows claimed by {threads} "
                  "threads from an atomic counter as the reference CLI does",
        "cpu_model": cpu["model"], "sockets": cpu["sockets"], "logical_cpus": cpu["logical_cpus"],
    }


def batch_leg(pt, n_jobs, steps, tune=True, graph=True, compare=True):
    """The batched workload: n_jobs seeded jobs, generated and packed on host
    threads, merged into one device scene, tuned, then rendered in one call
    per step; the same jobs one pt_render call at a time for comparison."""
    from concurrent.futures import ThreadPoolExecutor

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=16) as ex:
        jobs = list(ex.map(lambda s: pt.generate_job(s, 1.0), range(n_jobs)))
        ptgps = list(ex.map(lambda j: pt.pack(j.ptsd), jobs))
    t1 = time.perf_counter()
    n_pix = sum(j.width * j.height for j in jobs)
    n_paths = sum(j.width * j.height * j.spp for j in jobs)
    n_tris = sum(j.primitives for j in jobs if j.kind == 1)
    log(f"batch: {n_jobs} jobs ({sum(j.kind == 1 for j in jobs)} meshes, {n_tris} triangles), {n_pix} pixels, "
        f"{n_paths} paths; generated and packed in {t1 - t0:.1f} s")
    bjobs = [pt.BatchJob(k, j.width, j.height, j.spp, seed=j.seed) for k, j in enumerate(jobs)]
    out = {}
    with pt.Batch(ptgps) as b:
        t2 = time.perf_counter()
        tuned = b.autotune(bjobs) if tune else None
        if tuned:
            log(f"batch: autotuned in {time.perf_counter() - t2:.1f} s: stack {tuned.stack_entries}, "
                f"blocks/CU {tuned.blocks_per_cu}, streams {tuned.streams}, paths {tuned.max_paths}")
        out_buf = pt.DeviceBuffer(12 * n_pix)   # the images stay in HBM, as in the timed frames above
        _, st = b.render(bjobs, device_out=out_buf, graph=graph)   # warmup (captures the graphs)
        rays, ms, walls = 0, 0.0, []
        for k in range(steps):
            t3 = time.perf_counter()
            _, st = b.render(bjobs, device_out=out_buf, graph=graph)
            walls.append(time.perf_counter() - t3)
            rays += st.rays
            ms += st.render_ms
            log(f"batch step {k}: {st.rays} rays, {st.render_ms:.1f} ms on the device, {walls[-1] * 1e3:.1f} ms wall")
        out = {
            "workload": f"{n_jobs} seeded jobs (config-0/1 types), {n_pix} pixels, {n_paths} paths, 10 bounces",
            "mrays_per_s": round(rays / (ms * 1e-3) / 1e6, 1),
            "ms_per_batch": round(ms / steps, 2),
            "wall_mrays_per_s": round(rays / sum(walls) / 1e6, 1),
        }
        out["mode"] = "shared-queue passes: slices of every job in one bounce loop per pass"
        out["passes"] = st.batches
        if tuned:
            out["tuned"] = {"stack_entries": tuned.stack_entries, "blocks_per_cu": tuned.blocks_per_cu,
                            "max_paths": tuned.max_paths}
            out["tuner_tried"] = [{"stack_entries": t[0], "blocks_per_cu": t[1], "max_paths": t[3],
                                   "mrays_per_s": round(t[4], 1)} for t in tuned.tried]
        if compare:
            # The same jobs as per-job bounce loops over 4 HIP streams (PT_BATCH_STREAMS), for comparison.
            b.render(bjobs, device_out=out_buf, graph=graph, shared=False, streams=4)
            srays, sms = 0, 0.0
            for k in range(steps):
                _, sst = b.render(bjobs, device_out=out_buf, graph=graph, shared=False, streams=4)
                srays += sst.rays
                sms += sst.render_ms
            out["per_job_streams_mrays_per_s"] = round(srays / (sms * 1e-3) / 1e6, 1)
        out_buf.free()
    if not compare:
        return out
    # The same jobs one pt_render call at a time, each its own scene, resident
    # and warmed up; device time summed and wall time of the calls.
    scenes = [pt.Scene(g) for g in ptgps]
    images = [pt.DeviceBuffer(12 * j.width * j.height) for j in jobs]
    for s, j, im in zip(scenes, jobs, images):
        s.render(j.width, j.height, j.spp, seed=j.seed, device_out=im)
    rays, ms = 0, 0.0
    t4 = time.perf_counter()
    for s, j, im in zip(scenes, jobs, images):
        _, _, st = s.render(j.width, j.height, j.spp, seed=j.seed, device_out=im)
        rays += st.rays
        ms += st.render_ms
    wall = time.perf_counter() - t4
    for s, im in zip(scenes, images):
        im.free()
        s.close()
    out["one_call_per_job_mrays_per_s"] = round(rays / (ms * 1e-3) / 1e6, 1)
    out["one_call_per_job_wall_mrays_per_s"] = round(rays / wall / 1e6, 1)
    out["batched_over_one_call_per_job"] = round(out["mrays_per_s"] / out["one_call_per_job_mrays_per_s"], 3)
    log(f"batch: {out['mrays_per_s']} Mrays/s batched ({out['wall_mrays_per_s']} wall), "
        f"{out['one_call_per_job_mrays_per_s']} one call per job ({out['one_call_per_job_wall_mrays_per_s']} wall)")
    return out


# The guide network's FLOPs per evaluated hit: four 64x64 layers and a 64x32
# output layer, two FLOPs a multiply-add (biases not counted).
MLP_FLOPS = 2 * (4 * 64 * 64 + 64 * 32)
MFMA_F16_PEAK_TFLOPS = 2500.0   # MI355X dense FP16 matrix peak


def network_roofline(s, W, H, spp):
    """The guide's inference kernel against the FP16 MFMA peak: one counting
    render (diffuse hits evaluated) and one timing render (HIP events around
    each inference launch) of the same guided frame."""
    _, _, cst = s.render(W, H, spp, seed=3, guide=True, count=True)
    _, _, tst = s.render(W, H, spp, seed=3, guide=True, timing=True)
    achieved = cst.guide_queries * MLP_FLOPS / (tst.guide_ms * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": round(achieved, 1), "peak": MFMA_F16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / MFMA_F16_PEAK_TFLOPS, 4), "kernel": "k_guide_infer (all guided bounces)",
            "workload": f"{W}x{H}, {spp} spp guided", "queries": cst.guide_queries,
            "flops_per_query": MLP_FLOPS, "ms": round(tst.guide_ms, 3), "launches": 10 * tst.batches}


def guided_leg(pt, ref_spp, spp_plain=1024, steps_per_batch=128, guide_seeds=(2, 3, 4)):
    """Equal-time relMSE of the guided and plain tracers on the hard-lighting
    room at 1080p. The reference is ref_spp plain samples in two independent
    halves; relMSE subtracts the reference's own variance (estimated per pixel
    from the halves), so a reference only a few times the test spp does not
    pull both tracers' errors together. Budget: the plain render at
    spp_plain; the guided tracer trains for a quarter of it (1-spp batches,
    steps_per_batch Adam steps after each) and renders for the rest. Online
    training is noisy (its result varies with the network's initial weights),
    so the protocol runs once per guide seed and reports each run and the
    median."""
    import numpy as np

    W, H = 1920, 1080
    t0 = time.perf_counter()
    ptgp = pt.pack(pt.generate(4, 1, 1.0))
    log(f"guided: room scene packed in {time.perf_counter() - t0:.1f} s")
    chunk = 256
    with pt.Scene(ptgp) as s:
        halves = []
        for h in range(2):
            acc = np.zeros((H, W, 3), np.float64)
            n = max(1, ref_spp // (2 * chunk))
            for k in range(n):
                img, _, st = s.render(W, H, chunk, seed=90000 + 1000 * h + k)
                acc += img
            halves.append(acc / n)
            log(f"guided: reference half {h}: {n * chunk} spp ({st.render_ms:.0f} ms a chunk)")
        ref = 0.5 * (halves[0] + halves[1])
        ref_var = 0.25 * (halves[0] - halves[1]) ** 2   # per pixel: Var of the mean of the two halves
        den = ref ** 2 + 0.01

        def rel_mse(img):
            return float(np.mean(((img - ref) ** 2 - ref_var) / den))

        plain, _, sp = s.render(W, H, spp_plain, seed=11)
        e_plain = rel_mse(plain)
        budget = sp.render_ms
        runs = []
        for gs in guide_seeds:
            g = pt.Guide(s, seed=gs, learning_rate=3e-3, steps_per_batch=steps_per_batch)
            s.set_guide(g)
            train_ms, n = 0.0, 0
            while train_ms < 0.25 * budget:
                _, _, st = s.render(W, H, 1, seed=5000 + n, train=True, samples_per_batch=1)
                train_ms += st.render_ms
                n += 1
            steps, used, loss = g.stats()
            # The guided render's cost per spp, measured at the batch size the
            # plain render ran at (the library's choice for this frame) once
            # that shape's graph is captured (the capture's host time falls
            # inside the first render's device-time events), then the rest of
            # the budget in whole batches of that size, so both tracers run
            # the same launch shapes.
            spb = max(16, sp.samples_per_batch)
            s.render(W, H, spb, seed=6, guide=True)   # captures the guided graph of this batch shape
            _, _, st1 = s.render(W, H, spb, seed=7, guide=True)
            per_spp = st1.render_ms / spb
            spp_g = max(spb, int((budget - train_ms) / per_spp) // spb * spb)
            guided, _, sg = s.render(W, H, spp_g, seed=21, guide=True, samples_per_batch=spb)
            e_guided = rel_mse(guided)
            guided_eq, _, _ = s.render(W, H, spp_plain, seed=41, guide=True)
            e_guided_eq = rel_mse(guided_eq)
            if not runs:
                net = network_roofline(s, W, H, 16)
                log(f"guided: network roofline {json.dumps(net)}")
            # The time to reach the plain tracer's relMSE, measured: a guided
            # render (new seed) at the sample count where relMSE ~ 1/spp
            # predicts the match, a multiple of 16 spp in batches of the
            # library's choice (as the plain render's), and 25% more samples
            # at a time (at most three times) until its relMSE is at most
            # plain's. Training time counts.
            t_pred = train_ms + sg.render_ms * e_guided / e_plain
            spp_m = max(16, math.ceil(spp_g * e_guided / e_plain / 16) * 16)
            for attempt in range(4):
                gm, _, sm = s.render(W, H, spp_m, seed=61 + attempt, guide=True)
                e_m = rel_mse(gm)
                if e_m <= e_plain or attempt == 3:
                    break
                spp_m = math.ceil(spp_m * 1.25 / 16) * 16
            s.set_guide(None)
            g.close()
            run = {"guide_seed": gs, "train_ms": round(train_ms, 1), "train_batches": n, "adam_steps": steps,
                   "spp": spp_g, "render_ms": round(sg.render_ms, 1), "relmse": round(e_guided, 6),
                   "equal_time_relmse_ratio": round(e_guided / e_plain, 3),
                   "equal_spp_relmse_ratio": round(e_guided_eq / e_plain, 3),
                   "guided_over_plain_ms_per_spp": round(per_spp / (sp.render_ms / spp_plain), 3),
                   "time_to_plain_relmse": {"spp": spp_m, "render_ms": round(sm.render_ms, 1),
                                            "relmse": round(e_m, 6), "relmse_over_plain": round(e_m / e_plain, 3),
                                            "reached": bool(e_m <= e_plain)},
                   "time_to_plain_relmse_ratio": round((train_ms + sm.render_ms) / budget, 3),
                   "time_to_plain_relmse_ratio_extrapolated": round(t_pred / budget, 3)}
            log(f"guided: {json.dumps(run)}")
            runs.append(run)

    def med(k):
        return float(np.median([r[k] for r in runs]))

    out = {
        "workload": f"room with a 1M-triangle displaced mesh, lamps behind a 0.5 x 0.5 opening, {W}x{H}; "
                    f"reference {ref_spp} spp plain in two halves (relMSE less the reference's variance)",
        "plain": {"spp": spp_plain, "ms": round(budget, 1), "relmse": round(e_plain, 6)},
        "guided_runs": runs,
        "equal_time_relmse_ratio": med("equal_time_relmse_ratio"),
        "equal_spp_relmse_ratio": med("equal_spp_relmse_ratio"),
        "guided_over_plain_ms_per_spp": med("guided_over_plain_ms_per_spp"),
        "time_to_plain_relmse_ratio": med("time_to_plain_relmse_ratio"),
        "time_to_plain_relmse_measured": "guided renders until relMSE <= plain's; training time counted",
        "time_to_plain_relmse_ratio_extrapolated": med("time_to_plain_relmse_ratio_extrapolated"),
        "target_time_ratio": 0.5,
        "statistic": f"median over guide seeds {list(guide_seeds)}",
        "network_roofline": net,
    }
    log(f"guided: {json.dumps({k: v for k, v in out.items() if k != 'guided_runs'})}")
    return out


def configs_leg(pt, cpu_seconds, frames=3):
    """The other plain workloads, each against the reference renderer on the
    host cores: the sphere field (640x360, 16 spp, sky on) and the
    two-level instancing scene (three meshes of 50k, 200k and 500k
    triangles, 4096 instances; 1920x1080, 32 spp). GPU rate: rays over the
    device time of `frames` renders after one warmup, image kept in HBM."""
    out = {}
    for name, cfg, W, H, spp in (("spheres", 0, 640, 360, 16), ("instances", 2, 1920, 1080, 32)):
        t0 = time.perf_counter()
        ptsd = pt.generate(cfg, 1, 1.0)
        ptgp = pt.pack(ptsd)
        log(f"{name}: generated and packed in {time.perf_counter() - t0:.1f} s")
        image = pt.DeviceBuffer(12 * W * H)
        with pt.Scene(ptgp) as s:
            s.render(W, H, spp, seed=1, device_out=image)
            rays, ms = 0, 0.0
            for k in range(frames):
                _, _, st = s.render(W, H, spp, seed=2 + k, device_out=image)
                rays += st.rays
                ms += st.render_ms
        image.free()
        gpu = rays / (ms * 1e-3) / 1e6
        cpu = cpu_baseline(ptsd, W, H, spp, cpu_seconds)
        out[name] = {"workload": f"{W}x{H}, {spp} spp, 10 bounces, NEE", "mrays_per_s": round(gpu, 1),
                     "ms_per_frame": round(ms / frames, 2), "rays_per_frame": rays // frames, "cpu_baseline": cpu,
                     "gpu_over_cpu": round(gpu / cpu["value"], 1)}
        log(f"{name}: {json.dumps(out[name])}")
    return out


def denoise_leg(pt):
    """The denoiser (the reference's filter as the CLI's --denoise step runs
    it) on a 1080p render of the reference's room at 16 spp with aux: the
    device time against the reference's own filter on the host's threads,
    and the largest relative difference between the two images."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    import render_check as rc

    W, H, spp = 1920, 1080, 16
    with pt.Scene(rc.load_scene("room")[1]) as s:
        rgb, aux, _ = s.render(W, H, spp, seed=3, aux=True)
    pt.denoise(rgb, aux, spp)   # warmup
    runs = [pt.denoise(rgb, aux, spp) for _ in range(3)]
    ours, ms = runs[-1][0], min(r[1] for r in runs)
    out = {"workload": f"the reference's room, {W}x{H}, {spp} spp with aux", "device_ms": round(ms, 3)}
    if ob.ref_available():
        # The CPU baseline: the reference's filter itself, on the same frame.
        threads = max(1, min(16, os.cpu_count() or 1))
        ref, secs = ob.ref_denoise(rgb, aux, spp, threads)
        out["cpu_baseline"] = {"value": round(secs * 1e3, 1), "unit": "ms", "cores": threads, "kind": "reference",
                               "sample": f"the full {W}x{H} frame, rows split over {threads} threads as "
                                         "denoise.hpp's for_rows does"}
        out["max_rel_diff_vs_reference"] = float(np.max(np.abs(ours - ref) / np.maximum(np.abs(ref), 1e-3)))
    log(f"denoise: {json.dumps(out)}")
    return out


def bvh_leg(pt):
    """BVH builds on the GPU against the host builder's threads: the whole
    scene pack of the 1M-triangle frame's scene (pt_pack against pt_pack_ex
    with the device builder; the rest of the pack is the same host code), and
    the builder alone over 1M uniform boxes. Both checked word for word."""
    import numpy as np
    ptsd = pt.generate(1, 1, 1.0)
    t = time.perf_counter()
    host = pt.pack(ptsd)
    host_ms = (time.perf_counter() - t) * 1e3
    dev_ms, dev = [], None
    for _ in range(3):   # the first call also loads the kernels
        t = time.perf_counter()
        dev = pt.pack(ptsd, device_bvh=True)
        dev_ms.append((time.perf_counter() - t) * 1e3)
    rng = np.random.default_rng(1)
    c = rng.uniform(0, 1, (1_000_000, 3))
    boxes = np.concatenate([c - 0.002, c + 0.002], 1).astype(np.float32)
    hn, ho, hb_ms = pt.bvh_build(boxes)
    db = [pt.bvh_build(boxes, device=True) for _ in range(2)]
    out = {"pack_workload": "displaced icosphere scene, 1,003,520 triangles (pt_pack vs pt_pack_ex)",
           "pack_host_ms": round(host_ms, 1), "pack_device_ms": round(min(dev_ms[1:]), 1),
           "pack_identical": bool(np.array_equal(host, dev)),
           "build_workload": "1,000,000 uniform boxes (pt_bvh_build)", "build_host_ms": round(hb_ms, 1),
           "build_device_ms": round(min(d[2] for d in db), 1),
           "build_identical": bool(np.array_equal(db[-1][0], hn) and np.array_equal(db[-1][1], ho)),
           "host_threads": os.cpu_count()}
    log(f"bvh: {json.dumps(out)}")
    return out


def guided_kernels(pt):
    """Plain and guided 16-spp 1080p renders of the room, twice each, after
    four training batches: the kernels a profiler should see per render."""
    ptgp = pt.pack(pt.generate(4, 1, 1.0))
    W, H, out = 1920, 1080, {}
    with pt.Scene(ptgp) as s:
        g = pt.Guide(s, seed=2, learning_rate=3e-3, steps_per_batch=16)
        s.set_guide(g)
        for k in range(4):
            _, _, st = s.render(W, H, 1, seed=100 + k, train=True, samples_per_batch=1)
        for k in range(2):
            _, _, sp = s.render(W, H, 16, seed=1)
            _, _, sg = s.render(W, H, 16, seed=2, guide=True)
            log(f"room 1080p 16 spp: plain {sp.render_ms:.1f} ms, guided {sg.render_ms:.1f} ms")
        out = {"plain_ms": round(sp.render_ms, 2), "guided_ms": round(sg.render_ms, 2),
               "plain_rays": sp.rays, "guided_rays": sg.rays}
        g.close()
    return out


def probe(pt, args):
    """One frame of the default workload and a 1 GiB device copy: the
    program the counter passes profile."""
    ptgp = pt.pack(pt.generate(1, 1, args.scale))
    image = pt.DeviceBuffer(args.width * args.height * 12)
    with pt.Scene(ptgp) as scene:
        _, _, st = scene.render(args.width, args.height, args.spp, seed=1, device_out=image, stack_entries=args.stack)
    image.free()
    nbytes = args.copy_probe << 20
    if nbytes:
        a, b = pt.DeviceBuffer(nbytes), pt.DeviceBuffer(nbytes)
        pt._check(pt.library().pt_device_copy(pt.ctypes.c_void_p(b.ptr), pt.ctypes.c_void_p(a.ptr), nbytes, 2))
        a.free()
        b.free()
    return {"rays": st.rays, "copy_bytes": nbytes}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--scale", type=float, default=1.0, help="triangle-count scale (1.0 = 1,003,520)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=6.0, help="CPU work per baseline run (best of three)")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--stack", type=int, default=0, help="LDS stack entries per lane (0: the library's choice)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-workload leg")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--batch-no-tune", action="store_true", help="batch leg with the default launch shape")
    ap.add_argument("--batch-alone", action="store_true",
                    help="batch leg without the per-job-stream and one-call-per-job comparisons (for profiles)")
    ap.add_argument("--batch-eager", action="store_true", help="batch leg without captured graphs")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided equal-time leg")
    ap.add_argument("--guided-ref-spp", type=int, default=4096)
    ap.add_argument("--guided-plain-spp", type=int, default=1024,
                    help="the equal-time budget: the plain render at this many spp")
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoiser leg")
    ap.add_argument("--no-bvh", action="store_true", help="skip the BVH build leg")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip pt_scene_autotune on the headline frame (run after its timed frames)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (HBM traffic and VALU issue of the traversal stage)")
    ap.add_argument("--no-configs", action="store_true", help="skip the sphere-field and instancing legs")
    ap.add_argument("--only", choices=["batch", "guided", "guided-kernels", "configs", "denoise", "bvh", "probe"],
                    default=None,
                    help="run one leg alone and print its JSON (for profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also copy MIB MiB device to device (a kernel of known traffic, to scale HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the path shards by samples or image tiles, see the design notes", file=sys.stderr)
        sys.exit(2)

    import photon_tracer_amd as pt

    if args.only:
        leg = {"batch": lambda: batch_leg(pt, args.batch_jobs, max(1, args.steps), not args.batch_no_tune,
                                          not args.batch_eager, not args.batch_alone),
               "guided": lambda: guided_leg(pt, args.guided_ref_spp, args.guided_plain_spp),
               "guided-kernels": lambda: guided_kernels(pt),
               "configs": lambda: configs_leg(pt, args.cpu_seconds),
               "denoise": lambda: denoise_leg(pt),
               "bvh": lambda: bvh_leg(pt),
               "probe": lambda: probe(pt, args)}[args.only]
        print(json.dumps({args.only: leg()}), flush=True)
        return

    t0 = time.perf_counter()
    ptsd = pt.generate(1, 1, args.scale)
    t1 = time.perf_counter()
    ptgp = pt.pack(ptsd)
    t2 = time.perf_counter()
    n_tris = int(ptsd[13])
    log(f"scene: {n_tris} triangles, generated in {t1 - t0:.2f} s, BVH built and packed in {t2 - t1:.2f} s "
        f"(depth {pt.ptgp_depth(ptgp)}), {ptgp.size * 4 / 1e6:.1f} MB")
    W, H, spp = args.width, args.height, args.spp
    scene = pt.Scene(ptgp)
    arch, cus = pt.device_info()
    log(f"device: {arch}, {cus} CUs")

    # The image stays in HBM during the timed frames (the scene is resident
    # from the start); the host copy is timed separately below.
    image = pt.DeviceBuffer(W * H * 12)
    for k in range(args.warmup):
        _, _, st = scene.render(W, H, spp, seed=1000 + k, device_out=image, stack_entries=args.stack)
        log(f"warmup {k}: {st.render_ms:.1f} ms, {st.mrays_per_s:.0f} Mrays/s")

    rays = 0
    gpu_ms = 0.0
    t_start = time.perf_counter()
    for k in range(args.steps):
        _, _, st = scene.render(W, H, spp, seed=1 + k, device_out=image, stack_entries=args.stack)
        rays += st.rays
        gpu_ms += st.render_ms
        log(f"step {k}: {st.rays} rays ({st.primary_rays} primary, {st.shadow_rays} shadow), "
            f"{st.render_ms:.1f} ms on the device")
    wall = time.perf_counter() - t_start
    value = rays / wall / 1e6
    samples = W * H * spp * args.steps
    log(f"timed: {args.steps} frames in {wall:.3f} s wall ({gpu_ms:.1f} ms device): {value:.1f} Mrays/s, "
        f"{samples / wall / 1e6:.1f} Msamples/s")
    t_h = time.perf_counter()
    _, _, hst = scene.render(W, H, spp, seed=1, stack_entries=args.stack)
    host_wall = time.perf_counter() - t_h
    log(f"with the image copied to host memory: {hst.rays / host_wall / 1e6:.1f} Mrays/s ({host_wall * 1e3:.1f} ms)")

    roofline = None
    if not args.no_roofline:
        # The same rays as step 0 (seed 1): one render counting the traversal
        # stage's visits, one timing each of its launches with HIP events on
        # the scene's stream.
        _, _, cst = scene.render(W, H, spp, seed=1, count=True, stack_entries=args.stack)
        _, _, tst = scene.render(W, H, spp, seed=1, timing=True, stack_entries=args.stack)
        queries = cst.visits[6]  # closest-hit and shadow queries traced by the traversal-stage launches
        launches = 10 * tst.batches
        launch_s = tst.extend_ms * 1e-3 / launches
        served = pt.QUERY_BYTES * queries + sum(v * b for v, b in zip(cst.visits, pt.VISIT_BYTES))
        scene_bytes = traversal_scene_bytes(ptgp)
        algo = pt.QUERY_BYTES * queries / launches + scene_bytes   # per launch
        achieved = algo / launch_s / 1e9
        roofline = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "traversal stage: k_camera_packet (bounce 0's camera rays, 64-ray packets) and k_extend "
                      "(closest-hit rays + the previous bounce's shadow rays)",
            "launches": launches, "avg_launch_ms": round(launch_s * 1e3, 4),
            "algorithmic_bytes_per_launch": int(algo),
            "algorithmic_model": f"{pt.QUERY_BYTES} B per query (ray in, hit out) + the scene's nodes, leaf records "
                                 f"and instances once ({scene_bytes} B)",
            "bytes_served_per_launch": int(served / launches),
            "bytes_served_gbs": round(served / launches / launch_s / 1e9, 1),
            "per_query": {"queries": queries, "nodes": cst.visits[0] / queries, "leafrefs": cst.visits[1] / queries,
                          "tris": cst.visits[2] / queries, "spheres": cst.visits[3] / queries},
        }
        log(f"roofline: {json.dumps(roofline)}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ptsd, W, H, spp, args.cpu_seconds)
        log(f"cpu baseline: {json.dumps(cpu)}")
    if roofline is not None and not args.no_counters:
        # The counter passes run as child processes after the timed frames,
        # while this process holds its scene but runs nothing on the GPU.
        counts = counter_passes(args)
        tc = traversal_counters(counts, cus) if counts else None
        if tc:
            roofline["traffic"] = tc["traffic"]
            roofline["traffic_source"] = ("rocprofv3 FETCH_SIZE and WRITE_SIZE passes in this run (scaled by "
                                          f"{tc['fetch_scale']} and {tc['write_scale']} from a 1 GiB copy)")
            roofline["hbm_measured_gbs"] = round(tc["traffic"] / (roofline["avg_launch_ms"] * 1e-3) / 1e9, 1)
            roofline["hbm_measured_frac"] = round(roofline["hbm_measured_gbs"] / HBM_PEAK_GBS, 4)
            if "valu_busy" in tc:
                roofline["issue"] = {"valu_busy": tc["valu_busy"], "lane_use": tc["lane_use"],
                                     "valu_instructions_per_launch": tc["valu_instructions_per_launch"],
                                     "source": "rocprofv3 SQ counter pass in this run"}
        log(f"counters: {json.dumps(tc)}")
    if args.copy_probe:
        nbytes = args.copy_probe << 20
        a, b = pt.DeviceBuffer(nbytes), pt.DeviceBuffer(nbytes)
        pt._check(pt.library().pt_device_copy(pt.ctypes.c_void_p(b.ptr), pt.ctypes.c_void_p(a.ptr), nbytes, 2))
        log(f"copy probe: {nbytes} bytes device to device")
        a.free()
        b.free()
    autotune = None
    if not args.no_tune:
        # pt_scene_autotune on the same frame, last, since the scene keeps the
        # shape it picks: the candidates' rates, and the tuned shape's frame
        # time against the default shape's timed frames above.
        t_t = time.perf_counter()
        r = scene.autotune(W, H, spp, seed=1)
        _, _, tst2 = scene.render(W, H, spp, seed=1, device_out=image)
        autotune = {"stack_entries": r.stack_entries, "blocks_per_cu": r.blocks_per_cu, "max_paths": r.max_paths,
                    "tuned_frame_ms": round(tst2.render_ms, 2), "default_frame_ms": round(gpu_ms / args.steps, 2),
                    "seconds": round(time.perf_counter() - t_t, 1),
                    "tried": [{"stack_entries": t[0], "blocks_per_cu": t[1], "max_paths": t[3],
                               "mrays_per_s": round(t[4], 1)} for t in r.tried]}
        log(f"autotune: {json.dumps(autotune)}")
    image.free()
    scene.close()
    batch = None if args.no_batch else batch_leg(pt, args.batch_jobs, max(1, args.steps), not args.batch_no_tune,
                                                  not args.batch_eager)
    guided = None if args.no_guided else guided_leg(pt, args.guided_ref_spp, args.guided_plain_spp)
    denoised = None if args.no_denoise else denoise_leg(pt)
    bvh = None if args.no_bvh else bvh_leg(pt)
    configs = None if args.no_configs else configs_leg(pt, args.cpu_seconds)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"displaced icosphere, {n_tris} triangles, {W}x{H}, {spp} spp, 10 bounces, NEE, "
                               "RR from bounce 3", "width": W, "height": H, "spp": spp, "triangles": n_tris,
                   "parallelism": "single GPU"},
        "samples_per_s": round(samples / wall, 1),
        "host_output_mrays_per_s": round(hst.rays / host_wall / 1e6, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": autotune,
        "batch": batch,
        "guided": guided,
        "denoise": denoised,
        "bvh_build": bvh,
        "configs": configs,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
