#!/usr/bin/env python3
"""Throughput of one rank's render call on one MI355X.

Workload (the metric's configuration): a 256^3 local volume with three sources
-- float32 density (seeded Gaussian mixture + 3-octave value noise), float32x3
curl-noise vector field shown through the `length` functor, 2,000,000 uniform
particles on a 64^3 cell grid -- iso-surface at threshold 0.5, one clipping
plane, SSAO, trilinear sampling at step 0.5, 1920x1080.  A step is one
doVisualization(redraw): particle cell lists + classification + the draw
callback, with the camera alternating between the default pose and a
volume-filling pose.  Inputs are resident in HBM before timing starts and the
framebuffer stays on the device (no PCIe in the timed region); fps_host_output
repeats the frames with the framebuffer copied to pinned host memory each step.
The timed frames record HIP events around the march only (its time for the
roofline); the per-kernel breakdown comes from the same frames run just before
with events between all kernels (ms_per_step_all_timing_events).

Metric: trilinear samples per second (volume + iso march samples actually
executed, after early ray termination), with frames/s alongside.  The same run
also times the brick render forms (FP32 / FP16 / BF16 field bricks,
post-classification) on the same workload and reports them under "forms".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD-published
# measured DRAM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a
# plain copy of known size); written from the profile summaries in profiles/
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")

FORMS = {"combined": (0, 0), "bricks-f32": (1, 0), "bricks-f16": (1, 1), "bricks-bf16": (1, 2)}
VALUE_BYTES = {0: 4, 1: 2, 2: 2}  # bytes per stored brick value


def kernel_bytes(counts, spec, n_vox, n_particles, n_visible, form="combined"):
    """Algorithmic bytes per launch of each kernel (read + write), from the
    executed work counts of one frame."""
    import numpy as np

    W, H = spec.framebuffer
    attr_dim = spec.particles[0].feature_dim if spec.particles else 1
    n_cells = int(np.prod(spec.particle_grid)) if spec.particles else 0
    npx = W * H
    src_bytes = sum(4 * v.feature_dim for v in spec.volumes)
    comps = sum(v.feature_dim for v in spec.volumes)
    out = {
        # 49 depth taps + ao write
        "ssao": npx * (49 * 4 + 4),
        # G-buffer read colour 4 + normal 12 + depth 4 + ao 4 + volume colour 16, colour write 4
        "shade": 44 * npx,
        # the visible-records build: every particle's attribute in; per visible particle its
        # position 12 in, cell id, rank, slot index and permutation entry 4 each written and
        # read back, its slot record 16 written and read back and its render record 16 written;
        # per cell the count (fill 4, scan read 2 x 4), count and offset 8 written, offsets 8
        # read and the record range 8 written
        "binning": n_particles * 4 * attr_dim + n_visible * (12 + 4 * 2 * 4 + 2 * 16 + 16) + n_cells * (4 + 8 + 8 + 8 + 8),
    }
    particles = 8 * counts["cells_visited"] + 16 * counts["sphere_tests"]
    if form == "combined":
        # 16 B of sources in, two RGBA8 texels out per voxel
        out["classify"] = n_vox * (src_bytes + 8)
        # fused march: 8 RGBA8 taps per volume sample and per iso sample, 7 iso samples at a hit;
        # G-buffer writes colour 4 + normal 12 + depth 4 + volume colour 16
        out["march"] = (32 * (counts["vol_samples"] + counts["iso_samples"]) + 7 * 32 * counts["iso_hits"]
                        + particles + 36 * npx)
    else:
        vb = VALUE_BYTES[FORMS[form][1]]
        # brick build: sources in; 648 pairs of 2 values per 512 voxels and component out, + scales
        out["classify"] = n_vox * src_bytes + (n_vox // 512) * comps * (648 * 2 * vb + 4)
        # one classification per step: 8 taps x stored bytes per component (+ the brick scale
        # for 16-bit forms); 7 more at an iso hit
        per = comps * (8 * vb + (4 if vb < 4 else 0))
        out["march"] = per * counts["march_steps"] + 7 * per * counts["iso_hits"] + particles + 36 * npx
    return out


def host_cpus():
    """What the host offers this process: `nproc` (the CPUs it may use; GNU nproc also
    honours OMP_NUM_THREADS), `nproc --all`, and lscpu's model name."""
    import subprocess

    def run(cmd):
        try:
            return subprocess.run(cmd, capture_output=True, text=True, timeout=20).stdout
        except (OSError, subprocess.SubprocessError):
            return ""

    try:
        nproc = int(run(["nproc"]).strip())
    except ValueError:
        nproc = len(os.sched_getaffinity(0))
    try:
        nall = int(run(["nproc", "--all"]).strip())
    except ValueError:
        nall = os.cpu_count() or nproc
    model = ""
    for ln in run(["lscpu"]).splitlines():
        if ln.lower().startswith("model name"):
            model = ln.split(":", 1)[1].strip()
            break
    return {"nproc": nproc, "nproc_all": nall, "affinity": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline(spec, poses, threads, frames=3):
    """The CPU restatement of the reference path (classification + cell lists + draw
    of the full frame) at each pose, like the GPU line: with `threads` OpenMP threads
    (the reference's AccCpuOmp2Blocks backend: one pixel or voxel per block), the
    median of `frames` frames per pose; and with one thread (AccCpuSerial), one frame
    per pose.  Frames/s and samples/s average the poses as the GPU line does."""
    import numpy as np

    from oracle import oracle

    def run(th, reps):
        per = []
        for p in poses:
            s = spec.with_camera(p)
            ps = oracle.PreparedScene(s, threads=th)
            secs, samples = [], 0
            for _ in range(reps):
                res = oracle.render(s, 32, threads=th, prepared=ps)
                secs.append(res.seconds["classify"] + res.seconds["bin"] + res.seconds["render"])
                samples = res.counters["vol_samples"] + res.counters["iso_samples"]
            per.append((float(np.median(secs)), samples, res.seconds))
        t = sum(x[0] for x in per)
        smp = sum(x[1] for x in per)
        return {
            "value": smp / t,
            "fps": len(poses) / t,
            "frame_s": {name: x[0] for name, x in zip(("default", "fill"), per)},
            "phase_s": {name: {k: float(v) for k, v in x[2].items()} for name, x in zip(("default", "fill"), per)},
            "samples": {name: x[1] for name, x in zip(("default", "fill"), per)},
        }

    host = host_cpus()
    allc = run(threads, frames)
    serial = run(1, 1)
    return {
        "value": allc["value"],
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full 1920x1080 frames (classification + cell lists + draw) at the default and the fill pose; "
        f"{threads} threads: median of {frames} frames per pose; 1 thread: one frame per pose; "
        "FP32 C++ restatement of the reference CPU path (OpenMP, one pixel or voxel per iteration)",
        "fps": allc["fps"],
        "frame_s": allc["frame_s"],
        "phase_s": allc["phase_s"],
        "samples": allc["samples"],
        "serial": {"value": serial["value"], "unit": "samples/s", "cores": 1, "fps": serial["fps"],
                   "frame_s": serial["frame_s"]},
        "host": host,
    }


def set_pose_of(vis, p):
    import numpy as np

    vis.look_at = np.asarray(p.look_at, np.float64)
    vis.rotation = np.asarray(p.rotation, np.float64)
    vis.distance = float(p.distance)
    vis._update_modelview()


def measure(vis, poses, out, steps, warmup, instrumented=None):
    """Counting frames per pose (untimed), warmup, `steps` frames with timing events
    between all kernels (the per-kernel breakdown), then `steps` timed frames with timing
    events around the march only (each event costs a few us on the stream; the march's
    HIP-event time in avg_us comes from the timed frames).  `instrumented`, a dict,
    receives the frame time of the fully instrumented frames."""
    import numpy as np
    import torch

    def set_pose(p):
        set_pose_of(vis, p)

    counts = []
    vis.set_counting(True)
    for p in poses:
        set_pose(p)
        vis.update_buffers()
        vis.draw(out=out)
        counts.append(vis.counters())
    vis.set_counting(False)
    for k in range(warmup):
        set_pose(poses[k % 2])
        vis.update_buffers()
        vis.draw(out=out)
    vis.synchronize()
    torch.cuda.synchronize()
    per = {"classify": [], "binning": [], "march": [], "ssao": [], "shade": []}
    t0 = time.perf_counter()
    for k in range(steps):
        set_pose(poses[k % 2])
        vis.update_buffers()
        vis.draw(out=out)
        t = vis.timings()  # waits for the step: each frame completes before the next (as the reference)
        per["classify"].append(t["optimization_buffer_us"])
        per["binning"].append(t["binning_us"])
        per["march"].append(t["march_us"])
        per["ssao"].append(t["ssao_us"])
        per["shade"].append(t["shade_us"])
    vis.synchronize()
    if instrumented is not None:
        instrumented["ms_per_step"] = (time.perf_counter() - t0) / steps * 1e3
        instrumented["march_us"] = float(np.mean(per["march"]))
    vis.set_tuning(vis.TUNE_TIMING_EVENTS, 2)
    march = []
    total_samples = 0
    vis.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        pi = k % 2
        set_pose(poses[pi])
        vis.update_buffers()
        vis.draw(out=out)
        march.append(vis.timings()["march_us"])  # waits for the step
        total_samples += counts[pi]["vol_samples"] + counts[pi]["iso_samples"]
    vis.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    vis.set_tuning(vis.TUNE_TIMING_EVENTS, 1)
    avg_us = {k: float(np.mean(v)) for k, v in per.items()}
    avg_us["march"] = float(np.mean(march))
    avg_counts = {k: float(np.mean([c[k] for c in counts])) for k in counts[0]}
    return wall, total_samples, avg_us, avg_counts, counts


def tex_layout_study(vis, poses, frames=6):
    """March time and classification time (which holds the brick copy, or the binary16
    copy's stores) of each texture layout at the tuned schedule, per pose; medians over
    `frames` updated frames."""
    import numpy as np

    keep = vis.tex_layout
    res = {}
    for layout in (0, 8, 16):
        vis.set_tuning(vis.TUNE_TEX_LAYOUT, layout)
        vis.update_buffers()
        per = {}
        for name, p in zip(("default", "fill"), poses):
            set_pose_of(vis, p)
            vis.update_buffers()
            vis.draw(out=False)
            m, c = [], []
            for _ in range(frames):
                vis.update_buffers()
                vis.draw(out=False)
                t = vis.timings()
                m.append(t["march_us"])
                c.append(t["optimization_buffer_us"])
            per[name] = {"march_us": float(np.median(m)), "classify_us": float(np.median(c))}
        res[f"layout={layout}"] = per
    vis.set_tuning(vis.TUNE_TEX_LAYOUT, keep)
    vis.update_buffers()
    return res


def learned_ao(vis, poses, out, args):
    """Train the occlusion network on the GPU from plain-SSAO triples of 16 scenes of
    this workload's style and size with other seeds (fields, particles, clipping
    planes) at random poses at the workload's frame size, crops x steps triples
    generated on the fly; gate it on this scene's poses against the plain SSAO pass;
    time the frame with each occlusion pass in the combined and FP16-brick forms."""
    import numpy as np
    import torch

    from isaac_amd import aonet, scenes

    vis.set_render_form(*FORMS["combined"])
    vis.update_buffers()
    vis.set_ao_mode(vis.AO_PLAIN)
    W, H = vis.framebuffer_size
    t0 = time.perf_counter()
    n = int(vis.global_size[0])
    n_particles = int(vis.particle_sources[0].pos.shape[0]) if vis.particle_sources else 0
    ts = aonet.TrainingSet(0, fb=(W, H), n=n, particles=max(1, n_particles))
    ts.vis.aonet_set_params(aonet.init_params(0))
    losses = aonet.train(ts.vis, steps=args.ao_train_steps, seed=1, crops=16, prepare=ts.prepare)
    vis.aonet_set_params(ts.vis.aonet_get_params())
    ts.close()
    del ts
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    held_out = [scenes.DEFAULT_POSE, scenes.FILL_POSE, scenes.random_pose(np.random.default_rng(99))]
    gate = aonet.gate(vis, held_out, floor_db=40.0, min_gain_db=3.0)
    timing = {}
    for name in ("combined", "bricks-f32", "bricks-f16"):
        vis.set_render_form(*FORMS[name])
        for mode, label in ((vis.AO_PLAIN, "plain"), (vis.AO_LEARNED, "learned")):
            # the learned pass is timed whether or not the gate kept it (the gate result is reported)
            if mode == vis.AO_LEARNED and name == "bricks-f32":
                continue
            vis.set_ao_mode(mode)
            w2, s2, us2, c2, _ = measure(vis, poses, out, args.form_steps, 2)
            timing[f"{name}/{label}"] = {
                "ms_per_step": w2 / args.form_steps * 1e3,
                "occlusion_ms": us2["ssao"] / 1e3,
                "march_ms": us2["march"] / 1e3,
            }
    vis.set_ao_mode(vis.AO_PLAIN)
    vis.set_render_form(*FORMS["combined"])
    res = {
        "train_steps": args.ao_train_steps,
        "train_s": train_s,
        "train_mse_first20": float(np.mean(losses[:20])),
        "train_mse_last20": float(np.mean(losses[-20:])),
        "training_triples": args.ao_train_steps * 16,
        "training_set": f"16 seeded {n}^3 scenes of the workload's style and other seeds (Gaussian mixture + "
                        f"value noise, iso 0.5, {n_particles} particles, random clipping plane) at random poses, "
                        f"{W}x{H}, 16 crops of 128^2 per step",
        "gate_floor_db": 40.0,
        "gate_min_gain_db": 3.0,
        "gate_passed": bool(gate["passed"]),
        "image_psnr_db": gate["image_psnr"],
        "occlusion_map_psnr_db": gate["ao_psnr"],
        "image_psnr_without_occlusion_db": gate["image_psnr_no_ao"],
        "gain_over_no_occlusion_db": gate["gain_db"],
        "frames": timing,
    }
    if "bricks-f16/learned" in timing:
        # the plain FP32 form: the combined (pre-classified) form, and the FP32 brick form
        res["speedup_learned_f16_vs_plain_f32"] = timing["combined/plain"]["ms_per_step"] / timing["bricks-f16/learned"]["ms_per_step"]
        res["speedup_learned_f16_vs_plain_f32_bricks"] = (timing["bricks-f32/plain"]["ms_per_step"]
                                                          / timing["bricks-f16/learned"]["ms_per_step"])
    if "combined/learned" in timing:
        # the network's dense roofline: FP16 MFMA, 2.5 PFLOP/s (MI355X dense peak)
        W, H = vis.framebuffer_size
        gf = aonet.flops_per_frame(W, H) / 1e9
        tf = gf / timing["combined/learned"]["occlusion_ms"]
        res["network"] = {"gflop_per_frame": gf, "tflops": tf, "bound": "mfma", "peak_tflops": 2500.0,
                          "frac": tf / 2500.0}
    return res


def precision_workload(steps: int = 5):
    """The precision form's own workload: the 512^3 Gaussian-mixture density at
    3840x2160, trilinear, step 0.25, in the combined FP32 form and the FP32 / FP16 /
    BF16 brick forms: frame and march times, and the march's algorithmic bytes per
    launch against the HBM peak."""
    import numpy as np
    import torch

    from isaac_amd import scenes
    from isaac_amd import visualization as V

    spec = scenes.precision_scene()
    W, H = spec.framebuffer
    n_vox = int(np.prod(spec.local_size))
    vis = V.from_scene(spec, 0)
    out = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda")
    res = {"workload": f"{spec.local_size[0]}^3 f32 Gaussian mixture, {W}x{H}, trilinear, step {spec.step}",
           "forms": {}}
    for name in ("combined", "bricks-f32", "bricks-f16", "bricks-bf16"):
        f, d = FORMS[name]
        vis.set_render_form(f, d)
        wall, samples, us, counts, _ = measure(vis, [spec.camera, spec.camera], out, steps, 1)
        kb = kernel_bytes(counts, spec, n_vox, 0, 0, name)
        ach = kb["march"] / (us["march"] * 1e-6) / 1e9
        res["forms"][name] = {
            "ms_per_step": wall / steps * 1e3,
            "samples_per_s": samples / wall,
            "kernel_ms": {k: v / 1e3 for k, v in us.items()},
            "march_steps_per_frame": counts["march_steps"],
            "roofline": {"kernel": "march", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": kb["march"]},
        }
    vis.close()
    del out
    torch.cuda.synchronize()
    return res


def compositing(n_images: int = 8, width: int = 1920, height: int = 1080, reps: int = 20):
    """The per-GPU compositing step (isaac_hip_composite): n_images RGBA8 frames of
    the sub-volumes one GPU renders, blended in composite order into one; timed with
    HIP events on the stream it runs on.  Algorithmic bytes: every input pixel read
    once, the output written once."""
    import torch

    from isaac_amd import composite as K

    px = width * height
    g = torch.Generator(device="cuda").manual_seed(5)
    imgs = [torch.randint(0, 256, (px, 4), dtype=torch.uint8, device="cuda", generator=g) for _ in range(n_images)]
    order = list(range(n_images))
    out = torch.empty_like(imgs[0])
    for _ in range(3):
        K.composite(imgs, order, out)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        K.composite(imgs, order, out)
    stop.record()
    torch.cuda.synchronize()
    us = start.elapsed_time(stop) * 1e3 / reps
    nbytes = (n_images + 1) * px * 4
    return {"images": n_images, "framebuffer": [width, height], "us": us, "algorithmic_bytes": nbytes,
            "GBs": nbytes / (us * 1e-6) / 1e9, "hbm_frac": nbytes / (us * 1e-6) / 8.0e12}


def jpeg_frames(vis_frame, quality: int = 90, reps: int = 10):
    """The default video-frame path (isaac_hip_jpeg_encode, libjpeg-turbo's baseline
    JPEG reproduced on the GPU) on one rendered 1080p frame: wall time per frame
    including the copy of the file to the host, and the file size."""
    import torch

    from isaac_amd import frame as K

    dev = torch.from_numpy(vis_frame).cuda()
    jpg = K.jpeg(dev, quality)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        jpg = K.jpeg(dev, quality)
    ms = (time.perf_counter() - t0) / reps * 1e3
    return {"framebuffer": [int(vis_frame.shape[1]), int(vis_frame.shape[0])], "quality": quality, "ms": ms,
            "bytes": len(jpg)}


def batched_workload(args):
    """48 sub-volumes of side U{32..256} (FP16 bricks: Gaussian-mixture density with an
    iso-surface + float3 noise via length) x 8 seeded views at 960x540 with learned
    occlusion: the whole set as one persistent launch sequence replayed from a hipGraph,
    against the same frames rendered one launch sequence per frame."""
    import numpy as np
    import torch

    from isaac_amd import aonet
    from isaac_amd import batch as B

    t0 = time.perf_counter()
    vises, sides = B.make_subvolumes(0, count=args.batch_count, seed=5, fb=(960, 540), dtype=1)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    projs, mvs = B.views_for(vises, args.batch_views, seed=5)
    nframes = len(vises) * args.batch_views
    # occlusion network: trained at 960x540 on the same kind of triples as the 1080p one
    t1 = time.perf_counter()
    ts = aonet.TrainingSet(0, fb=(960, 540), n=128, particles=250_000)
    ts.vis.aonet_set_params(aonet.init_params(0))
    losses = aonet.train(ts.vis, steps=args.ao_train_steps, seed=2, crops=16, prepare=ts.prepare)
    params = ts.vis.aonet_get_params()
    ts.close()
    del ts
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    vises[0].aonet_set_params(params)
    batch = B.Batch(vises, 960, 540)
    bg = vises[0].background_color
    res = {"sub_volumes": len(vises), "views_each": args.batch_views, "frames": nframes, "framebuffer": [960, 540],
           "voxels": int(sum(s ** 3 for s in sides)), "rays": nframes * 960 * 540, "dtype": "f16 bricks",
           "setup_s": t_build, "train_s": t_train, "train_mse_last20": float(np.mean(losses[-20:]))}
    # plain occlusion images, then learned: PSNR of the whole set
    vises[0].set_ao_mode(vises[0].AO_PLAIN)
    batch.set_views(projs, mvs, bg=bg)
    batch.render(B.GRAPH)
    plain = batch.images()
    vises[0].set_ao_mode(vises[0].AO_LEARNED)
    batch.set_views(projs, mvs, bg=bg)
    res["tuning"] = B.autotune(batch)
    batch.render(B.GRAPH)
    learned = batch.images()
    res["learned_vs_plain_psnr_db"] = aonet.psnr(learned, plain, 255.0)
    del plain, learned
    times = {}
    for mode, name in ((B.GRAPH, "graph"), (B.PERSISTENT, "persistent"), (B.PER_FRAME, "per_frame")):
        batch.render(mode)
        batch.synchronize()
        ts = []
        for _ in range(args.batch_repeats):
            batch.render(mode)
            ts.append(batch.last_time_us())
        times[name] = float(np.median(ts)) / 1e3
    res["ms"] = times
    res["frames_per_s"] = nframes / (times["graph"] * 1e-3)
    res["rays_per_s"] = res["rays"] / (times["graph"] * 1e-3)
    res["speedup_vs_per_frame_launches"] = times["per_frame"] / times["graph"]
    # the march's work over all frames (each frame once through the counting per-context
    # render), and the batched sequence without the occlusion pass (march + shading)
    tot = {"march_steps": 0, "iso_hits": 0, "cells_visited": 0, "sphere_tests": 0}
    for f in range(nframes):
        v = vises[f // args.batch_views]
        v.set_counting(True)
        v.draw(out=False, projection=projs[f], modelview_=mvs[f])
        c = v.counters()
        v.set_counting(False)
        for k in tot:
            tot[k] += int(c[k])
    for v in vises:
        v.ambient_occlusion = {"isEnabled": False, "weight": 0.5}
        v.update_buffers()
    batch.set_views(projs, mvs, bg=bg)
    batch.render(B.GRAPH)
    batch.synchronize()
    ts = []
    for _ in range(args.batch_repeats):
        batch.render(B.GRAPH)
        ts.append(batch.last_time_us())
    t_noao = float(np.median(ts)) * 1e-6
    npx = nframes * 960 * 540
    per = 4 * (8 * 2 + 4)  # 4 components (density + float3 field), 8 FP16 taps + the brick scale each
    march_b = per * tot["march_steps"] + 7 * per * tot["iso_hits"] + 8 * tot["cells_visited"] + 16 * tot["sphere_tests"] \
        + 36 * npx
    seq_b = march_b + 44 * npx
    ach = seq_b / t_noao / 1e9
    res["march_steps"] = tot["march_steps"]
    res["roofline"] = {"kernel": "batched march + shading (occlusion off), one graph replay", "bound": "hbm",
                       "ms": t_noao * 1e3, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": seq_b}
    batch.close()
    for v in vises:
        v.close()
    torch.cuda.synchronize()
    return res


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256, help="local volume side (the workload is 256)")
    ap.add_argument("--form", default="combined", choices=sorted(FORMS), help="render form of the headline line")
    ap.add_argument("--no-forms", action="store_true", help="skip timing the other render forms")
    ap.add_argument("--form-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--prefetch", type=int, default=0, help="march kernel: fetch step i+1 before compositing step i")
    ap.add_argument("--march-waves", type=int, default=0, help="march occupancy target (0, 6, 8)")
    ap.add_argument("--march-threads", type=int, default=256, help="march workgroup: 256 (4 waves) or 64 (one wave)")
    ap.add_argument("--tex-layout", type=int, default=0,
                    help="texture the march reads: 0 linear, 8 bricks of 8^3, 16 binary16 copy of the volume RGBA")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep --prefetch / --march-waves / --march-threads instead of autotuning")
    ap.add_argument("--no-learned-ao", action="store_true", help="skip training / timing the learned occlusion pass")
    ap.add_argument("--ao-train-steps", type=int, default=1250, help="16 crops each: 20,000 triples")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sub-volumes x views workload")
    ap.add_argument("--batch-count", type=int, default=48)
    ap.add_argument("--batch-views", type=int, default=8)
    ap.add_argument("--batch-repeats", type=int, default=5)
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 512^3, 3840x2160, step 0.25 precision-form workload")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this size first (scales profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py --gpus {args.gpus}: the library renders one rank's sub-volume on one GPU; "
            "the exchange of multi-GPU sort-last compositing is not implemented, so only --gpus 1 is measured.",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    from isaac_amd import scenes
    from isaac_amd import visualization as V

    if args.copy_calibration_mib:
        nbytes = args.copy_calibration_mib << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    n = args.size
    spec = scenes.prod_scene(n=n)
    poses = [scenes.DEFAULT_POSE, scenes.FILL_POSE]
    vis = V.from_scene(spec, 0)
    vis.set_tuning(vis.TUNE_MARCH_PREFETCH, args.prefetch)
    out = torch.empty((spec.framebuffer[1], spec.framebuffer[0], 4), dtype=torch.uint8, device="cuda")
    n_vox = n ** 3
    n_particles = int(spec.particles[0].pos.shape[0])
    n_visible = int(np.count_nonzero(spec.particles[0].attr[:, 0] >= 0.995))

    vis.set_tuning(vis.TUNE_MARCH_WAVES, args.march_waves)
    vis.set_tuning(vis.TUNE_MARCH_THREADS, args.march_threads)
    vis.set_tuning(vis.TUNE_TEX_LAYOUT, args.tex_layout)
    form, dt = FORMS[args.form]
    vis.set_render_form(form, dt)
    tuning = None
    if not args.no_tune and args.form == "combined":
        from isaac_amd import autotune

        vis.update_buffers()
        tuning = autotune.tune_march(vis, poses)
        tuning["tex_layout_study"] = tex_layout_study(vis, poses)
    instrumented = {}
    wall, total_samples, avg_us, avg_counts, counts = measure(vis, poses, out, args.steps, args.warmup, instrumented)
    ms_per_step = wall / args.steps * 1e3
    value = total_samples / wall
    # the same frames with the framebuffer copied to (pinned) host memory each step, as
    # the reference's drawCallBack copies it into IceT's image (lib/isaac.hpp:2530-2539)
    host_out = torch.empty((spec.framebuffer[1], spec.framebuffer[0], 4), dtype=torch.uint8, pin_memory=True)
    wall_h, _, _, _, _ = measure(vis, poses, host_out, args.steps, 2)
    kb = kernel_bytes(avg_counts, spec, n_vox, n_particles, n_visible, args.form)
    dom = max(avg_us, key=lambda k: avg_us[k])
    achieved = kb[dom] / (avg_us[dom] * 1e-6) / 1e9
    traffic = None
    traffic_src = None
    if os.path.exists(TRAFFIC_FILE) and args.size == 256:
        with open(TRAFFIC_FILE) as f:
            tf = json.load(f)
        entry = tf.get("kernels", {}).get(f"{args.form}:{dom}")
        if entry:
            traffic = entry["bytes_per_launch"]
            traffic_src = tf.get("source")
    line = {
        "metric": "samples/s + fps @256^3x3,1080p",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.form in ("combined", "bricks-f32") else args.form.split("-")[1],
        "data": "synthetic (seeded Gaussian mixture + value-noise density, curl-noise vector field, uniform particles)",
        "config": {
            "workload": f"{n}^3 local volume x 3 sources (f32 density, f32x3 curl field via length, "
            f"{n_particles} particles on a {spec.particle_grid[0]}^3 cell grid), iso 0.5, 1 clipping plane, "
            "SSAO, trilinear step 0.5, 1920x1080, poses default/fill alternating",
            "frame": "cell lists + classification + draw per step",
            "form": args.form,
        },
        "fps": args.steps / wall,
        "fps_host_output": args.steps / wall_h,
        "ms_per_step_host_output": wall_h / args.steps * 1e3,
        # the same frames with timing events between all kernels (per-kernel breakdown below)
        "ms_per_step_all_timing_events": instrumented["ms_per_step"],
        "samples_per_frame": {"default": counts[0]["vol_samples"] + counts[0]["iso_samples"],
                              "fill": counts[1]["vol_samples"] + counts[1]["iso_samples"]},
        "kernel_ms": {k: v / 1e3 for k, v in avg_us.items()},
        "kernel_algorithmic_GBs": {k: kb[k] / (avg_us[k] * 1e-6) / 1e9 if avg_us[k] > 0 else None for k in kb},
        "tuning": tuning,
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": kb[dom],
        },
    }
    if not args.no_forms:
        forms = {}
        for name in ("combined", "bricks-f32", "bricks-f16", "bricks-bf16"):
            if name == args.form:
                continue
            f, d = FORMS[name]
            vis.set_render_form(f, d)
            w2, s2, us2, c2, _ = measure(vis, poses, out, args.form_steps, 2)
            kb2 = kernel_bytes(c2, spec, n_vox, n_particles, n_visible, name)
            forms[name] = {
                "ms_per_step": w2 / args.form_steps * 1e3,
                "fps": args.form_steps / w2,
                "samples_per_s": s2 / w2,
                "kernel_ms": {k: v / 1e3 for k, v in us2.items()},
                "march_algorithmic_GBs": kb2["march"] / (us2["march"] * 1e-6) / 1e9,
                "march_steps_per_frame": c2["march_steps"],
            }
        line["forms"] = forms
    if not args.no_learned_ao:
        line["learned_ao"] = learned_ao(vis, poses, out, args)
    vis.set_render_form(*FORMS["combined"])
    vis.update_buffers()
    line["jpeg_frame"] = jpeg_frames(vis.draw())
    vis.close()
    if not args.no_batch:
        line["batched"] = batched_workload(args)
    if not args.no_precision:
        line["precision"] = precision_workload()
    line["compositing"] = compositing()
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cpus()["nproc"]
        cb = cpu_baseline(spec, poses, threads)
        line["cpu_baseline"] = cb
        line["gpu_vs_cpu"] = {"samples_per_s": value / cb["value"], "fps": line["fps"] / cb["fps"],
                              "fps_vs_serial": line["fps"] / cb["serial"]["fps"]}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
