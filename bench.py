"""Benchmark of the EDXRaster hot path on one MI355X.

One step = one full frame (Renderer::RenderMesh: transform, clip/setup,
binning, raster, depth test, shading, write-back) of the 2,097,152-triangle
16x16 sphere grid at 1920x1080 with the mesh resident in HBM.  Prints one JSON
line with the frame throughput, the roofline of the dominant kernel (timed
live with HIP events on the library's stream) and the CPU oracle timed on
this host's cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SHADER_NAMES = {0: "Lambertian, constant albedo", 1: "Blinn-Phong s=200", 2: "vertex-colour Blinn-Phong s=32",
                3: "textured Lambertian, trilinear"}


def device_mesh(mesh):
    """Copies the mesh into HBM (torch allocations) and returns (edxr_mesh, keepalive)."""
    import torch
    from edxraster_amd import _lib as L
    planes = [torch.from_numpy(a).to("cuda") for a in mesh.planes()]
    idx = torch.from_numpy(mesh.indices.view(np.int32)).to("cuda")
    uv = [None if a is None else torch.from_numpy(a).to("cuda") for a in (mesh.tu, mesh.tv)]
    tex = None if mesh.tex_ids is None else torch.from_numpy(mesh.tex_ids.view(np.int32)).to("cuda")
    torch.cuda.synchronize()
    ptr = lambda t: None if t is None else C.cast(C.c_void_p(t.data_ptr()), L.F)
    uptr = lambda t: None if t is None else C.cast(C.c_void_p(t.data_ptr()), L.U32)
    m = L.Mesh(*[ptr(t) for t in planes], uptr(idx), mesh.num_vertices, mesh.num_triangles, L.MEM_DEVICE,
               ptr(uv[0]), ptr(uv[1]), uptr(tex))
    return m, (planes, idx, uv, tex)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, budget_s):
    from oracle.oracle import Oracle
    try:
        allowed = len(os.sched_getaffinity(0))
    except Exception:
        allowed = os.cpu_count() or 1
    nproc = os.cpu_count() or 1
    # At most 16 threads: a one-GPU share of the host's cores (the count and
    # the host's total are both reported).
    cores = max(1, min(allowed, 16))
    o = Oracle()

    def median_frame(threads, budget, max_frames):
        o.render_workload(w, threads=threads)  # warm-up (thread pool, page-in)
        times = []
        t_end = time.perf_counter() + budget
        while len(times) < max_frames and (not times or time.perf_counter() < t_end):
            t0 = time.perf_counter()
            o.render_workload(w, threads=threads)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), len(times)

    med, n = median_frame(cores, budget_s, 20)
    # 12 threads: the reference's per-tile reference lists are per core, at most 12
    med12, n12 = median_frame(12, budget_s / 2, 10)
    return {"value": w.mesh.num_triangles / med / 1e6, "unit": "Mtris/s", "cores": cores, "kind": "port",
            "sample": "%d full frames of %s (median %.1f ms/frame), C++/SSE restatement of the reference "
                      "CPU pipeline built -O3 -msse4.1, %d threads on %s" % (n, w.name, med * 1e3, cores, cpu_model()),
            "nproc": nproc, "cpus_allowed": allowed, "cpu_model": cpu_model(),
            "value_12_threads": w.mesh.num_triangles / med12 / 1e6,
            "sample_12_threads": "%d frames, median %.1f ms/frame" % (n12, med12 * 1e3)}


def copy_bandwidth_gbs(nbytes=1 << 30, reps=10):
    """Achievable HBM bandwidth on this GPU: a device-to-device copy of nbytes,
    counted as nbytes read + nbytes written per copy, timed with HIP events."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


# Streaming kernel of known size for the counter calibration: out = -in over
# CALIBRATION_BYTES of float32 reads 2^28 bytes and writes 2^28 bytes.
CALIBRATION_BYTES = 1 << 28
CALIBRATION_KERNEL = "neg_kernel"


def calibration_probe():
    """One launch of the calibration kernel (profiled beside the frames)."""
    import torch
    a = torch.ones(CALIBRATION_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.neg(a, out=b)
    torch.cuda.synchronize()
    del a, b


# HBM bytes per launch of the raster/shade pass (k_raster + k_shade), keyed by
# (workload, samples per pixel): FETCH_SIZE and WRITE_SIZE of the pass from two
# rocprofv3 --pmc runs of this bench with --calibrate, each scaled by what the
# same counter reads for the calibration kernel (bench.py --pmc-summary prints
# the table; the commands and raw KiB are in profiles/pmc_raster_shade_traffic.md).
# None where not measured.
RASTER_SHADE_TRAFFIC = {
    ("sphere_grid_16x16x64_1920x1080", 1): 57721519,        # tuning 1024,4
    ("heightfield_2236_3840x2160", 1): 710691778,           # tuning 4096,8
    ("sphere_grid_16x16x64_1920x1080", 4): 194235727,       # tuning 1024,4
    ("heightfield_2236_textured_3840x2160", 1): 771988753,  # tuning 4096,8
}


def pmc_summary(fetch_dir, write_dir):
    """Per-kernel mean FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of two
    rocprofv3 --pmc runs, the calibration factors (known bytes / counter KiB
    of the calibration kernel) and the raster/shade pass's HBM bytes."""
    import csv
    import glob

    def means(d, counter):
        acc = {}
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row["Kernel_Name"]
                    key = ("k_raster<true>" if "k_rasterILb1" in name or "k_raster<true" in name else
                           "k_raster<false>" if "k_raster" in name else
                           "k_shade_tex" if "k_shade_tex" in name else
                           "k_shade" if "k_shade" in name else
                           "calibration" if CALIBRATION_KERNEL in name else None)
                    if key:
                        acc.setdefault(key, []).append(float(row["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}

    fetch, nf = means(fetch_dir, "FETCH_SIZE")
    write, nw = means(write_dir, "WRITE_SIZE")
    known_kib = CALIBRATION_BYTES / 1024.0
    ff = known_kib / fetch["calibration"]
    fw = known_kib / write["calibration"]
    rows = [k for k in ("k_raster<false>", "k_raster<true>", "k_shade", "k_shade_tex") if k in fetch]
    traffic_kib = sum(fetch[k] * ff + write.get(k, 0.0) * fw for k in rows)
    return {"fetch_kib": fetch, "write_kib": write, "dispatches": {"fetch": nf, "write": nw},
            "fetch_factor": ff, "write_factor": fw, "kernels": rows,
            "traffic_bytes": int(round(traffic_kib * 1024))}


def frame_leg(w, steps, warmup, brief=False, tuning=None, msaa=0):
    """Frames of one workload: the timed loop replays the captured frame
    (hipGraph); a second loop of eager frames records HIP events around each
    stage on the library's stream, which time the raster/shade pass for the
    roofline (algorithmic bytes: 4 B bin entry + 40 B minimal setup record per
    tile reference, 44 B per referenced vertex, and per pixel 4 B of colour
    plus 8 B of depth and triangle ID per sample).  msaa = log2 of the samples
    per pixel (Renderer::SetMSAAMode)."""
    import torch
    from edxraster_amd import Renderer
    T, P, S = w.mesh.num_triangles, w.width * w.height, 1 << msaa
    r = Renderer(w.width, w.height)
    r.set_msaa_mode(msaa)
    w.apply(r)
    dm, keep = device_mesh(w.mesh)
    # Launch tuning (triangles per binning chunk, raster work items per CU):
    # given, or picked by the library's autotuner on this frame before timing.
    if tuning:
        r.set_tuning(*tuning)
        tuning = {"tris_per_chunk": tuning[0], "raster_blocks_per_cu": tuning[1], "frame_ms": None}
    else:
        tuning = r.autotune([w.job(mesh=dm)])
    for _ in range(warmup):
        r.render_mesh(dm)
    st = r.stats()
    recs = r.debug_records()
    v_ref = int(np.unique(recs[:, 13:16]).size) if recs.size else 0

    r.set_graph(True)
    for _ in range(max(1, warmup)):
        r.render_mesh(dm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render_mesh(dm)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    r.set_graph(False)

    r.set_profiling(True)
    r.profile(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render_mesh(dm)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / steps
    stage_ms, frames = r.profile()
    r.set_profiling(False)
    r.close()
    stage_avg = {k: v / max(frames, 1) for k, v in stage_ms.items()}
    n_ref = int(st["refs"])
    b_rs = n_ref * 44 + 44 * v_ref + (4 + 8 * S) * P
    raster_ms = stage_avg["raster"]
    achieved = b_rs / (raster_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": RASTER_SHADE_TRAFFIC.get((w.name, S)),
            "kernel": "k_raster+k_shade (raster/shade pass)", "algorithmic_bytes": b_rs, "kernel_ms": raster_ms}
    # Compulsory bytes per stage (achieved GB/s = bytes / stage time):
    # setup: 36 B in + 64 B out (shading and screen vertex) per vertex, 12 B
    # indices + 4 B descriptor per triangle; bin count: 4 B descriptor per
    # triangle + the chunk x tile count matrix written (4 B per entry); scans:
    # the matrix read twice and written once; scatter: 6 B (descriptor and
    # count) per triangle, the matrix read, 4 B per reference written.
    V, C = w.mesh.num_vertices, int(st["bin_chunks"])
    tiles = ((w.width + 31) // 32) * ((w.height + 31) // 32)
    stage_bytes = {"setup": 100 * V + 16 * T, "bin_count": 4 * T + 4 * C * tiles, "bin_scan": 12 * C * tiles,
                   "bin_scatter": 6 * T + 4 * C * tiles + 4 * n_ref, "raster": b_rs}
    stage_gbs = {k: stage_bytes[k] / (stage_avg[k] * 1e-3) / 1e9 for k in stage_bytes if stage_avg.get(k)}
    out = {"workload": w.name, "ms": ms, "eager_ms": eager_ms, "stage_ms": stage_avg, "stage_gbs": stage_gbs,
           "roofline": roof, "stats": st, "v_ref": v_ref, "tuning": tuning}
    if brief:
        out = {"workload": w.name, "samples": S, "triangles": T, "frame_ms": ms, "mtris_per_s": T / (ms * 1e-3) / 1e6,
               "stage_ms": stage_avg, "stage_gbs": stage_gbs, "roofline": roof, "tuning": tuning}
    return out


def batched_leg(steps, warmup, tuning=None, plain_renders=True):
    """256 independent jobs (19.2M triangles, 256^2-1024^2): one batched,
    graph-captured launch sequence vs 256 plain per-job renders.  A given
    tuning replaces the autotuner; plain_renders=False times the batch only
    (profiling runs)."""
    import torch
    from edxraster_amd import configs, Renderer
    from edxraster_amd import _lib as L
    jobs = configs.batch_jobs()
    dm, keep = [], []
    for w in jobs:
        m, k = device_mesh(w.mesh)
        dm.append(m)
        keep.append(k)
    arr = (L.Job * len(jobs))(*[w.job(mesh=m) for w, m in zip(jobs, dm)])
    r = Renderer(64, 64)
    if tuning:
        r.set_tuning(*tuning)
    else:
        tuning = r.autotune(arr)
    r.set_graph(True)
    for _ in range(warmup):
        r.render_batch(arr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render_batch(arr)
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / steps
    st = r.stats()
    r.close()
    T = sum(w.mesh.num_triangles for w in jobs)
    if not plain_renders:
        return {"metric": "batched frames/s", "jobs": len(jobs), "triangles": T,
                "frames_per_s": len(jobs) / batch_s, "batch_ms": batch_s * 1e3, "tuning": tuning}
    p = Renderer(64, 64)

    def plain():
        for w, m in zip(jobs, dm):
            p.resize(w.width, w.height)
            w.apply(p)
            p.render_mesh(m)
    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = max(1, steps // 2)
    for _ in range(reps):
        plain()
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / reps
    p.close()
    return {"metric": "batched frames/s", "jobs": len(jobs), "triangles": T,
            "pixels": int(sum(w.width * w.height for w in jobs)),
            "frames_per_s": len(jobs) / batch_s, "batch_ms": batch_s * 1e3,
            "plain_frames_per_s": len(jobs) / plain_s, "plain_ms": plain_s * 1e3,
            "speedup_vs_plain": plain_s / batch_s, "mtris_per_s": T / batch_s / 1e6,
            "tuning": tuning, "refs": int(st["refs"]), "graph": True}


def diff_leg(steps, warmup):
    """64 views at 512x512 of the 20k-triangle sphere-on-heightfield scene:
    forward (float RGB) + hand-written adjoint, L2 loss against targets with
    perturbed colours, light and shininess."""
    import torch
    from edxraster_amd import configs, Renderer, diff
    mesh, views = configs.diff_scene()
    idx, planes = diff.mesh_tensors(mesh)
    rng = np.random.default_rng(5)
    light0 = np.array(views[0].light, np.float64)
    axis = rng.normal(size=3)
    r = Renderer(64, 64)
    with torch.no_grad():
        tgt_cols = [(planes[k] + torch.from_numpy(rng.normal(0, 0.1, mesh.num_vertices).astype(np.float32)).cuda())
                    for k in (6, 7, 8)]
        tl = torch.tensor(configs.rotate(light0, axis, 5.0), dtype=torch.float32)
        tm = torch.tensor([0.2, 3.0, 3.0, 48.0])
        target = diff.render_views(r, views, idx, *planes[:6], *tgt_cols, tl, tm).clone()
    for t in planes:
        t.requires_grad_(True)
    light = torch.tensor(light0, dtype=torch.float32, requires_grad=True)
    material = torch.tensor([0.2, 3.0, 3.0, 32.0], requires_grad=True)

    def step():
        img = diff.render_views(r, views, idx, *planes, light, material)
        loss = ((img - target) ** 2).sum()
        loss.backward()
        return img

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    fwd = bwd = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        img = diff.render_views(r, views, idx, *planes, light, material)
        loss = ((img - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd += t1 - t0
        bwd += t2 - t1
    fwd, bwd = fwd / steps * 1e3, bwd / steps * 1e3
    check = diff_checks(r, views, idx, planes, light, material, target)
    r.close()
    return {"metric": "fwd+bwd ms/view", "views": len(views), "res": views[0].width,
            "triangles": mesh.num_triangles, "fwd_ms": fwd, "bwd_ms": bwd,
            "ms_per_view": (fwd + bwd) / len(views), "bwd_over_fwd": bwd / fwd, "checks": check,
            "note": "fwd includes the L2 loss; bwd the loss gradient + edxr_backward (adjoint kernels)"}


def diff_checks(r, views, idx, planes, light, material, target):
    """Correctness of the timed differentiable pass, from the library alone:
    the float image quantises to the byte image of every view, and the
    adjoint's d loss / d light_x agrees with a central difference of the
    float64-summed loss (h = 1e-2; visibility does not depend on the light)."""
    import torch
    from edxraster_amd import diff
    for p in planes:
        p.grad = None
    light.grad = material.grad = None
    img = diff.render_views(r, views, idx, *planes, light, material)
    ((img.double() - target.double()) ** 2).sum().backward()
    analytic = float(light.grad[0])
    quantised = all(np.array_equal((img[k].detach().cpu().numpy() * np.float32(255.0) + np.float32(0.5))
                                   .astype(np.int32), r.batch_read(k)[0][..., :3].astype(np.int32))
                    for k in range(len(views)))
    h = 1e-2
    with torch.no_grad():
        def loss_at(dx):
            l2 = light.detach().clone()
            l2[0] += dx
            im = diff.render_views(r, views, idx, *[p.detach() for p in planes], l2, material.detach())
            return float(((im.double() - target.double()) ** 2).sum())
        fd = (loss_at(h) - loss_at(-h)) / (2 * h)
    rel = abs(analytic - fd) / max(abs(fd), 1e-30)
    assert quantised, "float image does not quantise to the byte image"
    assert rel <= 1e-2, "adjoint d loss / d light_x %g vs central difference %g" % (analytic, fd)
    return {"float_image_quantises_to_bytes": quantised, "param": "light[0]", "adjoint": analytic,
            "central_difference": fd, "rel_err": rel}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="grid1m")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 10M-triangle, multisampled, batched and differentiable legs")
    ap.add_argument("--msaa", type=int, default=0, help="log2 samples per pixel of the timed frame (profiling runs)")
    ap.add_argument("--tuning", default=None,
                    help="TRIS_PER_CHUNK,RASTER_BLOCKS_PER_CU instead of autotuning (e.g. for profiling runs)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also launch the counter-calibration kernel once (for rocprofv3 --pmc runs)")
    ap.add_argument("--leg", choices=["batched", "differentiable"], default=None,
                    help="run only the batched (batch timed alone) or the differentiable leg and print its JSON "
                         "(profiling runs)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print the raster/shade traffic from two rocprofv3 --pmc output folders and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the multi-GPU path is not built",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from edxraster_amd import configs, Renderer

    tuning = tuple(int(x) for x in args.tuning.split(",")) if args.tuning else None
    if args.leg == "batched":
        print(json.dumps(batched_leg(args.steps, args.warmup, tuning, plain_renders=False)), flush=True)
        return
    if args.leg == "differentiable":
        print(json.dumps(diff_leg(args.steps, args.warmup)), flush=True)
        return
    w = configs.WORKLOADS[args.workload]()
    T, V, P = w.mesh.num_triangles, w.mesh.num_vertices, w.width * w.height
    fr = frame_leg(w, args.steps, args.warmup, tuning=tuning, msaa=args.msaa)
    ms, stage_avg, st = fr["ms"], fr["stage_ms"], fr["stats"]
    b_frame = 36 * V + 12 * T + 12 * P

    out = {
        "metric": "frame ms + Mtris/s (1M tris @1080p)",
        "value": T / (ms * 1e-3) / 1e6,
        "unit": "Mtris/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int32",
        "data": "synthetic (seeded SplitMix64 scene generated on the host, resident in HBM)",
        "config": {"workload": w.name, "samples_per_pixel": 1 << args.msaa, "triangles": T, "vertices": V, "width": w.width, "height": w.height,
                   "shader": SHADER_NAMES.get(w.shader, str(w.shader)), "parallelism": "single GPU",
                   "launch": "hipGraph replay per frame"},
        "frame_ms": ms,
        "gpix_s": P / (ms * 1e-3) / 1e9,
        "eager_frame_ms": fr["eager_ms"],
        "tuning": fr["tuning"],
        "stage_ms": stage_avg,
        "stage_gbs": fr["stage_gbs"],
        "dominant_stage": max(stage_avg, key=stage_avg.get),
        "roofline": fr["roofline"],
        "frame_roofline_frac": b_frame / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "hbm_copy_gbs": None,
        "stats": {k: int(v) for k, v in st.items()},
        "v_ref": fr["v_ref"],
    }
    if not args.no_extras:
        if args.workload != "hf10m":
            h = configs.WORKLOADS["hf10m"]()
            out["frame_10m"] = frame_leg(h, max(5, args.steps // 2), 2, brief=True)
        # the same frame with 4 samples per pixel, at the tuning picked above
        t = fr["tuning"]
        out["msaa_4x"] = frame_leg(w, max(5, args.steps // 2), 2, brief=True, msaa=2,
                                   tuning=(t["tris_per_chunk"], t["raster_blocks_per_cu"]))
        # the 10M frame with the textured Lambertian shader (trilinear, two 1024^2 slots)
        if args.workload != "hf10m_tex":
            ht = configs.WORKLOADS["hf10m_tex"]()
            out["textured_10m"] = frame_leg(ht, max(5, args.steps // 2), 2, brief=True)
        out["batched"] = batched_leg(max(3, args.steps // 4), 2)
        out["differentiable"] = diff_leg(max(3, args.steps // 4), 2)
    if args.calibrate:
        calibration_probe()
    # the achievable peak beside the published one: a 1 GiB device copy
    copy_gbs = copy_bandwidth_gbs()
    out["hbm_copy_gbs"] = copy_gbs
    for leg in [out] + [out[k] for k in ("frame_10m", "msaa_4x", "textured_10m") if k in out]:
        leg["roofline"]["frac_of_copy"] = leg["roofline"]["achieved"] / copy_gbs
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
