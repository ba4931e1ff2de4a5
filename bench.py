"""Benchmark: one step = one full frame of the 1080p workload.

Workload (the configuration the metric is quoted on): 1920x1080, 1.04M
triangles (1000 instanced UV spheres + a 200x200-quad ground grid), 8
Blinn-Phong materials, 4 point + 1 directional light all casting 2048^2 shadow
maps, AO on, sRGB.  A frame is (*Renderer).Render(): shadow passes, forward
raster, deferred shading with shadows and AO, gamma.  The scene is resident
in HBM before the timed region; each step is one synchronous prh_render.

Prints one JSON line.  `value` is frames/s on the one GPU; Mtris/s and Mpix/s
ride along.  After the timed region the same frames run again with every
pass on one stream, so each kernel's HIP-event time is its own: `roofline` is
the longest of them (the dominant kernel), `roofline_shade_pass` the whole
shading pass (AO + deferred shading) against the survey's byte formula, each
with its VALU issue rate beside the HBM one.  `cpu_baseline` is the C
restatement of the reference CPU renderer (the parity oracle) timed on this
host in the reference's throughput mode, with the one-thread mode beside it;
`fp64_err` is the shading's largest error against an FP64 re-evaluation.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "1080p frames/s + Mtris/s + HBM GB/s; batched & fwd+adjoint vs plain 1-view"


EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)
VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions/s: 256 CUs x 4 SIMDs, 4 cycles each at 2.4 GHz


def shade_bytes(scene, vis: np.ndarray) -> dict:
    """Algorithmic (distinct) DRAM bytes of one k_shade launch.  Per pixel:
    the 8-byte visibility key and the 4-byte RGBA8 output.  Per covered pixel:
    the 4-byte AO factor if its material has AO, and one 4-byte shadow texel
    per shadow map if it receives shadows.  Per distinct visible triangle: its
    16-byte shading record (vertex indices, material), and per visible clip
    piece its 144-byte clip record.  Per distinct vertex of the visible
    unclipped triangles: the 3 x 16-byte transformed vertex (screen xyz + 1/w,
    normal, back-transformed position) and its 4-byte colour."""
    P = vis.size
    cov = vis != EMPTY
    order = (vis[cov] & np.uint64(0xFFFFFFFF)).astype(np.int64)
    g, piece = order >> 3, order & 7
    tri_counts = np.array([scene.meshes[m].tri.shape[0] for m in scene.obj_mesh])
    vtx_counts = np.array([scene.meshes[m].pos.shape[0] for m in scene.obj_mesh])
    obj_tri_start = np.concatenate([[0], np.cumsum(tri_counts)])
    obj_vtx_start = np.concatenate([[0], np.cumsum(vtx_counts)])
    mesh_tri_off = np.concatenate([[0], np.cumsum([m.tri.shape[0] for m in scene.meshes])])
    tri_v_all = np.concatenate([m.tri for m in scene.meshes]).astype(np.int64)
    tri_mat_all = np.concatenate([m.mat for m in scene.meshes]).astype(np.int64)
    o = np.searchsorted(obj_tri_start, g, side="right") - 1
    mesh = np.asarray(scene.obj_mesh)[o]
    t = mesh_tri_off[mesh] + (g - obj_tri_start[o])
    direct = piece == 7
    ug = np.unique(g[direct])
    n_clip = np.unique(order[~direct]).size
    od = np.searchsorted(obj_tri_start, ug, side="right") - 1
    td = mesh_tri_off[np.asarray(scene.obj_mesh)[od]] + (ug - obj_tri_start[od])
    verts = np.unique((obj_vtx_start[od][:, None] + tri_v_all[td]).ravel())
    mat = tri_mat_all[t] + np.where(tri_mat_all[t] >= 0, np.asarray(scene.obj_mat_base)[o], 0)
    ok = (mat >= 0) & (mat < len(scene.materials))
    ao_m = np.array([m.ambient_occlusion for m in scene.materials] + [False])
    rs_m = np.array([m.receive_shadow for m in scene.materials] + [False])
    mi = np.where(ok, mat, len(scene.materials))
    n_shadow = sum(1 for l in scene.lights if l.cast_shadow) if scene.shadow_map else 0
    n_ao = int(ao_m[mi].sum())
    n_rs = int(rs_m[mi].sum()) if n_shadow else 0
    b = (P * (8 + 4) + n_ao * 4 + n_rs * 4 * n_shadow + np.unique(g).size * 16 + n_clip * 144 + verts.size * 52)
    return dict(bytes=int(b), pixels=P, covered=int(cov.sum()), visible_tris=int(ug.size + n_clip),
                visible_verts=int(verts.size), ao_px=n_ao)


def ao_bytes(scene, vis: np.ndarray, n_ao: int) -> int:
    """Algorithmic DRAM bytes of one k_ao launch: every pixel's 8-byte key,
    each covered pixel's 4-byte depth (the taps re-read the depth buffer,
    which stays in cache) and each AO pixel's 4-byte factor."""
    return int(vis.size * 8 + int((vis != EMPTY).sum()) * 4 + n_ao * 4)


def shade_pass_bytes(P: int, ao: bool, n_shadow: int, visible_tris: int) -> int:
    """Algorithmic DRAM bytes of the whole shading pass by the survey's
    formula: per pixel the 8-byte visibility key, the 4-byte RGBA8 output,
    the 4-byte depth if AO is on and one 4-byte shadow texel per shadowed
    light; per distinct visible triangle 12 bytes of indices and 72 of
    positions and normals."""
    return int(P * (8 + 4 + (4 if ao else 0) + 4 * n_shadow) + visible_tris * (12 + 72))


def pmc_summary(path: str) -> dict:
    """Per-kernel figures from a committed rocprofv3 --pmc summary: lines
    `kernel.<name>.<field>: <value>` (traffic_bytes, valu_insts per
    dispatch, with FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel
    calibration), and `source_commit: <hash>`, the tree it was recorded on."""
    out = {}
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("kernel.") and ":" in line:
                    k, v = line.split(":", 1)
                    _, name, field_ = k.strip().split(".", 2)
                    out.setdefault(name, {})[field_] = float(v)
                elif line.startswith("source_commit:"):
                    out["_commit"] = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return out


PASSES = ("total", "shadow", "vertex", "setup", "raster", "ao", "shade")


def run_frames(r, steps: int):
    """`steps` synchronous frames timed by the wall clock, then `steps` more
    whose per-pass times are read from each frame's own HIP events (reading
    them is host work the timed frames leave out): (wall seconds, mean
    per-pass ms)."""
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render_device()  # synchronous: returns when the frame is complete
    dt = time.perf_counter() - t0
    acc = dict.fromkeys(PASSES, 0.0)
    for _ in range(steps):
        r.render_device()
        st = r.stats()
        for k in PASSES:
            acc[k] += getattr(st, "ms_" + k)
    return dt, {k: v / steps for k, v in acc.items()}


def adjoint_bench(dev, steps: int) -> dict:
    """Forward + adjoint against a plain single view: the 512x512 ~100k-
    triangle 4-light scene, target rendered from seed parameters, gradients
    at a 10 % Gaussian perturbation.  Device times from HIP events: one
    frame (prh_render), one adjoint launch (loss + all gradients)."""
    from polyred_amd import diff, render, scenes

    s = scenes.c4_adjoint()
    r = render.renderer_for(s, dev)
    r.render_device()
    p0 = diff.scene_params(s)
    truth = diff.perturb(p0, np.random.default_rng(4), 0.1)
    target = np.nan_to_num(r.ForwardLinear(truth))
    r.Backward(p0, target)
    frame = adj = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render_device()
        frame += r.stats().ms_total
    t1 = time.perf_counter()
    for _ in range(steps):
        loss, _ = r.Backward(p0, target)
        adj += r.stats().ms_diff
    t2 = time.perf_counter()
    r.free()
    frame, adj = frame / steps, adj / steps
    return {"workload": s.name, "triangles": s.n_triangles, "ms_frame": round(frame, 4), "ms_adjoint": round(adj, 4),
            "ms_fwd_adjoint": round(frame + adj, 4), "fwd_adjoint_vs_plain": round((frame + adj) / frame, 3),
            "fwd_adjoint_per_s": round(1000.0 / (frame + adj), 1),
            "wall_ms_backward_call": round(1000.0 * (t2 - t1) / steps, 3), "loss": loss}


def batch_bench(dev, reps: int) -> dict:
    """Batched forward: 64 independent views (each its own scene: 256^2 to
    1920x1080, 10k-500k triangles, 1-8 shadow-casting lights with 1024^2
    maps), rendered as 64 sequential single-view prh_render calls, as one
    prh_render_batch of wide launches (one per pass over every view), as
    that batch captured in a HIP graph and replayed, and (for comparison)
    as each view's own launches over the batch streams, eager and as a
    graph.  Wall-clock per batch, scenes resident."""
    from polyred_amd import render, scenes

    views = scenes.c3_views()
    rs = [render.renderer_for(s, dev) for s in views]
    for r in rs:
        r.render_device()
    # the wide batch's knobs, per size class, timed before the measurement
    # (results unchanged: every view stays bit-identical)
    render.render_batch(rs)
    ms_default = min(render.render_batch(rs) for _ in range(3))
    small, large = render.autotune_batch(rs, frames=3)
    knobs = ("item_prims", "raster_wg_per_cu", "shadow_direct_max", "main_direct_max", "raster_tile")
    tuned = {"small_views": {k: getattr(small, k) for k in knobs}, "large_views": {k: getattr(large, k) for k in knobs},
             "ms_batch_default_knobs": round(ms_default, 3), "ms_batch_tuned": round(small.ms, 3)}
    for streams in (False, True):
        render.render_batch(rs, streams=streams)
        render.render_batch(rs, graph=True, streams=streams)

    def timed(fn):
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        return 1000.0 * (time.perf_counter() - t) / reps

    seq = timed(lambda: [r.render_device() for r in rs])
    bat = timed(lambda: render.render_batch(rs))
    gr = timed(lambda: render.render_batch(rs, graph=True))
    sbat = timed(lambda: render.render_batch(rs, streams=True))
    sgr = timed(lambda: render.render_batch(rs, graph=True, streams=True))
    for r in rs:
        r.free()
    n_tri = sum(s.n_triangles for s in views)
    n_px = sum(s.n_pixels for s in views)
    best = min(bat, gr)
    return {"views": len(views), "triangles": n_tri, "pixels": n_px,
            "shadow_maps": sum(len(s.lights) for s in views), "shadow_size": views[0].shadow_size,
            "ms_sequential": round(seq, 3), "ms_batch": round(bat, 3), "ms_graph": round(gr, 3),
            "ms_streams": round(sbat, 3), "ms_streams_graph": round(sgr, 3),
            "speedup_streams_graph": round(seq / sgr, 3),
            "speedup_batch": round(seq / bat, 3), "speedup_graph": round(seq / gr, 3),
            "views_per_s": round(1000.0 * len(views) / best, 1), "mtris_per_s": round(n_tri / best / 1e3, 1),
            "mpix_per_s": round(n_px / best / 1e3, 1), "tuning": tuned}


def fit_bench(dev, steps: int = 200) -> dict:
    """Batched inverse rendering: 32 views (256^2 to 1024^2) of one 250k-
    triangle scene with 8 materials and 4 shadow-casting point lights
    (1024^2 maps) at fixed visibility.  Targets from the scene's own
    materials and lights; start 30 % off (Gaussian, per value); 200 Adam
    steps at lr 1e-2 on Kd, Ks, shininess, light positions and intensities,
    each step the adjoint of all 32 views in one launch plus the update (a
    replayed HIP graph).  Run twice: per-pixel state rebuilt from the
    visibility buffer at every step, and read from a stored G-buffer."""
    from polyred_amd import diff, render, scenes

    s, cams = scenes.c5_inverse()
    gs, rs = render.views_of(s, cams, dev)
    render.render_batch(rs)  # every view's frame (visibility and shadow maps): one wide batch
    ms_render = render.render_batch(rs)
    truth = diff.scene_params(s)
    rng = np.random.default_rng(5)
    start = truth.copy()
    groups = ("kd", "ks", "shininess", "light_pos", "light_int")
    for n in groups:
        a = getattr(start, n)
        a *= (1 + 0.3 * rng.standard_normal(a.shape)).astype(a.dtype)
    out = {"workload": s.name, "views": len(rs), "triangles": s.n_triangles,
           "pixels": int(sum(w * h for _, w, h in cams)), "steps": steps, "lr": 1e-2,
           "ms_render_views_batch": round(ms_render, 3), "rmse_start": diff.rmse(start, truth),
           "gauge_free_start": diff.gauge_free_error(start, truth)}
    for mode, stored in (("rebuilt", False), ("stored", True)):
        fit = diff.Fit(rs, stored_gbuffer=stored)
        fit.targets(truth)
        fit.run(start, steps=2)  # warm-up
        p, hist = fit.run(start, steps=steps)
        st = fit.stats()
        out[mode] = {"ms_per_step": round(st.ms_steps / steps, 4), "ms_total": round(st.ms_steps, 2),
                     "ms_gbuffer": round(st.ms_gbuffer, 3),
                     "visibility_mb": round(st.visibility_bytes / 2 ** 20, 1),
                     "loss_start": float(hist[0]), "loss_end": float(hist[-1]),
                     "rmse_end": diff.rmse(p, truth), "gauge_free_end": diff.gauge_free_error(p, truth)}
        fit.free()
    gs.free()
    return out


def cpu_baseline(scene, frames: int) -> dict:
    """The C restatement of the reference CPU renderer (the parity oracle)
    timed on this host's cores in the reference's two modes: B, throughput
    (every raster pass and the deferred pass on all the threads, per-pixel
    locks, as the default Workers(nproc)), and A, one thread (Workers(1)).
    Also returns the parity-order frame (for the parity check and the FP64
    shading error)."""
    import oracle  # the parity oracle, timed as the CPU baseline only

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fr = oracle.render(scene)  # the parity order; also warms the page cache
    try:
        oracle.set_workers(threads)
        t = time.perf_counter()
        for _ in range(frames):
            oracle.render(scene)
        dt_b = time.perf_counter() - t
        oracle.set_workers(1)
        t = time.perf_counter()
        oracle.render(scene)
        dt_a = time.perf_counter() - t
    finally:
        oracle.set_workers(0)
    omit = ("the reference also walks the scene graph for its lights at every pixel (render/raster.go:334), "
            "which the restatement does once per frame, so both modes are faster than the Go renderer would be")
    return dict(value=frames / dt_b, unit="frames/s", cores=threads, kind="port",
                sample=f"Mode B: {frames} full frames of the same 1080p workload through the C restatement of the "
                       f"reference CPU renderer, every pass on {threads} threads with per-pixel locks, "
                       f"{dt_b:.1f} s; {omit}",
                mode_a={"value": round(1.0 / dt_a, 4), "unit": "frames/s", "cores": 1,
                        "sample": f"Mode A: 1 full frame on one thread (Workers(1)), {dt_a:.1f} s"},
                _frame=fr)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-view batched measurement")
    ap.add_argument("--no-fit", action="store_true", help="skip the batched inverse-rendering measurement")
    ap.add_argument("--small", action="store_true", help="reduced scene for a quick check")
    ap.add_argument("--only-batch", action="store_true",
                    help="run the 64-view batched measurement alone (for a profile of the batch)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default knobs instead of prh_autotune's choice (untimed, before the warm-up)")
    ap.add_argument("--overlap", choices=("auto", "0", "1"), default="auto",
                    help="the timed frames' shadow passes beside the main view's passes on a second stream: "
                         "the autotuner's choice (auto), off or on")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run 3 elementwise copies of MIB MiB (torch.neg), the known-size kernel that "
                         "calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build renders on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    from polyred_amd import abi, render, scenes

    if args.copy_probe > 0:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)  # an elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        del src, dst

    if args.only_batch:
        dev = render.Device(0)
        print(json.dumps({"batched": batch_bench(dev, 3)}))
        dev.close()
        return 0

    if args.small:
        scene = scenes.c2_scale(width=480, height=270, n_spheres=200, grid=50, shadow_size=512)
    else:
        scene = scenes.c2_scale()
    dev = render.Device(0)
    r = render.renderer_for(scene, dev)
    tuned = None
    if not args.no_autotune:
        t = r.Autotune(frames=5)  # occupancy, bin capacity, small-triangle thresholds; results unchanged
        tuned = {"raster_wg_per_cu": t.raster_wg_per_cu, "item_prims": t.item_prims,
                 "shadow_direct_max": t.shadow_direct_max, "main_direct_max": t.main_direct_max,
                 "overlap_shadows": t.overlap_shadows, "raster_tile": t.raster_tile}
    if args.overlap != "auto":
        r.set_tuning(overlap_shadows=int(args.overlap))
    overlap = bool(r.tuning().overlap_shadows)
    for _ in range(args.warmup):
        r.render_device()
    dt, ms = run_frames(r, args.steps)  # the timed region
    # (outside the timed region) the same frames with every pass on one
    # stream: each pass's time is its kernels' own, and the passes add up
    r.set_tuning(overlap_shadows=0)
    for _ in range(2):
        r.render_device()
    _, ms1 = run_frames(r, args.steps)
    r.set_tuning(overlap_shadows=int(overlap))
    # (outside the timed region) the frame plus its image copied to the host
    # over PCIe, as a caller that reads every frame back sees it
    n_rb = max(3, args.steps // 4)
    host = render.pinned_array((scene.height, scene.width, 4))  # page-locked: the link's full rate
    r.image(host)
    t_rb = time.perf_counter()
    for _ in range(n_rb):
        r.render_device()
        r.image(host)
    ms_readback = 1000.0 * (time.perf_counter() - t_rb) / n_rb
    vis = r.read(abi.PRH_BUF_VIS)
    img = r.image()
    st = r.stats()
    n_shadow = sum(1 for l in scene.lights if l.cast_shadow) if scene.shadow_map else 0
    any_ao = any(m.ambient_occlusion for m in scene.materials)
    sb = shade_bytes(scene, vis)
    ab = ao_bytes(scene, vis, sb["ao_px"])
    ms_step = 1000.0 * dt / args.steps
    fps = args.steps / dt
    pmc_path = os.path.join("profiles", "c2_1080p_serial_pmc.md")
    pmc = pmc_summary(os.path.join(ROOT, pmc_path))

    def roof(kernel: str, nbytes: int, ms_kernel: float, pmc_kernels=()) -> dict:
        achieved = nbytes / (ms_kernel * 1e-3) / 1e9
        ks = [pmc.get(k, {}) for k in (pmc_kernels or (kernel,))]
        traffic = sum(p["traffic_bytes"] for p in ks) if ks and all("traffic_bytes" in p for p in ks) else None
        insts = sum(p["valu_insts"] for p in ks) if ks and all("valu_insts" in p for p in ks) else None
        out = {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
               "algorithmic_bytes": nbytes, "ms": round(ms_kernel, 4)}
        if insts is not None:
            rate = insts / (ms_kernel * 1e-3)
            out["valu"] = {"insts_per_launch": int(insts), "achieved": round(rate / 1e9, 1), "peak": VALU_PEAK / 1e9,
                           "unit": "G wave64 insts/s", "frac": round(rate / VALU_PEAK, 3)}
        return out

    pmc_note = (f"traffic and VALU instruction counts per launch from {pmc_path} (rocprofv3 --pmc, recorded on "
                f"commit {pmc.get('_commit', '?')}); times from this run's HIP events with every pass on one stream")
    # the dominant kernel: the longest of the frame's kernels run alone
    ao_roof = dict(roof("k_ao", ab, ms1["ao"]), standalone=True, source=pmc_note)
    shade_roof = dict(roof("k_shade", sb["bytes"], ms1["shade"]), standalone=True, covered_px=sb["covered"],
                      visible_tris=sb["visible_tris"], visible_verts=sb["visible_verts"], source=pmc_note)
    dominant = ao_roof if ms1["ao"] >= ms1["shade"] else shade_roof
    # the whole shading pass (the reference's shade: Blinn-Phong, shadows,
    # AO, gamma) against the survey's byte formula
    pass_bytes = shade_pass_bytes(scene.n_pixels, any_ao, n_shadow, sb["visible_tris"])
    pass_roof = dict(roof("k_ao + k_shade", pass_bytes, ms1["ao"] + ms1["shade"], ("k_ao", "k_shade")),
                     standalone=True, formula="P*(8 vis + 4 out + 4 depth [AO] + 4 per shadowed light) + "
                                              "T_vis*(12 + 72)", source=pmc_note)
    other = ms1["total"] - sum(ms1[k] for k in PASSES[1:])

    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, no assets)",
        "config": {"workload": scene.name, "width": scene.width, "height": scene.height,
                   "triangles": scene.n_triangles, "lights": len(scene.lights), "shadow_maps": n_shadow,
                   "shadow_size": scene.shadow_size, "ao": any_ao, "srgb": True},
        "mtris_per_s": round(scene.n_triangles * fps / 1e6, 2),
        "mpix_per_s": round(scene.n_pixels * fps / 1e6, 2),
        "roofline": dominant,
        "roofline_shade_pass": pass_roof,
        "roofline_kernels": {"k_ao": ao_roof, "k_shade": shade_roof},
        # per-pass device times of the timed frames; with the shadow passes
        # on the second stream (overlap) the shadow span and the shading's
        # first half run beside the main view's passes, so they do not add up
        "kernel_ms": dict({k: round(v, 4) for k, v in ms.items()}, overlap=overlap),
        # every pass on one stream: the passes and "other" (clears, the
        # background colour) add up to total
        "kernel_ms_serial": dict({k: round(v, 4) for k, v in ms1.items()}, other=round(other, 4)),
        "frame": {"prims": st.n_prims, "clipped": st.n_clipped, "bin_pairs": st.n_bin_pairs},
        "tuning": tuned,
        "ms_per_frame_with_readback": round(ms_readback, 4),
    }
    r.free()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(dev, max(5, args.steps // 2))
    if not args.no_batch:
        out["batched"] = batch_bench(dev, 3)
    if not args.no_fit:
        out["inverse"] = fit_bench(dev)
    dev.close()
    if not args.no_cpu_baseline:
        cb = cpu_baseline(scene, args.cpu_frames)
        fr = cb.pop("_frame")
        order = (vis & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        order[vis == np.uint64(0xFFFFFFFFFFFFFFFF)] = 0xFFFFFFFF
        out["parity_vs_cpu"] = bool(np.array_equal(order, fr.vis_order) and np.array_equal(img, fr.image))
        # the shading's largest |float32 - float64| over the frame (0..255
        # scale of the pre-Round sums; libprh's sums equal the oracle's)
        out["fp64_err"] = float(fr.err64.value)
        cb["value"] = round(cb["value"], 4)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
