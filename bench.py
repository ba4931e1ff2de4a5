"""Benchmark: one step = one full frame of the 1080p workload.

Workload (the configuration the metric is quoted on): 1920x1080, 1.04M
triangles (1000 instanced UV spheres + a 200x200-quad ground grid), 8
Blinn-Phong materials, 4 point + 1 directional light all casting 2048^2 shadow
maps, AO on, sRGB.  A frame is (*Renderer).Render(): shadow passes, forward
raster, deferred shading with shadows and AO, gamma.  The scene is resident
in HBM before the timed region; each step is one synchronous prh_render.

Prints one JSON line.  `value` is frames/s on the one GPU; Mtris/s and Mpix/s
ride along.  `roofline` is for the dominant kernel, timed with HIP events on
libprh's stream over the timed region; `cpu_baseline` is the C restatement of
the reference CPU renderer (the parity oracle) timed on this host.
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


def pmc_summary(path: str) -> dict:
    """Per-kernel figures from a committed rocprofv3 --pmc summary: lines
    `kernel.<name>.<field>: <value>` (traffic_bytes, valu_insts per
    dispatch, with FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel
    calibration)."""
    out = {}
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("kernel.") and ":" in line:
                    k, v = line.split(":", 1)
                    _, name, field_ = k.strip().split(".", 2)
                    out.setdefault(name, {})[field_] = float(v)
    except OSError:
        pass
    return out


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
            "mpix_per_s": round(n_px / best / 1e3, 1)}


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
    import oracle  # the parity oracle, timed as the CPU baseline only

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.render(scene)  # warm the page cache / lazy init
    t = time.perf_counter()
    fr = None
    for _ in range(frames):
        fr = oracle.render(scene)
    dt = time.perf_counter() - t
    return dict(value=frames / dt, unit="frames/s", cores=threads, kind="port",
                sample=f"{frames} full frames of the same 1080p workload through the C restatement of the "
                       f"reference CPU renderer (raster passes sequential, deferred pass on {threads} OpenMP "
                       f"threads), {dt:.1f} s", _frame=fr)


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
                 "overlap_shadows": t.overlap_shadows}
    for _ in range(args.warmup):
        r.render_device()
    acc = dict(total=0.0, shadow=0.0, vertex=0.0, setup=0.0, raster=0.0, ao=0.0, shade=0.0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render_device()  # synchronous: returns when the frame is complete
        st = r.stats()
        acc["total"] += st.ms_total
        acc["shadow"] += st.ms_shadow
        acc["vertex"] += st.ms_vertex
        acc["setup"] += st.ms_setup
        acc["raster"] += st.ms_raster
        acc["shade"] += st.ms_shade
        acc["ao"] += st.ms_ao
    dt = time.perf_counter() - t0
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
    ms = {k: v / args.steps for k, v in acc.items()}
    vis = r.read(abi.PRH_BUF_VIS)
    img = r.image()
    st = r.stats()
    n_shadow = sum(1 for l in scene.lights if l.cast_shadow)
    sb = shade_bytes(scene, vis)
    split_shade = bool(r.tuning().overlap_shadows and scene.shadow_map and n_shadow > 0
                       and any(m.ambient_occlusion for m in scene.materials))
    ab = ao_bytes(scene, vis, sb["ao_px"])
    ms_step = 1000.0 * dt / args.steps
    fps = args.steps / dt
    pmc = pmc_summary(os.path.join(ROOT, "profiles", "c2_1080p_pmc.md"))

    def roof(kernel: str, nbytes: int, ms_kernel: float) -> dict:
        achieved = nbytes / (ms_kernel * 1e-3) / 1e9
        p = pmc.get(kernel, {})
        r = {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
             "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": p.get("traffic_bytes"),
             "algorithmic_bytes": nbytes, "ms": round(ms_kernel, 4)}
        if "valu_insts" in p:
            r["valu_issue_frac"] = round(p["valu_insts"] / (ms_kernel * 1e-3) / VALU_PEAK, 3)
        return r

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
                   "shadow_size": scene.shadow_size, "ao": True, "srgb": True},
        "mtris_per_s": round(scene.n_triangles * fps / 1e6, 2),
        "mpix_per_s": round(scene.n_pixels * fps / 1e6, 2),
        "kernel_ms": {k: round(v, 4) for k, v in ms.items()},
        # the dominant kernel (AO, VALU-bound: its HBM fraction is small by
        # construction) and the deferred-shading kernel the HBM target names
        # (with the shadow passes on the second stream and AO in the scene, the
        # shading's first half runs beside k_ao: both times are then spans of
        # kernels sharing the GPU, and "overlapped" says so)
        "roofline": dict(roof("k_ao", ab, ms["ao"]), overlapped=split_shade),
        "roofline_shade": dict(roof("k_shade", sb["bytes"], ms["shade"]), covered_px=sb["covered"],
                               visible_tris=sb["visible_tris"], visible_verts=sb["visible_verts"],
                               overlapped=split_shade),
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
        cb["value"] = round(cb["value"], 4)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
