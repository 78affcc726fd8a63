"""Benchmark of the geometric-query path on one MI355X.

Headline workload (one "step" = one pass of the hot path over one batch):
  rays_scene -- procedural terrain 708x708 (999,698 tris) + 64 icospheres L5
  (1,310,720 tris) = 2,310,418 FP32 triangles; 1,048,576 incoherent rays traced for the
  closest hit (tmin 1e-5, tmax 1e7) and again for any hit (tmax U[1,50]).  value = rays
  per second over both passes, inputs resident in HBM, BVH built before timing.

The same JSON line carries:
  roofline      achieved algorithmic GB/s of the closest-hit traversal kernel over its
                mean duration (HIP events on the launch stream), against the 8 TB/s HBM3E
                peak.  Algorithmic bytes, the same formula for every traversal leg:
                  query I/O + node_bytes per wide node visited + 36 per triangle tested
                (node_bytes = the traversed node encoding: 64 for the 8-bit 4-wide nodes
                rays use, 128 for the FP32 4-wide nodes closest points use; 36 = three
                float3 vertices), visits counted by the kernel's counting variant on the
                same BVH and inputs.  I/O: closest hit 32 in + 16 out, any hit 32 + 1,
                closest point 12 in + 20 out + 12 barycentrics.  Backward: 144 B per
                closest-point query, 180 B per ray.
  cpu_baseline  the oracle's CPU LBVH (C, pthreads on every host core) on a bounded
                sample of the same rays.
  extra         the other workloads, each with the same roofline fields: closest-point
                queries/s (5,120 and 5,242,880 triangles), any hit, forward and backward
                on the differentiable 1.3M-triangle workload, the 256-mesh batch against
                the plain per-mesh path (default and autotuned parameters), the build
                (bytes: what a build must read and write, 100 B per triangle + the wide
                nodes), the batch replay (its builds + its queries' visits).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-rt_amd"))
sys.path.insert(0, ROOT)

METRIC = "closest-pt queries/s + rays/s, 1M-tri, fwd & fwd+bwd, batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
TRI_BYTES = 36  # three float3 vertices per triangle tested
CPQ_BWD_BYTES = 144  # g, prim, bary, pt, dist, q, index gather, g_q out, 3 x 2 x 12 vertex RMW
RAY_BWD_BYTES = 180  # the same for a ray, plus 36 B of vertex gathers for the normal


def roof(byts, ms, kernel):
    """Roofline entry: algorithmic bytes per launch over the launch's mean time."""
    gbs = byts / (ms * 1e-3) / 1e9
    return {"kernel": kernel, "bound": "hbm", "bytes_per_launch": int(byts), "kernel_ms": ms,
            "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}


def cpq_bytes(n, nodes, tris, node_bytes):
    """Algorithmic bytes of a closest-point launch: query I/O (12 in, 20 out, 12
    barycentrics) + node_bytes per node visited + 36 per triangle tested."""
    return n * (12 + 20 + 12) + nodes * node_bytes + tris * TRI_BYTES


def ray_bytes(n, nodes, tris, node_bytes, any_hit=False):
    """Algorithmic bytes of a ray launch: 32 in + 16 out (1 for any hit) per ray, plus
    node_bytes per node visited and 36 per triangle tested."""
    return n * (32 + (1 if any_hit else 16)) + nodes * node_bytes + tris * TRI_BYTES


def build_bytes(info):
    """The least a BVH build must move: read each triangle's indices and vertices (12 +
    36 B), write its packed record and leaf-order entry (48 + 4 B) and every wide node in
    each encoding kept (cpq_node_bytes, plus node_bytes when rays use 8-bit nodes)."""
    per_node = info["cpq_node_bytes"] + (info["node_bytes"] if info["node_bytes"] != info["cpq_node_bytes"] else 0)
    return info["n_tris"] * (12 + 36 + 48 + 4) + info["n_nodes"] * per_node


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, steps, warmup, stream=None):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps  # ms


def measured_traffic(args, n_tris, n_rays, node_format):
    """HBM bytes per closest-hit launch from the committed counter summary
    (profiles/rays_scene_traffic.json: FETCH_SIZE and WRITE_SIZE passes, each scaled by
    what the same counter reads for an elementwise kernel that reads and writes 256 MiB),
    when it was measured on this exact configuration; None otherwise."""
    path = os.path.join(ROOT, "profiles", "rays_scene_traffic.json")
    try:
        m = json.load(open(path))
    except (OSError, ValueError):
        return None
    keys = ("n_tris", "n_rays", "node_width", "stack_depth", "treelet_size", "node_format",
            "fetch_kb_raw", "write_kb_raw", "calib_kib", "calib_fetch_kb_raw", "calib_write_kb_raw")
    if any(k not in m for k in keys):
        return None
    if (m["n_tris"], m["n_rays"], m["node_width"], m["stack_depth"], m["treelet_size"],
            m["node_format"], m.get("leaf_size", 4), m.get("waves_per_cu", 0), m.get("leaf_trigger", 0),
            m.get("refill_trigger", 0)) != \
            (n_tris, n_rays, args.width, args.stack, args.treelet, node_format, args.leaf, args.waves,
             args.leaf_trigger, args.refill_trigger):
        return None
    rd = m["fetch_kb_raw"] * m["calib_kib"] / m["calib_fetch_kb_raw"]
    wr = m["write_kb_raw"] * m["calib_kib"] / m["calib_write_kb_raw"]
    return (rd + wr) * 1024.0


def fp64_rays(O, V, F, R, prim, t, every):
    """FP32 closest hits against the oracle's FP64 brute force on every `every`-th ray:
    id agreement, and the relative error of t where the ids agree."""
    idx = np.arange(0, R.shape[0], every)
    ep, et, _ = O.ray_closest_brute_f64(V.astype(np.float64), F, R[idx].astype(np.float64))
    p, tt = prim[idx], t[idx].astype(np.float64)
    same = (p == ep)
    hit = same & (ep >= 0)
    rel = np.abs(tt[hit] - et[hit]) / np.maximum(np.abs(et[hit]), 1e-30)
    return {"sample": int(idx.size), "id_agreement": float(same.mean()),
            "t_max_rel_err": float(rel.max()) if rel.size else 0.0,
            "t_p999_rel_err": float(np.quantile(rel, 0.999)) if rel.size else 0.0}


def fp64_cpq(O, V, F, Q, prim, dist, every):
    """FP32 closest points against the oracle's FP64 brute force on every `every`-th
    query: id agreement, and the relative error of the distance where the ids agree."""
    idx = np.arange(0, Q.shape[0], every)
    ep, _, ed, _ = O.cpq_brute_f64(V.astype(np.float64), F, Q[idx].astype(np.float64))
    p, d = prim[idx], dist[idx].astype(np.float64)
    same = p == ep
    rel = np.abs(d[same] - ed[same]) / np.maximum(ed[same], 1e-30)
    return {"sample": int(idx.size), "id_agreement": float(same.mean()),
            "dist_max_rel_err": float(rel.max()) if rel.size else 0.0,
            "dist_p999_rel_err": float(np.quantile(rel, 0.999)) if rel.size else 0.0}


def roof_kernel(args, info, kind):
    q8 = info["node_format"] in (2, 3) if kind != "cpq" else info["node_format"] == 2
    name = {"closest": "k_ray<W=%d,%s,closest>", "any": "k_ray<W=%d,%s,any>", "cpq": "k_cpq<W=%d,%s>+k_cpq_finish"}[kind]
    return name % (args.width, "q8" if q8 else "f32")


def ray_bench(args, torch, gpurt, W, O, dev, ctx):
    t0 = time.time()
    w = W.rays_scene()
    log(f"[bench] rays_scene generated in {time.time() - t0:.1f}s: {w['meta']}")
    V, F = w["verts"], w["tris"]
    tv = torch.from_numpy(V).to(dev)
    tf = torch.from_numpy(F).to(dev)
    rays = torch.from_numpy(w["rays"]).to(dev)
    rays_any = torch.from_numpy(w["rays_any"]).to(dev)
    n = rays.shape[0]
    # The process's first build also loads the build kernels' code objects (one-time
    # cost); the build leg reports the second, warm build and the first one beside it.
    build_kw = dict(node_width=args.width, treelet_size=args.treelet, node_format=args.format,
                    leaf_size=args.leaf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bvh = gpurt.Bvh(ctx, tv, tf, **build_kw)
    build_cold_ms = (time.perf_counter() - t0) * 1e3
    bvh.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bvh = gpurt.Bvh(ctx, tv, tf, **build_kw)
    build_ms = (time.perf_counter() - t0) * 1e3
    info = bvh.info()
    log(f"[bench] GPU build {build_ms:.1f} ms (first build of the process {build_cold_ms:.1f} ms), {info}")
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    uv = torch.empty((n, 2), dtype=torch.float32, device=dev)
    occ = torch.empty(n, dtype=torch.uint8, device=dev)

    def step_serial():
        bvh.trace_closest(rays, out=(prim, t, uv), stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger)
        bvh.trace_any(rays_any, out=occ, stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger)

    # The two passes are independent, so a step issues the any-hit pass on a second
    # stream beside the closest-hit pass (the way the reference's ray-generation shader
    # keeps both ray kinds in flight) and joins before the next step; the end of one
    # launch, where few lanes are busy, then overlaps the other's work.
    # Each concurrent pass launches 16 waves per CU (half of a CU's 32 wave slots) so that
    # both are resident together, and refills at 16 idle lanes; alone, a pass keeps the
    # library defaults.
    s_any = torch.cuda.Stream(dev)
    waves_overlap = args.waves or 16
    refill_overlap = args.refill_trigger or 16

    def step_overlapped():
        fork = torch.cuda.Event()
        fork.record()
        bvh.trace_closest(rays, out=(prim, t, uv), stack_depth=args.stack, waves_per_cu=waves_overlap, leaf_trigger=args.leaf_trigger, refill_trigger=refill_overlap)
        with torch.cuda.stream(s_any):
            s_any.wait_event(fork)
            bvh.trace_any(rays_any, out=occ, stack_depth=args.stack, waves_per_cu=waves_overlap, leaf_trigger=args.leaf_trigger, refill_trigger=refill_overlap)
            join = torch.cuda.Event()
            join.record(s_any)
        torch.cuda.current_stream().wait_event(join)

    ms_serial = timed(step_serial, args.steps, args.warmup)
    ms = ms_serial if args.serial else timed(step_overlapped, args.steps, args.warmup)
    ms_closest = timed(lambda: bvh.trace_closest(rays, out=(prim, t, uv), stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger),
                       args.steps, 1)
    ms_any = timed(lambda: bvh.trace_any(rays_any, out=occ, stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger), args.steps, 1)
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    bvh.trace_closest(rays, out=(prim, t, uv), stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger, stats=st)
    torch.cuda.synchronize()
    nodes, tris, spills, _ = st.cpu().tolist()
    bytes_closest = ray_bytes(n, nodes, tris, info["node_bytes"])
    st.zero_()
    bvh.trace_any(rays_any, out=occ, stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger, stats=st)
    torch.cuda.synchronize()
    nodes_a, tris_a, spills_a, _ = st.cpu().tolist()
    hit_rate = float((prim >= 0).float().mean().item())
    occ_rate = float(occ.float().mean().item())

    # parity of every ray against the oracle's CPU LBVH (not timed)
    cpu_bvh_t0 = time.perf_counter()
    cbvh = O.CpuBvh(V, F, 4)
    cpu_build_s = time.perf_counter() - cpu_bvh_t0
    ep, et, euv = cbvh.trace_closest(w["rays"])
    eo = cbvh.trace_any(w["rays_any"])
    parity = bool(np.array_equal(prim.cpu().numpy(), ep) and np.array_equal(t.cpu().numpy(), et)
                  and np.array_equal(uv.cpu().numpy(), euv) and np.array_equal(occ.cpu().numpy(), eo))
    fp64 = fp64_rays(O, V, F, w["rays"], prim.cpu().numpy(), t.cpu().numpy(), 1024)

    # CPU baseline: the oracle LBVH on every host core, full passes over the same rays
    # (closest hit + any hit) repeated until about --cpu-seconds of CPU work
    threads = O.num_threads()
    passes, cpu_s = 0, 0.0
    while cpu_s < args.cpu_seconds and passes < 1000:
        t0 = time.perf_counter()
        cbvh.trace_closest(w["rays"])
        cbvh.trace_any(w["rays_any"])
        cpu_s += time.perf_counter() - t0
        passes += 1
    cpu_rate = 2 * n * passes / cpu_s
    # the achievable HBM rate on this box: a device copy of 1 GiB (read + write)
    src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    copy_ms = timed(lambda: dst.copy_(src), 10, 2)
    copy_gbs = 2 * src.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del src, dst
    rl = roof(bytes_closest, ms_closest, roof_kernel(args, info, "closest"))
    rl.update({"nodes_per_ray": nodes / n, "tris_per_ray": tris / n, "stack_spills_per_ray": spills / n})
    rl["copy_achievable_GBps"] = copy_gbs
    rl["frac_of_copy"] = rl["achieved"] / copy_gbs
    rl["traffic"] = measured_traffic(args, int(F.shape[0]), n, info["node_format"])
    build_roof = roof(build_bytes(info), build_ms, "BVH build, all kernels (host-timed warm call)")
    res = {
        "value": 2 * n / (ms * 1e-3),
        "ms_per_step": ms,
        "config": {"workload": "rays_scene", "n_tris": int(F.shape[0]), "n_rays": n,
                   "passes": ["closest_hit", "any_hit"], "streams": 1 if args.serial else 2,
                   "node_width": args.width,
                   "leaf_size": args.leaf, "stack_depth": args.stack, "treelet_size": args.treelet,
                   "node_format": info["node_format"], "waves_per_cu": args.waves,
                   "leaf_trigger": args.leaf_trigger, "refill_trigger": args.refill_trigger},
        "roofline": rl,
        "cpu_baseline": {"value": cpu_rate, "unit": "rays/s", "cores": threads, "kind": "port",
                         "sample": f"{passes} full passes of the {n} closest-hit + {n} any-hit "
                                   f"rays of rays_scene on the oracle's C LBVH (leaf 4, "
                                   f"{threads} threads); its single-thread build "
                                   f"({cpu_build_s:.1f}s) is not included",
                         "seconds": cpu_s, "build_s": cpu_build_s},
        "extra_rays": {"step": "closest-hit and any-hit passes one after the other" if args.serial else
                       "any-hit pass on a second stream beside the closest-hit pass, joined every step; "
                       "each pass at %d waves per CU, refill trigger %d" % (waves_overlap, refill_overlap),
                       "serial_rays_per_s": 2 * n / (ms_serial * 1e-3), "ms_per_step_serial": ms_serial,
                       "closest_rays_per_s": n / (ms_closest * 1e-3),
                       "any_rays_per_s": n / (ms_any * 1e-3), "gpu_build_ms": build_ms,
                       "gpu_build_first_call_ms": build_cold_ms,
                       "build_roofline": build_roof,
                       "hit_rate": hit_rate, "occluded_rate": occ_rate,
                       "any_nodes_per_ray": nodes_a / n, "any_tris_per_ray": tris_a / n,
                       "any_roofline": roof(ray_bytes(n, nodes_a, tris_a, info["node_bytes"], any_hit=True),
                                            ms_any, roof_kernel(args, info, "any")),
                       "bvh": info, "parity_all_rays_vs_cpu_lbvh": parity, "fp64_check": fp64},
    }
    del bvh
    return res


def cpq_bench(name, w, args, torch, gpurt, dev, ctx, O=None, fp64_every=0, parity_every=0):
    V, F, Q = w["verts"], w["tris"], w["queries"]
    tv = torch.from_numpy(V).to(dev)
    tf = torch.from_numpy(F).to(dev)
    q = torch.from_numpy(Q).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bvh = gpurt.Bvh(ctx, tv, tf, node_width=args.width, treelet_size=args.treelet,
                    node_format=args.format, leaf_size=args.leaf)
    build_ms = (time.perf_counter() - t0) * 1e3
    n = q.shape[0]
    out = (torch.empty(n, dtype=torch.int32, device=dev), torch.empty((n, 3), device=dev),
           torch.empty(n, device=dev), torch.empty((n, 3), device=dev))
    ms = timed(lambda: bvh.closest_point(q, out=out, stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger), args.steps, 2)
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    bvh.closest_point(q, out=out, stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, stats=st)
    torch.cuda.synchronize()
    nodes, tris, _, _ = st.cpu().tolist()
    info = bvh.info()
    byts = cpq_bytes(n, nodes, tris, info["cpq_node_bytes"])
    prim = out[0].cpu().numpy()
    del bvh
    parity = None
    if parity_every:
        # every parity_every-th query against the oracle's CPU LBVH (equal to brute force)
        idx = np.arange(0, n, parity_every)
        cb = O.CpuBvh(V, F, 4)
        ep, ept, ed, eb = cb.closest_point(Q[idx])
        parity = {"sample": int(idx.size),
                  "bitexact_vs_cpu_lbvh": bool(np.array_equal(prim[idx], ep) and
                                               np.array_equal(out[2].cpu().numpy()[idx], ed) and
                                               np.array_equal(out[1].cpu().numpy()[idx], ept) and
                                               np.array_equal(out[3].cpu().numpy()[idx], eb))}
    fp64 = fp64_cpq(O, V, F, Q, prim, out[2].cpu().numpy(), fp64_every) if fp64_every else None
    return {"workload": name, "n_tris": int(F.shape[0]), "n_queries": n,
            "queries_per_s": n / (ms * 1e-3), "ms": ms, "gpu_build_ms": build_ms,
            "nodes_per_query": nodes / n, "tris_per_query": tris / n,
            "roofline": roof(byts, ms, roof_kernel(args, info, "cpq")),
            "parity": parity, "fp64_check": fp64}


def diff_bench(args, torch, gpurt, W, dev, ctx):
    w = W.diff()
    V, F = w["verts"], w["tris"]
    tv = torch.from_numpy(V).to(dev)
    tf = torch.from_numpy(F).to(dev)
    q0 = torch.from_numpy(w["queries"]).to(dev)
    o0 = torch.from_numpy(w["ray_o"]).to(dev)
    d0 = torch.from_numpy(w["ray_d"]).to(dev)
    bvh = gpurt.Bvh(ctx, tv, tf, node_width=args.width, treelet_size=args.treelet,
                    node_format=args.format, leaf_size=args.leaf)
    nq, nr = q0.shape[0], o0.shape[0]

    def fwd():
        with torch.no_grad():
            dist, _ = gpurt.closest_point_distance(bvh, q0, tv)
            t, _ = gpurt.ray_hit_t(bvh, o0, d0, tv)
        return dist, t

    def fwd_bwd():
        q = q0.detach().requires_grad_(True)
        o = o0.detach().requires_grad_(True)
        v = tv.detach().requires_grad_(True)
        dist, _ = gpurt.closest_point_distance(bvh, q, v)
        t, _ = gpurt.ray_hit_t(bvh, o, d0, v)
        loss = dist.sum() + torch.where(torch.isfinite(t), t, torch.zeros_like(t)).sum()
        loss.backward()
        return loss

    ms_f = timed(fwd, args.steps, 2)
    ms_fb = timed(fwd_bwd, args.steps, 2)
    # the backward launches alone, on saved forward outputs
    prim, pt, dist, bary = bvh.closest_point(q0)
    rays = gpurt.make_rays(o0, d0)
    rp, rt, ruv = bvh.trace_closest(rays)
    g1 = torch.ones(nq, device=dev)
    g2 = torch.ones(nr, device=dev)
    gq = torch.empty_like(q0)
    go = torch.empty((nr, 3), device=dev)
    gv = torch.zeros((tv.shape[0], 3), device=dev)
    # the backward launches in each vertex-sum mode: FP32 atomics (9 adds per query off
    # the surface and per hitting ray), the fixed order (sort by vertex, segmented sums),
    # and the default, which picks the faster for the call's size
    bwd = {}
    for mode in ("atomic", "fixed", "auto"):
        ctx.vertex_sums = mode
        bwd[mode] = {"cpq_ms": timed(lambda: bvh.closest_point_bwd(q0, prim, bary, pt, dist, g1, gq, gv), args.steps, 2),
                     "ray_ms": timed(lambda: bvh.trace_bwd(rays, rp, rt, ruv, g2, go, None, gv), args.steps, 2)}
    ctx.vertex_sums = "auto"
    ms_cb, ms_rb = bwd["auto"]["cpq_ms"], bwd["auto"]["ray_ms"]
    n_cq_atomics = 9 * int((dist > 0).sum())
    n_r_atomics = 9 * int((rp >= 0).sum())
    # visits of the forward launches (counting variants, untimed) for their bytes
    info = bvh.info()
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    bvh.closest_point(q0, stats=st)
    torch.cuda.synchronize()
    cn, ct = st[0].item(), st[1].item()
    st.zero_()
    bvh.trace_closest(rays, stats=st)
    torch.cuda.synchronize()
    rn, rt_ = st[0].item(), st[1].item()
    fwd_bytes = cpq_bytes(nq, cn, ct, info["cpq_node_bytes"]) + ray_bytes(nr, rn, rt_, info["node_bytes"])
    del bvh
    return {"workload": "diff", "n_tris": int(F.shape[0]), "n_queries": nq, "n_rays": nr,
            "fwd_units_per_s": (nq + nr) / (ms_f * 1e-3), "fwd_ms": ms_f,
            "fwd_bwd_units_per_s": (nq + nr) / (ms_fb * 1e-3), "fwd_bwd_ms": ms_fb,
            "fwd_bwd_over_fwd": ms_fb / ms_f,
            "fwd_roofline": roof(fwd_bytes, ms_f, "forward: closest-point + closest-hit launches (autograd calls)"),
            "fwd_bwd_roofline": roof(fwd_bytes + nq * CPQ_BWD_BYTES + nr * RAY_BWD_BYTES, ms_fb,
                                     "forward + backward (autograd calls)"),
            "cpq_bwd_roofline": roof(nq * CPQ_BWD_BYTES, ms_cb, "closest-point backward (default vertex sums)"),
            "ray_bwd_roofline": roof(nr * RAY_BWD_BYTES, ms_rb, "hit-distance backward (default vertex sums)"),
            "bwd_ms_by_vertex_sums": bwd,
            "cpq_bwd_atomic_rate_per_s": n_cq_atomics / (bwd["atomic"]["cpq_ms"] * 1e-3),
            "ray_bwd_atomic_rate_per_s": n_r_atomics / (bwd["atomic"]["ray_ms"] * 1e-3)}


def batched_bench(args, torch, gpurt, W, dev, ctx):
    """256 meshes (log-uniform 1K..524K triangles), each with its own closest-point
    queries and rays: one batch replay (all builds + the persistent query launches, a
    HIP graph) against the plain path (per mesh: build, closest-point launch, closest-hit
    launch), both rebuilding every BVH each step; then the same with the backward pass
    (batched backward graph vs the plain per-mesh backward launches)."""
    t0 = time.time()
    w = W.batched()
    log(f"[bench] batched generated in {time.time() - t0:.1f}s: {w['meta']}")
    meshes = [{k: torch.from_numpy(v).to(dev) for k, v in m.items()} for m in w["meshes"]]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch = gpurt.Batch(ctx, meshes, node_width=args.width, stack_depth=args.stack,
                        waves_per_cu=args.waves, leaf_size=args.leaf, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger,
                        treelet_size=args.batch_treelet, node_format=args.format)
    create_s = time.perf_counter() - t0
    ms_b = timed(batch.run, max(3, args.steps // 4), 1)
    res = batch.run()

    plain_out = {}

    def plain():
        for i, m in enumerate(meshes):
            bvh = gpurt.Bvh(ctx, m["verts"], m["tris"], node_width=args.width, leaf_size=args.leaf,
                            treelet_size=args.batch_treelet, node_format=args.format)
            p, _, d, _ = bvh.closest_point(m["queries"], stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger)
            rp, rt, _ = bvh.trace_closest(m["rays"], stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger)
            plain_out[i] = (p, d, rp, rt)
            bvh.close()

    ms_p = timed(plain, 2, 1)
    torch.cuda.synchronize()
    parity = all(torch.equal(plain_out[i][0], r["q_prim"]) and torch.equal(plain_out[i][1], r["q_dist"])
                 and torch.equal(plain_out[i][2], r["r_prim"]) and torch.equal(plain_out[i][3], r["r_t"])
                 for i, r in enumerate(res))

    # forward + backward: gradients w.r.t. query points, ray origins/directions and all
    # vertices (distance + hit-t terms), batched graph vs the plain per-mesh path
    grads = batch.enable_backward()

    def batched_fb():
        batch.run()
        batch.backward()

    ms_bfb = timed(batched_fb, max(3, args.steps // 4), 1)
    plain_g = {}

    def plain_fb():
        for i, m in enumerate(meshes):
            bvh = gpurt.Bvh(ctx, m["verts"], m["tris"], node_width=args.width, leaf_size=args.leaf,
                            treelet_size=args.batch_treelet, node_format=args.format)
            p, pt, d, b = bvh.closest_point(m["queries"], stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger)
            rp, rt, ruv = bvh.trace_closest(m["rays"], stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger)
            g = grads[i]
            gq = torch.empty_like(m["queries"])
            go = torch.empty((rp.shape[0], 3), device=dev)
            gd = torch.empty((rp.shape[0], 3), device=dev)
            gv = torch.zeros((m["verts"].shape[0], 3), device=dev)
            bvh.closest_point_bwd(m["queries"], p, b, pt, d, g["g_dist"], gq, gv)
            bvh.trace_bwd(m["rays"], rp, rt, ruv, g["g_t"], go, gd, gv)
            plain_g[i] = (gq, go, gv)
            bvh.close()

    ms_pfb = timed(plain_fb, 2, 1)
    torch.cuda.synchronize()
    parity_bwd = all(torch.equal(plain_g[i][0], g["g_q"]) and torch.equal(plain_g[i][1], g["g_origin"])
                     and float((plain_g[i][2] - g["g_verts"]).abs().max()) <=
                     1e-5 * (float(plain_g[i][2].abs().max()) + 1e-6)
                     for i, g in enumerate(grads))
    # algorithmic bytes of one replay: every mesh's build (build_bytes) and its queries'
    # visits, counted by the counting variants on the same trees built the plain way
    # (untimed)
    replay_bytes = 0
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    for m in meshes:
        bvh = gpurt.Bvh(ctx, m["verts"], m["tris"], node_width=args.width, leaf_size=args.leaf,
                        treelet_size=args.batch_treelet, node_format=args.format)
        info = bvh.info()
        st.zero_()
        bvh.closest_point(m["queries"], stats=st)
        torch.cuda.synchronize()
        replay_bytes += cpq_bytes(m["queries"].shape[0], st[0].item(), st[1].item(), info["cpq_node_bytes"])
        st.zero_()
        bvh.trace_closest(m["rays"], stats=st)
        torch.cuda.synchronize()
        replay_bytes += ray_bytes(m["rays"].shape[0], st[0].item(), st[1].item(), info["node_bytes"])
        replay_bytes += build_bytes(info)
        bvh.close()
    # the batched backward with FP32 atomic vertex sums (the default picks the fixed order
    # for a batch this size)
    ctx.vertex_sums = "atomic"
    dgrads = batch.enable_backward()
    ctx.vertex_sums = "auto"
    ms_bfb_atomic = timed(batched_fb, max(3, args.steps // 4), 1)
    torch.cuda.synchronize()
    parity_det = all(torch.equal(a["g_q"], b["g_q"]) and
                     float((a["g_verts"] - b["g_verts"]).abs().max()) <= 1e-5 * (float(a["g_verts"].abs().max()) + 1e-6)
                     for a, b in zip(grads, dgrads))
    # the autotuner over the space the batched workload names (node width {2,4,8}, stack
    # depth {16,32,64}, waves per CU), untimed; then the tuned batch is timed like the
    # default one.  It runs after the plain path's timings: the tuner's many batch
    # creations and destructions leave the HIP runtime slow to service host calls for a
    # few hundred milliseconds, which would inflate the plain path's host-bound time.
    tuned = None
    if not args.no_tune:
        from gpurt import autotune
        t0 = time.perf_counter()
        tuned = autotune.tune_batch(ctx, meshes, use_cache=False, reps=3,
                                    space={"node_width": (2, 4, 8), "stack_depth": (16, 32, 64),
                                           "waves_per_cu": (8, 16, 24), "treelet_size": (args.batch_treelet,),
                                           "node_format": (args.format,)})
        tune_s = time.perf_counter() - t0
        tbatch = gpurt.Batch(ctx, meshes, leaf_size=args.leaf, **tuned["build"], **tuned["query"])
        ms_t = timed(tbatch.run, max(3, args.steps // 4), 1)
        tres = tbatch.run()
        torch.cuda.synchronize()
        parity_tuned = all(torch.equal(a["q_prim"], b["q_prim"]) and torch.equal(a["r_t"], b["r_t"])
                           for a, b in zip(tres, res))
        tbatch.close()
    nq = sum(m["queries"].shape[0] for m in meshes)
    nr = sum(m["rays"].shape[0] for m in meshes)
    batch.close()
    return {"workload": "batched", "n_meshes": len(meshes), "n_tris": w["meta"]["n_tris_total"],
            "n_queries": nq, "n_rays": nr,
            "batched_units_per_s": (nq + nr) / (ms_b * 1e-3), "batched_ms": ms_b,
            "plain_units_per_s": (nq + nr) / (ms_p * 1e-3), "plain_ms": ms_p,
            "batched_over_plain": ms_p / ms_b,
            "batched_roofline": roof(replay_bytes, ms_b, "batched replay: 256 builds + closest-point and closest-hit launches"),
            "fwd_bwd_batched_units_per_s": (nq + nr) / (ms_bfb * 1e-3), "fwd_bwd_batched_ms": ms_bfb,
            "fwd_bwd_batched_roofline": roof(replay_bytes + nq * CPQ_BWD_BYTES + nr * RAY_BWD_BYTES, ms_bfb,
                                             "batched forward + backward replays"),
            "fwd_bwd_plain_units_per_s": (nq + nr) / (ms_pfb * 1e-3), "fwd_bwd_plain_ms": ms_pfb,
            "fwd_bwd_batched_over_plain": ms_pfb / ms_bfb, "batch_create_s": create_s,
            "fwd_bwd_batched_atomic_sums_ms": ms_bfb_atomic, "parity_bwd_atomic_vs_fixed_sums": bool(parity_det),
            "autotuned": None if tuned is None else {
                "build": tuned["build"], "query": tuned["query"], "candidates": tuned["candidates"],
                "tuner_s": tune_s, "tuner_default_ms": tuned["default_ms"], "tuner_best_ms": tuned["objective_ms"],
                "batched_ms": ms_t, "batched_units_per_s": (nq + nr) / (ms_t * 1e-3),
                "batched_over_plain": ms_p / ms_t, "parity_vs_default": bool(parity_tuned)},
            "parity_batched_vs_plain": bool(parity), "parity_bwd_batched_vs_plain": bool(parity_bwd),
            "treelet_size": args.batch_treelet,
            "note": "each step rebuilds all 256 BVHs and answers every query; fwd+bwd adds the "
                    "gradients of sum(dist) and sum(t) w.r.t. queries, rays and vertices"}


def tlas_bench(args, torch, gpurt, W, O, dev, ctx):
    """rays_scene laid out as the reference builds it: a terrain BLAS and one icosphere
    BLAS placed by 65 instances (TLAS), the same 1,048,576 + 1,048,576 rays; closest and
    any hit through two levels, checked bit for bit against the oracle's instanced brute
    force on a sample of rays."""
    sc = W.instanced_scene()
    blas = [gpurt.Bvh(ctx, torch.from_numpy(V).to(dev), torch.from_numpy(F).to(dev), node_width=args.width,
                      treelet_size=args.treelet, node_format=args.format, leaf_size=args.leaf)
            for V, F in sc["meshes"]]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tlas = gpurt.Tlas(ctx, [(blas[m], sc["xf"][i].reshape(3, 4)) for i, m in enumerate(sc["mesh_of"])])
    build_ms = (time.perf_counter() - t0) * 1e3
    rays = torch.from_numpy(sc["rays"]).to(dev)
    rays_any = torch.from_numpy(sc["rays_any"]).to(dev)
    n = rays.shape[0]
    out = (torch.empty(n, dtype=torch.int32, device=dev), torch.empty(n, dtype=torch.int32, device=dev),
           torch.empty(n, device=dev), torch.empty((n, 2), device=dev))
    occ = torch.empty(n, dtype=torch.uint8, device=dev)
    q = dict(stack_depth=args.stack, waves_per_cu=args.waves, leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger)
    ms_c = timed(lambda: tlas.trace_closest(rays, out=out, **q), args.steps, 1)
    ms_a = timed(lambda: tlas.trace_any(rays_any, out=occ, **q), args.steps, 1)
    st = torch.zeros(4, dtype=torch.int64, device=dev)
    tlas.trace_closest(rays, out=out, stats=st, **q)
    torch.cuda.synchronize()
    nodes, tris, _, _ = st.cpu().tolist()
    idx = np.arange(0, n, 4096)
    eo, ep, et, euv = O.inst_ray_brute_f32(sc["xf"], sc["mesh_of"], sc["meshes"], sc["rays"][idx])
    go, gp, gt, guv = (x.cpu().numpy()[idx] for x in out)
    parity = bool(np.array_equal(go, eo) and np.array_equal(gp, ep) and np.array_equal(gt, et)
                  and np.array_equal(guv, euv))
    tlas.close()
    return {"workload": "instanced_scene", **{k: v for k, v in sc["meta"].items() if k != "workload"},
            "closest_rays_per_s": n / (ms_c * 1e-3), "any_rays_per_s": n / (ms_a * 1e-3),
            "closest_ms": ms_c, "any_ms": ms_a, "tlas_build_ms": build_ms,
            "nodes_per_ray": nodes / n, "tris_per_ray": tris / n,
            "parity_sample": int(idx.size), "parity_vs_instanced_brute_force": parity}


def tracer_bench(args, torch, gpurt, W, O, dev, ctx):
    """The wavefront path tracer on a closed Cornell-style box (two 81,920-triangle
    spheres, a block, five walls, one area light; 163,878 triangles over 9 BLASes) at
    1024 x 1024 with the reference's defaults (8 samples per pixel per frame, max depth 8,
    Russian roulette, Blinn-Phong), for direct light sampling and MIS.  Paths/s counts
    camera paths; rays/s every ray traced (bounce rays through the TLAS plus shadow / light
    rays).  Parity: a 32 x 32 render against the oracle's per-pixel restatement (fraction of
    pixels within 1e-4 relative).  CPU: the oracle on a 64 x 64, 2-sample sample."""
    from gpurt import tracer as T
    sc = W.tracer_box(sphere_level=6)
    scene = T.Scene(ctx, sc["objects"], node_width=args.width)
    cam = sc["camera"]
    w = h = 1024
    iv, ip = T.camera(cam["eye"], cam["target"], yfov_deg=cam["yfov_deg"], aspect=1.0)
    res = {"workload": "tracer_box", "triangles": int(sum(len(o["indices"]) for o in sc["objects"])),
           "width": w, "height": h, "samples": 8, "max_depth": 8}
    for name, integ in (("direct", 0), ("mis", 2)):
        kw = dict(samples=8, max_depth=8, integrator=integ, brdf=0, use_rr=True)
        img = torch.zeros((h, w, 4), dtype=torch.float32, device=dev)
        counts = np.zeros(2, np.uint64)
        scene.render(w, h, iv, ip, image=img, seed=1, ray_counts=counts, **kw)
        steps = max(2, args.steps // 4)
        ms = timed(lambda: scene.render(w, h, iv, ip, image=img, seed=1, **kw), steps, 1)
        rays = int(counts.sum())
        small = dict(width=32, height=32, inv_view=iv, inv_proj=ip, clear_col=(0.3,) * 3, env_light=(0.0,) * 3,
                     frame=0, samples=2, max_frame=256, qmc=0, max_depth=8, use_metalness=0, integrator=integ, brdf=0,
                     use_rr=1, seed=5)
        g = scene.render(32, 32, iv, ip, samples=2, max_depth=8, integrator=integ, seed=5)
        torch.cuda.synchronize()
        ref = O.render(sc["objects"], small)
        g = g.cpu().numpy()[..., :3]
        agree = float((np.abs(g - ref[..., :3]) <= 1e-4 * (1 + np.abs(ref[..., :3]))).all(-1).mean())
        cpu = dict(small, width=64, height=64)
        t0 = time.perf_counter()
        O.render(sc["objects"], cpu)
        cpu_s = time.perf_counter() - t0
        res[name] = {"ms_per_frame": ms, "paths_per_s": w * h * 8 / (ms * 1e-3),
                     "rays_per_frame": rays, "rays_per_s": rays / (ms * 1e-3),
                     "parity_pixels_agreeing_32x32": agree,
                     "cpu_oracle_paths_per_s": 64 * 64 * 2 / cpu_s, "cpu_threads": O.load().orc_num_threads()}
    scene.close()
    return res


def profile_mode(args, torch, gpurt, W, dev, ctx):
    """Short run for hardware-counter passes: `--profile-reps` launches of the kernel
    `--profile-workload` names (closest-hit rays on rays_scene, or closest-point queries
    on cpq_small / cpq_large), then the same number of elementwise adds over 256 MiB (a
    plain streaming kernel that reads 256 MiB and writes 256 MiB, whose
    FETCH_SIZE/WRITE_SIZE calibrate the counters)."""
    kind = args.profile_workload
    w = W.rays_scene() if kind == "rays_scene" else W.diff() if kind.startswith("diff_bwd") else getattr(W, kind)()
    if kind.startswith("diff_bwd"):
        ctx.vertex_sums = "fixed" if kind == "diff_bwd_fixed" else "atomic"
    tv = torch.from_numpy(w["verts"]).to(dev)
    tf = torch.from_numpy(w["tris"]).to(dev)
    bvh = gpurt.Bvh(ctx, tv, tf, node_width=args.width, treelet_size=args.treelet,
                    node_format=args.format, leaf_size=args.leaf)
    if kind.startswith("diff_bwd"):
        q = torch.from_numpy(w["queries"]).to(dev)
        rays = gpurt.make_rays(torch.from_numpy(w["ray_o"]).to(dev), torch.from_numpy(w["ray_d"]).to(dev))
        n = q.shape[0] + rays.shape[0]
        prim, pt, dist, bary = bvh.closest_point(q)
        rp, rt, ruv = bvh.trace_closest(rays)
        g1, g2 = torch.ones(q.shape[0], device=dev), torch.ones(rays.shape[0], device=dev)
        gq, go = torch.empty_like(q), torch.empty((rays.shape[0], 3), device=dev)
        gv = torch.zeros((tv.shape[0], 3), device=dev)
        for _ in range(args.profile_reps):
            bvh.closest_point_bwd(q, prim, bary, pt, dist, g1, gq, gv)
            bvh.trace_bwd(rays, rp, rt, ruv, g2, go, None, gv)
    elif kind == "rays_scene":
        rays = torch.from_numpy(w["rays"]).to(dev)
        n = rays.shape[0]
        out = (torch.empty(n, dtype=torch.int32, device=dev), torch.empty(n, device=dev),
               torch.empty((n, 2), device=dev))
        for _ in range(args.profile_reps):
            bvh.trace_closest(rays, out=out, stack_depth=args.stack, waves_per_cu=args.waves,
                              leaf_trigger=args.leaf_trigger, refill_trigger=args.refill_trigger)
    else:
        q = torch.from_numpy(w["queries"]).to(dev)
        n = q.shape[0]
        out = (torch.empty(n, dtype=torch.int32, device=dev), torch.empty((n, 3), device=dev),
               torch.empty(n, device=dev), torch.empty((n, 3), device=dev))
        for _ in range(args.profile_reps):
            bvh.closest_point(q, out=out, stack_depth=args.stack, waves_per_cu=args.waves,
                              leaf_trigger=args.leaf_trigger)
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(args.profile_reps):
        torch.add(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"profile": kind, "launches": args.profile_reps,
                      "units": n, "calib_bytes_read": src.numel() * 4,
                      "calib_bytes_written": dst.numel() * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=4, choices=[2, 4, 8])
    ap.add_argument("--stack", type=int, default=0, choices=[0, 8, 16, 32, 64],
                    help="LDS stack entries per lane (0 = library default: 8 for rays, 16 for "
                         "closest points)")
    ap.add_argument("--leaf", type=int, default=4, choices=range(1, 9), metavar="1..8",
                    help="max triangles per leaf")
    ap.add_argument("--waves", type=int, default=0, help="resident waves per CU (0 = library choice)")
    ap.add_argument("--leaf-trigger", type=int, default=0,
                    help="lanes (of 64) with a postponed leaf that switch a wave to leaf tests "
                         "(0 = library default)")
    ap.add_argument("--serial", action="store_true",
                    help="run the headline step's two passes one after the other on one stream")
    ap.add_argument("--refill-trigger", type=int, default=0,
                    help="idle lanes (of 64) at which a ray wave writes its finished rays and "
                         "fetches new ones (0 = library default)")
    ap.add_argument("--treelet", type=int, default=7, choices=[0, 5, 6, 7],
                    help="SAH treelet restructuring of the single-mesh BVHs (leaves per "
                         "treelet), 0 = off; the batched leg rebuilds every step and uses "
                         "--batch-treelet")
    ap.add_argument("--batch-treelet", type=int, default=0, choices=[0, 5, 6, 7])
    ap.add_argument("--format", type=int, default=0, choices=[0, 1, 2],
                    help="node encoding: 0 = library default, 1 = FP32 child boxes, "
                         "2 = 8-bit child boxes")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: repeat full passes until this much CPU time")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 5.2M-triangle clustered closest-point workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-mesh batched leg")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched-vs-plain leg (for profiling it)")
    ap.add_argument("--only-tracer", action="store_true", help="print only the path-tracer leg")
    ap.add_argument("--profile", action="store_true",
                    help="only run closest-hit launches + calibration copies (for rocprofv3)")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--profile-workload", default="rays_scene",
                    choices=["rays_scene", "cpq_small", "cpq_large", "diff_bwd", "diff_bwd_fixed"],
                    help="kernel of the --profile run: closest-hit rays, closest-point queries, or the "
                         "backward passes of the differentiable workload (FP32 atomic or "
                         "fixed-order vertex sums)")
    ap.add_argument("--no-tune", action="store_true", help="batched leg: skip the autotuner sweep")
    ap.add_argument("--only", choices=["diff", "cpq_large", "batched", "tracer"],
                    help="print only this leg (for profiling it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import gpurt
    from gpurt import workloads as W
    from oracle import oracle as O
    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda", 0)
    ctx = gpurt.Context(0)
    if args.profile:
        profile_mode(args, torch, gpurt, W, dev, ctx)
        return
    if args.only_batched or args.only == "batched":
        print(json.dumps(batched_bench(args, torch, gpurt, W, dev, ctx)), flush=True)
        return
    if args.only == "diff":
        print(json.dumps(diff_bench(args, torch, gpurt, W, dev, ctx)), flush=True)
        return
    if args.only == "cpq_large":
        print(json.dumps(cpq_bench("cpq_large", W.cpq_large(), args, torch, gpurt, dev, ctx, O, fp64_every=1024,
                                   parity_every=64)), flush=True)
        return
    if args.only_tracer or args.only == "tracer":
        print(json.dumps(tracer_bench(args, torch, gpurt, W, O, dev, ctx)), flush=True)
        return
    r = ray_bench(args, torch, gpurt, W, O, dev, ctx)
    extra = {"rays": r.pop("extra_rays")}
    if not args.no_extra:
        log("[bench] cpq_small")
        extra["cpq_small"] = cpq_bench("cpq_small", W.cpq_small(), args, torch, gpurt, dev, ctx, O, 1)
        log("[bench] instanced scene (TLAS)")
        extra["tlas"] = tlas_bench(args, torch, gpurt, W, O, dev, ctx)
        log("[bench] wavefront path tracer")
        extra["tracer"] = tracer_bench(args, torch, gpurt, W, O, dev, ctx)
        log("[bench] diff fwd / fwd+bwd")
        extra["diff"] = diff_bench(args, torch, gpurt, W, dev, ctx)
        if not args.no_batched:
            log("[bench] batched vs plain")
            extra["batched"] = batched_bench(args, torch, gpurt, W, dev, ctx)
        if not args.no_large:
            log("[bench] cpq_large")
            extra["cpq_large"] = cpq_bench("cpq_large", W.cpq_large(), args, torch, gpurt, dev, ctx, O,
                                           fp64_every=1024, parity_every=64)
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": r["config"],
        "roofline": r["roofline"],
        "cpu_baseline": r["cpu_baseline"],
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
