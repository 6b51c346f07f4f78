#!/usr/bin/env python3
"""Benchmark of the volume integrator on one MI355X.

Workload (default "c2"): 256^3 anisotropic SGGX fibre volume (seeded 4-octave
value noise in [0,8], von Mises-Fisher fibre field, albedo 0.9) under a
procedural 512x256 sky + sun environment (1e4:1), 1024x1024, 64 spp per step,
max depth 16, FP32.  One step = one r2_render call: the next 64 samples of the
frame (progressive accumulation), so the K timed steps together form a
K*64-spp frame whose relMSE against the FP64 4096-spp oracle (1/64 strided
pixel subset, committed under tests/golden/) is reported.

Prints ONE JSON line.  Inputs (grid, env) and outputs (per-step frames) stay in
HBM during the timed region.  The roofline entry uses the integrator kernel's
HIP-event duration measured inside the library on its own stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "M paths/sec at 1024² on 256³ SGGX grid @relMSE≤1e-3; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per lookup kind: 32-byte voxel record (24 B packed),
# 4-byte majorant, brick density = 4 B indirection + 4 B range + 1 B atlas,
# env = 2 CDF reads (8 B) + 4 RGB texels (48 B); +12 B framebuffer per path
B_VOX_FP32, B_VOX_PACKED, B_MAJ, B_BRICK, B_ENV, B_FB = 32, 24, 4, 9, 56, 12
GOLDEN = ROOT / "tests" / "golden"


def relmse(x, r):
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def host_cpus() -> dict:
    """The host cores this process may use: nproc, the affinity mask, the cgroup CPU
    quota (cpu.max) when one is set, and the CPU model.  The CPU baseline runs one
    OpenMP thread per usable core: the affinity count, capped by the quota."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cgroup_quota_cpus": None,
            "model": None, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        q, per = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            info["cgroup_quota_cpus"] = round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                info["model"] = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    n = info["affinity"]
    if info["cgroup_quota_cpus"]:
        n = min(n, max(1, int(info["cgroup_quota_cpus"] + 0.5)))
    info["threads_used"] = n
    return info


def cpu_baseline(w, threads: int, stride: int, spp: int):
    """Time the FP32 oracle (CPU restatement, OpenMP) on a strided pixel subset."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O  # the CPU oracle: checker and baseline only

    cfg = dict(w["render"])
    W, H = cfg["width"], cfg["height"]
    ys, xs = np.meshgrid(np.arange(0, H, stride), np.arange(0, W, stride), indexing="ij")
    pix = (ys * W + xs).reshape(-1).astype(np.int32)
    P = np.repeat(pix, spp)
    S = np.tile(np.arange(spp, dtype=np.int32), len(pix))
    sc = O.Scene(w["grids"], w["instances"], w["env"], "f32", surfaces=cfg.get("surfaces", ()))
    t0 = time.perf_counter()
    _, cnt = sc.render_paths(cfg, P, S, threads=threads, counters=True)
    dt = time.perf_counter() - t0
    mean_cnt = cnt.astype(np.float64).mean(axis=0)
    return len(P) / dt / 1e6, dt, len(P), dict(zip(O.CNT_NAMES, mean_cnt.tolist()))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--no-serving", action="store_true", help="skip the render-server latency probe")
    ap.add_argument("--surfaces", action="store_true",
                    help="render workloads: the volume on the forest scenes' ground circle with a glossy sphere")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=None,
                    help="pixel stride of the CPU baseline sample (default: 1 for c1, 2 otherwise)")
    ap.add_argument("--cpu-spp", type=int, default=None,
                    help="samples per pixel of the CPU baseline sample (default: the workload's, at most 64)")
    ap.add_argument("--quality-spp", type=int, default=2048,
                    help="samples per pixel of the frame whose relMSE against the FP64 oracle is reported; "
                         "rendered after the timed region (the timed steps' samples count towards it)")
    ap.add_argument("--no-brick-grid", action="store_true")
    ap.add_argument("--precision", choices=["fp32", "packed"], default=None,
                    help="voxel record precision (default: the workload's own)")
    ap.add_argument("--requests", type=int, default=80, help="requests of the serving workload (--workload c5)")
    ap.add_argument("--serving-requests", type=int, default=40,
                    help="requests of the serving block of the render workloads' line")
    ap.add_argument("--rate", type=float, default=5.0, help="Poisson arrival rate (requests/s) of the serving workload")
    ap.add_argument("--cpu-voxels", type=int, default=1_500_000,
                    help="random voxels per LOD of the filter CPU baseline sample")
    ap.add_argument("--render-steps", type=int, default=2,
                    help="c3: 1024^2 x 64 spp renders of each filtered LOD (0: skip)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="MiB copied device-to-device before rendering (reference kernel for PMC byte counters)")
    a = ap.parse_args(argv)

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(rendering would shard by pixel tiles or sample ranges with one framebuffer reduce)",
              file=sys.stderr)
        return 2

    if a.workload == "c3":
        return bench_filter(a)
    if a.workload == "c5":
        return bench_serving(a)

    import renderer2_amd as r2
    from renderer2_amd import _hip, scenes

    kw = {"precision": a.precision} if a.precision else {}
    w = scenes.mixed_workload(a.workload, **kw) if a.surfaces else scenes.make_workload(a.workload, seed=1, **kw)
    cfg = dict(w["render"], use_brick_grid=not a.no_brick_grid)
    W, H, spp = cfg["width"], cfg["height"], cfg["spp"]
    ctx = r2.Context(0)
    grids = [ctx.upload_grid(g) for g in w["grids"]]
    env = ctx.upload_env(w["env"])
    frame_bytes = W * H * 3 * 4
    bufs = [_hip.DeviceBuffer(frame_bytes) for _ in range(a.steps)]
    warm = _hip.DeviceBuffer(frame_bytes)

    copy_calibration(a.copy_calib)
    for k in range(a.warmup):
        ctx.render_device(dict(cfg, sample_begin=1_000_000 + k * spp), grids, w["instances"], env, warm.addr)
    _hip.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        ctx.render_device(dict(cfg, sample_begin=k * spp), grids, w["instances"], env, bufs[k].addr)
        kernel_ms.append(ctx.stats()["kernel_ms"])
    _hip.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    paths_per_step = W * H * spp
    value = paths_per_step * a.steps / elapsed / 1e6
    avg_kernel_ms = float(np.mean(kernel_ms))

    # quality frame: the timed steps' frames plus further launches of the following
    # samples (outside the timed region) up to --quality-spp samples per pixel; each
    # frame is already divided by its own spp, so the frame is their mean
    quality = None
    gold = GOLDEN / f"{a.workload}_oracle_fp64.npz"
    if gold.exists() and not a.surfaces:
        z = np.load(gold, allow_pickle=False)
        st = int(z["stride"])
        ref = z["image"].astype(np.float64)
        n_q = max(a.steps, -(-a.quality_spp // spp))
        acc = np.zeros(ref.shape, np.float64)
        curve = {}
        for k in range(n_q):
            if k < a.steps:
                fr = bufs[k].to_host(np.float32, (H, W, 3))
            else:
                ctx.render_device(dict(cfg, sample_begin=k * spp), grids, w["instances"], env, warm.addr)
                fr = warm.to_host(np.float32, (H, W, 3))
            acc += fr[::st, ::st]
            if ((k + 1) & k) == 0 or k + 1 == n_q:  # 1, 2, 4, ... launches and the last
                curve[(k + 1) * spp] = relmse(acc / (k + 1), ref)
        from renderer2_amd.flip import flip_hdr

        img = acc / n_q
        reach = [n for n, e in sorted(curve.items()) if e <= 1e-3]
        quality = dict(relmse=relmse(img, ref), flip=round(flip_hdr(z["image"], img), 5), spp=n_q * spp,
                       oracle_spp=int(z["spp"]),
                       oracle="FP64 CPU oracle, %s" % ("full frame" if st == 1 else "1/%d strided pixel subset" % (st * st)),
                       relmse_by_spp={str(k): round(v, 7) for k, v in sorted(curve.items())},
                       spp_to_relmse_1e3=reach[0] if reach else None,
                       timed_frame_spp=a.steps * spp,
                       relmse_timed_frame=curve.get(a.steps * spp))

    serving = None
    if not a.no_serving:
        serving = serve_c5(ctx, a, n_req=a.serving_requests)

    cpu = None
    counts = None
    host = host_cpus()
    if not a.no_cpu_baseline:
        threads = host["threads_used"]
        stride = a.cpu_stride or (1 if a.workload == "c1" else 2)
        cspp = a.cpu_spp or min(spp, 64)
        v, dt, n, counts = cpu_baseline(w, threads, stride, cspp)
        cpu = {"value": round(v, 4), "unit": "M paths/s", "cores": threads, "kind": "port",
               "sample": f"FP32 oracle, {'full frame' if stride == 1 else f'1/{stride ** 2} strided pixel subset'} "
                         f"x {cspp} spp = {n} paths in {dt:.1f} s",
               "host": host}
    roof = None
    if counts is not None:
        b_vox = B_VOX_PACKED if w["grids"][0].get("precision") == "packed" else B_VOX_FP32
        b_path = (counts["vox"] * b_vox + counts["maj"] * B_MAJ + counts["brick"] * B_BRICK + counts["env"] * B_ENV
                  + B_FB)
        algo_bytes = b_path * paths_per_step
        achieved = algo_bytes / (avg_kernel_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "bytes_per_path": round(b_path, 1), "kernel": "k_paths_sm", "kernel_ms": round(avg_kernel_ms, 3)}
        if a.workload == "c2" and w["grids"][0].get("precision", "fp32") == "fp32" and not a.surfaces:
            roof["traffic"], roof["traffic_source"] = profile_traffic("k_paths_sm_c2_256cube_1024sq_64spp")
            roof["valu_issue"] = profile_valu_issue("k_paths_sm_c2_256cube_1024sq_64spp")
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "M paths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if w["grids"][0].get("precision", "fp32") == "fp32" else "fp16 records / fp32 transport",
        "data": "synthetic",
        "config": {"workload": {"c1": "64^3 isotropic SGGX value-noise volume, 256x256, 16 spp, depth 8",
                                "c2": "256^3 anisotropic SGGX fibre volume, procedural sky+sun env, "
                                      "1024x1024, 64 spp per step, depth 16",
                                "c4": "512^3 isotropic value-noise volume, 1024x1024, 256 spp per step, depth 8"}
                               [a.workload] + (" on the forest ground circle with a glossy sphere"
                                               if a.surfaces else ""),
                   "precision": w["grids"][0].get("precision", "fp32"),
                   "grid": w["params"]["n"], "width": W, "height": H, "spp_per_step": spp,
                   "max_depth": cfg["max_depth"], "brick_grid": cfg["use_brick_grid"],
                   "paths_per_step": paths_per_step},
        "roofline": roof,
        "cpu_baseline": cpu,
        "quality": quality,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


PROFILES = Path(__file__).resolve().parent / "profiles"


def profile_traffic(name: str):
    """HBM bytes per launch of the profiled kernel, from the rocprofv3 summary of the
    same workload (calibrated FETCH_SIZE + WRITE_SIZE passes), or None."""
    import re

    f = PROFILES / f"{name}.md"
    if not f.exists():
        return None, None
    m = re.search(r"\*\*([0-9.]+) MB per launch\*\*", f.read_text())
    return (int(float(m.group(1)) * 1e6), f"profiles/{f.name}") if m else (None, None)


def profile_valu_issue(name: str):
    """VALU issue rate of the profiled kernel (SQ_INSTS_VALU per CU per cycle, from
    the rocprofv3 summary of the same workload) against the rate a CU sustains on
    independent FP32 FMAs (renderer2_amd/bin/valu_peak), or None."""
    import re

    f = PROFILES / f"{name}.md"
    if not f.exists():
        return None
    m = re.search(r"VALU issue: .*?= ([0-9.]+) wavefront instructions per CU per cycle, [0-9.]+ % of the ([0-9.]+)",
                  f.read_text())
    if not m:
        return None
    achieved, ceiling = float(m.group(1)), float(m.group(2))
    return {"achieved": achieved, "ceiling": ceiling, "frac": round(achieved / ceiling, 3),
            "unit": "wavefront instructions per CU per cycle",
            "source": f"profiles/{f.name}, profiles/valu_peak_fma_exp2.md"}


def copy_calibration(mib: int) -> None:
    """One device-to-device copy of a known size, so that a PMC byte counter can be
    scaled against it (with HSA_ENABLE_SDMA=0 the copy runs as a blit kernel)."""
    if mib <= 0:
        return
    from renderer2_amd import _hip

    nb = mib << 20
    src, dst = _hip.DeviceBuffer(nb), _hip.DeviceBuffer(nb)
    _hip.check(_hip.hip().hipMemset(src.ptr, 1, nb), "hipMemset")
    _hip.check(_hip.hip().hipMemcpy(dst.ptr, src.ptr, nb, _hip.D2D), "hipMemcpy D2D")
    _hip.synchronize()
    src.free()
    dst.free()


def run_server(ctx, grids, env, reqs, budget=16 << 30, max_inflight=16, slice_paths=1 << 24):
    """Submit each request at its arrival time, wait for all; latency is per request
    from submission to the frame in host memory (the server's own clock)."""
    srv = ctx.server(hbm_budget_bytes=budget, max_inflight=max_inflight, slice_paths=slice_paths)
    t0 = time.perf_counter()
    ids = []
    for t, cfg, inst in reqs:
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        ids.append(srv.submit(cfg, grids, inst, env))
    for i in ids:
        srv.wait(i)
    elapsed = time.perf_counter() - t0
    st = srv.stats()
    srv.stop()
    paths = sum(c["width"] * c["height"] * c["spp"] for _, c, _ in reqs)
    return {"requests": len(reqs), "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
            "throughput_mpaths_s": round(paths / elapsed / 1e6, 2),
            "server_mpaths_s": round(st["paths_per_sec"] / 1e6, 2), "completed": st["completed"],
            "rejected": st["rejected"], "peak_gib": round(st["peak_bytes"] / 2 ** 30, 3), "wall_s": round(elapsed, 2),
            "paths": paths}


def serve_c5(ctx, a, n_req: int) -> dict:
    """Render server under a seeded Poisson stream: device-resident LOD grids of the
    256^3 anisotropic field at 128^3, 256^3 and 512^3 plus the 64^3 value-noise
    grid; resolution uniform over {256^2, 512^2, 1024^2}, spp over {16, 64, 256},
    orbit camera and LOD chosen per request.  The same request set is first run
    closed-loop (all submitted at once) to measure the server's capacity; the
    Poisson run's offered load is stated against it."""
    from renderer2_amd import scenes

    grids = [ctx.upload_grid(g) for g in scenes.serving_grids()]
    env = ctx.upload_env(scenes.synth_env("sky", 512, 256))
    warm = scenes.serving_requests(1, None, 0)[0]
    ctx.render(dict(warm[1], spp=1), grids, warm[2], env)  # warm-up
    cap = run_server(ctx, grids, env, scenes.serving_requests(n_req, None, 11))
    mean_paths = cap["paths"] / n_req
    capacity_rps = cap["throughput_mpaths_s"] * 1e6 / mean_paths
    res = run_server(ctx, grids, env, scenes.serving_requests(n_req, a.rate, 11))
    offered = a.rate * mean_paths / 1e6
    res.update(arrival_rate_per_s=a.rate, offered_mpaths_s=round(offered, 2),
               capacity_mpaths_s=cap["throughput_mpaths_s"], capacity_req_s=round(capacity_rps, 2),
               utilisation=round(offered / cap["throughput_mpaths_s"], 3), closed_loop_p50_ms=cap["p50_ms"],
               closed_loop_p99_ms=cap["p99_ms"], sizes=[256, 512, 1024], spps=[16, 64, 256],
               grids=["c2 field 128^3", "c2 field 256^3", "c2 field 512^3", "c1 field 64^3"], max_depth=16,
               workload="seeded Poisson stream, resolution uniform over {256,512,1024}^2, spp over {16,64,256}, "
                        "orbit camera and LOD uniform per request")
    for g in grids:
        g.free()
    env.free()
    return res


def bench_serving(a) -> int:
    import renderer2_amd as r2

    ctx = r2.Context(0)
    res = serve_c5(ctx, a, a.requests)
    line = {
        "metric": "render server: M paths/s with p50/p99 request latency under a Poisson stream",
        "value": res["throughput_mpaths_s"],
        "unit": "M paths/s",
        "n_gpus": 1,
        "steps": a.requests,
        "warmup": 0,
        "ms_per_step": round(res["wall_s"] * 1e3 / a.requests, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "serving: 256^3 anisotropic field at 128^3/256^3/512^3 + 64^3 value-noise grid "
                               "resident, Poisson arrivals, resolution {256,512,1024}^2, spp {16,64,256}, "
                               "orbit camera, depth 16",
                   "requests": a.requests, "arrival_rate_per_s": a.rate},
        "serving": res,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_filter_baseline(mesh, lods, w, threads: int, per_lod: int):
    """Time the FP32 filter oracle (OpenMP over voxels) on uniform random voxel
    samples of each LOD and extrapolate the 3-LOD job; also returns the records
    of the LOD-0 sample for a parity spot check.

    The oracle casts every voxel's rays, as the reference does; the GPU skips the
    voxels that no candidate triangle can reach (their no-hit record is exact).
    For a like-for-like figure the sampled voxels whose rays hit something
    (density > 0) are timed again on their own: a CPU that skipped every voxel
    without a hit (a superset of the GPU's skipped voxels) would take that
    time, extrapolated.  Returns (rate, voxels sampled, BVH build s, spot
    check, rate of the hit-voxels-only CPU, fraction of voxels with a hit)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O  # the CPU oracle: checker and baseline only

    t0 = time.perf_counter()
    of = O.MeshFilter(mesh)
    t_bvh = time.perf_counter() - t0
    rng = np.random.default_rng(12345)
    total_s, total_vox, sampled, spot = 0.0, 0, 0, None
    hit_s, hits = 0.0, 0
    for lod, vs in enumerate(lods):
        g = of.setup(vs)
        nvox = int(g.nv[0]) * int(g.nv[1]) * int(g.nv[2])
        ids = rng.integers(0, nvox, min(per_lod, nvox)).astype(np.int64)
        t0 = time.perf_counter()
        rec = of.run(g, samples=w["samples"], iterations=w["iterations"], sigma_t=w["sigma_t"], alpha=w["alpha"],
                     clip_rays=w["clip_rays"], seed=1, lod=lod, ids=ids, threads=threads)
        dt = time.perf_counter() - t0
        total_s += dt / len(ids) * nvox
        total_vox += nvox
        sampled += len(ids)
        hit = rec[:, :4].copy().view(np.float32)[:, 0] > 0
        if hit.any():
            t0 = time.perf_counter()
            of.run(g, samples=w["samples"], iterations=w["iterations"], sigma_t=w["sigma_t"], alpha=w["alpha"],
                   clip_rays=w["clip_rays"], seed=1, lod=lod, ids=ids[hit], threads=threads)
            hit_s += (time.perf_counter() - t0) / len(ids) * nvox
        hits += int(hit.sum()) / len(ids) * nvox
        if lod == 0:
            spot = (ids, rec)
    return (total_vox / total_s / 1e6, sampled, t_bvh, spot, total_vox / hit_s / 1e6 if hit_s > 0 else None,
            hits / total_vox)


def bench_filter(a) -> int:
    """Mesh -> SGGX filter of the 2.4 M-triangle mesh to LODs 512^3, 256^3, 128^3.
    One step filters all three LODs into device-resident record arrays."""
    import renderer2_amd as r2
    from renderer2_amd import _hip, scenes

    w = scenes.FILTER_WORKLOADS["c3"]
    mesh = scenes.make_filter_mesh("c3", seed=1)
    ctx = r2.Context(0)
    copy_calibration(a.copy_calib)
    gm = ctx.upload_mesh(mesh)
    vs0 = scenes.finest_voxel_size(mesh, w["finest"])
    lods = [vs0 * 2.0 ** l for l in range(w["lods"])]
    grids = [gm.setup(vs) for vs in lods]
    nvox = [int(g.num_voxels[0]) * int(g.num_voxels[1]) * int(g.num_voxels[2]) for g in grids]
    bufs = [_hip.DeviceBuffer(n * 32) for n in nvox]
    params = [r2.filter_params(vs, w["samples"], w["iterations"], w["sigma_t"], w["alpha"], w["clip_rays"], seed=1,
                               lod=l) for l, vs in enumerate(lods)]

    def step():
        st = []
        for p, b in zip(params, bufs):
            ctx.filter_lod(gm, p, out_dev_ptr=b.addr)
            st.append(ctx.filter_stats())
        return st

    for _ in range(a.warmup):
        step()
    _hip.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        stats.append(step())
    _hip.synchronize()
    elapsed = time.perf_counter() - t0
    value = sum(nvox) * a.steps / elapsed / 1e6
    last = stats[-1]
    kms = [float(np.mean([s[l]["filter_kernel_ms"] for s in stats])) for l in range(len(lods))]
    # algorithmic bytes of k_filter per launch: the active voxels' records (32 B),
    # their (offset, count, id) entries (12 B), and per candidate its index (4 B)
    # plus its three world-space vertices (48 B)
    algo = [s["active_voxels"] * (32 + 12) + s["pairs"] * (4 + 48) for s in last]
    achieved = algo[0] / (kms[0] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": "k_filter",
            "kernel_ms": round(kms[0], 3), "lod": 0}
    roof["traffic"], roof["traffic_source"] = profile_traffic("k_filter_c3_mesh_512_256_128")
    cpu = None
    parity = None
    if not a.no_cpu_baseline:
        host = host_cpus()
        threads = host["threads_used"]
        v, n, t_bvh, (ids, ref), v_hit, f_hit = cpu_filter_baseline(mesh, lods, w, threads, a.cpu_voxels)
        cpu = {"value": round(v, 4), "unit": "M voxels/s", "cores": threads, "kind": "port",
               "sample": f"FP32 filter oracle on {n} uniform random voxels over the 3 LODs, "
                         f"extrapolated to the whole job (BVH build {t_bvh:.1f} s excluded)",
               "host": host,
               "value_hit_voxels_only": round(v_hit, 4) if v_hit else None,
               "hit_voxel_fraction": round(f_hit, 5),
               "note": "value: every voxel's rays cast (the reference's algorithm); value_hit_voxels_only: "
                       "the same oracle timed on the sampled voxels whose rays hit something, as if the CPU "
                       "skipped every other voxel (more than the GPU skips)"}
        got = ctx.filter_voxels(gm, params[0], ids.astype(np.uint32))
        parity = {"records_bitexact": float(np.mean(np.all(got == ref, axis=1))), "voxels": int(len(ids)),
                  "nonempty": int(np.sum(ref[:, :4].copy().view(np.float32)[:, 0] > 0))}
    # each LOD rendered at 1024^2, 64 spp under the constant white environment of
    # the 64^3 workload (its camera and depth); the mesh is scaled to unit extent
    # and the media take sigma_a = 0, sigma_s = the filter's SigmaT, as the
    # reference's example scene does for filtered meshes
    render = None
    if a.render_steps > 0:
        rw = scenes.make_workload("c1")
        env = ctx.upload_env(rw["env"])
        ext = float((mesh["positions"].max(axis=0) - mesh["positions"].min(axis=0)).max())
        sc = 1.0 / ext
        inst = [dict(grid=0, pos=(0, 0, 0), orientation=(0, 1, 0), scaling=(sc, sc, sc))]
        rcfg = dict(rw["render"], width=1024, height=1024, spp=64)
        frame = _hip.DeviceBuffer(1024 * 1024 * 12)
        mps, kern = [], []
        for fg, vs, b in zip(grids, lods, bufs):
            gd = scenes.filtered_grid(fg, vs, records_dev=b.addr)
            gd["sigma_a"], gd["sigma_s"] = (0.0,) * 3, (float(w["sigma_t"]),) * 3
            g = ctx.upload_grid(gd)
            ctx.render_device(dict(rcfg, sample_begin=1_000_000), [g], inst, env, frame.addr)
            _hip.synchronize()
            t0 = time.perf_counter()
            km = 0.0
            for k in range(a.render_steps):
                ctx.render_device(dict(rcfg, sample_begin=k * 64), [g], inst, env, frame.addr)
                km += ctx.stats()["kernel_ms"]
            _hip.synchronize()
            mps.append(round(1024 * 1024 * 64 * a.render_steps / (time.perf_counter() - t0) / 1e6, 2))
            kern.append(round(km / a.render_steps, 3))
            g.free()
        frame.free()
        env.free()
        render = {"mpaths_s": mps, "kernel_ms": kern, "frame": "1024x1024, 64 spp per launch, constant white env, "
                  "depth 8", "steps": a.render_steps}
    line = {
        "metric": "M voxels/s mesh->SGGX filter (3 LODs 512^3/256^3/128^3, 64 rays/voxel)",
        "value": round(value, 3),
        "unit": "M voxels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "2.4 M triangle noise-displaced sphere + 200 k quads filtered to 3 LODs, "
                               "64 rays/voxel, 8 iterations, clipped rays",
                   "triangles": int(mesh["indices"].shape[0]),
                   "lods": [list(map(int, g.num_voxels)) for g in grids],
                   "active_voxels": [s["active_voxels"] for s in last], "pairs": [s["pairs"] for s in last],
                   "filter_kernel_ms": [round(k, 3) for k in kms],
                   "lod_total_ms": [round(s["total_ms"], 2) for s in last]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "quality": parity,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "gpu_vs_cpu_hit_voxels_only": (round(value / cpu["value_hit_voxels_only"], 1)
                                       if cpu and cpu["value_hit_voxels_only"] else None),
        "lod_render": render,
    }
    print(json.dumps(line), flush=True)
    for b in bufs:
        b.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
