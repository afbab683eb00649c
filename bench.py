#!/usr/bin/env python3
"""Traversal microbenchmark on one MI355X: Mrays/s of the ray-query hot path.

Workload (seeded): 410 level-4 icospheres = 2,099,200 FP32 triangles in one
mesh inside a 100-unit cube; ray sets 1920x1080 primary closest-hit rays,
2,000,000 cosine-hemisphere bounce closest-hit rays from primary hits, and
2,000,000 shadow rays to uniform points on 64 random triangles (unoccluded
query with target triangle).  One step = the three launches over all
6,073,600 rays, inputs resident in HBM.

Prints ONE JSON line (see the keys below).  ``roofline`` is priced for the
dominant kernel from its algorithmic bytes (per-ray work counts from the
counting variant of the same traversal) over its average launch duration,
measured with HIP events on the stream the kernels run on.  ``cpu_baseline``
times the oracle (CPU restatement of the reference traversal) on a bounded
sample of the same rays with all host threads.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "toy-path-tracing_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
B_RAY_IN = 32
B_HIT_OUT = 24
B_OCC_OUT = 1
B_NODE_FP32 = 112  # 96 B child boxes + 16 B child refs
B_NODE_Q8 = 64
B_TRI = 40  # 36 B positions + 4 B triangle id
B_INST = 64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nodes", choices=["fp32", "q8", "both"], default="both",
                    help="node format(s) to time; the reported value is the faster")
    ap.add_argument("--cpu-sample", type=int, default=2_100_000,
                    help="rays per set for the CPU baseline (0: skip; the default covers every ray)")
    ap.add_argument("--spheres", type=int, default=410)
    ap.add_argument("--rays", type=int, default=2_000_000, help="bounce and shadow ray counts")
    ap.add_argument("--render", type=int, default=1, help="also time the Cornell path tracer (1) or not (0)")
    ap.add_argument("--render-scenes", default="openpbr,openpbr_glossy,lambert_nee,textured",
                    help="comma-separated render lines to time (openpbr, openpbr_glossy, lambert_nee, textured; "
                         "also glass, film)")
    ap.add_argument("--adjoint", type=int, default=1, help="also time the path-replay adjoint (1) or not (0)")
    ap.add_argument("--city", type=int, default=1,
                    help="also time the four-frame city batch (1) or not (0): megakernel vs wavefront fp32 / q8")
    ap.add_argument("--city-spp", type=int, default=4, help="samples per pixel of the city frames")
    ap.add_argument("--city-adjoint", type=int, default=1,
                    help="also time the batched multi-view adjoint on the inverse-rendering city (1) or not (0)")
    ap.add_argument("--city-adjoint-spp", type=int, default=32, help="samples per pixel of the batched adjoint views")
    ap.add_argument("--nee-proxy", type=int, default=1,
                    help="also measure the OpenPBR proxy NEE policy against the reference policy (1) or not (0)")
    ap.add_argument("--refill", type=int, default=0, help="trace-kernel refill threshold (0: library default)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="resident blocks per CU (0: occupancy)")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="tpt_device_set_param before the traversal runs (repeatable)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of this many MiB three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes) and exit")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: only one GPU is used in this build (asked for --gpus {args.gpus}); "
              "the ray batch would shard by rays with the scene replicated per GPU", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_probe:
        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "launches": 3}), flush=True)
        return

    import tpt
    from tpt import scenes

    t0 = time.time()
    sf = scenes.sphere_field(n_spheres=args.spheres)
    sc = scenes.single_mesh_scene(sf.mesh)
    log(f"scene: {sf.mesh.triangle_count()} triangles, {len(sf.mesh.qbvh.nodes)} nodes, build+gen "
        f"{time.time() - t0:.1f}s")
    dev = tpt.Device(0)
    if args.refill:
        dev.set_param("refill_below", args.refill)
    if args.blocks_per_cu:
        dev.set_param("blocks_per_cu", args.blocks_per_cu)
    for kv in args.set:
        name, value = kv.split("=", 1)
        dev.set_param(name, int(value))
    ds = sc.upload(dev)
    primary = scenes.primary_rays(1920, 1080, **scenes.TRAVERSAL_CAMERA)
    hp = ds.closest_hit(primary)
    bounce = scenes.bounce_rays(primary, hp, sf.mesh, args.rays)
    shadow = scenes.shadow_rays(primary, hp, sf.mesh, args.rays, n_lights=64)
    log(f"rays: primary {len(primary)} (hit {np.mean(hp['triangle'] != tpt.INVALID_INDEX):.3f}), "
        f"bounce {len(bounce)}, shadow {len(shadow.rays)}")

    def to_dev(a):
        return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).to("cuda")

    sets = {
        "primary": dict(kind="closest", n=len(primary), rays=to_dev(primary), host=primary),
        "bounce": dict(kind="closest", n=len(bounce), rays=to_dev(bounce), host=bounce),
        "shadow": dict(kind="any", n=len(shadow.rays), rays=to_dev(shadow.rays), host=shadow.rays,
                       target=to_dev(shadow.target), target_host=shadow.target),
    }
    for s in sets.values():
        s["out"] = torch.empty(s["n"] * (B_HIT_OUT if s["kind"] == "closest" else B_OCC_OUT), dtype=torch.uint8,
                               device="cuda")
    # a dedicated (default-priority) stream: the kernels and the timing events
    # go on it; torch's legacy default stream would map to the library's own.
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    dev.set_stream(stream.cuda_stream)

    def launch(name, q8):
        s = sets[name]
        if s["kind"] == "closest":
            ds.closest_hit_device(s["rays"], s["out"], s["n"], quantized=q8)
        else:
            ds.unoccluded_device(s["rays"], s["target"], None, s["out"], s["n"], quantized=q8)

    # algorithmic bytes per launch from per-ray work counts (untimed)
    def algo_bytes(name, q8):
        s = sets[name]
        c = ds.counts(s["host"], "closest" if s["kind"] == "closest" else "any",
                      target=s.get("target_host"), quantized=q8).astype(np.float64).sum(0)
        out_b = B_HIT_OUT if s["kind"] == "closest" else B_OCC_OUT + 8
        return s["n"] * (B_RAY_IN + out_b) + c[0] * (B_NODE_Q8 if q8 else B_NODE_FP32) + c[1] * B_TRI + c[2] * B_INST, c / s["n"]

    results = {}
    modes = {"fp32": [False], "q8": [True], "both": [False, True]}[args.nodes]
    for q8 in modes:
        label = "q8" if q8 else "fp32"
        for _ in range(args.warmup):
            for name in sets:
                launch(name, q8)
        torch.cuda.synchronize()
        ev = {name: [] for name in sets}
        t_start = time.perf_counter()
        for _ in range(args.steps):
            for name in sets:
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(stream)
                launch(name, q8)
                b.record(stream)
                ev[name].append((a, b))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        kern_ms = {name: float(np.mean([a.elapsed_time(b) for a, b in ev[name]])) for name in sets}
        results[label] = dict(wall_s=wall, kern_ms=kern_ms, q8=q8)
        log(f"{label}: {args.steps} steps in {wall:.3f}s; per-kernel ms " +
            ", ".join(f"{k}={v:.3f}" for k, v in kern_ms.items()))

    best = min(results, key=lambda k: results[k]["wall_s"])
    r = results[best]
    rays_per_step = sum(s["n"] for s in sets.values())
    ms_per_step = r["wall_s"] / args.steps * 1e3
    value = rays_per_step / (ms_per_step / 1e3) / 1e6

    # roofline of the dominant kernel of the reported configuration
    dom = max(r["kern_ms"], key=lambda k: r["kern_ms"][k])
    bytes_launch, per_ray = algo_bytes(dom, r["q8"])
    achieved = bytes_launch / (r["kern_ms"][dom] / 1e3) / 1e9
    per_set = {}
    for name in sets:
        per_set[name] = round(sets[name]["n"] / (r["kern_ms"][name] / 1e3) / 1e6, 1)
    log(f"dominant {dom}: {bytes_launch / 1e6:.1f} MB algorithmic per launch, per-ray "
        f"nodes/tris/insts {per_ray.round(2).tolist()}, {achieved:.0f} GB/s")

    cpu = None
    if args.cpu_sample > 0:
        cpu = cpu_baseline(sc, sets, args.cpu_sample)

    render = render_bench(dev, stream, args) if args.render else None
    nee_proxy = nee_proxy_bench(dev, stream, args) if args.nee_proxy else None
    adjoint = adjoint_bench(dev, stream, args) if args.adjoint else None
    city = city_bench(dev, stream, args) if args.city else None
    city_adjoint = city_adjoint_bench(dev, stream, args) if args.city_adjoint else None

    line = {
        "metric": "Mrays/s",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "traversal microbenchmark: 410 level-4 icospheres (2,099,200 triangles, one mesh, "
                        "100-unit cube); 1920x1080 primary + 2M bounce closest-hit, 2M shadow any-hit",
            "nodes": best,
            "rays_per_step": rays_per_step,
            "mrays_per_s_by_set": per_set,
            "kernel_ms": {k: round(v, 4) for k, v in r["kern_ms"].items()},
            "fp32_vs_q8_wall_s": {k: round(v["wall_s"], 4) for k, v in results.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"{dom} ({'closest' if sets[dom]['kind'] == 'closest' else 'any'}, {best})",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(sets[dom]["kind"], best, sets[dom]["n"]),
        },
        "cpu_baseline": cpu,
        "render": render,
        "nee_proxy": nee_proxy,
        "adjoint": adjoint,
        "city": city,
        "city_adjoint": city_adjoint,
    }
    # the three throughputs side by side: samples/s of the wavefront path
    # tracer with 8-bit nodes against the plain FP32 megakernel on the city
    # batch (equal relMSE: the same estimator on the same samples), Mrays/s,
    # and gradients/s of the full-parameter adjoint
    summary = {"mrays_per_s": line["value"]}
    if city:
        summary["city_samples_per_s"] = {
            "megakernel_fp32": city["megakernel"]["samples_per_s"],
            "wavefront_fp32": city["wavefront"]["samples_per_s"],
            "wavefront_q8": city["wavefront_q8"]["samples_per_s"],
            "q8_wavefront_vs_fp32_megakernel": city["wavefront_q8"]["speedup_vs_megakernel"],
            "fp32_wavefront_vs_fp32_megakernel": city["wavefront"]["speedup_vs_megakernel"],
        }
        summary["city_qbvh_build"] = city["qbvh_build"]
    if nee_proxy:
        summary["nee_proxy_relmse_gain_at_equal_time"] = {k: v["relmse_gain_at_equal_time"]
                                                          for k, v in nee_proxy.items()}
    if city and "nee_proxy" in city:
        for k, v in city["nee_proxy"].items():
            summary.setdefault("nee_proxy_relmse_gain_at_equal_time", {})["city_" + k] = v["relmse_gain_at_equal_time"]
    if adjoint:
        summary["grad_per_s"] = {k: v["grad_per_s"] for k, v in adjoint.items()}
    if city_adjoint:
        summary.setdefault("grad_per_s", {})["city_32_views"] = city_adjoint["grad_per_s"]
    line["summary"] = summary
    print(json.dumps(line), flush=True)
    ds.close()
    dev.close()


def measured_traffic(kind, nodes, n_rays):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (each scaled by the same counter's
    reading for a plain copy of known size), when those runs profiled this
    kernel on this ray count; otherwise None."""
    want = {("any", "fp32"): "trace_kernel<false,1,false,false,false>",
            ("closest", "fp32"): "trace_kernel<false,0,false,false,false>"}.get((kind, nodes))
    path = ROOT / "profiles" / "trace_any_fp32_T2M_pmc.txt"
    if want is None or not path.exists():
        return None
    for ln in path.read_text().splitlines():
        m = re.match(r"traffic_bytes_per_launch kernel=(\S+) rays=(\d+) bytes=(\d+)", ln)
        if m and m.group(1) == want and int(m.group(2)) == n_rays:
            return int(m.group(3))
    return None


def render_bench(dev, stream, args):
    """FP32 path tracer (megakernel, then the autotuned wavefront) on the
    Cornell configuration (512x512, 64 spp, max depth 8): the configuration's
    diffuse OpenPBR walls (new(c).with_specular_weight(0); the reference
    gives OpenPBR surfaces no light-tree NEE), the same walls with the
    reference's OpenPbrMaterial::new defaults (specular weight 1, roughness
    0.3, IOR 1.5: the full BSDF with its GGX tables), and the same box with
    the reference's NormalizedLambert walls + Emissive light, where light-tree
    NEE with MIS runs; and the textured box (every texture slot, EWA lookups
    with ray differentials)."""
    import torch

    from tpt import camera, scenes

    w = h = 512
    spp, depth = 64, 8
    out = {}
    for key, make, label in (
        ("openpbr", scenes.cornell, "OpenPBR diffuse (new(c).with_specular_weight(0)), BSDF sampling + emission"),
        ("openpbr_glossy", lambda: scenes.cornell(specular_weight=1.0),
         "OpenPBR defaults (EON base + GGX specular with its albedo and energy-compensation tables), "
         "BSDF sampling + emission"),
        ("lambert_nee", scenes.cornell_lambert, "NormalizedLambert + Emissive light, light-tree NEE + MIS"),
        ("textured", scenes.cornell_textured,
         "OpenPBR with base-colour, roughness, metalness, fuzz and emission textures (EWA over mip pyramids, "
         "ray differentials and cones), a Mirror box and a NormalizedLambert box with light-tree NEE"),
        ("glass", scenes.cornell_glass,
         "OpenPBR transmission: smooth glass, rough tinted glass, a thin-walled frosted pane (not run by default)"),
        ("film", scenes.cornell_film,
         "OpenPBR thin film: rough metal, smooth coated dielectric and a thin-walled pane under films "
         "(not run by default)"),
    ):
        if key not in args.render_scenes.split(","):
            continue
        cs = make()
        rds = cs.scene.upload(dev)
        cam = camera.pinhole(**cs.camera, width=w, height=h)
        img_dev = torch.empty(w * h * 3, dtype=torch.float32, device="cuda")
        rds.render_device(cam, w, h, spp, depth, img_dev)
        torch.cuda.synchronize()
        reps = max(1, min(args.steps, 5))
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            rds.render_device(cam, w, h, spp, depth, img_dev)
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        img = img_dev.view(h, w, 3).cpu().numpy()
        res = {
            "workload": f"Cornell box 512x512, 64 spp, max depth 8, {label}",
            "ms_per_frame": round(ms, 3),
            "samples_per_s": round(w * h * spp / (ms / 1e3), 1),
            "image_mean": [round(float(x), 4) for x in img.reshape(-1, 3).mean(0)],
        }
        log(f"render {key}: {ms:.2f} ms/frame = {res['samples_per_s'] / 1e6:.1f} Msamples/s")
        # wavefront path tracer on the same frame: FP32 and 8-bit nodes, after
        # the autotuner has picked paths in flight and the trace launch shape
        wf_img = torch.empty_like(img_dev)
        # 32 samples per pixel: 8 Mi paths, two passes at the largest paths-in-flight candidate
        tune = rds.autotune([(cam, w, h, spp, depth)], tune_spp=32)
        res["autotune"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tune.items()}
        log(f"render {key} autotune: {res['autotune']}")
        for q8 in (False, True):
            rds.render_batch_device([(cam, w, h, spp, depth)], [wf_img], quantized=q8)
            torch.cuda.synchronize()
            a.record(stream)
            for _ in range(reps):
                rds.render_batch_device([(cam, w, h, spp, depth)], [wf_img], quantized=q8)
            b.record(stream)
            torch.cuda.synchronize()
            wms = a.elapsed_time(b) / reps
            tag = "wavefront_q8" if q8 else "wavefront"
            res[tag] = {"ms_per_frame": round(wms, 3), "samples_per_s": round(w * h * spp / (wms / 1e3), 1),
                        "speedup_vs_megakernel": round(ms / wms, 3)}
            if not q8:
                res[tag]["identical_to_megakernel"] = bool(torch.equal(wf_img, img_dev))
            log(f"render {key} {tag}: {wms:.2f} ms/frame = {w * h * spp / (wms / 1e3) / 1e6:.1f} Msamples/s "
                f"({ms / wms:.2f}x megakernel)")
        if args.cpu_sample > 0:
            sys.path.insert(0, str(ROOT / "tests"))
            import oracle  # noqa: E402  (test infrastructure: CPU baseline leg only)

            threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
            r = oracle.OracleRenderer(cs.scene, cs.materials)
            rows = 64
            t = time.perf_counter()
            r.render(cam, w, h, spp, depth, rows=(h // 2 - rows // 2, h // 2 + rows // 2), threads=threads)
            dt = time.perf_counter() - t
            res["cpu_baseline"] = {"value": round(rows * w * spp / dt, 1), "unit": "samples/s", "cores": threads,
                                   "kind": "port",
                                   "sample": f"{rows} middle rows of the frame ({rows * w * spp} samples, {dt:.1f}s)"}
            log(f"render {key} cpu baseline: {rows * w * spp / dt / 1e6:.2f} Msamples/s on {threads} threads")
        rds.close()
        out[key] = res
        dev.set_param("wf_paths", 1 << 21)  # the traversal defaults for whatever runs next
        dev.set_param("blocks_per_cu", 4)
        dev.set_param("refill_below", 24)
        dev.set_param("wf_sort", 0)
    return out


def relmse(x, ref):
    """Mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2) (tpt.imageio.relmse's
    definition, here on device tensors)."""
    return float(((x - ref) ** 2 / (ref * ref + 1e-2)).mean())


def policy_comparison(rds, stream, frame, render, ref_spp):
    """The two NEE policies on one frame: time `render(frame)` (a device
    render of one frame) under each, then the relMSE of each image against a
    ref_spp-sample image of the proxy policy (both policies converge to the
    same image; the proxy's has the lower variance).  relMSE falls as
    1 / samples, so at the reference policy's frame time the proxy's relMSE
    is its own times its time over the reference's: the gain at equal time is
    the reference policy's relMSE over that."""
    import torch

    cam, w, h, spp, depth = frame
    img = torch.empty(w * h * 3, dtype=torch.float32, device="cuda")
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    imgs = {}
    for policy in ("reference", "openpbr_proxy"):
        rds.set_nee_policy(policy)
        render(frame, img)
        torch.cuda.synchronize()
        a.record(stream)
        render(frame, img)
        b.record(stream)
        torch.cuda.synchronize()
        out[policy] = {"ms_per_frame": round(a.elapsed_time(b), 3)}
        imgs[policy] = img.clone()
    ref = torch.empty_like(img)
    render((cam, w, h, ref_spp, depth), ref)  # proxy policy (set last above)
    torch.cuda.synchronize()
    rds.set_nee_policy("reference")
    for policy in imgs:
        out[policy]["relmse"] = relmse(imgs[policy], ref)
        out[policy]["image_mean"] = round(float(imgs[policy].mean()), 5)
    t_ref, t_prox = out["reference"]["ms_per_frame"], out["openpbr_proxy"]["ms_per_frame"]
    e_ref, e_prox = out["reference"]["relmse"], out["openpbr_proxy"]["relmse"]
    out["reference_image"] = f"{ref_spp} spp, proxy policy (mean {float(ref.mean()):.5f})"
    out["relmse_gain_at_equal_spp"] = round(e_ref / max(e_prox, 1e-30), 2)
    out["relmse_gain_at_equal_time"] = round(e_ref / max(e_prox * t_prox / t_ref, 1e-30), 2)
    for policy in imgs:
        out[policy]["relmse"] = round(out[policy]["relmse"], 6)
    return out


def nee_proxy_bench(dev, stream, args):
    """The OpenPBR proxy NEE policy (tpt_scene_set_nee_policy) against the
    reference policy on the Cornell configuration (512x512, 64 spp, max depth
    8; diffuse OpenPBR walls, and the reference's OpenPBR defaults): frame
    time of the megakernel under each policy and relMSE against a
    2048-spp image, and the gain at equal time."""
    from tpt import camera, scenes

    w = h = 512
    out = {}
    for key, make in (("cornell_openpbr", scenes.cornell),
                      ("cornell_openpbr_glossy", lambda: scenes.cornell(specular_weight=1.0))):
        cs = make()
        rds = cs.scene.upload(dev)
        cam = camera.pinhole(**cs.camera, width=w, height=h)
        res = policy_comparison(rds, stream, (cam, w, h, 64, 8),
                                lambda f, o: rds.render_device(f[0], f[1], f[2], f[3], f[4], o), 2048)
        res["workload"] = f"Cornell box 512x512, 64 spp, max depth 8 ({key}), megakernel"
        log(f"nee proxy {key}: reference {res['reference']} proxy {res['openpbr_proxy']} gain at equal time "
            f"{res['relmse_gain_at_equal_time']}x")
        rds.close()
        out[key] = res
    return out


def city_bench(dev, stream, args):
    """The wavefront-batching workload: the procedural city (about 9.4 M
    triangles, 2048 emissive triangles in the light tree, a 2048x1024 sky
    with a sun, eight OpenPBR variants) rendered as four frames of uneven
    resolution (1920x1080, 1280x720, 640x360, 3840x2160: 11,520,000 pixels)
    at max depth 16.  The plain FP32 megakernel renders the frames one launch
    each; the wavefront path tracer renders all four in one batch, with FP32
    and with 8-bit quantised nodes, after the autotuner.  Rendered again with
    the exact shadow rule, both wavefront images are compared with the
    megakernel's bit for bit (identical images: equal relMSE); with 8-bit
    nodes a path can take the other triangle of an exact tie, so the line
    reports how many pixels differ."""
    import torch

    from tpt import camera, scenes

    spp, depth = args.city_spp, 16
    t = time.perf_counter()
    cs = scenes.city(build=False)
    t_gen = time.perf_counter() - t
    # the QBVH build of every mesh and of the instances: host threads, then
    # the device build (tpt_qbvh_build_device), which must give the same trees
    t = time.perf_counter()
    cs.scene.build_qbvh()
    t_host = time.perf_counter() - t
    host_trees = [(m.qbvh.nodes, m.qbvh.primitive_indices) for m in cs.scene.meshes]
    host_trees.append((cs.scene.qbvh.nodes, cs.scene.qbvh.primitive_indices))
    t = time.perf_counter()
    cs.scene.build_qbvh(device=dev)
    t_gpu = time.perf_counter() - t
    gpu_trees = [(m.qbvh.nodes, m.qbvh.primitive_indices) for m in cs.scene.meshes]
    gpu_trees.append((cs.scene.qbvh.nodes, cs.scene.qbvh.primitive_indices))
    same = all(a[0].tobytes() == b[0].tobytes() and np.array_equal(a[1], b[1]) for a, b in zip(host_trees, gpu_trees))
    t_build = t_gen + t_host
    tris = sum(m.triangle_count() for m in cs.scene.meshes)
    log(f"city: {tris} triangles, {cs.emissive_triangles} emissive, generated in {t_gen:.1f}s; QBVH build "
        f"host {t_host * 1e3:.0f} ms, GPU {t_gpu * 1e3:.0f} ms (bounds in host memory, trees copied back), "
        f"identical trees: {same}")
    t = time.perf_counter()
    rds = cs.scene.upload(dev)
    t_upload = time.perf_counter() - t
    log(f"city upload (layout, 8-bit nodes encoded on the device, light tree, textures): {t_upload:.2f}s")
    frames = [(camera.pinhole(**v, width=w, height=h), w, h, spp, depth) for v, w, h in cs.cameras]
    pixels = sum(w * h for _, w, h, _, _ in frames)
    samples = pixels * spp
    mega = [torch.empty(w * h * 3, dtype=torch.float32, device="cuda") for _, w, h, _, _ in frames]
    reps = max(1, min(args.steps, 3))
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    def megakernel():
        for (cam, w, h, s_, d), o in zip(frames, mega):
            rds.render_device(cam, w, h, s_, d, o)

    ms = timed(megakernel)
    res = {
        "workload": f"procedural city, {tris} triangles, {cs.emissive_triangles} emissive triangles + 2048x1024 sky "
                    f"with sun, 8 OpenPBR variants; 4 frames (1920x1080, 1280x720, 640x360, 3840x2160), "
                    f"{spp} spp, max depth {depth}",
        "scene_build_s": round(t_build, 2),
        "qbvh_build": {"host_ms": round(t_host * 1e3, 1), "gpu_ms": round(t_gpu * 1e3, 1),
                       "speedup": round(t_host / t_gpu, 2), "identical_trees": bool(same)},
        "upload_s": round(t_upload, 2),
        "megakernel": {"ms_per_batch": round(ms, 3), "samples_per_s": round(samples / (ms / 1e3), 1)},
    }
    log(f"city megakernel fp32: {ms:.1f} ms per 4 frames = {samples / (ms / 1e3) / 1e6:.1f} Msamples/s")
    # the workload itself (46 M paths at 4 spp); at more samples per pixel the
    # trials run at 4 spp, the same per-sample work
    tune = rds.autotune(frames, quantized=True, tune_spp=min(spp, 4))
    res["autotune"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tune.items()}
    log(f"city autotune: {res['autotune']}")
    wf = [torch.empty_like(o) for o in mega]
    # the material sort of the shade queue, both settings (the autotuner keeps the faster)
    chosen = tune["wf_sort"]
    res["material_sort_ms"] = {}
    for v in (0, 1):
        dev.set_param("wf_sort", v)
        res["material_sort_ms"]["on" if v else "off"] = round(timed(lambda: rds.render_batch_device(frames, wf)), 3)
    dev.set_param("wf_sort", chosen)
    log(f"city wavefront fp32 with / without the material sort: {res['material_sort_ms']}")
    for q8 in (False, True):
        wms = timed(lambda: rds.render_batch_device(frames, wf, quantized=q8))
        tag = "wavefront_q8" if q8 else "wavefront"
        res[tag] = {"ms_per_batch": round(wms, 3), "samples_per_s": round(samples / (wms / 1e3), 1),
                    "speedup_vs_megakernel": round(ms / wms, 3)}
        # the timed runs use the early-exit shadow walk; with the exact shadow
        # rule (closest hit, as the megakernel) the images must be the megakernel's
        rds.render_batch_device(frames, wf, quantized=q8, exact_shadow=True)
        torch.cuda.synchronize()
        res[tag]["identical_to_megakernel"] = all(bool(torch.equal(x, y)) for x, y in zip(wf, mega))
        # 8-bit nodes change the child visit order: at exact ties (a shared
        # edge) another triangle can be the hit, and that pixel's path differs
        differ = sum(int((x != y).view(-1, 3).any(1).sum()) for x, y in zip(wf, mega))
        res[tag]["pixels_differing_from_megakernel"] = differ
        res[tag]["mean_relative_difference"] = round(
            sum(float((x - y).abs().sum()) for x, y in zip(wf, mega)) / max(sum(float(y.sum()) for y in mega), 1e-30),
            8)
        log(f"city {tag}: {wms:.1f} ms per 4 frames = {samples / (wms / 1e3) / 1e6:.1f} Msamples/s "
            f"({ms / wms:.2f}x megakernel), exact-shadow image identical: {res[tag]['identical_to_megakernel']} "
            f"({differ} of {pixels} pixels differ)")
    res["image_means"] = [[round(float(x), 4) for x in o.view(-1, 3).mean(0).tolist()] for o in mega]
    if args.nee_proxy:
        # the NEE policies on the 640x360 frame: the city as configured (sky
        # and sun) and at night (the same geometry without the sky: lamps only)
        cam3, w3, h3, _, _ = frames[2]

        def wave(f, o):
            rds_cur.render_batch_device([f], [o.view(f[2], f[1], 3)])

        res["nee_proxy"] = {}
        rds_cur = rds
        res["nee_proxy"]["day"] = policy_comparison(rds, stream, (cam3, w3, h3, spp, depth), wave, 1024)
        res["nee_proxy"]["day"]["workload"] = f"city frame 640x360, {spp} spp, max depth {depth}, wavefront, sky + lamps"
        log(f"city nee proxy day: {res['nee_proxy']['day']}")
    rds.close()
    if args.nee_proxy:
        cs.scene.environment = None
        rds_cur = cs.scene.upload(dev)
        res["nee_proxy"]["night"] = policy_comparison(rds_cur, stream, (cam3, w3, h3, spp, depth), wave, 1024)
        res["nee_proxy"]["night"]["workload"] = (f"city frame 640x360, {spp} spp, max depth {depth}, wavefront, "
                                                 "lamps only (the sky removed)")
        log(f"city nee proxy night: {res['nee_proxy']['night']}")
        rds_cur.close()
    dev.set_param("wf_paths", 1 << 21)
    dev.set_param("blocks_per_cu", 4)
    dev.set_param("refill_below", 24)
    dev.set_param("wf_sort", 0)
    return res


def city_adjoint_bench(dev, stream, args):
    """Batched differentiable rendering at scale: the inverse-rendering city
    (scenes.city_inverse: the city's geometry, 64 learnable OpenPBR
    materials, four 1024x1024 FP32 base-colour textures from seeded Perlin
    noise, the lamps and the sky), 32 views whose side is log-uniform in
    256..1024, 32 spp, depth 8.  One gradient = the forward render of all
    views in one wavefront batch (the autotuner's schedule) + one batched
    path-replay adjoint launch over every view (tpt_render_adjoint_batch)
    with respect to 320 parameters and 12.58 M texels.  The FP32 compensated
    accumulation's error against FP64 accumulation is measured at 2 spp."""
    import torch

    from tpt import camera, scenes

    spp, depth = args.city_adjoint_spp, 8
    t = time.perf_counter()
    c = scenes.city_inverse()
    t_build = time.perf_counter() - t
    ds = c.scene.upload(dev)
    frames = [(camera.pinhole(**v, width=w, height=h), w, h, spp, depth) for v, w, h in c.views]
    pixels = sum(w * h for _, w, h, _, _ in frames)
    n_tex = sum(int(np.prod(tx.pixels.shape)) for tx in c.scene.textures)
    log(f"city adjoint: {len(frames)} views, {pixels} pixels, {len(c.materials)} materials, {n_tex} texels, "
        f"generated and built in {t_build:.1f}s")
    tune = ds.autotune([(cam, w, h, s_, d) for cam, w, h, s_, d in frames], tune_spp=2)
    outs = [torch.empty(w * h * 3, dtype=torch.float32, device="cuda") for _, w, h, _, _ in frames]
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    ds.render_batch_device(frames, outs)
    b.record(stream)
    torch.cuda.synchronize()
    fwd_ms = a.elapsed_time(b)
    # dL/dI of 0.5 |I - 0.5|^2 for every view
    views = [f + ((o - 0.5).contiguous(),) for f, o in zip(frames, outs)]
    d_par = torch.zeros(5 * len(c.materials), dtype=torch.float32, device="cuda")
    d_tex = torch.zeros(n_tex, dtype=torch.float32, device="cuda")
    a.record(stream)
    ds.render_adjoint_batch_device(views, d_par, d_tex)
    b.record(stream)
    torch.cuda.synchronize()
    adj_ms = a.elapsed_time(b)
    ms = fwd_ms + adj_ms
    log(f"city adjoint: forward {fwd_ms:.0f} ms + batched adjoint {adj_ms:.0f} ms = {1e3 / ms:.3f} grad/s")
    # FP32 compensated vs FP64 accumulation, 2 spp
    small = [(cam, w, h, 2, d, g) for (cam, w, h, _, d), g in zip(frames, [v[5] for v in views])]
    p32, p64 = torch.zeros_like(d_par), torch.zeros_like(d_par)
    t32, t64 = torch.zeros_like(d_tex), torch.zeros_like(d_tex)
    ds.render_adjoint_batch_device(small, p32, t32)
    ds.render_adjoint_batch_device(small, p64, t64, fp64=True)
    torch.cuda.synchronize()
    err_p = float((p32 - p64).abs().max() / p64.abs().max().clamp_min(1e-30))
    err_t = float((t32 - t64).abs().max() / t64.abs().max().clamp_min(1e-30))
    ds.close()
    log(f"city adjoint: FP32 compensated vs FP64 accumulation, max error / max gradient: parameters {err_p:.2e}, "
        f"texels {err_t:.2e}")
    return {
        "workload": f"inverse-rendering city ({sum(m.triangle_count() for m in c.scene.meshes)} triangles, "
                    f"{len(c.materials) - 1} learnable OpenPBR materials + the lamp emitter, 4 x 1024x1024 FP32 "
                    f"base-colour textures, 2048x1024 sky), {len(frames)} views of side 256..1024 (log-uniform), "
                    f"{pixels} pixels, {spp} spp, depth {depth}, L2 image loss",
        "parameters": 5 * (len(c.materials) - 1) + n_tex,
        "forward_ms": round(fwd_ms, 1),
        "adjoint_ms": round(adj_ms, 1),
        "ms_per_gradient": round(ms, 1),
        "grad_per_s": round(1e3 / ms, 4),
        "adjoint_samples_per_s": round(pixels * spp / (adj_ms / 1e3), 1),
        "forward_schedule": tune,
        "fp32_vs_fp64_max_error": {"parameters": err_p, "texels": err_t},
    }


def adjoint_bench(dev, stream, args):
    """Path-replay adjoint on the inverse-rendering Cornell box, 256x256,
    16 spp, depth 6 (an L2 loss's image gradient as input): gradient
    evaluations per second, for (1) 16 diffuse OpenPBR materials with
    learnable base colours plus a learnable 512x512 RGB base-colour texture
    (exact base-colour ratios), and (2) the configuration's full parameter
    set: 16 glossy materials with learnable base colour, specular roughness
    and metalness (seeded), plus the texture.  One evaluation = primal +
    replay of every pixel sample, parameter and texel gradients (EWA and
    pyramid adjoints) on the device."""
    import torch

    from tpt import camera, scenes

    w = h = 256
    spp, depth = 16, 6
    reps = max(1, min(args.steps, 5))
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    out = {}
    for key, glossy in (("base_color_texels", False), ("params_texels", True)):
        cs = scenes.cornell_inverse(glossy=glossy)
        ds = cs.scene.upload(dev)
        cam = camera.pinhole(**cs.camera, width=w, height=h)
        img = torch.empty(w * h * 3, dtype=torch.float32, device="cuda")
        fms = timed(lambda: ds.render_device(cam, w, h, spp, depth, img))
        dloss = (img - 0.5).contiguous()  # dL/dI of 0.5 * |I - 0.5|^2
        n_tex = sum(int(np.prod(t.pixels.shape)) for t in cs.scene.textures)
        d_tex = torch.empty(n_tex, dtype=torch.float32, device="cuda")
        rows = 5 if glossy else 3
        d_par = torch.empty(rows * len(cs.materials), dtype=torch.float32, device="cuda")
        if glossy:
            ms = timed(lambda: ds.render_adjoint_params_device(cam, w, h, spp, depth, dloss, d_par, d_tex))
        else:
            ms = timed(lambda: ds.render_adjoint_texels_device(cam, w, h, spp, depth, dloss, d_par, d_tex))
        ds.close()
        label = ("16 glossy OpenPBR materials (learnable base_color, specular_roughness, base_metalness)" if glossy
                 else "16 OpenPBR materials (learnable base_color)")
        out[key] = {
            "workload": f"inverse-rendering Cornell box: {label} + learnable 512x512 RGB base-colour texture, "
                        "256x256, 16 spp, depth 6, L2 image loss",
            "parameters": rows * 16 + n_tex,
            "ms_per_gradient": round(ms, 3),
            "grad_per_s": round(1e3 / ms, 2),
            "samples_per_s": round(w * h * spp / (ms / 1e3), 1),
            "forward_ms": round(fms, 3),
        }
        log(f"adjoint {key}: {ms:.2f} ms per gradient = {1e3 / ms:.1f} grad/s (forward {fms:.2f} ms)")
    return out


def cpu_baseline(sc, sets, sample):
    """The oracle on a seeded sample of each ray set, all host threads."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle  # noqa: E402  (test infrastructure: CPU baseline leg only)

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    t = time.time()
    o = oracle.oracle_scene_from(sc)
    log(f"cpu baseline: oracle build {time.time() - t:.1f}s (untimed), {threads} threads")
    rng = np.random.default_rng(123)
    total_rays, total_s = 0, 0.0
    for name, s in sets.items():
        idx = np.sort(rng.choice(s["n"], min(sample, s["n"]), replace=False))
        rays = s["host"][idx]
        t = time.perf_counter()
        if s["kind"] == "closest":
            o.closest_hit(rays, threads=threads)
        else:
            o.unoccluded(rays, s["target_host"][idx], threads=threads)
        dt = time.perf_counter() - t
        total_rays += len(rays)
        total_s += dt
        log(f"cpu baseline {name}: {len(rays)} rays {dt:.2f}s = {len(rays) / dt / 1e6:.2f} Mrays/s")
    return {
        "value": round(total_rays / total_s / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} seeded rays from each of the primary, bounce and shadow sets "
                  f"({total_rays} rays, {total_s:.1f}s); shadow answered as the reference does "
                  "(closest hit + target test)",
    }


if __name__ == "__main__":
    main()
