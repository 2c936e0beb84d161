"""Benchmark: the 1080p clipmap GI frame on one MI355X.

Workload (6 x 128^3 clipmap GI): the seeded procedural room scaled to 16
units (21 260 triangles), RGBA16F radiance + RGBA8 opacity clipmaps of
6 levels x 128^3 (6 faces, 1-texel border), a 2048^2 fp32 shadow map from a
seeded directional light, and a 1920x1080 G-buffer ray cast from a seeded
camera with per-pixel perceptual roughness U[0.1, 0.8] and metallic > 0 so
that every pixel traces the specular cone.

One step = one GI frame through the reference's passes: VoxelizationPass
(clear, 6 levels of opacity voxelization, opacity mips, border wrap),
RadianceInjectionPass with every level due (clear, injection with 16-tap PCF,
copy-alpha, radiance mips) and VoxelConeTracingPass in CombinedGI mode
(16 diffuse cones + 1 specular cone per pixel). The G-buffer and shadow map
are frame inputs, produced once before timing.

Prints one JSON line (see the repository's benchmark contract).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time
from pathlib import Path

# the CPU baseline's OpenMP threads stay on neighbouring cores (read when the
# OpenMP runtime starts, so it is set before anything loads one)
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np  # noqa: E402

ROOT = Path(__file__).resolve().parent
METRIC = "1080p GI ms/frame + cone samples/s; batched & adjoint vs plain, same GPU"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (dual-issue / packed; 78.6 without packing)
FLOP_PER_SAMPLE = 384  # trilinear filtering of one cone sample: 2 levels x 3 faces x 8 texels x 4 channels x FMA
B_SAMPLE_RGBA16F = 384  # 2 levels x 3 faces x 8 texels x 8 B gathered per sample (mostly L2 hits)
B_PIXEL = 124  # G-buffer 28 B + outputs 32 B + 16 PCF taps x 4 B


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=32, help="rows of the 1080p image traced by the CPU baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the plain-VXAO, SVO, batched and adjoint workloads reported beside the frame")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates the HBM byte counters) and exit")
    return ap.parse_args()


def copy_probe(mib: int) -> None:
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE of
    this kernel divided by the bytes it must move gives the counters' scale."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}), flush=True)


def setup():
    from vk_voxel_cone_tracing_amd import workloads

    return workloads.clipmap_frame()


def frame(st, ev=None):
    """One GI frame; with `ev` records per-pass events."""
    from vk_voxel_cone_tracing_amd import workloads

    workloads.clipmap_frame_run(st, ev)


def extras() -> dict:
    """The other measured workloads (plain dense VXAO 640x360, the 512^3
    octree from 1M triangles, 64 batched scenes vs their per-scene loop, the
    256^3 adjoint vs its forward), each on its own synthetic inputs."""
    from vk_voxel_cone_tracing_amd import workloads as wl

    res = {}
    for name, fn in (("plain_vxao_640x360", wl.plain_vxao), ("svo_512", wl.svo_build), ("batched_64", wl.batched),
                     ("adjoint_256", wl.adjoint)):
        t0 = time.perf_counter()
        try:
            res[name] = fn()
        except Exception as e:  # reported, not fatal: the frame metric above stands on its own
            res[name] = {"error": f"{type(e).__name__}: {e}"}
        res[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        print(f"[bench] {name}: {json.dumps(res[name])}", file=sys.stderr, flush=True)
    return res


def host_cpu() -> dict:
    """CPU model and the threads an OpenMP region of this process gets."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", avail) or avail)
    return dict(model=model, nproc=os.cpu_count(), affinity=avail, omp_threads=threads,
                omp_proc_bind=os.environ.get("OMP_PROC_BIND", ""))


def _rel_l2(got, ref) -> float:
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))


def cpu_baseline(st, rows: int, runs: int = 5):
    """The oracle (the CPU restatement of the reference shaders) on the
    frame's own inputs: all voxel passes of the frame (opacity voxelization
    and injection of the 6 levels run as 6 concurrent threads, one per level;
    the mips are OpenMP over destination texels) plus `rows` evenly spaced
    rows of the trace (OpenMP over rows). One warm-up, then the median of
    `runs` runs of each part; the frame time is the voxel passes plus the
    traced rows scaled to 1080. The traced rows are then compared with the
    GPU frame (relative L2 against the FP32 oracle, and against the FP64
    oracle on a subset of them)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_api as O

    ctx = st["ctx"]
    scene = st["scene"]
    RES, LV = 128, 6
    depth = st["shadow"].cpu().numpy()
    regions = ctx.regions
    descs = [O.voxel_desc_from_library(ctx.voxelization_desc(lvl), regions[lvl]) for lvl in range(LV)]
    g = {k: v.cpu().numpy() for k, v in st["g"].items()}
    g = {k: (v.view(np.uint32) if v.dtype == np.int32 else (v.view(np.uint16) if v.dtype == np.int16 else v))
         for k, v in g.items()}
    vdesc = st["vct"].desc()
    fields = dict(volume_center=list(vdesc.volume_center), rendering_mode=8, voxel_size=vdesc.voxel_size,
                  volume_dimension=128.0, trace_start_offset=1.0, indirect_diffuse_intensity=8.0,
                  ambient_occlusion_factor=0.5, min_trace_step_factor=1.0, indirect_specular_intensity=3.0,
                  occlusion_decay=2.0, enable_32_cones=0)
    p = O.trace_params(fields, st["vp_inv"], st["eye"], LV)
    W, H = st["W"], st["H"]
    step = max(1, H // rows)
    row_list = np.asarray(list(range(step // 2, H, step))[:rows], dtype=np.int32)

    def voxel_passes():
        t0 = time.perf_counter()
        _, rad = O.clipmap_voxel_passes(scene, regions, descs, depth, st["light"], RES, LV, 10)
        return time.perf_counter() - t0, rad

    def trace_rows(rad):
        t0 = time.perf_counter()
        rd, rs, n = O.trace_rows(rad, RES, LV, g, p, depth, st["light"], row_list)
        return time.perf_counter() - t0, n, rd, rs

    t_pass, t_rows = [], []
    for run in range(runs + 1):  # run 0 is the warm-up
        tp, rad = voxel_passes()
        tt, nsamp, rd, rs = trace_rows(rad)
        if run:
            t_pass.append(tp)
            t_rows.append(tt)
    tp, tt = statistics.median(t_pass), statistics.median(t_rows)
    full_ms = (tp + tt * H / len(row_list)) * 1e3
    # the GPU frame on the same rows: FP32 oracle, and FP64 oracle on every 4th row
    dout = st["dout"].cpu().numpy()[row_list]
    sout = st["sout"].cpu().numpy()[row_list]
    parity = dict(rows=int(len(row_list)), diffuse_rel_l2_vs_fp32=_rel_l2(dout, rd[row_list]),
                  specular_rel_l2_vs_fp32=_rel_l2(sout, rs[row_list]))
    sub = row_list[::4]
    d64, s64, _ = O.trace_rows(rad, RES, LV, g, p, depth, st["light"], sub, fp64=True)
    pos_sub = [int(np.where(row_list == r)[0][0]) for r in sub]
    parity.update(fp64_rows=int(len(sub)), diffuse_rel_l2_vs_fp64=_rel_l2(dout[pos_sub], d64[sub]),
                  specular_rel_l2_vs_fp64=_rel_l2(sout[pos_sub], s64[sub]),
                  oracle_fp32_diffuse_rel_l2_vs_fp64=_rel_l2(rd[sub], d64[sub]),
                  oracle_fp32_specular_rel_l2_vs_fp64=_rel_l2(rs[sub], s64[sub]))
    hc = host_cpu()
    return dict(value=round(full_ms, 1), unit="ms/frame", cores=hc["omp_threads"], kind="port",
                sample=f"all 6-level voxel passes (median {tp:.2f} s; voxelization and injection one thread per "
                       f"level, mips {hc['omp_threads']} OpenMP threads) + {len(row_list)} of {H} trace rows "
                       f"(median {tt:.2f} s on {hc['omp_threads']} OpenMP threads, {nsamp} cone samples), scaled "
                       f"to the full 1920x1080 frame; 1 warm-up + median of {runs} runs",
                cone_samples_per_s=round(nsamp / tt, 1) if tt > 0 else None,
                runs_s=dict(voxel_passes=[round(x, 3) for x in t_pass], trace_rows=[round(x, 3) for x in t_rows]),
                host=hc, frame_parity=parity)


def _crop_rows(g: dict, vp_inv, r0: int, h: int):
    """Rows [r0, r0 + h) of a host G-buffer as a G-buffer of their own, with
    the inverse view-projection that gives its pixels the full image's rays
    (NDC y of the crop maps to (h y + 2 r0 + h) / H - 1 of the full image)."""
    H = g["depth"].shape[0]
    sub = {k: np.ascontiguousarray(v[r0:r0 + h]) for k, v in g.items()}
    S = np.eye(4)
    S[1, 1] = h / H
    S[1, 3] = (2 * r0 + h) / H - 1.0
    return sub, np.asarray(vp_inv, np.float64) @ S


def _host_g(g) -> dict:
    out = {}
    for k, v in g.items():
        a = v.cpu().numpy()
        out[k] = a.view(np.uint32) if a.dtype == np.int32 else (a.view(np.uint16) if a.dtype == np.int16 else a)
    return out


def workload_cpu_baselines() -> dict:
    """The oracle on bounded samples of the other workloads' own inputs
    (OpenMP where the oracle has it), each scaled to the workload's unit:
    plain VXAO (voxelize + mips + the full 640x360 FP64 trace), the 512^3
    octree (fragment list, pool and bricks of the 1M-triangle scene), 4 of the
    64 batched scenes (voxelize + mips + 64 of 512 trace rows each, scaled to
    all rows and scenes) and the adjoint (level 0 + mips + one view's forward
    and backward over 32 of 512 rows, scaled to 16 views)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_api as O
    from vk_voxel_cone_tracing_amd import workloads as wl

    hc = host_cpu()
    out = {}

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return time.perf_counter() - t0, r

    def dense_frame(scene, gmin, ext, res, g, vp_inv, td, rows=None):
        pos, nrm, mat = scene.world_triangles()
        d, _lv = O.new_dense(res, O.FMT_RGBA8, gmin, ext)
        tv, _ = timed(lambda: O.lib().orc_dense_voxelize(C.byref(d), C.byref(O.tris_struct(pos, nrm, mat)),
                                                         O.materials_struct(scene), 1))
        tm, _ = timed(lambda: O.lib().orc_dense_mip(C.byref(d)))
        H = g["depth"].shape[0]
        gg, vpi = (g, vp_inv) if rows is None else _crop_rows(g, vp_inv, rows[0], rows[1])
        p = O.dense_params(vpi, aperture=td.aperture, max_dist=td.max_distance, decay=td.occlusion_decay,
                           ao_factor=td.ambient_occlusion_factor, diffuse_intensity=td.indirect_diffuse_intensity)
        h = gg["depth"].shape[0]
        r64 = np.zeros((h, gg["depth"].shape[1], 4), np.float64)
        n = C.c_uint64(0)
        tt, _ = timed(lambda: O.lib().orc_dense_trace(C.byref(d), C.byref(O.gbuf_struct(gg)), C.byref(p), 0, h,
                                                      O.vptr(r64), C.byref(n)))
        return tv + tm + tt * H / h, n.value * H / h

    try:
        inp = wl.plain_vxao_inputs()
        g = _host_g(inp["g"])
        t, n = dense_frame(inp["scene"], inp["gmin"], inp["ext"], inp["res"], g, inp["vp_inv"], inp["td"])
        out["plain_vxao_640x360"] = dict(value=round(t * 1e3, 1), unit="ms/frame", cores=hc["omp_threads"],
                                         kind="port", cone_samples=int(n),
                                         sample="the whole frame: dense voxelize + mips (one thread), FP64 trace "
                                                "of all 640x360 pixels (OpenMP over rows)")
    except Exception as e:
        out["plain_vxao_640x360"] = {"error": f"{type(e).__name__}: {e}"}
    try:
        inp = wl.svo_inputs()
        scene = inp["scene"]
        pos, nrm, mat = scene.world_triangles()
        depth = inp["depth"].cpu().numpy()
        osm = O.Shadow(depth.ctypes.data, depth.shape[1], depth.shape[0])
        level, cap = 9, 16_000_000
        fr = np.zeros((cap, 2), np.uint32)
        tf, nf = timed(lambda: O.lib().orc_svo_fragments(
            C.byref(O.tris_struct(pos, nrm, mat)), O.materials_struct(scene), C.byref(osm),
            C.byref(O.light_struct(inp["light"])), (C.c_float * 3)(*scene.bb_min.tolist()),
            (C.c_float * 3)(*scene.bb_max.tolist()), 1 << level, 1, 0x3F, O.vptr(fr), cap))
        pool = np.zeros((max(1_000_000, 8 * nf), 2), np.uint32)
        cnt = C.c_uint32(0)
        tb, _ = timed(lambda: O.lib().orc_svo_build(O.vptr(fr), nf, level, 0x3F, O.vptr(pool), len(pool),
                                                    C.byref(cnt)))
        O.lib().orc_svo_bricks.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
        bricks = np.zeros((cnt.value, 27), np.uint32)
        tk, _ = timed(lambda: O.lib().orc_svo_bricks(pool.ctypes.data, cnt.value, level, bricks.ctypes.data))
        inp["ctx"].close()
        out["svo_512"] = dict(value=round((tf + tb + tk) * 1e3, 1), unit="ms/build", cores=hc["omp_threads"],
                              kind="port", fragments=int(nf), nodes=int(cnt.value),
                              parts_s=dict(fragments=round(tf, 2), pool=round(tb, 2), bricks=round(tk, 2)),
                              sample="the whole build: fragment list (one thread, fragment order), node pool "
                                     "(one thread), 3^3 bricks (OpenMP over nodes)")
    except Exception as e:
        out["svo_512"] = {"error": f"{type(e).__name__}: {e}"}
    try:
        from vk_voxel_cone_tracing_amd import scenes as sc

        items, td = wl.batched_inputs()
        order = sorted(range(len(items)), key=lambda i: items[i]["scene"].triangle_count)
        pick = [order[k] for k in (0, len(order) // 3, 2 * len(order) // 3, len(order) - 1)]
        tot, ns = 0.0, 0.0
        for i in pick:
            it = items[i]
            g = _host_g(it["g"])
            vp_inv = sc.camera(3000 + i, 512, 512)[1]  # the camera batched_inputs gives scene i
            t, n = dense_frame(it["scene"], it["gmin"], it["ext"], it["res"], g, vp_inv, td, rows=(224, 64))
            tot += t
            ns += n
        out["batched_64"] = dict(value=round(tot * len(items) / len(pick) * 1e3, 1), unit="ms/batch",
                                 cores=hc["omp_threads"], kind="port",
                                 sample=f"scenes {pick} (smallest, thirds and largest by triangle count): "
                                        "voxelize + mips + trace of rows 224-287 of 512 each (FP64, OpenMP over "
                                        "rows), scaled to 512 rows and to 64 scenes")
    except Exception as e:
        out["batched_64"] = {"error": f"{type(e).__name__}: {e}"}
    try:
        inp = wl.adjoint_inputs()
        res = inp["res"]
        a = O.Adjoint(res, 1.0, (0.0, 0.0, 0.0))
        rgba = np.concatenate([inp["albedo"].cpu().numpy(), inp["opacity"].cpu().numpy()[..., None]], -1)
        tl, _ = timed(lambda: a.set_level0(rgba))
        g = {k: _host_g({k: inp[k][0]})[k] for k in ("diffuse", "normal", "depth")}
        g["specular"] = g["diffuse"]
        g["emission"] = g["normal"]
        vp_inv = inp["cams"][0][1]
        gg, vpi = _crop_rows(g, vp_inv, 240, 32)
        pp = inp["params"]
        P = O.dense_params(vpi, aperture=pp.aperture, max_dist=pp.max_distance, step_factor=pp.min_trace_step_factor,
                           start_offset=pp.trace_start_offset, decay=pp.occlusion_decay,
                           ao_factor=pp.ambient_occlusion_factor, diffuse_intensity=pp.indirect_diffuse_intensity)
        tr, img = timed(lambda: a.render(gg, P))
        tg, _ = timed(lambda: a.backward(gg, P, np.ones_like(img)))
        views, H = inp["depth"].shape[0], inp["depth"].shape[1]
        total = tl + (tr + tg) * views * H / gg["depth"].shape[0]
        out["adjoint_256"] = dict(value=round(total * 1e3, 1), unit="ms/(forward+backward)",
                                  cores=hc["omp_threads"], kind="port",
                                  parts_s=dict(level0_mips=round(tl, 2), render_rows=round(tr, 2),
                                               backward_rows=round(tg, 2)),
                                  sample="FP64 oracle adjoint: level 0 + mips once, then rows 240-271 of view 0 "
                                         "(render and backward, one thread), scaled to 512 rows x 16 views")
    except Exception as e:
        out["adjoint_256"] = {"error": f"{type(e).__name__}: {e}"}
    for k, v in out.items():
        print(f"[bench] cpu {k}: {json.dumps(v)}", file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    import torch

    torch.cuda.set_device(0)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    st = setup()
    for _ in range(args.warmup):
        frame(st)
    torch.cuda.synchronize()
    # timed region: exactly K frames, inputs resident in HBM
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame(st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_frame = elapsed * 1e3 / args.steps
    # per-pass device times of one more frame, and the samples it executed
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    trace_ms = []
    for _ in range(3):
        frame(st, ev)
        torch.cuda.synchronize()
        trace_ms.append(ev[2].elapsed_time(ev[3]))
    vox_ms = ev[0].elapsed_time(ev[1])
    inj_ms = ev[1].elapsed_time(ev[2])
    t_trace = sorted(trace_ms)[1]
    samples = int(st["vct"].samples.item())
    pixels = st["W"] * st["H"]
    # The trace pass's bound is the vector ALU: its texel gathers are served
    # by L2 (94-98 % hits), so the bytes they name are not HBM bytes. The
    # roofline is therefore FP32: the algorithmic FLOPs of the trilinear
    # filtering of every executed sample (counted by the kernels) over the
    # pass time, against the FP32 vector peak. The measured HBM bytes (from
    # rocprofv3 FETCH_SIZE / WRITE_SIZE) and the gathered bytes are reported
    # beside it, each against the 8 TB/s HBM peak.
    flops = samples * FLOP_PER_SAMPLE
    achieved_tflops = flops / (t_trace * 1e-3) / 1e12
    gather_bytes = samples * B_SAMPLE_RGBA16F + pixels * B_PIXEL
    traffic = None
    tf = ROOT / "profiles" / "trace_clipmap_1080p_traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    valu = None
    vf = ROOT / "profiles" / "trace_clipmap_1080p_valu.json"
    if vf.exists():
        try:
            valu = json.loads(vf.read_text()).get("summary")
        except Exception:
            valu = None
    vmem = None
    mf = ROOT / "profiles" / "trace_clipmap_1080p_vmem.json"
    if mf.exists():
        try:
            vmem = json.loads(mf.read_text()).get("summary")
        except Exception:
            vmem = None
    roofline = {
        "bound": "valu",
        "kernel": "cone-trace pass: k_spec_bin + k_trace<half,split> + k_trace_specular<half>",
        "achieved": round(achieved_tflops, 2),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4),
        "traffic": traffic,
        "algorithmic_flops_per_launch": flops,
        "flops_per_unit": f"{FLOP_PER_SAMPLE} FLOP per executed cone sample (trilinear FMAs of 2 levels x 3 faces "
                          f"x 8 texels x 4 channels)",
        "hbm": {
            "bytes_per_launch": traffic,
            "achieved_gbs": round(traffic / (t_trace * 1e-3) / 1e9, 1) if traffic else None,
            "frac": round(traffic / (t_trace * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if traffic else None,
            "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of the pass, scaled by a plain copy of known size",
        },
        "gather": {
            "bytes_per_launch": gather_bytes,
            "achieved_gbs": round(gather_bytes / (t_trace * 1e-3) / 1e9, 1),
            "note": f"{B_SAMPLE_RGBA16F} B of texels per sample + {B_PIXEL} B per pixel, requested from the "
                    f"cache hierarchy (98.5% L1 hits), not HBM bytes",
        },
        "valu_issue": valu,
        "vector_memory": vmem,
    }
    out = {
        "metric": METRIC,
        "value": round(ms_per_frame, 3),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "clipmap_gi_1080p: 6x128^3 clipmap (RGBA16F radiance, RGBA8 opacity), 2048^2 shadow map, "
                        "1920x1080 CombinedGI trace (16 diffuse + 1 specular cone), seeded room x16, 21260 tris",
            "frame_passes": "voxelize+opacity mips+border, inject+copy-alpha+radiance mips (all levels), trace",
        },
        "cone_samples_per_frame": samples,
        "cone_samples_per_s": round(samples / (t_trace * 1e-3), 1),
        "passes_ms": {"voxelization": round(vox_ms, 3), "radiance_injection": round(inj_ms, 3),
                      "cone_trace": round(t_trace, 3)},
        "roofline": roofline,
    }
    if not args.no_extras:
        out["workloads"] = extras()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(st, args.cpu_rows)
        if not args.no_extras:
            out["cpu_baseline"]["workloads"] = workload_cpu_baselines()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
