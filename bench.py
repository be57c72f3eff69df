"""Optical-stage throughput on one MI355X: sites/s for the 4,225-site plain stage.

One step = the reference's optical-stage timing boundary (maomi/fixed.py:337-378):
phase preparation + beam preparation + propagation of every site + mask assembly,
with the refractive-index volume already resident in HBM.  Workload: 512 x 512 x 60
index volume (n0 = 1.35 + Gaussian random field, std 0.02, correlation 3 voxels),
65 x 65 sites, 256 x 256 complex64 field per site, 60 slices, Zernike pupil, FP32.

Prints one JSON line with the throughput, the roofline of the persistent sites kernel
(timed with HIP events on its stream) and a same-run CPU baseline (the NumPy/SciPy
restatement of the reference CPU path, on a bounded sample of sites, extrapolated).
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
# HBM bytes per step of sites_kernel<256> on the 4225-site FP32 workload (its two launches: the
# whole rounds and the last round), from two rocprofv3 --pmc runs of this script: FETCH_SIZE x
# 2.0 + WRITE_SIZE x 1.0 (factors from a 1 GiB copy, --copy-probe); the runs and their commands
# are in profiles/sites_kernel_256_config2_pmc.md.
TRAFFIC_CONFIG2_FP32 = (2.0 * 1.108e8 + 1.0 * 1.958e8) * 1024
# The same for segmented_sites_kernel on the reference demo geometry (4225 sites), from
# profiles/segmented_sites_kernel_reference_geometry.md.
TRAFFIC_REF_DEMO = (2.0 * 1.945e7 + 1.0 * 3.474e7) * 1024
METRIC = "sites/s on 4225-site optical stage; surrogate vs propagator at ≤1% rel L2"


def affinity_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_cores():
    """Threads for the CPU baseline: every core this process may run on (sched_getaffinity),
    capped by the host's declared CPU share (OMP_NUM_THREADS) where one is set."""
    n = affinity_cores()
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def host_description():
    """CPU model (lscpu's "Model name", from /proc/cpuinfo) and the FFT library of the baseline."""
    import scipy
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "affinity_cores": affinity_cores(),
            "fft_library": f"scipy.fft (pocketfft) from scipy {scipy.__version__}, "
                           f"numpy {np.__version__}"}


def cpu_baseline(cfg, vol, offsets, zw, n_sites, q_sample, cores):
    """Time the oracle (complex64 restatement of the reference CPU path) on a sample."""
    import oracle as O
    O.set_workers(cores)
    t0 = time.perf_counter()
    base, _ = O.plain_beam(cfg, zw)
    k0dz = 2 * np.pi / cfg.wavelength_um * cfg.dz_um / cfg.voxels_per_slice
    phase = O.plain_phase(vol, cfg.n0, k0dz, cfg.voxels_per_slice)
    t_plan = time.perf_counter() - t0
    rng = np.random.default_rng(cfg.seed + 3)
    pick = np.sort(rng.choice(len(offsets), min(n_sites, len(offsets)), replace=False))
    t0 = time.perf_counter()
    masks = O.plain_stage_masks(base, phase, offsets[pick], wvl=np.float32(cfg.wavelength_um * 1e-6),
                                nr=np.float32(cfg.n0), dx=np.float32(cfg.dx_um * 1e-6),
                                dz=cfg.dz_um * 1e-6)
    t_sites = time.perf_counter() - t0
    energies = np.full(len(offsets), float(np.sum(masks[0], dtype=np.float64)))
    # v4: full FP64 solve, query evaluation timed on a sample and scaled to the FOV
    from oracle.optics_np import _green
    from scipy.spatial.distance import cdist
    p = offsets.astype(np.float64)
    t0 = time.perf_counter()
    coef = np.linalg.solve(_green(p, p), energies)
    t_solve = time.perf_counter() - t0
    q = np.stack(np.meshgrid(np.arange(cfg.fov) + 0.5, np.arange(cfg.fov) + 0.5, indexing="ij"),
                 -1).reshape(-1, 2)
    qs = q[:q_sample]
    t0 = time.perf_counter()
    chunk = max(1, 8_388_608 // len(p))
    for s in range(0, len(qs), chunk):
        d2 = cdist(qs[s:s + chunk], p, "sqeuclidean")
        with np.errstate(divide="ignore", invalid="ignore"):
            b = np.log(d2)
            b *= 0.5
            b -= 1
            b *= d2
        b[d2 == 0] = 0
        _ = b @ coef
    t_eval = (time.perf_counter() - t0) * len(q) / len(qs)
    total = t_plan + t_sites / len(pick) * len(offsets) + t_solve + t_eval
    host = host_description()
    return dict(value=len(offsets) / total, unit="sites/s", cores=cores, kind="port",
                sample=(f"{len(pick)} of {len(offsets)} seeded sites propagated ({cores} "
                        f"threads of {host['affinity_cores']} in the affinity mask, serial "
                        f"site loop, shifted FFT sequence), full plan ({t_plan:.2f} s) and v4 "
                        f"solve ({t_solve:.2f} s), v4 evaluation on {len(qs)} of {len(q)} "
                        f"pixels; sites {t_sites:.2f} s, per-site latency "
                        f"{t_sites / len(pick):.3f} s; extrapolated total {total:.1f} s"),
                measured_sites_seconds=t_sites, extrapolated_seconds=total,
                site_latency_seconds=t_sites / len(pick), **host)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="2",
                    choices=("1", "2", "3", "4", "5", "5infer", "ref", "inputs"),
                    help="1-5: the plain-stage workloads and the surrogate; ref: the reference's "
                         "own A/B/C geometry (demo scene) through run_fixed; inputs: raw index "
                         "fields and collection mask of a 256 x 256 x 50 um volume")
    ap.add_argument("--train-volumes", type=int, default=5,
                    help="surrogate (config 5): training volumes of 4225 sites each")
    ap.add_argument("--epochs", type=int, default=6, help="surrogate (config 5) epochs")
    ap.add_argument("--accuracy-sites", type=int, default=0,
                    help="FP64-oracle rel L2 on this many sampled sites (config 2/4)")
    ap.add_argument("--cpu-sites", type=int, default=64)
    ap.add_argument("--cpu-queries", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-overlap-v4", dest="overlap_v4", action="store_false",
                    help="configs 1/2/4: one sites launch, the whole v4 solve after it")
    ap.add_argument("--precision", default="fp32", choices=("fp32", "mixed"))
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run three device copies of GIB gibibytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the stage shards by sites; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    from maomi_amd.plain import PlainStage
    from maomi_amd.workloads import CONFIG1, CONFIG2, index_volume, site_offsets, zernike_weights

    if args.copy_probe > 0:
        n = int(args.copy_probe * (1 << 30)) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for i in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dst.copy_(src)
            torch.cuda.synchronize()
            print(f"[bench] copy {i}: {4 * n} bytes in {time.perf_counter() - t0:.6f} s",
                  file=sys.stderr, flush=True)
        return

    if args.config == "ref":
        return bench_reference(args)
    if args.config == "inputs":
        return bench_inputs(args)
    if args.config == "5infer":
        return bench_surrogate_inference(args)
    args.config = int(args.config)
    if args.config == 3:
        return bench_uneven(args)
    if args.config == 5:
        return bench_surrogate(args)
    if args.config == 4:
        args.precision = "mixed"
    cfg = CONFIG1 if args.config == 1 else CONFIG2
    vol_np = index_volume(cfg)
    offsets = site_offsets(cfg)
    dev = torch.device("cuda:0")
    vol = torch.as_tensor(vol_np, device=dev)
    st = PlainStage(cfg, vol, device=dev, precision=args.precision)
    print(f"[bench] config {args.config}: {len(offsets)} sites, N={cfg.n}, S={cfg.slices}, "
          f"{args.precision}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        st.run()
        print(f"[bench] warmup {i}: {st.timing}", file=sys.stderr, flush=True)

    # kernel time of the persistent launch, HIP events on the launch stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stage_s = {"plan": [], "propagation": [], "aggregation": []}
    for i in range(args.steps):
        ts = time.perf_counter()
        st.plan()
        torch.cuda.synchronize()
        tp = time.perf_counter()
        ev0.record()
        st.propagate_all(overlap=args.overlap_v4)
        ev1.record()
        torch.cuda.synchronize()
        tq = time.perf_counter()
        st.aggregate()
        torch.cuda.synchronize()
        te = time.perf_counter()
        kernel_ms.append(ev0.elapsed_time(ev1))
        for k, v in (("plan", tp - ts), ("propagation", tq - tp), ("aggregation", te - tq)):
            stage_s[k].append(v)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    P = len(offsets)
    value = P * args.steps / elapsed
    kms = float(np.mean(kernel_ms))
    # the same steps with the index volume uploaded from pinned host memory inside the plan,
    # as the reference's plan uploads its raw fields (batched.py:52-60); reported beside value
    pinned = torch.from_numpy(vol_np).pin_memory()
    torch.cuda.synchronize()
    up_ms = []
    t1 = time.perf_counter()
    for i in range(args.steps):
        ev0.record()
        st.volume.copy_(pinned, non_blocking=True)
        ev1.record()
        st.run()
        up_ms.append(ev0.elapsed_time(ev1))
    torch.cuda.synchronize()
    elapsed_up = time.perf_counter() - t1
    half = args.precision == "mixed"
    bytes_launch = cfg.bytes_per_site(half) * P
    achieved = bytes_launch / (kms * 1e-3) / 1e9
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s; sites kernel {kms:.2f} ms", file=sys.stderr,
          flush=True)

    accuracy = None
    if args.accuracy_sites > 0:
        accuracy = fp64_accuracy(cfg, vol_np, offsets, st, args.accuracy_sites)
        print(f"[bench] accuracy vs FP64 oracle: {accuracy}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cores = cpu_cores()
        cpu = cpu_baseline(cfg, vol_np, offsets, zernike_weights(cfg), args.cpu_sites,
                           args.cpu_queries, cores)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 screens, f32 fields" if half else "f32",
        "data": "synthetic (seeded Gaussian random index field; no downloads)",
        "config": {
            "workload": (f"plain optical stage: {cfg.padded - cfg.n}x{cfg.padded - cfg.n}x{cfg.slices} "
                         f"index volume, {P} sites, {cfg.n}x{cfg.n} complex64 field, "
                         f"{cfg.slices} slices, Zernike pupil" if cfg.zernike_std_rad > 0 else ""),
            "sites": P, "field": cfg.n, "slices": cfg.slices, "fov": cfg.fov,
            "precision": args.precision,
            "timing_boundary": "plan + propagation + aggregation (fixed.py:337-378)",
        },
        "stage_seconds": {k: round(float(np.mean(v)), 6) for k, v in stage_s.items()},
        "with_volume_upload": {
            "sites_per_s": P * args.steps / elapsed_up,
            "ms_per_step": elapsed_up / args.steps * 1e3,
            "upload_ms": float(np.mean(up_ms)), "bytes": int(vol_np.nbytes),
            "note": "index volume copied host (pinned) -> HBM at the start of every step"},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": TRAFFIC_CONFIG2_FP32 if (cfg is CONFIG2 and not half) else None,
            "kernel": "sites_kernel<%d>" % cfg.n, "kernel_ms": kms,
            "algorithmic_bytes_per_launch": bytes_launch,
        },
        "cpu_baseline": cpu,
    }
    if accuracy is not None:
        line["accuracy"] = accuracy
    print(json.dumps(line), flush=True)


def fp64_accuracy(cfg, vol_np, offsets, st, k):
    """rel L2 of k seeded sampled site masks against the FP64 (complex128) oracle."""
    import oracle as O
    from maomi_amd.workloads import zernike_weights
    O.set_workers(cpu_cores())
    rng = np.random.default_rng(cfg.seed + 3)
    pick = np.sort(rng.choice(len(offsets), min(k, len(offsets)), replace=False))
    ref = O.plain_stage(cfg, vol_np.astype(np.float64), offsets, zernike_weights(cfg),
                        dtype=np.complex128, sites=pick, with_masks=False)["masks"]
    got = st.masks[torch_index(pick)].cpu().numpy()
    per = [float(np.linalg.norm(g - r) / np.linalg.norm(r)) for g, r in zip(got, ref)]
    return {"reference": "FP64 oracle (complex128 restatement)", "sites": len(pick),
            "rel_l2_all": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
            "rel_l2_max_site": max(per)}


def torch_index(idx):
    import torch
    return torch.as_tensor(np.asarray(idx), dtype=torch.long)


def bench_uneven(args):
    """One persistent launch for 4225 sites of mixed field side and depth (captured graph)."""
    import torch
    from maomi_amd.uneven import CONFIG3, UnevenStage, bytes_per_site, draw_sites
    from maomi_amd.workloads import index_volume, site_offsets
    cfg = CONFIG3.base
    sizes, slices = draw_sites(CONFIG3)
    centres = site_offsets(cfg)
    vol_np = index_volume(cfg)
    vol = torch.as_tensor(vol_np, device="cuda:0")
    st = UnevenStage(cfg, vol, centres, sizes, slices).plan(overlap_v4=False)
    tuned = st.autotune()
    print(f"[bench] autotune: {st.tune_log}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        st.run()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kms = []
    stage_s = {"plan": [], "propagation": [], "aggregation": []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st.plan(overlap_v4=args.overlap_v4)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        ev0.record()
        st.propagate()
        ev1.record()
        torch.cuda.synchronize()
        tq = time.perf_counter()
        st.aggregate()
        torch.cuda.synchronize()
        te = time.perf_counter()
        kms.append(ev0.elapsed_time(ev1))
        for k, v in (("plan", tp - ts), ("propagation", tq - tp), ("aggregation", te - tq)):
            stage_s[k].append(v)
    elapsed = time.perf_counter() - t0
    P = len(sizes)
    kms = float(np.mean(kms))
    bl = float(bytes_per_site(sizes, slices).sum())
    ach = bl / (kms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_uneven(cfg, vol_np, st.windows, sizes, slices,
                                  st.centres)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)
    line = {
        "metric": METRIC, "value": P * args.steps / elapsed, "unit": "sites/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian random index field; no downloads)",
        "config": {"workload": ("uneven plain stage: 4225 sites, field side 128/256/512 "
                                "(p 0.5/0.4/0.1), depth U(20,120) um -> 10..60 slices, one "
                                "persistent launch in a captured graph"),
                   "sizes": {str(n): int(np.sum(sizes == n)) for n in (128, 256, 512)},
                   "autotuned": tuned},
        "stage_seconds": {k: round(float(np.mean(v)), 6) for k, v in stage_s.items()},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": ach / PEAK_HBM_GBS, "traffic": None, "kernel": "sites_mixed_kernel",
                     "kernel_ms": kms, "algorithmic_bytes_per_launch": bl},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_uneven(cfg, vol, windows, sizes, slices, centres, per_class=3):
    """The oracle on a stratified sample of the uneven sites: per field side, `per_class`
    seeded sites (the deepest among them), timed per slice; the per-side rate times each
    site's slice count gives the extrapolated total, plus the full v4 solve and a sampled
    evaluation as in cpu_baseline."""
    import oracle as O
    from dataclasses import replace
    from maomi_amd.workloads import zernike_weights
    cores = cpu_cores()
    O.set_workers(cores)
    zw = zernike_weights(cfg)
    t0 = time.perf_counter()
    k0dz = 2 * np.pi / cfg.wavelength_um * cfg.dz_um / cfg.voxels_per_slice
    phase = O.plain_phase(vol, cfg.n0, k0dz, cfg.voxels_per_slice)
    t_plan = time.perf_counter() - t0
    rng = np.random.default_rng(cfg.seed + 6)
    per_slice, sampled = {}, []
    for n in np.unique(sizes):
        cls = np.flatnonzero(sizes == n)
        pick = rng.choice(cls, min(per_class, len(cls)), replace=False)
        t, z = 0.0, 0
        for i in pick:
            sub = replace(cfg, n=int(n), slices=int(slices[i]), focus_um=slices[i] * cfg.dz_um)
            t1 = time.perf_counter()
            base, _ = O.plain_beam(sub, zw)
            O.plain_stage_masks(base, phase, windows[i:i + 1], wvl=np.float32(cfg.wavelength_um * 1e-6),
                                nr=np.float32(cfg.n0), dx=np.float32(cfg.dx_um * 1e-6),
                                dz=cfg.dz_um * 1e-6, sizes=[int(n)], slices=[int(slices[i])])
            t += time.perf_counter() - t1
            z += int(slices[i])
        per_slice[int(n)] = t / z
        sampled.append(f"{len(pick)} of {len(cls)} N={int(n)} sites ({t:.2f} s)")
    t_sites = float(sum(per_slice[int(n)] * int(s) for n, s in zip(sizes, slices)))
    from oracle.optics_np import _green
    p = centres.astype(np.float64)
    t0 = time.perf_counter()
    coef = np.linalg.solve(_green(p, p), np.ones(len(p)))
    t_solve = time.perf_counter() - t0
    total = t_plan + t_sites + t_solve
    host = host_description()
    return dict(value=len(sizes) / total, unit="sites/s", cores=cores, kind="port",
                sample=(f"stratified: {', '.join(sampled)}, timed per slice and extrapolated "
                        f"over every site's slice count ({t_sites:.1f} s), phase plan "
                        f"({t_plan:.2f} s) and v4 solve ({t_solve:.2f} s); {cores} threads"),
                extrapolated_seconds=total, **host)


def bench_reference(args):
    """run_fixed on the reference's own per-site geometry (demo scene: 16 x 16 x 10 um volume,
    20 um deep, 2 voxels/um, supersampling 2: fields 224^2 (A, C) and 144^2 (B), 4 + 20 + 2
    steps) at 65 x 65 = 4225 sites; one step = plan + propagation + aggregation."""
    import torch
    from maomi_amd._propagation import geometry
    from maomi_amd.fixed import run_fixed
    from maomi_amd.optics import OpticalConfig, prepare_fields
    kw = dict(volume_um=(16.0, 16.0, 10.0), depth_um=20.0, vres=2.0, tail_um=10.0)
    cfg = OpticalConfig(**kw)
    raw = prepare_fields(cfg, None)
    g = geometry(cfg, raw)
    steps = [len(z) - 1 for z in g.z]
    sides = (g.n, g.n2, g.n)
    for _ in range(args.warmup):
        run_fixed(cfg, None, raw_fields=raw, count_per_axis=65, render_image=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stages = []
    for _ in range(args.steps):
        res = run_fixed(cfg, None, raw_fields=raw, count_per_axis=65, render_image=False)
        stages.append(res["timing_seconds"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    P = len(res["points_um"])
    # the persistent launch's duration, HIP events on its stream
    kern = float(np.mean([t["propagation_kernel"] for t in stages]))
    # algorithmic bytes per site: each step reads and writes the field once and reads its
    # FP32 screen, plus the entry field; the focal PSF and both tail images are written once
    per_site = sum((20 * S + 8) * N * N for S, N in zip(steps, sides))
    per_site += 4 * int(np.prod(res["psf_3d"].shape[1:])) + 8 * int(np.prod(res["psfT"]["convmask"].shape))
    ach = per_site * P / kern / 1e9
    cpu = None
    if not args.no_cpu:
        import oracle as O
        O.set_workers(cpu_cores())
        ocfg = O.OracleConfig(**kw)
        pts = np.asarray(res["points_um"])[np.random.default_rng(5).choice(P, 6, replace=False)]
        t1 = time.perf_counter()
        O.propagate(ocfg, [np.asarray(r, np.float32) for r in raw], pts)
        per = (time.perf_counter() - t1) / len(pts)
        cpu = {"value": 1.0 / per, "unit": "sites/s", "cores": cpu_cores(), "kind": "port",
               "sample": f"6 of {P} sites through the oracle's propagate (scipy.fft, "
                         f"{cpu_cores()} threads, serial site loop), {per:.3f} s per site"}
    line = {
        "metric": METRIC, "value": P * args.steps / elapsed, "unit": "sites/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded raw index fields of the reference's generator; no downloads)",
        "config": {"workload": ("reference A/B/C geometry through run_fixed: demo scene "
                                f"{kw['volume_um']} um, {P} sites, fields {g.n}^2/{g.n2}^2, "
                                f"steps {steps[0]}/{steps[1]}/{steps[2]}, both tails, v4 masks"),
                   "reference_published": {"value": 1153.7, "unit": "sites/s",
                                           "hardware": "RTX 3080",
                                           "note": "4225 sites; the OpticalConfig behind it is "
                                                   "not published, so no ratio is claimed"}},
        "stage_seconds": {k: round(float(np.mean([t[k] for t in stages])), 6)
                          for k in ("plan", "propagation", "propagation_kernel",
                                    "aggregation")},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": ach / PEAK_HBM_GBS,
                     "traffic": TRAFFIC_REF_DEMO if P == 4225 else None,
                     "kernel": "segmented_sites_kernel",
                     "kernel_ms": kern * 1e3, "algorithmic_bytes_per_launch": per_site * P},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


INPUTS_SCENE = dict(volume_um=(256.0, 256.0, 50.0), depth_um=100.0, vres=1.0, tail_um=20.0)


def bench_inputs(args):
    """Input generation (maomi/optics.py:120-197, 384-415): the seeded raw A/B/C index fields
    with vessel occupancy and the collection mask for a 256 x 256 x 50 um volume 100 um deep at
    1 voxel/um (fields 532^2 x 24, 292^2 x 10, 532^2 x 4; vessels 532^2 x 150, 3% occupied,
    resident in HBM).  Noise is drawn on the GPU.  One step = prepare_fields + collection_mask;
    the unit is raw-field voxels per second."""
    import torch
    from maomi_amd.optics import OpticalConfig, collection_mask, prepare_fields
    cfg = OpticalConfig(**INPUTS_SCENE)
    shape = np.rint(np.array(cfg.parameters()[0].vasc_sz) * cfg.vres).astype(int)
    shape[2] = max(shape[2], 150)
    vessels_np = np.random.default_rng(7).random(shape) < 0.03
    vessels = torch.as_tensor(vessels_np, device="cuda:0")
    for _ in range(args.warmup):
        f = prepare_fields(cfg, vessels, device="cuda", out_device="cuda")
        collection_mask(vessels, cfg)
    voxels = sum(int(np.prod(a.shape)) for a in f)
    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f = prepare_fields(cfg, vessels, device="cuda", out_device="cuda", events=events)
        m = collection_mask(vessels, cfg, events=events)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # spectrum kernel: reads the two float32 draws, writes complex64; finish kernel: reads the
    # complex64 field, writes float32 (+ float32 occupancy read); the collection-mask launch
    # reports its own byte count (vessel slabs, FP64 prefix and partial tables, output)
    per_voxel = {"grf_spectrum": 16, "grf_finish": 12, "grf_finish_occupancy": 16,
                 "collection_mask": 1}
    kern = {}
    for name, dims, (a, b) in events:
        k = kern.setdefault(name, [0.0, 0])
        k[0] += a.elapsed_time(b) * 1e-3
        k[1] += per_voxel[name] * int(np.prod(dims))
    top = max(kern, key=lambda k: kern[k][0])
    ach = kern[top][1] / kern[top][0] / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import OracleConfig
        from oracle import inputs_np as I
        ocfg = OracleConfig(**INPUTS_SCENE)
        t = time.perf_counter()
        with torch.random.fork_rng(devices=[]):
            torch.manual_seed(cfg.seed)
            noises = [(torch.randn(s).numpy(), torch.randn(s).numpy())
                      for s in I.field_shapes(ocfg)]
        I.prepare_fields(ocfg, noises, vessels=vessels_np)
        I.collection_mask(vessels_np, ocfg)
        tc = time.perf_counter() - t
        cpu = dict(value=voxels / tc, unit="voxels/s", cores=1, kind="port",
                   sample=(f"the whole workload once: torch CPU normal draws, NumPy FP64 "
                           f"envelope and inverse FFT, occupancy mix, scipy fftconvolve "
                           f"collection disks ({tc:.2f} s)"))
    line = {
        "metric": "input generation voxels/s (raw A/B/C index fields + collection mask)",
        "value": voxels * args.steps / elapsed, "unit": "voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 disk sums)",
        "data": "synthetic (seeded normal draws and random vessel occupancy; no downloads)",
        "config": {"workload": ("raw A/B/C fields and collection mask, 256 x 256 x 50 um "
                                "volume 100 um deep, 1 voxel/um, vessels 3% occupied"),
                   "fields": [list(a.shape) for a in f], "vessels": [int(v) for v in shape],
                   "mask": list(m.shape)},
        "kernels_ms_per_step": {k: v[0] / args.steps * 1e3 for k, v in kern.items()},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": ach / PEAK_HBM_GBS, "traffic": None, "kernel": top,
                     "kernel_ms": kern[top][0] / args.steps * 1e3,
                     "algorithmic_bytes_per_step": kern[top][1] / args.steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_surrogate_inference(args):
    """The surrogate network's inference alone on the 4225 sites of the plain workload
    (features from the index volume, then the implicit-GEMM convolutions): random initial
    weights, since the arithmetic does not depend on their values.  One step = features +
    inference of every site; the roofline is the BF16 MFMA peak."""
    import torch
    from maomi_amd.plain import PlainStage
    from maomi_amd.surrogate import SurrogateNet, site_features
    from maomi_amd.workloads import CONFIG2, index_volume, site_offsets
    dev = torch.device("cuda:0")
    vol = torch.as_tensor(index_volume(CONFIG2), device=dev)
    offs = site_offsets(CONFIG2)
    net = SurrogateNet(seed=0, device=dev)
    P = len(offs)
    for _ in range(args.warmup):
        net.predict(site_features(vol, CONFIG2, offs, CONFIG2.focus))
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    inf = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x = site_features(vol, CONFIG2, offs, CONFIG2.focus)
        ev0.record()
        net.predict(x)
        ev1.record()
        torch.cuda.synchronize()
        inf.append(ev0.elapsed_time(ev1))
    elapsed = time.perf_counter() - t0
    ms = float(np.mean(inf))
    flops = net.forward_flops(64, 64) * P
    tf = flops / (ms * 1e-3) / 1e12
    line = {
        "metric": "surrogate inference sites/s (4225 sites)", "value": P * args.steps / elapsed,
        "unit": "sites/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 weights and activations, f32 accumulate",
        "data": "synthetic (seeded index volume, random initial weights)",
        "config": {"workload": "surrogate network inference on the 4225 sites of the "
                               "512x512x60 plain workload", "params": net.num_params()},
        "roofline": {"bound": "mfma", "achieved": tf, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tf / 2500.0, "traffic": None,
                     "kernel": "network inference (conv_implicit_kernel, every layer)",
                     "kernel_ms": ms, "algorithmic_flops_per_launch": flops},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_surrogate(args):
    """Learned surrogate on the 4225-site workload: train on propagator PSFs of independent
    volumes with the same statistics, then per step predict every site, audit a seeded 5%
    with the FP32 propagator and fall back where the audit fails.  The plain comparison is
    the propagator writing the same focal-window PSFs for all sites, timed in the same run."""
    import torch
    from maomi_amd.plain import PlainStage
    from maomi_amd.surrogate import (SurrogateNet, SurrogateStage, rel_l2_per_site,
                                     site_features, site_psfs, train)
    from maomi_amd.workloads import CONFIG2, index_volume
    from dataclasses import replace
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    xs, ts = [], []
    for i in range(args.train_volumes):
        c = replace(CONFIG2, seed=CONFIG2.seed + 100 + i)
        v = torch.as_tensor(index_volume(c), device=dev)
        s = PlainStage(c, v, device=dev).plan()
        xs.append(site_features(v, c, s.offsets, c.focus))
        ts.append(site_psfs(s))
        del s, v
        print(f"[bench] training volume {i}: {xs[-1].shape[1]} PSFs", file=sys.stderr, flush=True)
    x, t = torch.cat(xs, 1), torch.cat(ts, 1)
    del xs, ts
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    scale = float(t.mean() * 4)
    t /= scale
    net = SurrogateNet(seed=0, device=dev)
    t0 = time.perf_counter()
    hist = train(net, x, t, epochs=args.epochs, lr=1e-3, wd=0.05,
                 log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del x, t

    vol = torch.as_tensor(index_volume(CONFIG2), device=dev)
    st = PlainStage(CONFIG2, vol, device=dev)
    P = len(st.offsets)
    sur = SurrogateStage(net, scale, st, seed=CONFIG2.seed + 4)
    for _ in range(args.warmup):
        st.plan()
        sur.run()
        st.plan()
        site_psfs(st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reports = []
    for _ in range(args.steps):
        st.plan()
        sur.run()
        reports.append(dict(sur.report))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.plan()
        truth = site_psfs(st)
    torch.cuda.synchronize()
    plain = time.perf_counter() - t0
    # the network alone on every site: inference time (HIP events on the launch stream)
    # and accuracy against the propagator's PSFs
    feats = site_features(vol, CONFIG2, st.offsets, CONFIG2.focus)
    net.predict(feats)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    inf = []
    for _ in range(3):
        ev0.record()
        pred = net.predict(feats)
        ev1.record()
        torch.cuda.synchronize()
        inf.append(ev0.elapsed_time(ev1) * 1e-3)
    rel_all = rel_l2_per_site(pred * scale, truth).cpu().numpy()
    tf = net.forward_flops(64, 64) * P / np.mean(inf) / 1e12
    rep = reports[-1]
    value = P * args.steps / elapsed
    line = {
        "metric": METRIC, "value": value, "unit": "sites/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 weights, f32 accumulate; f32 propagator audit",
        "data": "synthetic (seeded Gaussian random index fields; no downloads)",
        "config": {"workload": ("learned surrogate + 5% audit on the 4225-site plain stage "
                                "(512x512x60, N=256, S=60, Zernike pupil); 33x33x9 focal-window "
                                "two-photon PSF per site"),
                   "train_psfs": P * args.train_volumes, "epochs": args.epochs,
                   "params": net.num_params()},
        "surrogate": {
            "mode": rep["mode"], "audited_sites": rep["audited"],
            "audit_fail_sites": rep["audit_fail"],
            "audit_rel_l2_median": rep["audit_rel_l2_median"],
            "audit_rel_l2_max": rep["audit_rel_l2_max"],
            "network_rel_l2_all_sites_median": float(np.median(rel_all)),
            "network_sites_within_1pct": float(np.mean(rel_all <= 0.01)),
            "plain_propagator_sites_per_s": P * args.steps / plain,
            "network_inference_sites_per_s": P / float(np.mean(inf)),
            "speed_vs_plain": value / (P * args.steps / plain),
            "train_seconds": t_train, "data_seconds": t_data,
            "train_loss_first_last": [hist[0], hist[-1]],
        },
        "roofline": {"bound": "mfma", "achieved": tf, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tf / 2500.0, "traffic": None,
                     "kernel": "network inference on all sites (conv_implicit_kernel)",
                     "kernel_ms": float(np.mean(inf)) * 1e3,
                     "algorithmic_flops_per_launch": net.forward_flops(64, 64) * P},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
