#!/usr/bin/env python3
"""Throughput of the polarized GRRT raymarch on one MI355X, in rays x frequencies per second.

Headline (one "step" = one full render of one image on a snapshot resident in HBM): the
2-level AMR torus snapshot (96x48x48 base cells in 12^3 blocks, refined where rho > 0.1 and
rho > 0.5 of the maximum: 464 leaves, 0.80M cells, spin 0.9375), a 1024 x 1024 image at
230 GHz, inclination 10 deg, m_unit 1.2e18, R_high 3, camera at r = 1e4, fov 40 x 40 Rg,
marched by the mixed-precision path (FP32 bulk; FP64 ray state, FP64 steps near the horizon
and the poles, FP64 cell location, Kahan Stokes sums).  The same image by the plain FP32 path
(the reference's FP32 method) and the FP64 path is timed beside it.  The snapshot is generated
(seeded) on first use, decoded on the device once and kept resident; the timed region renders
into device memory (no host copies).

Also measured (skip with --quick): the same image at 8 frequencies (86-690 GHz), two-phase
against in place; a 16-point sweep of m_unit and R_high through the same camera (rays marched once)
against the images rendered apart; the 64-image batch of uneven size in one persistent launch
(autotuned) against rendering the same images one after another with the plain FP32 path, and
the serving loop on two resident 1.9M-cell snapshots under Poisson arrivals at 80% of its
measured saturation (p50 / p99 latency against the same kind of requests served one at a time).

Prints one JSON line with the roofline of the march kernel (algorithmic flops per launch from
the counted operation mix in tests/golden/op_counts.json times the launch's own step / segment
counts, over its HIP-event duration) and a CPU baseline: the reference's CPU build (oracle/_ref,
when present) or the oracle, timed on a strided sample of the same image on the host's cores.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--precision mixed|fp32|fp64] [--quick]
                    [--only multifreq|sweep|batch|serve]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays·freq/s at 1024² (mixed vs plain FP32, same GPU); p50/p99 request latency"
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}  # MI355X vector peaks (FP32 counts packed dual-issue)
W = H = 1024
FREQ = 230e9


def log(msg: str) -> None:
    print("# " + msg, file=sys.stderr, flush=True)


def workload_config(tmpdir: str, preset: str = "c2", name: str = "bench"):
    from phoenix_amd import configs, snapshot as S

    dump, grid, info = S.ensure_snapshot(preset, S.default_dir())
    cfg = configs.write_config(os.path.join(tmpdir, f"{name}.toml"), dump, grid, camera__img_width=str(W),
                               camera__img_height=str(H), camera__frequency="230e9",
                               camera__inclination="10", plasma__m_unit="1.2e18", plasma__r_high="3")
    return cfg, dump, grid, info


def algorithmic_flops(info: dict, nf: int, precision: str, rays: int) -> tuple:
    """(FP32 flops, FP64 flops) of one launch, from the operation counts of the reference's
    algorithm (tests/golden/op_counts.json: the oracle's count for the FP32 and FP64 paths, the
    Hamiltonian twin's count for the mixed path, per phase on 2000 pixels of this workload):
    per ray the camera setup and the frequency-independent work of its samples (the same whether
    or not an escaping ray's march ends at rt_outer), per geodesic step the step phase (view, AMR
    lookup, step rule, RK4), per emitting segment x frequency the coefficients, operator and
    accumulation, times the launch's own ray / step / segment / dropped-segment counts.  The mixed
    path's FP64 steps are FP64 work."""
    with open(os.path.join(ROOT, "tests", "golden", "op_counts.json")) as f:
        o = json.load(f)["c2_1024_230GHz"]["mixed" if precision == "mixed" else precision]
    fl = {ph: sum(o["ops"][ph][k] for k in ("add", "mul", "div", "sqrt", "trans")) for ph in o["ops"]}
    per_ray = (fl["setup"] + fl["sample"]) / o["pixels"]
    per_step = fl["step"] / o["steps"]
    useful_segf = info["nplasma"] * nf - info["bad"]
    work = rays * per_ray + useful_segf * o["F_segf"]
    if precision == "fp64":
        return 0.0, work + info["steps"] * per_step
    s64 = info.get("fp64_steps", 0) if precision == "mixed" else 0
    return work + (info["steps"] - s64) * per_step, s64 * per_step


def measured_traffic(precision: str):
    """HBM bytes per launch of the march kernel on this workload, from the copy-calibrated
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary kept in profiles/, or None."""
    path = os.path.join(ROOT, "profiles", f"k_march_{precision}_c2_1024_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch"):
                return float(line.split("=", 1)[1].split()[0])
    return None


def roofline(info: dict, nf: int, precision: str, kernel_s: float, rays: int = W * H) -> dict:
    w32, w64 = algorithmic_flops(info, nf, precision, rays)
    achieved = (w32 + w64) / kernel_s / 1e12
    # the peak of this instruction mix: FP32 work at the FP32 peak, FP64 work at the FP64 peak
    t_peak = w32 / (PEAK_TFLOPS["fp32"] * 1e12) + w64 / (PEAK_TFLOPS["fp64"] * 1e12)
    peak = (w32 + w64) / t_peak / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": measured_traffic(precision)}


def cpu_baseline(cfg: str, dump: str, grid: str, threads: int, stride: int) -> dict:
    """Reference CPU build (or the oracle) on every stride-th pixel of the same image."""
    npix = (W * H + stride - 1) // stride
    sample = f"every {stride}th pixel of the 1024^2 image ({npix} rays x 1 frequency)"
    from oracle import oracle as O

    if O.have_ref():
        with tempfile.TemporaryDirectory() as td:
            _, st, secs = O.run_ref(32, cfg, os.path.join(td, "o.bin"), pixels=(0, npix, stride), threads=threads)
        kind = "reference"
    else:
        o = O.Oracle(dump, grid, 32, img_width=W, img_height=H, freqs=(FREQ,))
        t0 = time.time()
        o.render(np.arange(0, W * H, stride, dtype=np.int64), threads=threads)
        secs = time.time() - t0
        o.close()
        kind = "port"
    return {"value": npix / secs, "unit": "rays·freq/s", "cores": threads, "kind": kind,
            "sample": sample + f", {secs:.1f} s, FP32 build"}


def timed_renders(snap, prec, steps, warmup, sync, label):
    for i in range(warmup):
        _, info = snap.render(None, prec, to_host=False)
        log(f"{label} warmup {i}: kernel {info['kernel_s'] * 1e3:.1f} ms")
    sync()
    t0 = time.perf_counter()
    kern, infos = [], []
    for _ in range(steps):
        _, info = snap.render(None, prec, to_host=False)
        kern.append(info["kernel_s"])
        infos.append(info)
    sync()
    return time.perf_counter() - t0, kern, infos[-1]


def c4_requests(run, seed: int = 4):
    """64 images: 8 x 64^2, 16 x 128^2, 16 x 256^2, 16 x 512^2, 8 x 1024^2; inclination U(5, 175);
    frequency from {86, 230, 345} GHz; m_unit log-uniform in [5e17, 5e18]; R_high U(1, 40)."""
    rng = np.random.default_rng(seed)
    sizes = [64] * 8 + [128] * 16 + [256] * 16 + [512] * 16 + [1024] * 8
    return [run.request(img_width=n, img_height=n, inclination=float(rng.uniform(5, 175)),
                        freqs=(float(rng.choice([86e9, 230e9, 345e9])),),
                        m_unit=float(10 ** rng.uniform(np.log10(5e17), np.log10(5e18))),
                        r_high=float(rng.uniform(1, 40))) for n in sizes]


def c5_request(run, rng):
    """resolution from {128^2 .. 1024^2}, 1-4 frequencies log-uniform in 86-690 GHz, inclination
    U(5, 175), m_unit and R_high as in the 64-image batch"""
    n = int(rng.choice([128, 256, 512, 1024]))
    nf = int(rng.integers(1, 5))
    freqs = tuple(float(f) for f in np.sort(10 ** rng.uniform(np.log10(86e9), np.log10(690e9), nf)))
    return run.request(img_width=n, img_height=n, inclination=float(rng.uniform(5, 175)), freqs=freqs,
                       m_unit=float(10 ** rng.uniform(np.log10(5e17), np.log10(5e18))),
                       r_high=float(rng.uniform(1, 40)))


def bench_multifreq(run, snap, capi, prec, sync, steps: int = 3) -> dict:
    """The same 1024^2 image at the 8 frequencies log-spaced from 86 to 690 GHz (the precision-form
    frequency set): two-phase renders (each ray marched once, then every frequency folded over its
    recorded samples; the default from 2 frequencies) against the in-place march with 4 frequencies
    per lane and one march per chunk of 4 (PHX_TWO_PHASE_MIN_NF=0), and against the plain FP32 march
    in place (the reference's method: samples in place, whole-wave refill, every ray to r_out)."""
    freqs = tuple(float(f) for f in np.geomspace(86e9, 690e9, 8))
    rq = run.request(freqs=freqs)
    rays = W * H * len(freqs)
    plain = {"queue": 0, "refill": 64, "escape": 1}
    cases = [("two_phase", prec, None, None), ("in_place", prec, None, "0"),
             ("plain_fp32_two_phase", capi.FP32, plain, None), ("plain_fp32_in_place", capi.FP32, plain, "0")]
    out = {"frequencies_hz": list(freqs)}
    old = os.environ.get("PHX_TWO_PHASE_MIN_NF")
    try:
        for name, p, tune, env in cases:
            if env is None:
                os.environ.pop("PHX_TWO_PHASE_MIN_NF", None)
                if old is not None:
                    os.environ["PHX_TWO_PHASE_MIN_NF"] = old
            else:
                os.environ["PHX_TWO_PHASE_MIN_NF"] = env
            snap.render(rq, p, to_host=False, tune=tune)
            sync()
            t0 = time.perf_counter()
            ks = []
            for _ in range(steps):
                ks.append(snap.render(rq, p, to_host=False, tune=tune)[1]["kernel_s"])
            sync()
            out[f"{name}_rays_freq_per_s"] = steps * rays / (time.perf_counter() - t0)
            out[f"{name}_kernel_ms"] = float(np.mean(ks)) * 1e3
    finally:
        os.environ.pop("PHX_TWO_PHASE_MIN_NF", None)
        if old is not None:
            os.environ["PHX_TWO_PHASE_MIN_NF"] = old
    out["two_phase_vs_in_place"] = out["two_phase_rays_freq_per_s"] / out["in_place_rays_freq_per_s"]
    out["two_phase_vs_plain_fp32_in_place"] = out["two_phase_rays_freq_per_s"] / out["plain_fp32_in_place_rays_freq_per_s"]
    out["two_phase_vs_plain_fp32_two_phase"] = out["two_phase_rays_freq_per_s"] / out["plain_fp32_two_phase_rays_freq_per_s"]
    log(f"8 frequencies: two-phase {out['two_phase_kernel_ms']:.1f} ms, in place {out['in_place_kernel_ms']:.1f} ms, "
        f"plain FP32 in place {out['plain_fp32_in_place_kernel_ms']:.1f} ms")
    return out


def bench_sweep(run, snap, capi, prec, sync, npts: int = 16) -> dict:
    """A parameter sweep of the headline image: 16 points of m_unit (4, log-spaced 5e17-5e18) x
    R_high (4, 1-40) at 1024^2, 230 GHz, i = 10 deg, through one camera.  The sweep marches the rays
    once and folds every point from the unit-free sample records (one render_batch call); against
    the same 16 images as a batch of separate images (PHX_SWEEP=0, one launch) and against the plain
    FP32 march rendering them one after another (the reference's method)."""
    ms = np.geomspace(5e17, 5e18, 4)
    rh = np.linspace(1.0, 40.0, 4)
    pts = [run.request(m_unit=float(m), r_high=float(r)) for m in ms for r in rh][:npts]
    rays = sum(p.img_width * p.img_height * p.nfreqs for p in pts)
    out = {"points": len(pts), "rays_freq": rays}

    def timed(fn):
        fn()
        sync()
        t0 = time.perf_counter()
        k = fn()
        sync()
        return time.perf_counter() - t0, k

    w, k = timed(lambda: snap.render_batch(pts, prec, to_host=False)[1]["kernel_s"])
    out.update(sweep_rays_freq_per_s=rays / w, sweep_kernel_s=k)
    os.environ["PHX_SWEEP"] = "0"
    try:
        w, k = timed(lambda: snap.render_batch(pts, prec, to_host=False)[1]["kernel_s"])
    finally:
        os.environ.pop("PHX_SWEEP", None)
    out.update(separate_batch_rays_freq_per_s=rays / w, separate_batch_kernel_s=k)
    plain = {"queue": 0, "refill": 64, "escape": 1}
    w, k = timed(lambda: sum(snap.render(p, capi.FP32, to_host=False, tune=plain)[1]["kernel_s"] for p in pts))
    out.update(plain_fp32_sequential_rays_freq_per_s=rays / w, plain_fp32_sequential_kernel_s=k)
    out["sweep_vs_plain_fp32_sequential"] = out["sweep_rays_freq_per_s"] / out["plain_fp32_sequential_rays_freq_per_s"]
    out["sweep_vs_separate_batch"] = out["sweep_rays_freq_per_s"] / out["separate_batch_rays_freq_per_s"]
    log(f"sweep of {len(pts)} points: {out['sweep_rays_freq_per_s']:.4g} rays*freq/s (kernels {out['sweep_kernel_s']:.3f} s), "
        f"separate batch {out['separate_batch_rays_freq_per_s']:.4g}, plain FP32 one by one "
        f"{out['plain_fp32_sequential_rays_freq_per_s']:.4g}")
    return out


def bench_host(run, snap, capi, prec, sync, steps: int = 3) -> dict:
    """The headline image with its result handed to the host, as the drop-in boundary does: the
    resident snapshot's render with the 50 MB result copied over PCIe, and the one-shot path
    (phx_dev_render, device::render's contract: per-call device memory, table upload, the dump read
    and decoded on the device, march, copy back).  Neither is `value`."""
    rays = W * H * run.model.nfreqs
    out = {}
    snap.render(None, prec)
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        snap.render(None, prec)
    out["resident_to_host_rays_freq_per_s"] = steps * rays / (time.perf_counter() - t0)
    run.dev_render(prec)
    t0 = time.perf_counter()
    for _ in range(steps):
        _, info = run.dev_render(prec)
    w = time.perf_counter() - t0
    out.update(one_shot_rays_freq_per_s=steps * rays / w, one_shot_ms=w / steps * 1e3,
               one_shot_kernel_ms=info["kernel_s"] * 1e3, one_shot_decode_ms=info["decode_s"] * 1e3,
               one_shot_upload_ms=info["upload_s"] * 1e3)
    log(f"result to the host: resident {out['resident_to_host_rays_freq_per_s']:.4g} rays*freq/s, one-shot "
        f"phx_dev_render {out['one_shot_ms']:.1f} ms per image ({out['one_shot_rays_freq_per_s']:.4g} rays*freq/s)")
    return out


def bench_batch(run, snap, capi, prec, sync) -> dict:
    reqs = c4_requests(run)
    rays = sum(r.img_width * r.img_height * r.nfreqs for r in reqs)
    t0 = time.perf_counter()
    tune, sampled = snap.autotune(reqs, prec, sample_stride=32)
    t_tune = time.perf_counter() - t0
    log(f"c4 autotune {tune} ({t_tune:.1f} s)")
    sync()
    t0 = time.perf_counter()
    _, info = snap.render_batch(reqs, prec, to_host=False)
    sync()
    wall = time.perf_counter() - t0
    snap.set_tune()
    t0 = time.perf_counter()
    _, info32 = snap.render_batch(reqs, capi.FP32, to_host=False)
    sync()
    wall32 = time.perf_counter() - t0
    # the same images one after another by the plain FP32 path (samples processed in place, as the
    # reference's kernel does)
    seq_kernel = 0.0
    t0 = time.perf_counter()
    for r in reqs:
        _, i1 = snap.render(r, capi.FP32, to_host=False, plain=True)
        seq_kernel += i1["kernel_s"]
    sync()
    wall_seq = time.perf_counter() - t0
    log(f"c4 batch: {rays / wall:.4g} rays*freq/s batched, {rays / wall32:.4g} batched fp32, "
        f"{rays / wall_seq:.4g} one at a time plain fp32")
    return {"images": len(reqs), "rays": rays, "tune": tune, "autotune_s": t_tune,
            "batched_rays_per_s": rays / wall, "batched_kernel_s": info["kernel_s"],
            "batched_fp32_rays_per_s": rays / wall32,
            "sequential_plain_fp32_rays_per_s": rays / wall_seq, "sequential_kernel_s": seq_kernel,
            "batched_vs_sequential_plain_fp32": wall_seq / wall}


def queue_model(arrive, service, discipline: str) -> np.ndarray:
    """Latencies of an ideal single server with no overhead, fed the given arrival times and
    per-request service times: "fifo" serves one request at a time in arrival order, "ps" shares
    the server equally among the requests present (processor sharing), "srpt" serves the request
    with the least remaining service first, preempting at arrivals."""
    n = len(arrive)
    done = np.zeros(n)
    if discipline == "fifo":
        t = 0.0
        for i in range(n):
            t = max(t, arrive[i]) + service[i]
            done[i] = t
        return done - arrive
    if discipline == "srpt":
        rem, t, i, active = np.array(service, dtype=float), 0.0, 0, []
        while i < n or active:
            if not active:
                t = max(t, arrive[i])
                active.append(i)
                i += 1
                continue
            j = min(active, key=lambda q: (rem[q], q))
            finish = t + rem[j]
            if i < n and arrive[i] < finish:
                rem[j] -= arrive[i] - t
                t = arrive[i]
                active.append(i)
                i += 1
                continue
            t = finish
            rem[j] = 0.0
            done[j] = t
            active.remove(j)
        return done - arrive
    rem, t, i, active = np.array(service, dtype=float), 0.0, 0, []
    while i < n or active:
        if not active:
            t = max(t, arrive[i])
            active.append(i)
            i += 1
            continue
        k = len(active)
        finish = t + min(rem[j] for j in active) * k
        if i < n and arrive[i] < finish:
            for j in active:
                rem[j] -= (arrive[i] - t) / k
            t = arrive[i]
            active.append(i)
            i += 1
            continue
        for j in active:
            rem[j] -= (finish - t) / k
        t = finish
        for j in [j for j in active if rem[j] <= 1e-12]:
            done[j] = t
            active.remove(j)
    return done - arrive


def p99_lower_bound(arrive, service):
    """The least p99 latency (numpy's linear percentile) that any schedule of one server of this
    rate can reach on these arrivals, preemptive or not: with n requests the p99 is at least the
    latency ranked floor(0.99 (n - 1)), so some d = n - 1 - that rank requests may be sacrificed.
    Whatever the schedule, the other n - d requests' latencies are a schedule of those requests alone,
    whose largest latency no schedule brings below FIFO's (earliest deadline first with one common
    latency bound is FIFO, which minimises the largest latency).  So the bound is the least, over the
    d-sets left out, of FIFO's largest latency over the rest.  None when d > 2 (too many sets)."""
    n = len(arrive)
    d = n - 1 - int(np.floor(0.99 * (n - 1)))
    if d > 2 or n <= d:
        return None
    arrive, service = np.asarray(arrive, dtype=float), np.asarray(service, dtype=float)
    best = np.inf
    for drop in itertools.combinations(range(n), d):
        keep = np.ones(n, bool)
        keep[list(drop)] = False
        best = min(best, float(queue_model(arrive[keep], service[keep], "fifo").max()))
    return best


def bench_serve(capi, prec, tmpdir) -> dict:
    runs, snaps = [], []
    for name in ("c5", "c5_a05"):
        cfg, _, _, si = workload_config(tmpdir, name, name)
        r = capi.Run(cfg)
        runs.append(r)
        snaps.append(capi.Snapshot(r, capi.TABLE_FP32 | (capi.TABLE_FP64 if prec == capi.FP64 else 0)))
        log(f"serving snapshot {name}: {si['nleafs']} leaves, {si['cells']} cells")
    rng = np.random.default_rng(5)

    def draw(n):
        out = []
        for _ in range(n):
            k = int(rng.integers(0, 2))
            out.append((k, c5_request(runs[k], rng)))
        return out

    srv = capi.Server(snaps, prec)
    out = {"snapshots": ["c5 spin 0.9375", "c5 spin 0.5"]}
    try:
        # saturation: a closed burst of requests
        burst = draw(24)
        t0 = time.perf_counter()
        for t in [srv.submit(k, r) for k, r in burst]:
            srv.wait(t)
        sat_wall = time.perf_counter() - t0
        sat_rays = sum(r.img_width * r.img_height * r.nfreqs for _, r in burst)
        sat_req_rate = len(burst) / sat_wall
        log(f"serving saturation: {sat_rays / sat_wall:.4g} rays*freq/s, {sat_req_rate:.3g} requests/s")
        # the arrival sequence, each request first served alone (its single-request latency)
        arrivals = draw(120)
        single = np.array([srv.wait(srv.submit(k, r))[1] for k, r in arrivals])
        load = 0.8
        lam = load * sat_req_rate
        gaps = rng.exponential(1.0 / lam, len(arrivals))
        out.update({"load": load, "saturation_rays_per_s": sat_rays / sat_wall,
                    "saturation_requests_per_s": sat_req_rate, "requests": len(arrivals),
                    "single_p50_ms": float(np.percentile(single, 50)) * 1e3,
                    "single_p99_ms": float(np.percentile(single, 99)) * 1e3})
    finally:
        srv.stop()
    try:
        # Poisson arrivals at 80% of saturation, the same sequence and gaps for each job order
        for order, name in ((0, "oldest_first"), (1, "least_attained_service")):
            srv = capi.Server(snaps, prec, order=order)
            try:
                t_next = time.perf_counter()
                tickets = []
                for (k, r), g in zip(arrivals, gaps):
                    t_next += g
                    dt = t_next - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    tickets.append(srv.submit(k, r))
                lat = np.array([srv.wait(t)[1] for t in tickets])
                st = srv.stats()
            finally:
                srv.stop()
            slow = lat / single
            p50, p99 = (float(np.percentile(lat, q)) for q in (50, 99))
            res = {"p50_ms": p50 * 1e3, "p90_ms": float(np.percentile(lat, 90)) * 1e3, "p99_ms": p99 * 1e3,
                   "p99_vs_single_p99": p99 / (out["single_p99_ms"] * 1e-3),
                   "slowdown_p50": float(np.percentile(slow, 50)), "slowdown_p99": float(np.percentile(slow, 99)),
                   "throughput_rays_per_s": st["rays"] / max(st["wall_s"], 1e-9), "launches": st["launches"],
                   "peak_result_bytes": st["peak_bytes"]}
            if order == 0:
                # the same arrivals through ideal servers with no overhead that march rays at the
                # measured saturation rate (a request's service time: its rays / that rate): how
                # much of the latency the queueing itself accounts for at this load
                arrive = np.cumsum(gaps)
                service = np.array([r.img_width * r.img_height * r.nfreqs for _, r in arrivals]) / out["saturation_rays_per_s"]
                for disc in ("fifo", "ps", "srpt"):
                    ideal = queue_model(arrive, service, disc)
                    res[f"ideal_{disc}_p99_ms"] = float(np.percentile(ideal, 99)) * 1e3
                    res[f"ideal_{disc}_p99_vs_single_p99"] = res[f"ideal_{disc}_p99_ms"] / out["single_p99_ms"]
                # no schedule of a server at that rate gets a lower p99 on these arrivals
                lb = p99_lower_bound(arrive, service)
                if lb is not None:
                    res["any_schedule_p99_at_least_ms"] = lb * 1e3
                    res["any_schedule_p99_at_least_vs_single_p99"] = lb * 1e3 / out["single_p99_ms"]
            out[name] = res
            log(f"serving at {load:.0%} load, {name}: p50 {p50 * 1e3:.0f} ms, p99 {p99 * 1e3:.0f} ms "
                f"(x{res['p99_vs_single_p99']:.2f} of the one-at-a-time p99 {out['single_p99_ms']:.0f} ms), "
                f"slowdown p50 {res['slowdown_p50']:.2f} p99 {res['slowdown_p99']:.2f}")
    finally:
        for s in snaps:
            s.close()
        for r in runs:
            r.close()
    best = min(("oldest_first", "least_attained_service"), key=lambda n: out[n]["p99_vs_single_p99"])
    out["p99_vs_single_p99"] = out["oldest_first"]["p99_vs_single_p99"]
    out["best_order"] = best
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["mixed", "fp32", "fp64"], default="mixed")
    ap.add_argument("--quick", action="store_true", help="the headline image only (no batch or serving runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--only", choices=["multifreq", "sweep", "host", "batch", "serve"],
                    help="run only this measurement (8 frequencies, the 16-point sweep, the 64-image batch or "
                         "the serving loop) and print its figures as one JSON line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (scales the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * 2**20 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the path shards by image tiles with no "
              "exchange inside the march, and multi-GPU runs are not part of this build.", file=sys.stderr)
        return 2

    from phoenix_amd import capi

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only; the library synchronises its own stream
        sync = lambda: None  # noqa: E731

    precs = {"fp32": capi.FP32, "fp64": capi.FP64, "mixed": capi.MIXED}
    prec = precs[args.precision]
    extra = {}
    with tempfile.TemporaryDirectory() as td:
        cfg, dump, grid, sinfo = workload_config(td)
        log(f"snapshot c2: {sinfo['nleafs']} leaves, {sinfo['cells']} cells")
        run = capi.Run(cfg)
        snap = capi.Snapshot(run, capi.TABLE_FP32 | capi.TABLE_FP64)
        nf = run.model.nfreqs
        rays = W * H * nf
        log(f"resident snapshot {snap.bytes / 2**20:.1f} MiB")
        if args.only:
            legs = {"multifreq": lambda: bench_multifreq(run, snap, capi, prec, sync),
                    "sweep": lambda: bench_sweep(run, snap, capi, prec, sync),
                    "host": lambda: bench_host(run, snap, capi, prec, sync),
                    "batch": lambda: bench_batch(run, snap, capi, prec, sync),
                    "serve": lambda: bench_serve(capi, prec, td)}
            res = legs[args.only]()
            snap.close()
            run.close()
            print(json.dumps({"only": args.only, "precision": args.precision, args.only: res}))
            return 0
        wall, kern, info = timed_renders(snap, prec, args.steps, args.warmup, sync, args.precision)
        value = args.steps * rays / wall
        kernel_s = float(np.mean(kern))
        roof = roofline(info, nf, args.precision, kernel_s)
        extra.update({"kernel_ms_mean": kernel_s * 1e3, "kernel_ms_min": float(np.min(kern) * 1e3),
                      "steps_per_ray": info["steps"] / (W * H), "segments_per_ray": info["nplasma"] / (W * H),
                      "dropped_segments": info["bad"],
                      "fp64_step_fraction": info["fp64_steps"] / max(1, info["steps"]),
                      "vgprs": info["vgprs"], "scratch_bytes": info["scratch_bytes"], "lds_bytes": info["lds_bytes"],
                      "waves_per_cu": info["waves_per_cu"], "grid": [info["blocks"], info["block"]],
                      "device": info["device"], "decode_s": info["decode_s"],
                      "algorithmic_tflop_per_launch": sum(algorithmic_flops(info, nf, args.precision, W * H)) / 1e12,
                      "frac_of_unpacked_fp32_peak": roof["achieved"] / 78.6})
        log(f"timed {args.steps} renders ({args.precision}): {wall:.3f} s, {value:.4g} rays*freq/s")
        # the same image by the other paths.  "plain_fp32" is the reference's FP32 method as the
        # reference's kernel runs it: each lane marches its pixel with the samples processed in place,
        # a wave takes new pixels when all its lanes are done (kernel.cu:80-96 on a wave), and every
        # ray marches on to r_out.  "_escape" ends a ray once it moves outward beyond rt_outer (the
        # same results, fewer steps; the mixed path's default), "_full" marches it to r_out.  The
        # "_queue" variants refill each lane at once and process the samples through the per-wave
        # LDS queue.
        others = [("plain_fp32", capi.FP32, {"queue": 0, "refill": 64, "escape": 1}),
                  ("fp32_escape", capi.FP32, {"queue": 0, "refill": 64, "escape": 2}),
                  ("fp32_queue", capi.FP32, {"queue": 1, "escape": 2}),
                  ("fp64", capi.FP64, None), ("mixed", capi.MIXED, None), ("mixed_full", capi.MIXED, {"escape": 1}),
                  ("mixed_queue", capi.MIXED, {"queue": 1}), ("mixed_queue_lane", capi.MIXED, {"queue": 1, "refill": 1})]
        for name, p2, tune in others:
            if name == args.precision:
                continue
            steps = 1 if name == "fp64" else args.steps
            snap.render(None, p2, to_host=False, tune=tune)
            sync()
            t0 = time.perf_counter()
            k2, i2 = [], None
            for _ in range(steps):
                _, i2 = snap.render(None, p2, to_host=False, tune=tune)
                k2.append(i2["kernel_s"])
            sync()
            w2 = time.perf_counter() - t0
            extra[f"{name}_rays_freq_per_s"] = steps * rays / w2
            extra[f"{name}_kernel_ms"] = float(np.mean(k2)) * 1e3
            extra[f"{name}_steps_per_ray"] = i2["steps"] / (W * H)
            extra[f"{name}_roofline_frac"] = roofline(i2, nf, next(p for p in ("fp32", "fp64", "mixed") if p in name),
                                                      float(np.mean(k2)))["frac"]
        if "plain_fp32_rays_freq_per_s" in extra:
            extra["value_vs_plain_fp32"] = value / extra["plain_fp32_rays_freq_per_s"]
        if "fp32_escape_rays_freq_per_s" in extra:
            extra["value_vs_fp32_escape"] = value / extra["fp32_escape_rays_freq_per_s"]
        if not args.quick:
            try:
                extra["multifreq_8"] = bench_multifreq(run, snap, capi, prec, sync)
            except Exception as e:  # reported, not fatal: the headline is already measured
                extra["multifreq_8"] = {"error": str(e)}
            try:
                extra["sweep_16"] = bench_sweep(run, snap, capi, prec, sync)
            except Exception as e:
                extra["sweep_16"] = {"error": str(e)}
            try:
                extra["host_result"] = bench_host(run, snap, capi, prec, sync)
            except Exception as e:
                extra["host_result"] = {"error": str(e)}
            try:
                extra["batch_c4"] = bench_batch(run, snap, capi, prec, sync)
            except Exception as e:  # reported, not fatal: the headline is already measured
                extra["batch_c4"] = {"error": str(e)}
        snap.close()
        run.close()
        if not args.quick:
            try:
                extra["serving_c5"] = bench_serve(capi, prec, td)
            except Exception as e:
                extra["serving_c5"] = {"error": str(e)}
        cpu = None
        if not args.no_cpu_baseline:
            log("cpu baseline ...")
            cpu = cpu_baseline(cfg, dump, grid, args.cpu_threads, args.cpu_stride)
    label = {"mixed": "mixed precision (FP32 bulk, FP64 near horizon/poles and cell location, Kahan Stokes)",
             "fp32": "FP32 (the reference's FP32 method)", "fp64": "FP64"}[args.precision]
    out = {
        "metric": METRIC, "value": value, "unit": "rays·freq/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded Fishbone-Moncrief torus snapshot, BHAC format)",
        "config": {"workload": "c2 snapshot (96x48x48 in 12^3 blocks, 2 AMR levels, 0.80M cells), 1024x1024, "
                               "230 GHz, i=10, m_unit 1.2e18, R_high 3, " + label,
                   "image": [W, H], "frequencies_hz": [FREQ], "rays_per_step": rays},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
