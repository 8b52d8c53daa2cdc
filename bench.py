"""Benchmark of the line-intensity hot path on one MI355X.

Workload (default): C2v(M), Nvib = 2000, J 10 <-> 11 (blocks of 10,500 and
11,500 symmetrised functions, 88,000 levels), lower levels with
E <= 10,000 cm^-1, upper levels unrestricted, frequency window
0-10,000 cm^-1, T = 296 K. Synthetic seeded data: eigenvector blocks are QR
factors of N(0,1) blocks built on the GPU, the dipole is
10^(-|v-v'|/3) * LogNormal(0,1) with random signs.

One step = one pass of gi_compute over every lower level (half
linestrengths, correlation, intensity, filter, compaction and the copy of
the emitted lines to pinned host memory). Inputs are resident in HBM when
the timed region starts.

Prints one JSON line. --gpus above 1 is refused: this build measures one GPU.
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

METRIC = "transitions/s FP64 vs split-precision MFMA at ≤1e-6 rel; request p50/p99 ms"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
INT8_PEAK_TOPS = 5000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_c2_job(torch, nvib=2000, jvals=(10, 11), seed=20240, lower_max=10000.0,
                 window=(0.0, 10000.0), levels_below=3000, cache_dir=None):
    import __graft_entry__ as ge
    ge.load_package()
    from gpu_intensities_amd import synth
    nlev_total = nvib * (2 * jvals[0] + 1) + nvib * (2 * jvals[1] + 1)
    e_span = lower_max * nlev_total / levels_below
    t = time.time()
    job = synth.make_job("C2V", nvib=nvib, jvals=jvals, seed=seed, e_span=e_span,
                         freq_window=window, erange_lower=(0.0, lower_max),
                         erange_upper=(0.0, 1e9), vectors="none")
    log(f"host generation {time.time() - t:.1f}s; levels {job.nlevels}")
    t = time.time()
    dev = {}
    gen = torch.Generator(device="cuda")
    tag = f"{nvib}_{jvals[0]}_{jvals[1]}_{seed}"
    for jind in range(2):
        for g in range(job.sym["nrep"]):
            n = int(job.bsets[jind].nsize[g])
            path = os.path.join(cache_dir, f"u_{tag}_{jind}_{g}.npy") if cache_dir else None
            if path and os.path.exists(path):
                dev[(jind, g)] = torch.from_numpy(np.load(path)).cuda()
                continue
            gen.manual_seed(seed * 1000 + jind * 10 + g)
            a = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
            q, r = torch.linalg.qr(a)
            q *= torch.sign(torch.diagonal(r))[None, :]
            del a, r
            dev[(jind, g)] = q.t().contiguous()  # record irec0 = column irec0 of q
            del q
            if path:
                os.makedirs(cache_dir, exist_ok=True)
                np.save(path + ".tmp.npy", dev[(jind, g)].cpu().numpy())
                os.replace(path + ".tmp.npy", path)
    torch.cuda.synchronize()
    log(f"eigenvector blocks ready {time.time() - t:.1f}s (QR on the GPU, or the cache)")
    return job, dev


# HIP-event timer name -> kernel symbol in the rocprofv3 summaries
KERNEL_SYMBOL = {"half_ls_gemv": "k_gemv_sym", "half_ls_mfma": "k_halfls_jobs",
                 "half_ls_split": "k_halfls_split", "corr_split": "k_corr_split",
                 "corr_mfma": "k_corr_mfma", "corr_intensity": "k_corr_intensity",
                 "build_g": "k_build_g_jobs"}
PMC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "c2_pmc_traffic.json")


def measured_traffic(name):
    """HBM bytes per dispatch of the kernel behind timer `name`, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this workload,
    each scaled by the factor the same counter gives for a plain 1 GiB copy.
    None when no such profile is committed."""
    try:
        with open(PMC_FILE) as f:
            d = json.load(f)
        return d["kernels"][KERNEL_SYMBOL[name]]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def roofline_entry(name, kstats, work, nvib):
    """Roofline of one kernel: algorithmic work of all its launches in the last
    step divided by their HIP-event time (average per launch)."""
    r = _roofline(name, kstats, work, nvib)
    r["traffic"] = measured_traffic(name)
    if r["traffic"] is not None:
        r["traffic_unit"] = "bytes per launch (rocprofv3 PMC, copy-calibrated)"
    return r


def _roofline(name, kstats, work, nvib):
    ks = kstats[name]
    t = ks["total_ms"] / 1e3
    avg_ms = t / ks["launches"] * 1e3
    if name == "half_ls_gemv":
        # GEMV form: the needed mu blocks are read once per unit
        ach = work["mu_bytes"] / t / 1e9
        return dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=ach / HBM_PEAK_GBS, traffic=None, kernel=name, avg_launch_ms=avg_ms,
                    work="mu bytes of the symmetry-blocked GEMV, per unit")
    if name == "half_ls_mfma":
        ach = work["half_ls_flops"] / t / 1e12
        return dict(bound="mfma", achieved=ach, peak=FP64_PEAK_TFS, unit="TFLOP/s",
                    frac=ach / FP64_PEAK_TFS, traffic=None, kernel=name, avg_launch_ms=avg_ms,
                    work="FP64 flops of the symmetry-blocked products")
    if name == "corr_mfma":
        ach = work["corr_bytes"] / 4.0 / t / 1e12  # 2*nsize_F flops per candidate
        return dict(bound="mfma", achieved=ach, peak=FP64_PEAK_TFS, unit="TFLOP/s",
                    frac=ach / FP64_PEAK_TFS, traffic=None, kernel=name, avg_launch_ms=avg_ms,
                    work="2*nsize_F FP64 flops per candidate")
    if name in ("half_ls_split", "corr_split"):
        # int8 digit products the split method needs: S(S+1)/2 per logical
        # multiply-add of the FP64 product (S = 6 digits)
        pairs = 21
        logical = work["half_ls_flops"] if name == "half_ls_split" else work["corr_bytes"] / 4.0
        ach = logical * pairs / t / 1e12
        return dict(bound="mfma", achieved=ach, peak=INT8_PEAK_TOPS, unit="TOP/s",
                    frac=ach / INT8_PEAK_TOPS, traffic=None, kernel=name, avg_launch_ms=avg_ms,
                    work="int8 ops of the 6-digit split products (21 digit pairs per logical "
                         "multiply-add); logical FP64-equivalent rate "
                         f"{logical / t / 1e12:.1f} TFLOP/s")
    if name == "corr_intensity":
        ach = work["corr_bytes"] / t / 1e9
        return dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=ach / HBM_PEAK_GBS, traffic=None, kernel=name, avg_launch_ms=avg_ms,
                    work="8*nsize_F bytes (one eigenvector column) per candidate")
    return dict(bound="hbm", achieved=None, peak=HBM_PEAK_GBS, unit="GB/s", frac=None,
                traffic=None, kernel=name, avg_launch_ms=avg_ms)


def split_report(S, Sref, strong=1e-6):
    """Relative error of S over the lines with S_ref >= strong * max S_ref."""
    m = Sref >= strong * Sref.max()
    rel = np.abs(S[m] - Sref[m]) / Sref[m]
    return {"strong_lines": int(m.sum()), "max_rel_S_strong": float(rel.max()),
            "median_rel_S_strong": float(np.median(rel)), "within_1e-6": bool(rel.max() <= 1e-6)}


def split_at_temperature(eng, job, binding, oracle_py, T, nthr, elow=(0.0, 1000.0)):
    """Split-mode error at another temperature: one request (J 10 -> 11, A1
    lower levels with E - ZPE in elow) with Q(T) from the device, against the
    oracle run with the same T, Q and lower levels."""
    import dataclasses
    Q = eng.partition_function(T)
    eng.serve_start(8 << 30)
    try:
        got = eng.serve_fetch(eng.serve_submit(jindI=0, igammaI=0, jindF=1,
                                               mode=binding.GI_MODE_SPLIT, temperature=T,
                                               q_stat=Q, elow=elow))
    finally:
        eng.serve_stop()
    lv = job.levels
    e = lv["energy"] - job.zpe
    ids = np.nonzero((lv["jind"] == 0) & (lv["igamma"] == 0) & (e >= elow[0]) &
                     (e <= elow[1]))[0].astype(np.int32)
    ref = oracle_py.OracleJob(dataclasses.replace(job, temperature=T, q_stat=Q)).run(
        ids, nthreads=nthr)
    keep = lv["jind"][ref["ilevelF"]] == 1
    ref = {k: v[keep] for k, v in ref.items()}
    kg = np.lexsort((got["ilevelF"], got["ilevelI"]))
    kr = np.lexsort((ref["ilevelF"], ref["ilevelI"]))
    same = (len(kg) == len(kr) and np.array_equal(got["ilevelI"][kg], ref["ilevelI"][kr])
            and np.array_equal(got["ilevelF"][kg], ref["ilevelF"][kr]))
    out = {"T": T, "Q": Q, "lower_levels": int(len(ids)), "lines": int(len(kr)),
           "same_transition_set": bool(same)}
    if same and len(kr):
        out.update(split_report(got["S"][kg], ref["S"][kr]))
        relI = np.abs(got["I"][kg] - ref["I"][kr]) / np.maximum(np.abs(ref["I"][kr]), 1e-300)
        out["max_rel_I_strong"] = float(relI[ref["S"][kr] >= 1e-6 * ref["S"][kr].max()].max())
    return out


def serve_sweep(eng, job, binding, lambdas, seconds, mode, seed=7, budget=8 << 30):
    """Poisson request stream against the resident blocks: each request is a
    random (J, J') pair that the selection rules allow, a random lower
    symmetry, T ~ U[200, 2000] K with Q(T) from the device, a 1000 cm^-1
    window at a random start in [0, 9000] cm^-1 and an intensity threshold
    10^U(-32,-26). Latency = submit to lines ready in host memory (measured by
    the library); throughput = lines / (last completion - first submit)."""
    import threading
    rng = np.random.default_rng(seed)
    j0 = job.jvals[0]
    pairs = [(a, b) for a in range(2) for b in range(2)
             if abs(job.jvals[a] - job.jvals[b]) <= 1 and not (job.jvals[a] == job.jvals[b] == j0)]
    nrep = job.sym["nrep"]
    res = {}
    for lam in lambdas:
        n = max(int(lam * seconds), 4)
        gaps = rng.exponential(1.0 / lam, size=n)
        reqs = []
        for _ in range(n):
            a, b = pairs[int(rng.integers(len(pairs)))]
            nu0 = float(rng.uniform(0.0, 9000.0))
            reqs.append(dict(jindI=a, igammaI=int(rng.integers(nrep)), jindF=b, mode=mode,
                             temperature=float(rng.uniform(200.0, 2000.0)), q_stat=0.0,
                             nu=(nu0, nu0 + 1000.0), thresh_linestrength=0.0,
                             thresh_intensity=float(10.0 ** rng.uniform(-32, -26))))
        eng.serve_start(budget)
        tickets = []
        t0 = time.perf_counter()
        t_next = t0
        for gap, r in zip(gaps, reqs):
            t_next += gap
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(eng.serve_submit(**r))
        lat, lines = [], 0
        for t in tickets:
            nl, ms = eng.serve_wait(t)
            lat.append(ms)
            lines += nl
            eng.serve_fetch(t)
        t1 = time.perf_counter()
        st = eng.serve_stats()
        eng.serve_stop()
        lat = np.asarray(lat)
        res[str(lam)] = dict(requests=n, offered_s=round(float(gaps.sum()), 3),
                             p50_ms=float(np.percentile(lat, 50)),
                             p99_ms=float(np.percentile(lat, 99)),
                             lines=int(lines), lines_per_s=lines / (t1 - t0),
                             requests_per_s=n / (t1 - t0), batches=st["batches"],
                             max_batch=st["max_batch"])
        log(f"serve mode {mode} lambda {lam}: {res[str(lam)]}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-stride", type=int, default=50,
                    help="CPU baseline sample: every n-th lower level by energy")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,fp64_batched,split",
                    help="comma-separated modes to time; the fastest is reported as value")
    ap.add_argument("--cache-dir", default=None,
                    help="reuse the QR eigenvector blocks across processes (profiler runs)")
    ap.add_argument("--serve-lambdas", default="10,50,100,200,500",
                    help="request rates (req/s) of the serving sweep; empty to skip")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="offered duration of each serving rate")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device copy (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    if args.copy_probe:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copies": 5}))
        return

    import __graft_entry__ as ge
    ge.load_package()
    from gpu_intensities_amd import binding

    job, dev = build_c2_job(torch, cache_dir=args.cache_dir)
    eng = binding.Engine(0)
    eng.load(job, device_vectors=dev)
    cap = eng.count_candidates()
    bufs = eng.alloc_lines(cap, pinned=True)
    log(f"candidates {cap}")

    modes = {}
    if "split" in args.modes.split(","):
        t = time.time()
        eng.prepare_split()  # resident digit planes, built once like the uploads
        torch.cuda.synchronize()
        log(f"split digit planes {time.time() - t:.1f}s")
    for mname in args.modes.split(","):
        mode = binding.MODES[mname]
        for _ in range(args.warmup):
            eng.compute(bufs=bufs, mode=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = eng.compute(bufs=bufs, mode=mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) / args.steps * 1e3
        nl = int(len(res["S"]))
        kstats = eng.kernel_stats()
        work = eng.work_stats()
        modes[mname] = dict(ms_per_step=ms, transitions=nl, value=nl / (ms / 1e3),
                            kernels={k: dict(total_ms=round(v["total_ms"], 4),
                                             launches=v["launches"]) for k, v in kstats.items()},
                            work=work, _kstats=kstats)
        log(f"{mname}: {ms:.2f} ms/step, {nl} transitions, {nl / (ms / 1e3):.4g}/s; "
            f"kernels {modes[mname]['kernels']}")
    best = max(modes, key=lambda m: modes[m]["value"])
    ms_per_step = modes[best]["ms_per_step"]
    nlines = modes[best]["transitions"]
    value = modes[best]["value"]
    kstats = modes[best].pop("_kstats")
    work = modes[best]["work"]
    dom = max(kstats, key=lambda k: kstats[k]["total_ms"])
    roof = roofline_entry(dom, kstats, work, job.nvib)
    kern = modes[best]["kernels"]
    for m in modes.values():
        m.pop("_kstats", None)
    serving = None
    if args.serve_lambdas:
        lams = [float(x) for x in args.serve_lambdas.split(",") if x]
        serving = {}
        for mname in args.modes.split(","):
            serving[mname] = serve_sweep(eng, job, binding, lams, args.serve_seconds,
                                         binding.MODES[mname])
    lower_ok = np.nonzero((job.levels["energy"] - job.zpe >= job.erange_lower[0]) &
                          (job.levels["energy"] - job.zpe <= job.erange_lower[1]))[0]

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        sample = lower_ok[::args.cpu_stride].astype(np.int32)
        t = time.time()
        for key, tens in dev.items():
            job.vectors[key] = tens.cpu().numpy()
        log(f"vectors to host {time.time() - t:.1f}s")
        oj = oracle_py.OracleJob(job)
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or oracle_py.max_threads()
        t = time.perf_counter()
        ref = oj.run(sample, nthreads=nthr)
        tc = time.perf_counter() - t
        cpu = dict(value=len(ref["S"]) / tc, unit="transitions/s", cores=nthr, kind="port",
                   sample=f"every {args.cpu_stride}th lower level by energy "
                          f"({len(sample)} of {len(lower_ok)}), {len(ref['S'])} transitions, "
                          f"{tc:.1f} s, FP64 oracle (K-masked loops, OpenMP over lower levels)")
        log(f"cpu baseline {cpu}")
        parity = {}
        kr = np.lexsort((ref["ilevelF"], ref["ilevelI"]))
        for mname in modes:
            got = eng.compute(sample, bufs=bufs, mode=binding.MODES[mname])
            kg = np.lexsort((got["ilevelF"], got["ilevelI"]))
            same = (len(kg) == len(kr) and np.array_equal(got["ilevelI"][kg], ref["ilevelI"][kr])
                    and np.array_equal(got["ilevelF"][kg], ref["ilevelF"][kr]))
            p = dict(sample_lines=int(len(kr)), same_transition_set=bool(same))
            if same and len(kr):
                floor = np.maximum(np.abs(ref["amp"][kr]), ref["amp_abs"][kr])
                err = np.abs(got["amp"][kg] - ref["amp"][kr])
                rel = np.abs(got["S"][kg] - ref["S"][kr]) / np.maximum(ref["S"][kr], 1e-300)
                p.update({"max_rel_S": float(rel.max()), "median_rel_S": float(np.median(rel)),
                          "amp_within_1e-10": bool((err <= 1e-10 * floor).all()),
                          "lines_via_floor": int((rel > 1e-10).sum())})
            if same and len(kr) and mname == "split":
                p.update(split_report(got["S"][kg], ref["S"][kr]))
            parity[mname] = p
        if "split" in modes:
            parity["split_1500K"] = split_at_temperature(eng, job, binding, oracle_py, 1500.0,
                                                         nthr)
        log(f"parity {parity}")

    out = {
        "metric": METRIC, "value": value, "unit": "transitions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("i8 digits x6 (split products), f64 elsewhere" if best == "split" else "f64"),
        "data": "synthetic (seeded C2v generator, GPU QR eigenvectors)",
        "config": {"workload": "C2v Nvib=2000 J 10<->11, lower E<=10000 cm-1, window "
                               "0-10000 cm-1, T=296 K",
                   "nvib": job.nvib, "jvals": list(job.jvals), "lower_levels": int(len(lower_ok)),
                   "transitions": nlines, "mode": best, "parallelism": "1 GPU"},
        "speedup_vs_fp64_gemv": (value / modes["fp64"]["value"]) if "fp64" in modes else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kern,
        "work": work,
        "modes": modes,
        "parity": parity,
        "serving": serving,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
