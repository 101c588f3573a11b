#!/usr/bin/env python3
"""SHB deconvolution benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--no-cpu-baseline]

A step is one complete deconvolution job (set-up plus all SHB iterations)
of the workload, with the image and PSF already resident in GPU memory; K
steps are timed.  The workload (deconwolf_amd/workloads.py) is made with
numpy/scipy before the GPU is touched, the CPU baseline (the oracle's
float32 SHB with MKL's FFTW3-API transforms) runs on it in a child process,
then the GPU runs.

Prints one JSON line.  `value` is image voxel-iterations per second
(M*N*P*iterations per second of wall time over the timed steps).
`roofline` is for the kernel with the largest total time: its algorithmic
bytes per launch over its mean launch duration, both from HIP events
around every launch on the stream it runs on.  After the timed steps come
the c4 request stream (`served`), the c3 tiled image (`tiled`), a c5
sample (`c5`: the four channels of one multi-channel field, engine against
plain) and the c2 image at border quality 1 on the any-size path
(`anysize`), each with its own rates; none of them is `value`.
"""
import argparse
import copy
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# served stream: cap on the transform voxels (millions) iterating together,
# and the aging of least-work-first admission
SERVE_MVOX = 320
SERVE_AGING_MS = 1000
# job groups of the engine runs (--engine-groups): two groups on two streams
# measured 1-2 % faster on the c3 tiles and the FIFO stream
ENGINE_GROUPS = 2
# iterations of the job the CPU baseline times: the whole c2 job (about 35 s
# on 16 host threads); fewer would be extrapolated to the whole job
CPU_SAMPLE_ITERS = 50
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "shb_c2_pmc_traffic.md")
METRIC = "SHB voxel-iter/s batched-served vs plain per-image; p50/p99 job latency"
PLANNING = "measure"
ROUNDS_PER_STEP = None
STREAM_REL = None  # err_rel of the request stream's REL stopping (None: FIXED jobs)
PARITY_FILE = os.path.join(ROOT, "profiles", "parity_fullsize.md")


def host_cores():
    """(cores this process may use, os.cpu_count()): the CPU affinity mask,
    further limited by a cgroup CPU quota when one is set."""
    total = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = total
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, -(-int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return max(1, n), total


def cpu_threads():
    """All the host cores this process may use (OMP_NUM_THREADS, when set,
    says how many the host's operator gives a job)."""
    n, _ = host_cores()
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, env) if env > 0 else n)


def run_cpu_baseline(npz, iters_full, threads, timeout_s):
    env = dict(os.environ)
    env["MKL_THREADING_LAYER"] = "GNU"
    env["OMP_NUM_THREADS"] = str(threads)
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--inputs", npz, "--iters", str(CPU_SAMPLE_ITERS),
           "--full-iters", str(iters_full), "--threads", str(threads)]
    t = time.time()
    try:
        r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"value": None, "unit": "voxel-iter/s", "cores": threads, "kind": "port",
                "sample": "timed out after %d s" % timeout_s}
    if r.returncode != 0:
        return {"value": None, "unit": "voxel-iter/s", "cores": threads, "kind": "port",
                "sample": "failed: " + r.stderr.strip().splitlines()[-1] if r.stderr.strip() else "failed"}
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    d = json.loads(line)
    d["wall_s"] = time.time() - t
    usable, nproc = host_cores()
    d["host_cpus"] = nproc
    d["host_cpus_usable"] = usable
    return d


def parity_summary():
    """{config: {oracle: rel-L2}} from the committed full-size parity summary
    (the table of tests/test_gpu_parity_configs.py's figures), or None."""
    out = {}
    try:
        for line in open(PARITY_FILE):
            cols = [c.strip() for c in line.split("|")]
            if len(cols) < 9 or not cols[2].isdigit():
                continue
            row = {"iterations": int(cols[2]), "vs_float32_rel_l2": float(cols[5]), "vs_float32_max_rel": float(cols[6])}
            if cols[7] not in ("—", "-", ""):
                row["vs_float64_rel_l2"] = float(cols[7])
                row["vs_float64_max_rel"] = float(cols[8])
            out[cols[1]] = row
    except (OSError, ValueError):
        return None
    return out or None


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (scaled FETCH_SIZE + WRITE_SIZE), or None."""
    try:
        for line in open(TRAFFIC_FILE):
            cols = [c.strip() for c in line.split("|")]
            if len(cols) > 6 and cols[2] == kernel:
                return float(cols[6])
    except (OSError, ValueError):
        pass
    return None


def job_opts(**kw):
    """DwOpts of the benchmark's jobs (FIXED unless kw says otherwise):
    planning as --planning says."""
    import deconwolf_amd as dw
    kw.setdefault("iter_type", dw.ITER_FIXED)
    s = dw.DwOpts(**kw)
    if PLANNING == "rule":
        s.fftw3_planning = dw.method_shb.FFTW_ESTIMATE
    return s


def preplan(shapes):
    """Measures the transform plans of every distinct work size among
    (image shape, PSF shape) pairs before a timed region (dwhip_plan_tune,
    FFTW_MEASURE's analogue; cached, so the timed jobs find them); returns
    the seconds it took (0 with --planning rule)."""
    import deconwolf_amd as dw
    if PLANNING == "rule":
        return 0.0
    t = time.time()
    done = set()
    for ims, ps in shapes:
        w, _, rc = dw.work_size(ims, ps)
        if rc == 0 and w not in done:
            done.add(w)
            try:
                dw.plan_tune(w)
            except dw.DwhipError:
                pass  # sizes the fused passes cannot run keep the any-size path's one plan
    return time.time() - t


def _pct(xs, p):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(p / 100.0 * (len(xs) - 1))))]


def run_stream(n, rate, xy=None, fifo_leg=True, reuse=False):
    """The c4 stream served by the batched engine and, for comparison, by
    the plain per-image path (FIFO, one job at a time), both against the
    same arrival times; inputs are generated and resident first.  xy: the
    XY sizes to draw from (default the configuration's five)."""
    import torch
    import deconwolf_amd as dw
    from deconwolf_amd import workloads
    from deconwolf_amd import _lib
    from deconwolf_amd.serve import Engine, serve

    t = time.time()
    jobs = workloads.make_stream(n, rate, xy=xy, reuse=reuse)
    torch.cuda.synchronize()
    t_gen = time.time() - t
    t_plan = preplan([(tuple(j["im"].shape), tuple(j["psf"].shape)) for j in jobs])
    work = sum(j["shape"][0] * j["shape"][1] * j["shape"][2] * j["iters"] for j in jobs)
    arrivals = [j["arrival"] for j in jobs]

    def opts(j):
        if STREAM_REL:
            # REL stopping (dw_iterator_next) with the drawn count as maxiter
            return job_opts(nIter=j["iters"], iter_type=dw.ITER_REL, err_rel=STREAM_REL, maxiter=j["iters"])
        return job_opts(nIter=j["iters"])

    # plain per-image, FIFO
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat_p, last = [], 0.0
    ran = []  # iterations each job ran (the drawn count unless REL stops it)
    for j in jobs:
        now = time.perf_counter() - t0
        if j["arrival"] > now:
            time.sleep(j["arrival"] - now)
        s = opts(j)
        dw.deconvolve_shb_hip_device(j["im"], j["psf"], s)
        torch.cuda.synchronize()
        last = time.perf_counter() - t0
        lat_p.append(last - j["arrival"])
        ran.append(s.stats["iterations"])
    plain_s = last - arrivals[0]
    if STREAM_REL:
        # the engine's jobs stop after the same iterations (the same rule on
        # the same errors); the work is what ran
        work = sum(j["shape"][0] * j["shape"][1] * j["shape"][2] * r for j, r in zip(jobs, ran))

    free, total = torch.cuda.mem_get_info()
    budget = max(4.0, 0.8 * free / (1 << 30))

    def run_engine(order, mvox):
        eng = Engine(hbm_budget_gb=budget, max_active=64, order=order, active_mvox=mvox, groups=ENGINE_GROUPS,
                     aging_ms=SERVE_AGING_MS if order == _lib.ORDER_SHORTEST else 0, rounds_per_step=ROUNDS_PER_STEP)
        reqs = [(i, j["im"], j["psf"], opts(j)) for i, j in enumerate(jobs)]
        times = serve(eng, reqs, arrivals)
        eng.close()
        lat = [d - a for a, d in times.values()]
        return lat, max(d for _, d in times.values()) - arrivals[0]

    # served: least work first, at most SERVE_MVOX transform voxels iterating
    # together (a c2-sized job is 213 M); also FIFO with everything admitted
    lat_s, served_s = run_engine(_lib.ORDER_SHORTEST, SERVE_MVOX)
    lat_f, fifo_s = run_engine(_lib.ORDER_FIFO, 0) if fifo_leg else (lat_s, served_s)
    del jobs
    torch.cuda.empty_cache()
    return {
        "workload": workloads.describe_stream(n, rate, xy) + (
            "; one generated image per (XY, Z, wavelength) class, shared by its jobs" if reuse else "") + (
            "; REL stopping, err_rel %g, the drawn iteration count as maxiter (%d iterations run in all)" %
            (STREAM_REL, sum(ran)) if STREAM_REL else ""),
        "voxel_iters": work,
        "served_value": work / served_s,
        "plain_value": work / plain_s,
        "served_vs_plain": plain_s / served_s,
        "served_p50_ms": 1e3 * _pct(lat_s, 50), "served_p99_ms": 1e3 * _pct(lat_s, 99),
        "plain_p50_ms": 1e3 * _pct(lat_p, 50), "plain_p99_ms": 1e3 * _pct(lat_p, 99),
        "served_policy": "least work first, jobs waiting %d ms or more first (oldest first), at most %d M "
                         "transform voxels iterating together" % (SERVE_AGING_MS, SERVE_MVOX),
        "served_fifo_all_value": work / fifo_s,
        "served_fifo_all_p50_ms": 1e3 * _pct(lat_f, 50), "served_fifo_all_p99_ms": 1e3 * _pct(lat_f, 99),
        "engine_hbm_budget_gb": budget,
        "input_gen_s": t_gen,
        "planning_s": t_plan,
    }


def run_tiled():
    """c3: the 2048x2048x60 image in 16 overlapping tiles (tiling.c), the
    tiles batched through the engine against the same tiles run one after
    the other through the plain path; image voxel-iterations per second of
    the whole tiled deconvolution (extraction, all tiles, blend)."""
    import torch
    from deconwolf_amd import tiling, workloads
    from deconwolf_amd.serve import Engine

    t = time.time()
    im, psf, s = workloads.make_c3()
    s.fftw3_planning = job_opts().fftw3_planning
    torch.cuda.synchronize()
    t_gen = time.time() - t
    c = workloads.C3
    P, N, M = im.shape
    work = M * N * P * s.nIter
    shapes = []
    for tl in tiling.plan(M, N, P, c["tile"], c["overlap"]):
        ts = (tl.xsize[2], tl.xsize[1], tl.xsize[0])
        shapes.append((ts, tiling.tile_psf(psf, ts, s).shape))
    t_plan = preplan(shapes)
    res = {"workload": workloads.describe_c3(), "input_gen_s": t_gen, "planning_s": t_plan}
    for name, eng in (("plain_tiles", None), ("engine_tiles", "engine")):
        e = Engine(hbm_budget_gb=64.0, max_active=64, groups=ENGINE_GROUPS,
                   rounds_per_step=ROUNDS_PER_STEP) if eng else None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out, tiles, _ = tiling.deconvolve_tiles(im, psf, s, c["tile"], c["overlap"], engine=e)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        if e is not None:
            e.close()
        res[name + "_value"] = work / el
        res[name + "_s"] = el
        res["tiles"] = len(tiles)
        del out
    res["engine_vs_plain"] = res["engine_tiles_value"] / res["plain_tiles_value"]
    return res


def run_c5(fields=1, chunk=4):
    """c5: the four channels of `fields` field(s) (1024x1024x81, PSF
    181x181x81, 100 iterations), through the engine together and one after
    the other through the plain path; image voxel-iterations per second.
    Fields are generated and run `chunk` at a time (4 channels each), so the
    whole 64-field set fits in memory; only the deconvolutions are timed."""
    import torch
    import deconwolf_amd as dw
    from deconwolf_amd import workloads
    from deconwolf_amd.serve import Engine

    res = {"workload": workloads.describe_c5(fields), "input_gen_s": 0.0, "jobs": 0, "planning_s": 0.0,
           "plain_s": 0.0, "engine_s": 0.0, "chunk_fields": chunk}
    work_total = 0
    same = True
    for f0 in range(0, fields, chunk):
        t = time.time()
        jobs = []
        for f in range(f0, min(fields, f0 + chunk)):
            jobs += workloads.make_c5(f)
        for j in jobs:
            j["opts"].fftw3_planning = job_opts().fftw3_planning
        torch.cuda.synchronize()
        res["input_gen_s"] += time.time() - t
        res["jobs"] += len(jobs)
        res["planning_s"] += preplan([(tuple(j["im"].shape), tuple(j["psf"].shape)) for j in jobs])
        work = sum(j["im"].numel() * j["opts"].nIter for j in jobs)
        work_total += work
        outs = [torch.empty_like(j["im"]) for j in jobs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j, o in zip(jobs, outs):
            dw.deconvolve_shb_hip_device(j["im"], j["psf"], copy.copy(j["opts"]), out=o)
        torch.cuda.synchronize()
        res["plain_s"] += time.perf_counter() - t0
        e = Engine(hbm_budget_gb=64.0, max_active=64, groups=ENGINE_GROUPS, rounds_per_step=ROUNDS_PER_STEP)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i, j in enumerate(jobs):
            e.submit(i, j["im"], j["psf"], copy.copy(j["opts"]))
        done = set()
        while len(done) < len(jobs):
            done.update(e.step())
        torch.cuda.synchronize()
        res["engine_s"] += time.perf_counter() - t0
        same = same and all(bool(torch.equal(e.result(i)[0], outs[i])) for i in range(len(jobs)))
        e.close()
        del jobs, outs
        if fields > chunk:
            print("c5: fields %d-%d of %d done, plain %.1f s, engine %.1f s" %
                  (f0, min(fields, f0 + chunk) - 1, fields, res["plain_s"], res["engine_s"]), file=sys.stderr, flush=True)
    res["plain_value"] = work_total / res["plain_s"]
    res["engine_value"] = work_total / res["engine_s"]
    res["engine_vs_plain"] = res["engine_value"] / res["plain_value"]
    res["engine_equals_plain"] = same
    return res


def run_anysize():
    """c2 with --bq 1: deconwolf's exact work size (1115 = 5 * 223 along x
    and y, Bluestein there), timed at 3 and 13 iterations (best of two
    each); the difference gives the time per iteration."""
    import torch
    import deconwolf_amd as dw
    from deconwolf_amd import workloads

    im, psf, _, _ = workloads.make("c2")
    imt = torch.from_numpy(im).to("cuda:0")
    pst = torch.from_numpy(psf).to("cuda:0")
    best = {}
    for n in (3, 13, 3, 13):
        s = job_opts(nIter=n, borderQuality=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dw.deconvolve_shb_hip_device(imt, pst, s)
        torch.cuda.synchronize()
        best[n] = min(best.get(n, 1e9), time.perf_counter() - t0)
    it_ms = 1e3 * (best[13] - best[3]) / 10
    M, N, P = im.shape[2], im.shape[1], im.shape[0]
    return {"workload": "c2 at border quality 1: 1024x1024x60, PSF 181x181x79, work size %s (any-size path)" %
                        "x".join(str(v) for v in s.stats["work"]),
            "job_3_iterations_s": best[3], "job_13_iterations_s": best[13], "ms_per_iteration": it_ms,
            "image_voxel_iter_per_s": M * N * P / (it_ms * 1e-3)}


def main():
    global ENGINE_GROUPS, SERVE_AGING_MS, PLANNING, ROUNDS_PER_STEP, STREAM_REL
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-timeout", type=int, default=240)
    ap.add_argument("--stream", type=int, default=16, metavar="JOBS",
                    help="jobs of the mixed-size c4 request stream run after the c2 steps, served (batched "
                         "engine) and plain (one at a time); 0 skips it")
    ap.add_argument("--rate", type=float, default=20.0, help="Poisson arrival rate of the stream, jobs/s")
    ap.add_argument("--no-tiled", action="store_true", help="skip the c3 tiled run")
    ap.add_argument("--no-anysize", action="store_true", help="skip the c2 border-quality-1 (any-size path) timing")
    ap.add_argument("--c5-fields", type=int, default=1, help="fields of the c5 sample (4 channels each); 0 skips it")
    ap.add_argument("--engine-groups", type=int, default=ENGINE_GROUPS, choices=(1, 2),
                    help="job groups of the engine runs (2: two streams whose passes overlap)")
    ap.add_argument("--sweep", default="", metavar="RATES",
                    help="only run the c4 stream at these Poisson rates (jobs/s, comma-separated), "
                         "served and plain, with min(500, rate * --sweep-seconds) jobs each, and print one JSON "
                         "line with the table")
    ap.add_argument("--sweep-seconds", type=float, default=20.0)
    ap.add_argument("--sweep-xy", default="", help="XY sizes of the sweep's jobs (comma-separated; default all five)")
    ap.add_argument("--stream-rel", type=float, default=None, metavar="ERR_REL",
                    help="request-stream jobs stop by dw_iterator_next's REL rule (err_rel ERR_REL, the drawn "
                         "iteration count as maxiter) instead of running a fixed count")
    ap.add_argument("--rounds-per-step", type=int, default=None,
                    help="rounds an engine step may run (default: the library's)")
    ap.add_argument("--aging-ms", type=int, default=SERVE_AGING_MS,
                    help="waiting bound of the served stream's least-work-first admission")
    ap.add_argument("--planning", choices=("measure", "rule"), default="measure",
                    help="transform planning: measure a new size's plans before the timed region (the library's "
                         "default, FFTW_MEASURE's analogue) or plan by rule (FFTW_ESTIMATE; no trial launches, for "
                         "profiles of the timed passes alone)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only copy a GIB-sized float32 buffer on the GPU 3 times (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE/WRITE_SIZE readings) and exit")
    ap.add_argument("--anysize", action="store_true",
                    help="only time the c2 job at border quality 1 (deconwolf's exact work size 1115x1115x100, "
                         "the any-size path) at 3 and 13 iterations, print one JSON line and exit")
    a = ap.parse_args()
    ENGINE_GROUPS = a.engine_groups
    SERVE_AGING_MS = a.aging_ms
    PLANNING = a.planning
    ROUNDS_PER_STEP = a.rounds_per_step
    STREAM_REL = a.stream_rel
    if a.copy_calibration > 0:
        import torch
        n = int(a.copy_calibration * (1 << 30)) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": 4 * n, "launches": 3}))
        return
    if a.anysize:
        print(json.dumps(run_anysize()))
        return
    if a.sweep:
        if a.gpus != 1:
            sys.exit(2)
        xy = tuple(int(v) for v in a.sweep_xy.split(",")) if a.sweep_xy else None
        rows = []
        for r in (float(v) for v in a.sweep.split(",")):
            n = int(min(500, max(8, round(r * a.sweep_seconds))))
            t0 = time.time()
            row = run_stream(n, r, xy=xy, fifo_leg=False, reuse=True)
            row.update(rate=r, jobs=n, wall_s=time.time() - t0)
            print(json.dumps({"sweep_row": row}), flush=True)
            rows.append(row)
        print(json.dumps({"metric": METRIC, "sweep": rows, "engine_groups": ENGINE_GROUPS}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: libdwhip has no multi-GPU path (jobs would shard one per GPU "
              "with no communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from deconwolf_amd import workloads

    t0 = time.time()
    im, psf, s, raw_psf = workloads.make(a.config)
    t_gen = time.time() - t0
    P, N, M = im.shape
    iters = s.nIter

    cpu = None
    if not a.no_cpu_baseline:
        with tempfile.TemporaryDirectory() as td:
            npz = os.path.join(td, "inputs.npz")
            np.savez(npz, im=im, psf=psf)
            cpu = run_cpu_baseline(npz, iters, cpu_threads(), a.cpu_timeout)

    import torch
    import deconwolf_amd as dw
    from deconwolf_amd import _lib

    dev = torch.device("cuda:0")
    dim = torch.from_numpy(im).to(dev)
    dpsf = torch.from_numpy(psf).to(dev)
    out = torch.empty_like(dim)
    torch.cuda.synchronize()

    def step():
        st = job_opts(nIter=iters)
        dw.deconvolve_shb_hip_device(dim, dpsf, st, out=out)
        return st

    t_plan = preplan([(im.shape, psf.shape)])
    for _ in range(a.warmup):
        st = step()
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(a.steps):
        t1 = time.perf_counter()
        st = step()
        lat.append((time.perf_counter() - t1) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    prof = _lib.profile_read()
    _lib.profile_enable(False)

    value = M * N * P * iters * a.steps / elapsed
    kernels = {}
    for name, (n, ms, by) in prof.items():
        kernels[name] = {"launches": n, "avg_us": 1e3 * ms / n, "bytes_per_launch": by / n,
                         "GBps": (by / 1e9) / (ms / 1e3)}
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dname, (dn, dms, dby) = dom
    achieved = (dby / dn / 1e9) / (dms / dn / 1e3)
    all_ms = sum(v[1] for v in prof.values())
    all_by = sum(v[2] for v in prof.values())
    wx, wy, wz = st.stats["work"]
    lat_sorted = sorted(lat)

    def pct(p):
        return lat_sorted[min(len(lat_sorted) - 1, int(round(p / 100.0 * (len(lat_sorted) - 1))))]

    traffic = profiled_traffic(dname) if a.config == "c2" else None
    served = None
    if a.stream > 0:
        served = run_stream(a.stream, a.rate)
    tiled = None if a.no_tiled else run_tiled()
    c5 = run_c5(a.c5_fields) if a.c5_fields > 0 else None
    anysize = None if a.no_anysize else run_anysize()

    res = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "%s: %s" % (a.config, workloads.describe(a.config)),
            "path": "plain per-image (dwhip_shb_run_device, one job at a time)",
            "iterations": iters,
            "image": [M, N, P],
            "psf": [psf.shape[2], psf.shape[1], psf.shape[0]],
            "job": [wx, wy, wz],
            "reference_job": list(st.stats["ref_work"]),
        },
        "p50_job_ms": pct(50),
        "p99_job_ms": pct(99),
        "roofline": {
            "kernel": dname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "profiles/shb_c2_pmc_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)"
                              if traffic else None,
            "bytes_per_launch": dby / dn,
            "avg_launch_us": 1e3 * dms / dn,
            "all_passes_GBps": (all_by / 1e9) / (all_ms / 1e3),
        },
        "cpu_baseline": cpu,
        "parity": parity_summary(),
        "parity_source": "profiles/parity_fullsize.md (tests/test_gpu_parity_configs.py on MI355X, oracle at "
                         "deconwolf's work size)",
        "served": served,
        "tiled": tiled,
        "c5": c5,
        "anysize": anysize,
        "engine_groups": ENGINE_GROUPS,
        "kernels": kernels,
        "input_gen_s": t_gen,
        "planning_s": t_plan,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
