"""GPU tests of the batched/served engine (dwhip_engine_*): every job's
result in a mixed batch is bit-identical to the same job run alone through
dwhip_shb_run_device, whatever runs beside it, however it is admitted, with
the active jobs in one group or in two groups on two streams."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import synth
from deconwolf_amd.serve import Engine, serve

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def make_job(seed, shape, psfshape, niter, **opt):
    rng = np.random.default_rng(seed)
    P, N, M = shape
    pP, pN, pM = psfshape
    k = synth.gaussian_psf(pM, pN, pP, 1.2 + 0.1 * seed, 1.5, 1.8)
    V = synth.point_sources(shape, 25, rng=rng)
    im = synth.noisy_image(synth.blur(V, k), 40.0 + seed, rng=rng) + np.float32(5.0)
    s = dw.DwOpts(nIter=niter, iter_type=dw.ITER_FIXED)
    for a, v in opt.items():
        setattr(s, a, v)
    return torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV), s


JOBS = [
    (1, (16, 36, 40), (7, 9, 9), 12, {}),
    (2, (10, 24, 30), (5, 7, 7), 5, {"metric": dw.METRIC_MSE}),
    (3, (12, 40, 28), (7, 7, 5), 20, {"positivity": 0, "alphamax": 0.7}),
    (4, (8, 20, 24), (5, 5, 5), 1, {}),
    (5, (11, 26, 43), (7, 7, 9), 7, {"borderQuality": 1}),
    (6, (12, 27, 40), (5, 5, 5), 6, {"borderQuality": 0, "start_condition": dw.START_IDENTITY}),
]


def solo(im, k, s):
    import copy
    s2 = copy.copy(s)
    out = dw.deconvolve_shb_hip_device(im, k, s2)
    torch.cuda.synchronize()
    return out.cpu().numpy(), s2


def expected_steps(iters, cap):
    """Steps of an engine whose jobs are all admitted at once and all FIXED:
    each step runs rounds up to the next completion, at most cap."""
    t, steps = 0, 0
    while t < max(iters):
        t = min(min(i for i in iters if i > t), t + cap)
        steps += 1
    return steps


@pytest.mark.parametrize("groups,rounds_per_step", [(1, 1), (2, 1), (1, 8), (2, 8), (2, 3)])
def test_batch_equals_alone(groups, rounds_per_step):
    print("engine: mixed batch of", len(JOBS), "groups", groups, "rounds per step", rounds_per_step, flush=True)
    jobs = [make_job(sd, sh, ps, n, **o) for sd, sh, ps, n, o in JOBS]
    eng = Engine(hbm_budget_gb=2.0, max_active=16, groups=groups, rounds_per_step=rounds_per_step)
    for t, (im, k, s) in enumerate(jobs):
        eng.submit(t, im, k, s)
    done = []
    steps = 0
    while len(done) < len(jobs):
        done += eng.step()
        steps += 1
        assert steps < 100
    # all admitted at once; several rounds per step up to each completion
    assert steps == expected_steps([j[3] for j in JOBS], rounds_per_step)
    for t, (im, k, s) in enumerate(jobs):
        ref, s2 = solo(im, k, s)
        out, st = eng.result(t)
        assert st["iterations"] == s2.stats["iterations"]
        assert np.array_equal(out.cpu().numpy(), ref), t
        np.testing.assert_array_equal(st["err_trace"], s2.err_trace)
        assert st["bg"] == s2.bg
    eng.close()


@pytest.mark.parametrize("groups", [1, 2])
def test_budget_limited_admission_and_staggered_submission(groups):
    print("engine: budget-limited admission, groups", groups, flush=True)
    jobs = [make_job(10 + i, (16, 36, 40), (7, 9, 9), 3 + i) for i in range(5)]
    w, r, _ = dw.work_size((16, 36, 40), (7, 9, 9))
    V = w[0] * w[1] * w[2]
    # room for about two of these jobs (a job is ~5 V floats)
    # one round per step, so that the queue builds up while jobs iterate
    eng = Engine(hbm_budget_gb=2.5 * 5 * V * 4 / (1 << 30), max_active=4, groups=groups, rounds_per_step=1)
    for t in range(2):
        eng.submit(t, *jobs[t])
    seen_active = 0
    seen_queued = 0
    done = []
    t_next = 2
    for _ in range(200):
        done += eng.step()
        q, a, b = eng.status()
        seen_active = max(seen_active, a)
        seen_queued = max(seen_queued, q)
        if t_next < len(jobs):
            eng.submit(t_next, *jobs[t_next])
            t_next += 1
        if len(done) == len(jobs):
            break
    assert len(done) == len(jobs)
    assert seen_active <= 2 and seen_queued >= 1   # the budget, not max_active, limited admission
    for t, (im, k, s) in enumerate(jobs):
        ref, _ = solo(im, k, s)
        out, _ = eng.result(t)
        assert np.array_equal(out.cpu().numpy(), ref), t
    eng.close()


@pytest.mark.parametrize("groups", [1, 2])
def test_rel_stopping_in_engine(groups):
    print("engine: REL stopping, groups", groups, flush=True)
    im, k, s = make_job(21, (10, 24, 30), (5, 7, 7), 1)
    s.iter_type = dw.ITER_REL
    s.err_rel = 0.05
    s.maxiter = 40
    ref, s2 = solo(im, k, s)
    eng = Engine(hbm_budget_gb=0.5, max_active=2, groups=groups)
    eng.submit(0, im, k, s)
    im2, k2, s3 = make_job(22, (16, 36, 40), (7, 9, 9), 30)
    eng.submit(1, im2, k2, s3)
    done = []
    while len(done) < 2:
        done += eng.step()
    out, st = eng.result(0)
    assert st["iterations"] == s2.stats["iterations"]
    assert np.array_equal(out.cpu().numpy(), ref)
    eng.close()


@pytest.mark.parametrize("groups", [1, 2])
def test_device_stopping_multi_round_steps(groups):
    """REL and ABS jobs of the fused passes stop on the device (the
    end-of-round kernel applies dw_iterator_next, the passes skip a stopped
    job), so they run in graph-replayed multi-round steps beside FIXED jobs:
    each job's iteration count equals its solo run's and its result is
    bit-identical, and the engine takes fewer steps than rounds."""
    print("engine: device stopping in multi-round steps, groups", groups, flush=True)
    jobs = []
    im, k, s = make_job(31, (10, 24, 30), (5, 7, 7), 1, iter_type=dw.ITER_REL, err_rel=0.05, maxiter=40)
    jobs.append((im, k, s))
    im, k, s = make_job(32, (12, 30, 28), (5, 7, 7), 1, iter_type=dw.ITER_REL, err_rel=0.01, maxiter=40,
                        metric=dw.METRIC_MSE)
    jobs.append((im, k, s))
    im, k, s = make_job(33, (10, 26, 32), (5, 5, 5), 1, iter_type=dw.ITER_REL, err_rel=0.02, maxiter=40,
                        method=dw.METHOD_RL)
    jobs.append((im, k, s))
    # ABS: a bound crossed within the run (from a fixed run's error trace)
    im, k, s = make_job(34, (12, 28, 36), (7, 7, 7), 25)
    _, sf = solo(im, k, s)
    tr = sf.err_trace
    s.iter_type, s.err_abs, s.maxiter = dw.ITER_ABS, float(tr[9]) * 1.0001, 40
    jobs.append((im, k, s))
    # REL stopped by maxiter
    im, k, s = make_job(35, (8, 20, 24), (5, 5, 5), 1, iter_type=dw.ITER_REL, err_rel=1e-9, maxiter=11)
    jobs.append((im, k, s))
    im, k, s = make_job(36, (16, 36, 40), (7, 9, 9), 17)
    jobs.append((im, k, s))
    def alone(im, k, s):
        if s.method != dw.METHOD_RL:
            return solo(im, k, s)
        import copy
        s2 = copy.copy(s)
        out = dw.deconvolve_rl_hip_device(im, k, s2)
        torch.cuda.synchronize()
        return out.cpu().numpy(), s2

    refs = [alone(im, k, s) for im, k, s in jobs]
    eng = Engine(hbm_budget_gb=2.0, max_active=16, groups=groups, rounds_per_step=8)
    for i, (im, k, s) in enumerate(jobs):
        eng.submit(i, im, k, s)
    done, steps = [], 0
    while len(done) < len(jobs):
        done += eng.step()
        steps += 1
        assert steps < 200
    rounds = max(r[1].stats["iterations"] for r in refs)
    for i, (ref, s2) in enumerate(refs):
        out, st = eng.result(i)
        assert st["iterations"] == s2.stats["iterations"], (i, st["iterations"], s2.stats["iterations"])
        assert np.array_equal(out.cpu().numpy(), ref), i
    assert refs[3][1].stats["iterations"] < 25 and refs[4][1].stats["iterations"] == 11
    assert steps < rounds, (steps, rounds)
    eng.close()


def test_shortest_first_and_voxel_cap():
    """ORDER_SHORTEST admits the least work first; active_mvox caps the
    transform voxels iterating together (a job above the cap runs alone);
    results stay bit-identical to solo runs."""
    from deconwolf_amd import _lib
    specs = [(11, (16, 36, 40), (7, 9, 9), 9, {}), (12, (8, 20, 24), (5, 5, 5), 3, {}),
             (13, (10, 24, 30), (5, 7, 7), 6, {})]
    jobs = [make_job(sd, sh, ps, n, **o) for sd, sh, ps, n, o in specs]
    work = []
    for im, k, s in jobs:
        w, _, _ = dw.work_size(tuple(im.shape), tuple(k.shape), 2)
        work.append(w[0] * w[1] * w[2] * s.nIter)
    # one job at a time: completion order = increasing work
    eng = Engine(hbm_budget_gb=2.0, max_active=1, order=_lib.ORDER_SHORTEST)
    for t, (im, k, s) in enumerate(jobs):
        eng.submit(t, im, k, s)
    order = []
    while len(order) < len(jobs):
        order += eng.step()
        assert eng.status()[1] <= 1
    assert order == sorted(range(len(jobs)), key=lambda t: work[t])
    for t, (im, k, s) in enumerate(jobs):
        ref, _ = solo(im, k, s)
        out, _ = eng.result(t)
        assert np.array_equal(out.cpu().numpy(), ref), t
    eng.close()
    # the voxel cap (millions of transform voxels): two 1.2-Mvox jobs never
    # iterate together under a cap of 2, and do under a cap of 3
    big = [make_job(21 + i, (40, 150, 150), (9, 9, 9), 3) for i in range(2)]
    w, _, _ = dw.work_size((40, 150, 150), (9, 9, 9), 2)
    assert 1_000_000 < w[0] * w[1] * w[2] < 1_500_000
    for cap, most in ((2, 1), (3, 2)):
        eng = Engine(hbm_budget_gb=2.0, max_active=8, active_mvox=cap, rounds_per_step=1)
        for t, (im, k, s) in enumerate(big):
            eng.submit(t, im, k, s)
        n, seen = 0, 0
        while n < len(big):
            n += len(eng.step())
            seen = max(seen, eng.status()[1])
            assert eng.status()[1] <= most
        assert cap == 2 or seen == 2
        for t, (im, k, s) in enumerate(big):
            ref, _ = solo(im, k, s)
            out, _ = eng.result(t)
            assert np.array_equal(out.cpu().numpy(), ref), t
        eng.close()


def test_oversized_job_refused():
    im, k, s = make_job(30, (32, 90, 100), (9, 9, 9), 2)   # about 9 MB of job state
    eng = Engine(hbm_budget_gb=0.002, max_active=2)
    with pytest.raises(dw.DwhipError) as e:
        eng.submit(0, im, k, s)
    assert "NOMEM" in str(e.value)
    eng.close()


def test_serve_loop_latencies():
    jobs = [make_job(40 + i, (12, 30, 34), (5, 7, 7), 4 + (i % 3)) for i in range(8)]
    eng = Engine(hbm_budget_gb=0.5, max_active=8)
    reqs = [(i, im, k, s) for i, (im, k, s) in enumerate(jobs)]
    arrivals = [0.002 * i for i in range(len(jobs))]
    times = serve(eng, reqs, arrivals)
    assert len(times) == len(jobs)
    for i, (arr, dn) in times.items():
        assert dn is not None and dn >= arr
        ref, _ = solo(*jobs[i])
        out, _ = eng.result(i)
        assert np.array_equal(out.cpu().numpy(), ref)
    eng.close()
