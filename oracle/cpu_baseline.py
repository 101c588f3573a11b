"""CPU baseline of the benchmark: the oracle's float32 restatement of
deconwolf's deconvolve_shb (src/method_shb.c:77-424) with the FFTW3-API
float transforms of MKL (deconwolf's supported MKL build) when
libmkl_rt.so loads, on all the threads given.

TEST INFRASTRUCTURE: run by bench.py as a child process, before the
parent touches the GPU.

    python -m oracle.cpu_baseline --inputs FILE.npz --iters K --full-iters N --threads T

Times the set-up and K iterations of the job in FILE.npz and prints one
JSON line; the whole-job rate is extrapolated to N iterations:
value = M*N*P*N_iter / (t_setup + N_iter * t_iter / K).
"""
import argparse
import json
import os
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--inputs", required=True)
    ap.add_argument("--iters", type=int, default=2)
    ap.add_argument("--full-iters", type=int, default=50)
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from oracle import pyoracle as po

    d = np.load(a.inputs)
    im, psf = d["im"], d["psf"]
    threads = a.threads or (os.cpu_count() or 1)
    backend = "mkl-fftw3-float" if po.set_fft_backend(True, threads) else "oracle-fft"
    out, err, bg, n = po.shb(im, psf, niter=a.iters, nthreads=threads)
    t_setup, t_iter = po.last_times
    P, N, M = im.shape
    per_iter = t_iter / max(n, 1)
    job = t_setup + a.full_iters * per_iter
    print(json.dumps({
        "value": M * N * P * a.full_iters / job,
        "unit": "voxel-iter/s",
        "cores": threads,
        "kind": "port",
        "backend": backend,
        "setup_s": t_setup,
        "iter_s": per_iter,
        "iters_timed": n,
        "sample": ("the whole %d-iteration %dx%dx%d job (PSF %dx%dx%d, set-up included), %s, %d threads"
                   % (n, M, N, P, psf.shape[2], psf.shape[1], psf.shape[0], backend, threads))
                  if n >= a.full_iters else
                  ("set-up + %d of %d iterations of the %dx%dx%d job (PSF %dx%dx%d), %s, %d threads, "
                   "extrapolated to the %d-iteration job" % (n, a.full_iters, M, N, P, psf.shape[2], psf.shape[1],
                                                             psf.shape[0], backend, threads, a.full_iters)),
    }))


if __name__ == "__main__":
    main()
