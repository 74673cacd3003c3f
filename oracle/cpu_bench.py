"""One oracle chain timed on one host core -- the CPU baseline's worker.

Test infrastructure: bench.py's cpu_baseline leg starts one of these per core
(pinned with taskset) on a sample of the benchmark's data, written by the
parent as .npy files (genotypes [N][2][L] int8, NumAlleles [L] int32), and
sums the per-chain rates.  This process never touches the GPU.

    python3 oracle/cpu_bench.py DIR --K 5 --sweeps 10 --seed 3 [model options]

Prints one JSON line: {"sweeps": n, "seconds": t, "cpu": <cpu the process ran on>}.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import oracle  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--K", type=int, required=True)
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--chain", type=int, default=0)
    ap.add_argument("--freqscorr", type=int, default=1)
    ap.add_argument("--inferlambda", type=int, default=1)
    ap.add_argument("--metrofreq", type=int, default=10)
    a = ap.parse_args()
    geno = np.load(os.path.join(a.dir, "geno.npy"), allow_pickle=False)
    nall = np.load(os.path.join(a.dir, "nall.npy"), allow_pickle=False)
    N, _, L = geno.shape
    A = int(max(2, nall.max()))
    m = oracle.make_model(N, L, a.K, A, freqscorr=a.freqscorr, inferalpha=1,
                          inferlambda=a.inferlambda, metrofreq=a.metrofreq, burnin=0,
                          seed=a.seed, chain=a.chain)
    ch = oracle.Chain(m, geno, nall)
    ch.init()
    ch.sweep(0)   # untimed first sweep (a Metropolis one)
    t0 = time.perf_counter()
    for r in range(1, 1 + a.sweeps):   # reps 1..10 hold one Metropolis sweep, as 1 in 10 do
        ch.sweep(r)
    dt = time.perf_counter() - t0
    cpu = os.sched_getaffinity(0)
    print(json.dumps({"sweeps": a.sweeps, "seconds": dt, "cpu": sorted(cpu)}), flush=True)


if __name__ == "__main__":
    main()
