"""Autotuner for the tile shapes of a batch (``sgpu_tuning``).

For a given set of chains it times a few graph-replayed sweeps under each
candidate tuning and keeps the fastest, caching the choice per shape in a JSON
file.  The draws and counts do not depend on the tiles (only the grouping of
the FP64 lnP and Metropolis sums does), so tuning never changes a chain's
path beyond the last bits of those sums."""
from __future__ import annotations

import itertools
import json
import os
import time

from . import GENO_BYTE, GENO_SNP2, Batch, Tuning

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "structure_amd", "tuning.json")


def shape_key(models, datasets, data_of_chain) -> str:
    """Signature of a batch: per chain (format, N, L, K, A), sorted."""
    sig = sorted((datasets[j][0], m.N, m.L, m.K, m.A) for m, j in zip(models, data_of_chain))
    return json.dumps(sig)


def candidates(models, datasets, data_of_chain):
    fmts = {datasets[j][0] for j in data_of_chain}
    zb = [1536, 2048, 3072, 4096, 5120, 6144, 8192] if GENO_SNP2 in fmts else [0]
    mb = [512, 1024, 2048]
    bl = [16, 32, 64] if GENO_BYTE in fmts else [0]
    for z, m, b in itertools.product(zb, mb, bl):
        yield Tuning(z_blocks=z, metro_blocks=m, byte_loci=b, byte_chunk=0)


def _time(models, datasets, doc, tuning, warm, sweeps):
    b = Batch(models, datasets, doc, tuning=tuning)
    try:
        b.init()
        b.run(0, warm)
        b.synchronize()
        t0 = time.perf_counter()
        b.run(warm, sweeps)
        b.synchronize()
        return time.perf_counter() - t0
    finally:
        b.close()


def autotune(models, datasets, data_of_chain=None, sweeps=10, warm=None, cache=DEFAULT_CACHE,
             cands=None, log=None):
    """Fastest Tuning for this batch (cached per shape); returns (tuning, table)
    where table lists (tuning fields, seconds for `sweeps` sweeps).  `warm`
    defaults to one Metropolis period so that every sweep type is captured."""
    doc = list(data_of_chain) if data_of_chain is not None else [0] * len(models)
    key = shape_key(models, datasets, doc)
    store = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            store = json.load(f)
        if key in store:
            return Tuning(**store[key]["tuning"]), store[key]["table"]
    mf = max(m.metrofreq for m in models) or 1
    warm = warm if warm is not None else mf
    table = []
    best = None
    for t in (cands if cands is not None else candidates(models, datasets, doc)):
        dt = _time(models, datasets, doc, t, warm, sweeps)
        fields = {n: getattr(t, n) for n, _ in Tuning._fields_}
        table.append((fields, dt))
        if log:
            log(f"tune {fields}: {dt * 1e3 / sweeps:.3f} ms/sweep")
        if best is None or dt < best[1]:
            best = (t, dt)
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        store[key] = {"tuning": {n: getattr(best[0], n) for n, _ in Tuning._fields_},
                      "table": table}
        with open(cache, "w") as f:
            json.dump(store, f)
    return best[0], table
