"""Write tests/golden/chains_<case>.npz: chain-level statistics of the
reference's own CPU sweep (oracle/ref_chain.py over oracle/_ref), 8 seeds per
case of tests/chain_cases.py.

    python tests/golden/make_golden_chains.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import ref, ref_chain  # noqa: E402
from chain_cases import CASES, SEEDS, make_data  # noqa: E402


def main():
    if not ref.available():
        sys.exit("oracle/_ref/libstructref.so is not built")
    for case, c in CASES.items():
        geno, nall, A = make_data(case)
        stats = dict(qbar=[], lnp=[], alpha=[], fst=[], lam=[])
        for s in SEEDS:
            r = ref_chain.run_chain(geno, nall, c["K"], A, seed=1000 + s, burnin=c["burnin"],
                                    numreps=c["numreps"], **c["opts"])
            for k in stats:
                stats[k].append(r[k])
            print(case, s, round(r["lnp"], 1), flush=True)
        np.savez_compressed(os.path.join(HERE, f"chains_{case}.npz"), geno=geno, nall=nall,
                            A=np.array(A), **{k: np.array(v) for k, v in stats.items()})


if __name__ == "__main__":
    main()
