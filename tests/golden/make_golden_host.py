"""Write tests/golden/host/ fixtures from the reference's own parameter reader,
data-file reader and result writer (oracle/_ref).  testdata1.str beside them
is the reference's test data file.

    python tests/golden/make_golden_host.py
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import ref  # noqa: E402
from host_cases import (OUTPUT_VARIANTS, TESTDATA1, TESTDATA1_PARAMS, result_state,  # noqa: E402
                        write_param_files)
from structure_amd import host  # noqa: E402

OUT = os.path.join(HERE, "host")


def main():
    if not ref.available():
        sys.exit("oracle/_ref/libstructref.so is not built")
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        m, e = write_param_files(d)
        ref.set_int("ECHODATA", 0)
        p = ref.get_params(["structure", "-m", m, "-e", e, "-K", "4", "-D", "77", "-o",
                            "cmdout"])
        os.chdir(cwd)
    with open(os.path.join(OUT, "params_t1.json"), "w") as f:
        json.dump(p, f, indent=0, sort_keys=True)
    t = ref.read_input(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    np.savez_compressed(os.path.join(OUT, "input_testdata1.npz"), geno=t["geno"],
                        nall=t["nall"], translation=t["translation"],
                        population=t["population"], maxalleles=np.array(t["maxalleles"]))
    variant = OUTPUT_VARIANTS[0]
    K = variant["maxpops"]
    st = result_state(200, 5, K, t["maxalleles"], t["nall"], seed=0)
    pr = host.preset().as_dict()
    pr.update(TESTDATA1_PARAMS)
    pr.update(dict(datafile="testdata/testdata1.str", burnin=100, numreps=37, maxpops=K,
                   updatefreq=10, seed=1234))
    pr.update(variant)
    with tempfile.TemporaryDirectory() as d:
        pr["outfile"] = os.path.join(d, "theirs")
        ref.output_results(pr, t, 137, **st)
        txt = open(pr["outfile"] + "_f").read().replace(pr["outfile"], "X")
        q = open(pr["outfile"] + "_q").read()
    with open(os.path.join(OUT, "output_v0_f.txt"), "w") as f:
        f.write(txt)
    with open(os.path.join(OUT, "output_v0_q.txt"), "w") as f:
        f.write(q)
    # the ANCESTDIST variant: probability intervals in the ancestry table
    variant = OUTPUT_VARIANTS[4]
    K = variant["maxpops"]
    st = result_state(200, 5, K, t["maxalleles"], t["nall"], seed=4,
                      numboxes=variant["numboxes"])
    pr.update(variant)
    pr["maxpops"] = K
    with tempfile.TemporaryDirectory() as d:
        pr["outfile"] = os.path.join(d, "theirs")
        ref.output_results(pr, t, 137, **st)
        txt = open(pr["outfile"] + "_f").read().replace(pr["outfile"], "X")
    with open(os.path.join(OUT, "output_v4_f.txt"), "w") as f:
        f.write(txt)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
