"""The host side of the `structure` command (structure_amd/host/*.c) against
the reference's own parameter reader, data-file reader and result writer
(params.c, datain.c, output.c compiled into oracle/_ref), and against the
fixtures those produced (tests/golden/host/, tests/golden/make_golden_host.py).
CPU only."""
import json
import os

import numpy as np
import pytest

from host_cases import (OUTPUT_VARIANTS, TESTDATA1, TESTDATA1_PARAMS, input_params,
                        result_state, write_genotype_file, write_param_files)
from structure_amd import host

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "host")
SKIP_KEYS = {"origmissing"}   # set by the data reader, not the parameter reader


def _ours(args):
    return host.get_params(args).as_dict()


def _norm(d):
    return {k: (round(v, 12) if isinstance(v, float) else v) for k, v in d.items()
            if k not in SKIP_KEYS}


def _param_args(tmp_path):
    m, e = write_param_files(str(tmp_path))
    return ["structure", "-m", m, "-e", e, "-K", "4", "-D", "77", "-o", "cmdout"]


def test_params_match_reference(ref, tmp_path, monkeypatch):
    from oracle import ref as R
    monkeypatch.chdir(tmp_path)
    args = _param_args(tmp_path)
    R.set_int("ECHODATA", 0)
    theirs = R.get_params(args)
    ours = _ours(args)
    assert _norm(ours) == _norm({k: theirs[k] for k in ours})


def test_params_golden(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ours = _norm(_ours(_param_args(tmp_path)))
    with open(os.path.join(GOLD, "params_t1.json")) as f:
        gold = json.load(f)
    assert ours == {k: gold[k] for k in ours}
    # values the test file sets in unusual ways
    assert ours["datafile"] == "my data/file one.str"     # spaces kept, // comment cut
    assert ours["seed"] == 77 and ours["maxpops"] == 4     # command line wins
    assert ours["locispop"] == 8                           # %i reads octal
    assert ours["updatefreq"] == 14500                     # SetUpdateFreq with 0


def test_param_file_names_and_defaults():
    a = ["structure", "-m", "my", "params", "file", "-e", "\"q x\"", "-K", "3"]
    assert host.param_file_name(a, "-m") == "my params file"
    assert host.param_file_name(a, "-e") == "q x"
    assert host.param_file_name(["structure"], "-m") == "mainparams"
    assert host.param_file_name(["structure"], "-e") == "extraparams"
    p = host.preset()
    assert (p.burnin, p.numreps, p.metrofreq, p.alphapropsd, p.fpriormean) == \
        (10000, 20000, 20, 0.05, 0.1)


def test_missing_keys_are_errors(tmp_path, monkeypatch, capfd):
    monkeypatch.chdir(tmp_path)
    (tmp_path / "mainparams").write_text("#define NUMINDS 5\n")
    (tmp_path / "extraparams").write_text("\n")
    with pytest.raises(host.SgpuError):
        host.get_params(["structure"])
    out = capfd.readouterr().out
    assert "Need to specify the name of the data file" in out
    assert "Need to set NUMLOCI>0" in out


INPUT_CASES = [
    ("testdata1", None),
    ("synthetic", dict(N=9, L=6, seed=1)),
    ("onerow", dict(N=11, L=4, seed=2, onerow=True)),
    ("nomarkers", dict(N=7, L=9, seed=3, markernames=False, popflag=False, extracols=0,
                       missing=-1000)),
]


def _input_case(name, kw, tmp_path):
    if name == "testdata1":
        return dict(TESTDATA1_PARAMS, datafile=TESTDATA1)
    path = str(tmp_path / f"{name}.str")
    kw = dict(kw)
    N, L, seed = kw.pop("N"), kw.pop("L"), kw.pop("seed")
    write_genotype_file(path, N, L, seed, **kw)
    return input_params(path, N, L, **kw)


def _read_ours(pd):
    p = host.preset()
    for k, v in pd.items():
        setattr(p, k, v.encode() if isinstance(v, str) else v)
    return host.InputFile(p)


@pytest.mark.parametrize("name,kw", INPUT_CASES)
def test_input_matches_reference(ref, name, kw, tmp_path):
    from oracle import ref as R
    pd = _input_case(name, kw, tmp_path)
    theirs = R.read_input(pd)
    ours = _read_ours(pd)
    assert ours.A == theirs["maxalleles"]
    assert ours.missing == theirs["missing"]
    assert np.array_equal(ours.geno, theirs["geno"])
    assert np.array_equal(ours.nall, theirs["nall"])
    for l in range(ours.L):
        k = ours.nall[l]
        assert np.array_equal(ours.translation[l, :k], theirs["translation"][l, :k])
    assert ours.labels == theirs["labels"]
    assert np.array_equal(ours.population, theirs["population"])
    if pd["popflag"]:
        assert np.array_equal(ours.popflag, theirs["popflag"])
    if pd["markernames"]:
        assert ours.markernames == theirs["markernames"]


def test_input_golden_testdata1():
    g = np.load(os.path.join(GOLD, "input_testdata1.npz"))
    ours = _read_ours(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    assert ours.A == int(g["maxalleles"])
    assert np.array_equal(ours.geno, g["geno"])
    assert np.array_equal(ours.nall, g["nall"])
    assert np.array_equal(ours.translation, g["translation"])
    assert np.array_equal(ours.population, g["population"])
    # recoding invariants: codes are 0..k-1 by first appearance, translation is a bijection
    for l in range(ours.L):
        seen = []
        for i in range(ours.N):
            for c in range(2):
                v = ours.geno[i, c, l]
                if v != ours.missing and v not in seen:
                    seen.append(v)
        assert seen == list(range(ours.nall[l]))


def test_input_reports_short_file(tmp_path, capfd):
    path = str(tmp_path / "short.str")
    write_genotype_file(path, 5, 4, 9)
    pd = input_params(path, 6, 4)
    with pytest.raises(host.SgpuError):
        _read_ours(pd)
    assert "Unexpected end of input file" in capfd.readouterr().out


def _output_params(variant, K):
    p = host.preset()
    base = dict(TESTDATA1_PARAMS, datafile="testdata/testdata1.str", burnin=100, numreps=37,
                maxpops=K, updatefreq=10, seed=1234)
    base.update(variant)
    for k, v in base.items():
        setattr(p, k, v.encode() if isinstance(v, str) else v)
    p.origmissing = p.missing   # as reading the data file records it
    return p


@pytest.mark.parametrize("vi", range(len(OUTPUT_VARIANTS)))
def test_output_matches_reference(ref, vi, tmp_path):
    from oracle import ref as R
    variant = OUTPUT_VARIANTS[vi]
    K = variant["maxpops"]
    inp = _read_ours(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    st = result_state(inp.N, inp.L, K, inp.A, inp.nall, seed=vi,
                      numboxes=variant.get("numboxes") if variant.get("ancestdist") else None)
    p = _output_params(variant, K)
    p.outfile = str(tmp_path / "ours").encode()
    inp.write_results(137, **st, params=p)
    theirs_in = R.read_input(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    pd = p.as_dict()
    pd["outfile"] = str(tmp_path / "theirs")
    R.output_results(pd, theirs_in, 137, **st)
    a = (tmp_path / "ours_f").read_text()
    b = (tmp_path / "theirs_f").read_text()
    assert a.replace(str(tmp_path / "ours"), "X") == b.replace(str(tmp_path / "theirs"), "X")
    if variant.get("printqhat"):
        assert (tmp_path / "ours_q").read_text() == (tmp_path / "theirs_q").read_text()


@pytest.mark.parametrize("intermedsave,numreps", [(3, 37), (1, 36), (4, 40)])
def test_intermediate_files_match_reference(ref, intermedsave, numreps, tmp_path):
    """INTERMEDSAVE (structure.c:640-661): the save interval NUMREPS /
    (INTERMEDSAVE + 1) rounded up (Init.c:377-385), and each intermediate file
    <OUTFILE>_<(rep - BURNIN) / savefreq> byte for byte as OutPutResults(final=0)
    writes it, ANCESTDIST intervals included."""
    from oracle import ref as R
    variant = dict(OUTPUT_VARIANTS[4], numreps=numreps, intermedsave=intermedsave)
    K = variant["maxpops"]
    inp = _read_ours(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    p = _output_params(variant, K)
    savefreq = host.save_freq(p)
    expect = numreps // (intermedsave + 1)
    expect += 1 if expect * (intermedsave + 1) < numreps else 0
    assert savefreq == expect
    theirs_in = R.read_input(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    for idx in range(1, intermedsave + 1):
        rep = p.burnin + idx * savefreq
        if rep >= p.burnin + numreps:
            break
        st = result_state(inp.N, inp.L, K, inp.A, inp.nall, seed=idx, nrep=rep - p.burnin,
                          numboxes=variant["numboxes"])
        p.outfile = str(tmp_path / "ours").encode()
        inp.write_results(rep, **st, params=p, final=0, saveindex=(rep - p.burnin) // savefreq)
        pd = p.as_dict()
        pd["outfile"] = str(tmp_path / "theirs")
        R.output_results(pd, theirs_in, rep, **st, final=0, savefreq=savefreq)
        a = (tmp_path / f"ours_{idx}").read_text()
        b = (tmp_path / f"theirs_{idx}").read_text()
        assert a.replace(str(tmp_path / "ours"), "X") == b.replace(str(tmp_path / "theirs"), "X")


def test_host_calc_like_equals_oracle():
    """PRINTLIKES's burn-in lnP(D) on the host equals the oracle's CalcLike."""
    import oracle
    inp = _read_ours(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    rng = np.random.default_rng(2)
    K = 3
    Q = rng.dirichlet(np.ones(K), size=inp.N)
    P = np.zeros((inp.L, K, inp.A))
    for l in range(inp.L):
        P[l, :, :inp.nall[l]] = rng.dirichlet(np.ones(inp.nall[l]), size=K)
    ch = oracle.Chain(oracle.make_model(inp.N, inp.L, K, inp.A), inp.geno_int8(), inp.nall)
    assert inp.calc_like(Q, P) == ch.calc_like(Q, P)


def test_output_golden(tmp_path):
    variant = OUTPUT_VARIANTS[0]
    K = variant["maxpops"]
    inp = _read_ours(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    st = result_state(inp.N, inp.L, K, inp.A, inp.nall, seed=0)
    p = _output_params(variant, K)
    p.outfile = str(tmp_path / "ours").encode()
    inp.write_results(137, **st, params=p)
    got = (tmp_path / "ours_f").read_text().replace(str(tmp_path / "ours"), "X")
    with open(os.path.join(GOLD, "output_v0_f.txt")) as f:
        assert got == f.read()
    with open(os.path.join(GOLD, "output_v0_q.txt")) as f:
        assert (tmp_path / "ours_q").read_text() == f.read()


def test_output_golden_ancestdist(tmp_path):
    """The ANCESTDIST variant against the reference's file kept as a fixture."""
    variant = OUTPUT_VARIANTS[4]
    K = variant["maxpops"]
    inp = _read_ours(dict(TESTDATA1_PARAMS, datafile=TESTDATA1))
    st = result_state(inp.N, inp.L, K, inp.A, inp.nall, seed=4, numboxes=variant["numboxes"])
    p = _output_params(variant, K)
    p.outfile = str(tmp_path / "ours").encode()
    inp.write_results(137, **st, params=p, final=2)
    got = (tmp_path / "ours_f").read_text().replace(str(tmp_path / "ours"), "X")
    with open(os.path.join(GOLD, "output_v4_f.txt")) as f:
        assert got == f.read()


def test_est_log_prob_closed_form():
    x = np.array([-1000.0, -1002.5, -999.0, -1001.0])
    v = x.var(ddof=1)
    assert abs(host.est_log_prob(x.sum(), (x ** 2).sum(), 4) - (x.mean() - v / 2)) < 1e-9
