"""CPU tests of the product's host side: C-ABI exports, packers, generators."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import structure_amd as sa
from helpers import random_geno

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:sgpu|shost)_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(os.listdir(os.path.join(ROOT, "include"))))
def test_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(sa.LIB_PATH)
    names = _declared(header)
    assert len(names) >= 2
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/{header} but not exported"


def _unpack_snp2(words, N, L):
    lpad = words.size // ((N + 15) // 16)
    w = words.reshape(-1, lpad)[:, :L]
    out = np.empty((N, 2, L), np.int8)
    for i in range(N):
        code = (w[i >> 4] >> np.uint32(2 * (i & 15))) & np.uint32(3)
        out[i, 0] = np.where(code == 3, -9, code == 2)
        out[i, 1] = np.where(code == 3, -9, code >= 1)
    return out


def _unpack_byte(words, N, L):
    w = words.reshape(L, -1)   # [L][Gpad], locus-major
    out = np.empty((N, 2, L), np.int16)
    for i in range(N):
        for c in range(2):
            b = (w[:, i >> 1] >> np.uint32(16 * (i & 1) + 8 * c)) & np.uint32(0xff)
            out[i, c] = np.where(b == 0xff, -9, b).astype(np.int16)
    return out


def test_pack_snp2_roundtrip_and_padding():
    d = sa.synth_snp(37, 300, Ktrue=3, seed=4, miss=0.05, want_geno=True)
    fmt, words = sa.pack(d["geno"])
    assert fmt == sa.GENO_SNP2
    assert np.array_equal(words, d["words"])
    assert np.array_equal(_unpack_snp2(words, 37, 300), d["geno"])
    lpad = words.size // 3
    assert lpad % 256 == 0 and lpad >= 300
    pad = words.reshape(3, lpad)[:, 300:]
    assert np.all(pad == 0xffffffff)                       # padded loci are missing
    assert np.all((words.reshape(3, lpad)[2, :300] >> np.uint32(2 * 5)) ==
                  np.uint32(0xffffffff >> 10))              # individuals 37..47 missing


def test_pack_byte_roundtrip_keeps_copy_order():
    rng = np.random.default_rng(3)
    geno, nall = random_geno(rng, 21, 77, 9, miss=0.15)
    fmt, words = sa.pack(geno)
    assert fmt == sa.GENO_BYTE   # half-missing sites and (1,0) hets force the byte path
    assert np.array_equal(_unpack_byte(words, 21, 77), geno)
    assert words.size == 77 * 64                      # 11 word rows padded to 64
    w = words.reshape(77, 64)
    assert np.all(w[:, 11:] == 0xffffffff)            # padded rows are missing
    assert np.all(w[:, 10] >> np.uint32(16) == 0xffff)  # individual 21 of the last row missing


def test_pack_byte32_takes_up_to_255_alleles():
    """sgpu_pack_byte32 packs the reference's int32 Geno with its own missing code:
    the same words as the int8 packer where both apply, alleles up to 254 beyond."""
    rng = np.random.default_rng(8)
    geno, nall = random_geno(rng, 19, 40, 9, miss=0.1)
    g32 = geno.astype(np.int32)
    g32[g32 == -9] = -999
    out = np.empty(sa.packed_words(sa.GENO_BYTE, 19, 40), np.uint32)
    rc = sa.lib().sgpu_pack_byte32(g32.ctypes.data_as(ctypes.c_void_p), 19, 40, -999,
                                   out.ctypes.data_as(ctypes.c_void_p), out.size)
    assert rc == sa.OK and np.array_equal(out, sa.pack(geno, sa.GENO_BYTE)[1])
    wide, _ = random_geno(rng, 19, 40, 255, miss=0.1)
    assert wide.dtype == np.int16 and wide.max() > 200
    fmt, words = sa.pack(wide)
    assert fmt == sa.GENO_BYTE and np.array_equal(_unpack_byte(words, 19, 40), wide)
    bad = wide.astype(np.int32)
    bad[0, 0, 0] = 255                                   # 255 is the missing byte
    rc = sa.lib().sgpu_pack_byte32(bad.ctypes.data_as(ctypes.c_void_p), 19, 40, -9,
                                   out.ctypes.data_as(ctypes.c_void_p), out.size)
    assert rc == sa.ERR_ARG


@pytest.mark.parametrize("bad", [(1, 0), (0, -9), (-9, 1), (2, 2)])
def test_pack_snp2_rejects_unrepresentable_sites(bad):
    g = np.zeros((3, 2, 5), np.int8)
    g[1, 0, 2], g[1, 1, 2] = bad
    words = np.empty(sa.packed_words(sa.GENO_SNP2, 3, 5), np.uint32)
    rc = sa.lib().sgpu_pack_snp2(g.ctypes.data_as(ctypes.c_void_p), 3, 5,
                                 words.ctypes.data_as(ctypes.c_void_p), words.size)
    assert rc == sa.ERR_UNSUPPORTED
    with pytest.raises(sa.SgpuError):
        sa.pack(g, sa.GENO_SNP2)


def _check_first_appearance(geno, nall):
    N, _, L = geno.shape
    for l in range(L):
        seen = []
        for i in range(N):
            for c in range(2):
                x = int(geno[i, c, l])
                if x == -9:
                    continue
                if x not in seen:
                    assert x == len(seen), f"locus {l}: allele {x} appears before {len(seen)}"
                    seen.append(x)
        assert nall[l] == len(seen)


def test_synth_snp_is_recoded_by_first_appearance_and_deterministic():
    a = sa.synth_snp(50, 200, Ktrue=4, seed=9, want_geno=True, threads=1)
    b = sa.synth_snp(50, 200, Ktrue=4, seed=9, want_geno=True, threads=7)
    assert np.array_equal(a["words"], b["words"])
    _check_first_appearance(a["geno"], a["nall"])
    miss = (a["geno"][:, 0] == -9).mean()
    assert 0.002 < miss < 0.03
    het = a["geno"][:, 0] != a["geno"][:, 1]
    assert np.all(a["geno"][:, 0][het] == 0)   # heterozygotes listed (0, 1)


def test_synth_microsat_is_recoded_by_first_appearance():
    d = sa.synth_microsat(40, 60, amin=2, amax=10, Ktrue=3, seed=5)
    _check_first_appearance(d["geno"], d["nall"])
    assert d["A"] == d["nall"].max() <= 10
    assert d["nall"].min() >= 1


def test_create_rejects_bad_shapes_before_touching_the_device():
    d = sa.synth_snp(20, 30, Ktrue=2, seed=1)
    with pytest.raises(sa.SgpuError) as e:
        sa.Chain(sa.make_model(20, 30, 2, 3), sa.GENO_SNP2, d["words"], d["nall"])
    assert e.value.code == sa.ERR_ARG            # SNP2 needs MAXALLELES == 2
    with pytest.raises(sa.SgpuError) as e:
        sa.Chain(sa.make_model(20, 30, 2, 2), sa.GENO_SNP2, d["words"][:-1], d["nall"])
    assert e.value.code == sa.ERR_ARG
    with pytest.raises(sa.SgpuError) as e:
        sa.Chain(sa.make_model(20, 30, 21, 2), sa.GENO_SNP2, d["words"], d["nall"])   # K > 20
    assert e.value.code == sa.ERR_UNSUPPORTED
    bad = d["nall"].copy()
    bad[bad == 2] = 1                            # allele 1 present but NumAlleles says 1
    with pytest.raises(sa.SgpuError) as e:
        sa.Chain(sa.make_model(20, 30, 2, 2), sa.GENO_SNP2, d["words"], bad)
    assert e.value.code == sa.ERR_ARG


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_create_batch_checks_arguments_before_touching_the_device():
    d = sa.synth_snp(20, 30, Ktrue=2, seed=1)
    m0 = sa.make_model(20, 30, 2, 2, burnin=5)
    m1 = sa.make_model(20, 30, 3, 2, burnin=6)
    with pytest.raises(sa.SgpuError) as e:          # BURNIN differs across the batch
        sa.Batch([m0, m1], [(sa.GENO_SNP2, d["words"], d["nall"])])
    assert e.value.code == sa.ERR_ARG
    m2 = sa.make_model(21, 30, 3, 2, burnin=5)
    with pytest.raises(sa.SgpuError) as e:          # shares a data set with another N
        sa.Batch([m0, m2], [(sa.GENO_SNP2, d["words"], d["nall"])])
    assert e.value.code == sa.ERR_ARG
    with pytest.raises(sa.SgpuError) as e:          # data set index out of range
        sa.Batch([m0], [(sa.GENO_SNP2, d["words"], d["nall"])], data_of_chain=[1])
    assert e.value.code == sa.ERR_ARG


def test_host_struct_layouts_match_the_headers(tmp_path):
    """The ctypes mirrors of shost_params / shost_input / shost_results have
    the C layout (sizes and two field offsets, from a program compiled here)."""
    import ctypes as C
    from structure_amd import host
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "structure_host.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu", sizeof(shost_params),'
                   ' offsetof(shost_params, seed), offsetof(shost_params, emerror),'
                   ' sizeof(shost_input), sizeof(shost_results)); return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    want = [C.sizeof(host.Params), host.Params.seed.offset, host.Params.emerror.offset,
            C.sizeof(host.Input), C.sizeof(host.Results)]
    assert [int(x) for x in got] == want


def test_tuning_is_validated_before_touching_the_device():
    d = sa.synth_snp(20, 30, Ktrue=2, seed=1)
    m = sa.make_model(20, 30, 2, 2)
    for bad in (sa.Tuning(byte_loci=-1), sa.Tuning(byte_loci=512), sa.Tuning(z_blocks=-5)):
        with pytest.raises(sa.SgpuError) as e:
            sa.Batch([m], [(sa.GENO_SNP2, d["words"], d["nall"])], tuning=bad)
        assert e.value.code == sa.ERR_ARG
