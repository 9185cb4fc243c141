"""Host-side checks that need no GPU: the C ABI library, its q-gram front end,
.bin reader/writer, generator, error paths and the CLI's argument handling."""
import os
import re
import subprocess

import numpy as np
import pytest

from _util import dblp_text, filter_counts, golden_path, oracle_sets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(fg):
    header = open(os.path.join(ROOT, "include", "fgssjoin.h")).read()
    declared = set(re.findall(r"\b(fgss_[a-z_0-9]+)\s*\(", header))
    assert declared, "no declarations parsed"
    L = fg.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(fg.EXPORTS)


def test_ctypes_structs_match_the_header(fg, tmp_path):
    """The Python mirror's structures have the C header's sizes and field offsets
    (compiled with gcc against include/fgssjoin.h)."""
    import ctypes as C
    pairs = [("fgss_collection_info", fg.CollectionInfo), ("fgss_join_stats", fg.JoinStats),
             ("fgss_join_opts", fg.JoinOpts), ("fgss_batch_stats", fg.BatchStats),
             ("fgss_request_stats", fg.RequestStats), ("fgss_pairs", fg._Pairs)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fgssjoin.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                   check=True).stdout.splitlines())
    for cname, py in pairs:
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_text_front_end_matches_fixture(fg, tmp_path):
    coll = fg.Collection.from_text(golden_path("small.txt"), 3)
    out, loc = tmp_path / "s.bin", tmp_path / "s.bin.loc"
    coll.write_bin(str(out), str(loc))
    assert out.read_bytes() == open(golden_path("small.txt.3gram.bin"), "rb").read()
    assert loc.read_bytes() == open(golden_path("small.txt.3gram.bin.loc"), "rb").read()


def test_bin_reader_reverses_file_order(fg):
    coll = fg.Collection.from_bin(golden_path("small.txt.3gram.bin"))
    ids, lens, pos, tok = coll.arrays()
    assert ids.tolist() == [4, 2, 3, 1, 5]
    assert lens.tolist() == [24, 19, 18, 16, 11]
    assert pos.tolist() == [0, 24, 43, 61, 77]
    info = coll.info
    assert (info.n_sets, info.n_tokens, info.n_terms, info.max_len) == (5, 88, 42, 24)
    assert info.sorted_desc == 1 and info.on_device == 0
    assert np.array_equal(coll.bin_ints(), np.fromfile(golden_path("small.txt.3gram.bin"), "<i4"))


@pytest.mark.parametrize("q", [2, 3, 5])
def test_tokenizer_matches_oracle_dblp(fg, oracle_mod, q):
    text = dblp_text()
    b, _ = oracle_mod.tokenize(text, q)
    coll = fg.Collection.from_text_bytes(text, q)
    assert np.array_equal(coll.bin_ints(), b)


@pytest.mark.parametrize("kind,n", [(0, 4096), (1, 300)])
def test_tokenizer_matches_oracle_generated(fg, oracle_mod, kind, n):
    text = fg.generate_text(kind, n, 9)
    b, loc = oracle_mod.tokenize(text, 3)
    coll = fg.Collection.from_text_bytes(text, 3)
    assert np.array_equal(coll.bin_ints(), b)


def test_tokenizer_long_q_and_odd_bytes(fg, oracle_mod):
    text = b"ab\n\nA\x00B\x00CdEf\r\nzzzzzzzzzzzzzzzzzzzz\n\xff\xfe\xfd\xfc\xfb\xfa\xf9\xf8\xf7\n"
    for q in (1, 2, 3, 9, 12):
        b, _ = oracle_mod.tokenize(text, q)
        coll = fg.Collection.from_text_bytes(text, q)
        assert np.array_equal(coll.bin_ints(), b), q


def test_generator_is_seeded_and_shaped(fg):
    a = fg.generate_text(fg.GEN_TITLES, 5000, 2)
    assert a == fg.generate_text(fg.GEN_TITLES, 5000, 2)
    assert a != fg.generate_text(fg.GEN_TITLES, 5000, 3)
    lines = a.split(b"\n")[:-1]
    lens = np.array([len(x) for x in lines])
    assert len(lines) == 5000
    assert 55 < lens.mean() < 85 and lens.max() <= 125
    assert all(re.fullmatch(rb"[a-z ]+", x) for x in lines[:500])
    long = fg.generate_text(fg.GEN_LONG, 200, 3).split(b"\n")[:-1]
    ll = np.array([len(x) for x in long])
    assert 180 <= ll.min() and ll.max() <= 1215 and 550 < ll.mean() < 850


@pytest.mark.parametrize("kind,n,seed,t", [(0, 4096, 1, 0.6), (1, 3000, 13, 0.8), (0, 3000, 5, 0.9)])
def test_filter_restatement_candidates_match_oracle(fg, oracle_mod, kind, n, seed, t):
    """The host restatement of the join's filters (the checker of signature_checks and
    signature_checks_128 in the GPU tests) counts the oracle's candidates, and its later
    filters only narrow them."""
    coll = fg.Collection.generate(kind, n, seed)
    t = float(np.float32(t))
    cand, keep, pass64 = filter_counts(coll, t)
    _, ctr = oracle_mod.selfjoin(oracle_sets(oracle_mod, coll), t)
    assert cand == ctr["candidates"]
    assert ctr["intersections"] <= pass64 <= keep <= cand


def test_errors_are_status_codes(fg, tmp_path):
    with pytest.raises(fg.FgssError) as e:
        fg.Collection.from_bin(str(tmp_path / "missing.bin"))
    assert e.value.status == fg.FGSS_ERR_IO
    bad = tmp_path / "trunc.bin"
    bad.write_bytes(np.array([1, 5, 1, 2], "<i4").tobytes())
    with pytest.raises(fg.FgssError) as e:
        fg.Collection.from_bin(str(bad))
    assert e.value.status == fg.FGSS_ERR_FORMAT
    with pytest.raises(fg.FgssError) as e:
        fg.Collection.from_text_bytes(b"abc\n", 0)
    assert e.value.status == fg.FGSS_ERR_ARG
    coll = fg.Collection.from_bin(golden_path("small.txt.3gram.bin"))
    for t in (0.0, 1.0, -0.5, 1.5):
        with pytest.raises(fg.FgssError) as e:
            coll.selfjoin(t)
        assert e.value.status == fg.FGSS_ERR_ARG
    # a .bin in any record order loads (and joins: tests/test_gpu_join.py); non-ascending
    # tokens are refused before any device work
    unsorted = fg.Collection.from_bin_ints(np.array([1, 3, 1, 2, 3, 2, 1, 5], np.int32))
    assert unsorted.info.n_sets == 2 and unsorted.info.sorted_desc == 0
    dup = fg.Collection.from_bin_ints(np.array([1, 3, 1, 1, 3], np.int32))
    with pytest.raises(fg.FgssError) as e:
        dup.upload()
    assert e.value.status == fg.FGSS_ERR_UNSUPPORTED


def test_trailing_partial_int_is_ignored(fg, tmp_path):
    raw = open(golden_path("small.txt.3gram.bin"), "rb").read()
    p = tmp_path / "tail.bin"
    p.write_bytes(raw + b"\x01\x02")
    assert fg.Collection.from_bin(str(p)).info.n_sets == 5
    p.write_bytes(raw + b"\x01\x00\x00\x00\x02")   # an id without a size: read error
    with pytest.raises(fg.FgssError):
        fg.Collection.from_bin(str(p))


def _cli(fg, *args):
    return subprocess.run([fg.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def test_cli_argument_errors(fg, tmp_path):
    r = _cli(fg)
    assert r.returncode == 1 and "Usage" in r.stdout
    r = _cli(fg, str(tmp_path / "nope.bin"), "0.5")
    assert r.returncode == 1 and "does not exist" in r.stderr
    for t in ("0", "1", "1.5", "-2", "abc"):
        r = _cli(fg, golden_path("small.txt.3gram.bin"), t)
        assert r.returncode == 1 and "Threshold must be" in r.stderr, t
    r = _cli(fg, "-f", golden_path("small.txt"), "-q", "0", "-t", "0.5")
    assert r.returncode == 1


def test_token_id_range_rejected_before_device_work(fg):
    """A token id of 2^31-1 would make n_terms = max id + 1 wrap in int32 (the index's
    list offsets then hold fewer entries than a record's ids reach): the loader
    refuses it; ids up to 2^31-2 still load with n_terms = max id + 1."""
    with pytest.raises(fg.FgssError) as e:
        fg.Collection.from_bin_ints(np.array([7, 1, 2147483647, 8, 2, 3, 5], np.int32))
    assert e.value.status == fg.FGSS_ERR_UNSUPPORTED
    c = fg.Collection.from_bin_ints(np.array([7, 1, 2147483646, 8, 2, 3, 5], np.int32))
    assert c.info.n_terms == 2147483647
    with pytest.raises(fg.FgssError) as e:
        c.upload()   # n_terms + 1 list offsets would not fit the index's 32-bit term space
    assert e.value.status == fg.FGSS_ERR_UNSUPPORTED
    neg = fg.Collection.from_bin_ints(np.array([7, 2, -4, 1], np.int32))
    with pytest.raises(fg.FgssError) as e:
        neg.upload()   # validated on the host before the device is touched
    assert e.value.status == fg.FGSS_ERR_UNSUPPORTED


def test_noalloc_argument_and_size_checks(fg):
    """fgss_selfjoin_noalloc refuses bad arguments and collections beyond its static
    scratch before any device call (so these run without a GPU)."""
    def err(**kw):
        a = dict(tokens=0, lengths=0, offsets=0, n_sets=10, n_tokens=10, n_terms=10, threshold=0.5, out=0,
                 out_cap=0)
        a.update(kw)
        with pytest.raises(fg.FgssError) as e:
            fg.selfjoin_noalloc(**a)
        return e.value.status
    assert err() == fg.FGSS_ERR_ARG                      # null record arrays
    assert err(tokens=8, lengths=8, offsets=8, threshold=1.0) == fg.FGSS_ERR_ARG
    assert err(tokens=8, lengths=8, offsets=8, out_cap=5) == fg.FGSS_ERR_ARG   # null output
    assert err(tokens=8, lengths=8, offsets=8, n_sets=fg.NOALLOC_MAX_SETS + 1) == fg.FGSS_ERR_UNSUPPORTED
    assert err(tokens=8, lengths=8, offsets=8, n_tokens=fg.NOALLOC_MAX_TOKENS + 1) == fg.FGSS_ERR_UNSUPPORTED
    assert err(tokens=8, lengths=8, offsets=8, n_terms=fg.NOALLOC_MAX_TERMS + 1) == fg.FGSS_ERR_UNSUPPORTED
    assert err(tokens=8, lengths=8, offsets=8, table_slots=300) == fg.FGSS_ERR_ARG
    assert fg.selfjoin_noalloc(0, 0, 0, 0, 0, 0, 0.5, 0, 0) == (0, fg.JoinStats().as_dict())
    header = open(os.path.join(ROOT, "include", "fgssjoin.h")).read()
    for name, v in (("SETS", fg.NOALLOC_MAX_SETS), ("TOKENS", fg.NOALLOC_MAX_TOKENS), ("TERMS", fg.NOALLOC_MAX_TERMS)):
        assert re.search(rf"#define FGSS_NOALLOC_MAX_{name} {v}\b", header)
