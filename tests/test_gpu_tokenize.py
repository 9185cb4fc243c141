"""The device tokenizer equals the oracle's tokenizer (oracle/fgss_oracle.c, pinned to the
reference's fixture bytes and to the reference's own tokenizer sources) byte for byte in
the .bin and .loc it writes, and the host tokenizer field for field: ids, lengths,
tokens, q-gram offsets, term count and the dictionary (checked through probes
tokenized against it)."""
import sys

import numpy as np
import pytest

from _util import dblp_text, golden_path

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _assert_same(fg, host, dev, probe_text=None):
    for a, b, name in zip(host.arrays(), dev.arrays(), ("id", "len", "pos", "tok")):
        assert np.array_equal(a, b), name
    hi, di = host.info, dev.info
    assert (hi.n_sets, hi.n_tokens, hi.n_terms, hi.max_len, hi.sorted_desc) == \
        (di.n_sets, di.n_tokens, di.n_terms, di.max_len, di.sorted_desc)
    assert np.array_equal(host.bin_ints(), dev.bin_ints())
    if probe_text is not None and hi.n_terms > 0:
        ph = fg.Collection.from_text_dict(probe_text, host)
        pd = fg.Collection.from_text_dict(probe_text, dev)
        for a, b in zip(ph.arrays(), pd.arrays()):
            assert np.array_equal(a, b)


def _loc(c, tmp_path, name):
    p = str(tmp_path / name)
    c.write_bin(p, p + ".loc")
    return open(p, "rb").read(), open(p + ".loc", "rb").read()


def _assert_oracle(oracle_mod, dev, text, q, tmp_path):
    """the device collection's .bin and .loc bytes equal the oracle tokenizer's"""
    b, l = oracle_mod.tokenize(text, q)
    db, dl = _loc(dev, tmp_path, "o.bin")
    assert db == b.astype("<i4").tobytes(), "bin"
    assert dl == l.astype("<i4").tobytes(), "loc"


def test_fixture_bytes(fg, tmp_path):
    text = open(golden_path("small.txt"), "rb").read()
    dev = fg.Collection.from_text_bytes(text, 3, device=True)
    b, l = _loc(dev, tmp_path, "d.bin")
    assert b == open(golden_path("small.txt.3gram.bin"), "rb").read()
    assert l == open(golden_path("small.txt.3gram.bin.loc"), "rb").read()


@pytest.mark.parametrize("q", [1, 2, 3, 4, 5, 7, 8])
def test_dblp_all_q(fg, oracle_mod, tmp_path, q):
    text = dblp_text()
    _say(f"dblp q={q}")
    host = fg.Collection.from_text_bytes(text, q)
    dev = fg.Collection.from_text_bytes(text, q, device=True)
    _assert_oracle(oracle_mod, dev, text, q, tmp_path)
    _assert_same(fg, host, dev, probe_text=text[:20000] + b"\nnever seen zqxj words here\n")
    assert _loc(host, tmp_path, "h.bin") == _loc(dev, tmp_path, "d.bin")


@pytest.mark.parametrize("kind,n", [(0, 200000), (1, 20000)])
def test_generated(fg, oracle_mod, tmp_path, kind, n):
    text = fg.generate_text(kind, n, 31)
    _say(f"generated kind={kind} n={n}")
    host = fg.Collection.from_text_bytes(text, 3)
    dev = fg.Collection.from_text_bytes(text, 3, device=True)
    _assert_oracle(oracle_mod, dev, text, 3, tmp_path)
    _assert_same(fg, host, dev, probe_text=fg.generate_text(kind, 500, 99))


EDGE_TEXTS = [
    b"",
    b"\n\n\n",
    b"ab\ncd\n",                                   # every line shorter than q
    b"no trailing newline",
    b"\nLeading and Trailing\n\n",
    b"CRLF line\r\nanother one\r\n",
    b"MiXeD CaSe ABC abc AbC\n",
    b"aaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaa\n",     # one gram repeated: multi-digit occurrences
    b"x" * 5000 + b"\n" + b"ab" * 3000 + b"\n",    # long lines, occurrence numbers > 1000
    b"nul\x00inside\x00\x00 the text\n\x00\x00\x00\x00\n",
    b"\xff\xfe high bytes \x80\x81\n",
    b"same line\nsame line\nsame line\n",
]


@pytest.mark.parametrize("k", range(len(EDGE_TEXTS)))
@pytest.mark.parametrize("q", [1, 3, 6])
def test_edge_texts(fg, oracle_mod, tmp_path, k, q):
    text = EDGE_TEXTS[k]
    host = fg.Collection.from_text_bytes(text, q)
    dev = fg.Collection.from_text_bytes(text, q, device=True)
    _assert_oracle(oracle_mod, dev, text, q, tmp_path)
    _assert_same(fg, host, dev, probe_text=b"probe line\n" + text[:100])


def test_rejects(fg):
    with pytest.raises(fg.FgssError):
        fg.Collection.from_text_bytes(b"abc\n", 9, device=True)
    with pytest.raises(fg.FgssError):
        fg.Collection.from_text_bytes(b"abc\n", 0, device=True)


def test_device_tokens_join_like_host(fg):
    text = fg.generate_text(fg.GEN_TITLES, 30000, 8)
    dev = fg.Collection.from_text_bytes(text, 3, device=True)
    host = fg.Collection.from_text_bytes(text, 3)
    a, b = dev.selfjoin(0.7), host.selfjoin(0.7)
    assert np.array_equal(a.q, b.q) and np.array_equal(a.s, b.s) and np.array_equal(a.overlap, b.overlap)
