"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/sanitize/Makefile builds, with g++/gcc and no device code:
* collection_check -- the .bin reader on the reference's fixture, hand-made malformed streams
  and 4,000 seeded mutations of the fixture; the text front end at q = 1..12 on the fixture
  text, edge texts and generated text; probes against a dictionary; write + reload;
* fgssjoin_san -- the command-line tool (cli.cpp) whose device entry points report
  FGSS_ERR_DEVICE (tests/sanitize/no_device.cpp): argument handling, loading, the debug
  dump, the .bin/.loc writer and every error path up to the first device call;
* oracle_check -- every join of the oracle (closed form with and without self-checks,
  query subsets, the reference's tile loop at block sizes 1 / 7 / n, brute force, R-S),
  cross-checked as in test_oracle.py.
Any sanitizer report fails the run (-fno-sanitize-recover=all; leaks are reported at exit).
"""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

from _util import golden_path

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
BUILD = os.path.join(SAN, "build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
REPORTS = ("AddressSanitizer", "LeakSanitizer", "runtime error:", "UndefinedBehaviorSanitizer")


@pytest.fixture(scope="module")
def san_build():
    if not shutil.which("make") or not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("make/g++/gcc not available")
    r = subprocess.run(["make", "-C", SAN, "-j4"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return BUILD


def _run(args, **kw):
    r = subprocess.run(args, capture_output=True, env=ENV, timeout=300, **kw)
    err = r.stderr.decode("latin-1")
    for rep in REPORTS:
        assert rep not in err, err[-4000:]
    return r.returncode, r.stdout.decode("latin-1"), err


def test_collection_code(san_build, tmp_path):
    rc, out, err = _run([os.path.join(san_build, "collection_check"), golden_path("small.txt.3gram.bin"),
                         golden_path("small.txt"), str(tmp_path)])
    assert rc == 0, out + err
    assert "0 failed checks" in out


def test_oracle(san_build, tmp_path, fg):
    texts = [(golden_path("small.txt"), 3, ["0.1", "0.5", "0.8", "0.95"]),
             (golden_path("small.txt"), 1, ["0.3", "0.7"])]
    with gzip.open(golden_path("dblp_t_18k.txt.gz"), "rb") as f:
        dblp = b"\n".join(f.read().split(b"\n")[:600]) + b"\n"
    p = tmp_path / "dblp600.txt"
    p.write_bytes(dblp)
    texts.append((str(p), 3, ["0.5", "0.8"]))
    g = tmp_path / "titles.txt"
    g.write_bytes(fg.generate_text(fg.GEN_TITLES, 700, 5))
    texts.append((str(g), 3, ["0.6", "0.9"]))
    for path, q, ts in texts:
        rc, out, err = _run([os.path.join(san_build, "oracle_check"), path, str(q)] + ts)
        assert rc == 0, (path, q, out + err)
        assert "0 failed checks" in out


def test_cli(san_build, tmp_path):
    cli = os.path.join(san_build, "fgssjoin_san")
    fix_bin, fix_txt = golden_path("small.txt.3gram.bin"), golden_path("small.txt")
    rc, out, _ = _run([cli])
    assert rc == 1 and "Usage" in out
    # a .bin through loading and the debug dump to the first device call
    rc, out, err = _run([cli, fix_bin, "0.8", "debug"])
    assert rc == 1 and "# Records: 5" in err and "no device" in err
    assert out.count("id: [") == 5
    # the host tokenizer and writer: the reference's fixture bytes
    ob = str(tmp_path / "out.bin")
    rc, _, err = _run([cli, "-f", fix_txt, "-q", "3", "-t", "0.8", "--host-tokenizer", "-o", ob])
    assert rc == 1 and "no device" in err
    assert open(ob, "rb").read() == open(fix_bin, "rb").read()
    assert open(ob + ".loc", "rb").read() == open(fix_bin + ".loc", "rb").read()
    # error paths
    rc, _, err = _run([cli, str(tmp_path / "missing.bin"), "0.5"])
    assert rc == 1 and "does not exist" in err
    for bad in ("abc", "0", "1", "1.5", "-0.2"):
        rc, _, err = _run([cli, fix_bin, bad])
        assert rc == 1 and "Threshold must be" in err, bad
    for args in (["-f", fix_txt, "-q", "0", "-t", "0.5"], ["-f", fix_txt, "-t", "0.5"], ["-x"] * 3):
        rc, out, _ = _run([cli] + args)
        assert rc == 1 and "Usage" in out, args
    # malformed and mutated .bin files
    raw = np.fromfile(fix_bin, "<i4")
    rng = np.random.default_rng(12)
    for k in range(40):
        v = raw.copy()
        if k % 3 == 0:
            v = v[: int(rng.integers(0, v.size))]
        else:
            for _ in range(int(rng.integers(1, 4))):
                v[int(rng.integers(v.size))] = rng.choice([0, -1, 1, 65535, 2**31 - 1, -2**31, int(rng.integers(-2**31, 2**31))])
        b = tmp_path / f"m{k}.bin"
        v.astype("<i4").tofile(b)
        if k % 5 == 0:   # a stray trailing byte
            with open(b, "ab") as f:
                f.write(b"\x01")
        rc, out, err = _run([cli, str(b), "0.7"])
        assert rc == 1, k
        assert "Reading error." in out or "no device" in err or "ids must be below" in err, (k, out, err)
