"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The CPU oracle (every entry: assembly LINEAR and corotated with damping,
  plane collision, power method, monotone/classic/unpreconditioned solves,
  projector and preconditioner) built with gcc -fsanitize=address,undefined
  and run on a small beam.
* The library's QP-file reader (mprgp_qpfile.cpp + mprgp_error.cpp, which
  need no GPU) built the same way with g++ and fed a valid file, the
  malformed files of the reader's contract (including headers whose counts
  exceed the file) and a hypothesis fuzz of byte mutations: every file must
  give MPRGP_OK or MPRGP_EBADINPUT, with no sanitizer report.

Both harnesses abort on the first report (-fno-sanitize-recover=all)."""
import os
import struct
import subprocess

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from mprgp_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TMP = os.path.join(ROOT, "tests", "_tmp", "san")
SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _build(out, cmd):
    os.makedirs(TMP, exist_ok=True)
    path = os.path.join(TMP, out)
    r = subprocess.run(cmd + ["-o", path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return path


@pytest.fixture(scope="module")
def oracle_san():
    src = [os.path.join(ROOT, "oracle", f) for f in ("mprgp_oracle.c", "fem_oracle.c", "collide_oracle.c",
                                                      "general_oracle.c", "ica_oracle.c")]
    return _build("oracle_san", ["gcc", "-std=c99", "-ffp-contract=off", *SAN, *src,
                                 os.path.join(ROOT, "tests", "cpp", "oracle_san.c"), "-lm"])


@pytest.fixture(scope="module")
def qpfile_san():
    csrc = os.path.join(ROOT, "mprgp_amd", "csrc")
    return _build("qpfile_san", ["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
                                 os.path.join(csrc, "mprgp_qpfile.cpp"), os.path.join(csrc, "mprgp_error.cpp"),
                                 os.path.join(ROOT, "tests", "cpp", "qpfile_san.cpp")])


def test_oracle_under_sanitizers(oracle_san):
    X, tets = problems.box_mesh(6, 2, 2, 0.05)
    X0 = problems.place(X, 5.0, 0.005)
    rng = np.random.default_rng(3)
    xn = X0 + rng.normal(0.0, 1e-3, X0.shape)
    vn = np.tile([0.0, -1.0, 0.0], (len(X0), 1)) + rng.normal(0.0, 0.05, X0.shape)
    path = os.path.join(TMP, "beam.txt")
    os.makedirs(TMP, exist_ok=True)
    with open(path, "w") as f:
        f.write(f"{len(X0)} {len(tets)}\n")
        for a in (X0, xn, vn):
            np.savetxt(f, a, fmt="%.17g")
        np.savetxt(f, tets, fmt="%d")
    r = subprocess.run([oracle_san, path], capture_output=True, text=True, env=ENV, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "oracle_san ok" in r.stdout and "rows=0 " not in r.stdout, r.stdout


def _sparse(rows, cols, trips, nnz=None):
    out = struct.pack("<QQQ", rows, cols, len(trips) if nnz is None else nnz)
    for r, c, v in trips:
        out += struct.pack("<iid", r, c, v)
    return out


def _valid_file():
    a = [(i, i, 4.0) for i in range(6)] + [(0, 3, -1.0), (3, 0, -1.0), (1, 2, 0.5), (2, 1, 0.5)]
    raw = _sparse(6, 6, a) + struct.pack("<6d", *range(6))
    raw += _sparse(1, 6, [(0, 4, 1.0)]) + struct.pack("<d", 0.25) + struct.pack("<6d", *np.linspace(0, 1, 6))
    return raw


def _run_reader(binary, blobs):
    paths = []
    for i, blob in enumerate(blobs):
        p = os.path.join(TMP, f"f{i}.qp")
        with open(p, "wb") as f:
            f.write(blob)
        paths.append(p)
    r = subprocess.run([binary, *paths], capture_output=True, text=True, env=ENV, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    codes = [int(line.split()[0]) for line in r.stdout.splitlines()]
    assert len(codes) == len(blobs)
    return codes


def test_reader_contract_cases_under_sanitizers(qpfile_san):
    raw = _valid_file()
    head_a = 24 + 10 * 16
    cases = [
        raw,                                                       # valid
        raw[:-8],                                                  # x0 short
        struct.pack("<QQQ", 3, 3, 1 << 39),                        # nnz far beyond the file
        struct.pack("<QQQ", 1 << 30, 1 << 30, 0),                  # rows beyond the file (B short)
        raw[:head_a + 48] + _sparse(1, 6, [], nnz=(1 << 60)),      # J nnz beyond the file
        raw[:head_a + 48] + _sparse(1 << 31, 6, []),               # J rows over int range
        _sparse(6, 6, [(6, 0, 1.0)]) + raw[head_a:],               # row index out of range
        _sparse(6, 6, [(0, -1, 1.0)]) + raw[head_a:],              # negative column
        b"",                                                       # empty
        raw[:17],                                                  # torn header
    ]
    codes = _run_reader(qpfile_san, cases)
    assert codes[0] == 0
    assert codes[1:] == [-2] * (len(cases) - 1), codes


@settings(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.too_slow])
@given(st.lists(st.tuples(st.integers(0, 10_000), st.integers(0, 255)), min_size=1, max_size=8),
       st.integers(0, 400), st.sampled_from([0, 8, 16, 24, 208, 256, 280]),
       st.sampled_from([0, 1, 2, 3, 1 << 20, 1 << 33, (1 << 64) - 1]))
def test_reader_fuzz_under_sanitizers(qpfile_san, flips, cut, field, value):
    """Byte mutations, truncations and header fields set to extreme counts."""
    raw = bytearray(_valid_file())
    for pos, byte in flips:
        raw[pos % len(raw)] = byte
    if field + 8 <= len(raw):
        raw[field:field + 8] = struct.pack("<Q", value)
    blob = bytes(raw[:len(raw) - cut if cut < len(raw) else 0])
    codes = _run_reader(qpfile_san, [blob])
    assert codes[0] in (0, -2), codes
