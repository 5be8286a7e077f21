"""QP files in the reference's binary layout (writeQP / loadQP,
MPRGPUtility.h:406-487): A and J as {size_t rows, size_t cols, size_t nnz,
nnz x Eigen::Triplet<double> = (int32 row, int32 col, double value)}, then B,
c and x0 as doubles, in the order A, B, J, c, x0.  The layout below is built
byte by byte with struct, independently of the library's writer."""
import os
import shutil
import struct

import numpy as np
import pytest
import scipy.sparse as sp

import mprgp_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def qp_dir():
    d = os.path.join(ROOT, "tests", "_tmp", "qp")
    os.makedirs(d, exist_ok=True)
    yield d
    shutil.rmtree(d, ignore_errors=True)


def sparse_bytes(rows, cols, trips):
    out = struct.pack("<QQQ", rows, cols, len(trips))
    for r, c, v in trips:
        out += struct.pack("<iid", r, c, v)
    return out


def doubles(v):
    return struct.pack(f"<{len(v)}d", *v)


def tiny_qp():
    # 2 nodes (n = 6), A with an unsorted triplet order and a duplicate entry
    # split in two; one plane row on node 1 (J = n^T on its DoFs)
    a_trips = [(0, 0, 4.0), (1, 1, 4.0), (2, 2, 4.0), (3, 3, 5.0), (4, 4, 5.0), (5, 5, 2.5),
               (5, 5, 2.5), (0, 3, -1.0), (3, 0, -1.0), (2, 1, 0.5), (1, 2, 0.5)]
    B = [1.0, 2.0, 3.0, -1.0, -2.0, -3.0]
    j_trips = [(0, 4, 1.0), (0, 3, 0.0), (0, 5, 0.0)]
    c = [0.25]
    x0 = [0.0, 0.1, 0.2, 0.3, 0.4, 0.5]
    raw = sparse_bytes(6, 6, a_trips) + doubles(B) + sparse_bytes(1, 6, j_trips) + doubles(c) + doubles(x0)
    A = sp.coo_matrix(([t[2] for t in a_trips], ([t[0] for t in a_trips], [t[1] for t in a_trips])),
                      shape=(6, 6)).tocsr()
    J = sp.coo_matrix(([t[2] for t in j_trips], ([t[0] for t in j_trips], [t[1] for t in j_trips])),
                      shape=(1, 6)).tocsr()
    return raw, A, np.array(B), J, np.array(c), np.array(x0)


def test_reads_the_reference_layout(qp_dir):
    raw, A, B, J, c, x0 = tiny_qp()
    path = os.path.join(qp_dir, "tiny.qp")
    with open(path, "wb") as f:
        f.write(raw)
    assert M.qp_file_info(path) == (6, 1, A.nnz, J.nnz)
    A2, B2, J2, c2, x2 = M.read_qp(path)
    assert A2.has_sorted_indices and (A2 != A).nnz == 0
    assert A2[5, 5] == 5.0  # duplicate triplets add up (setFromTriplets)
    assert (J2 != J).nnz == 0
    assert np.array_equal(B2, B) and np.array_equal(c2, c) and np.array_equal(x2, x0)


def test_writer_emits_the_reference_layout(qp_dir):
    _, A, B, J, c, x0 = tiny_qp()
    path = os.path.join(qp_dir, "w.qp")
    M.write_qp(path, A, B, J, c, x0)
    Ac, Jc = A.tocoo(), J.tocsr().tocoo()
    a_trips = sorted(zip(Ac.row, Ac.col, Ac.data))
    j_trips = sorted(zip(Jc.row, Jc.col, Jc.data))
    expect = (sparse_bytes(6, 6, a_trips) + doubles(B) + sparse_bytes(1, 6, j_trips) + doubles(c)
              + doubles(x0))
    with open(path, "rb") as f:
        assert f.read() == expect


def test_round_trip_of_a_beam_qp(qp_dir):
    from mprgp_amd import problems
    from tests.common import scene_qp
    A, b, node, normal, c, x0 = scene_qp(problems.config1())
    J = problems.rows_to_J(node, normal, A.shape[0] // 3)
    path = os.path.join(qp_dir, "beam.qp")
    M.write_qp(path, A, b, J, c, x0)
    A2, b2, J2, c2, x2 = M.read_qp(path)
    assert (A2 != sp.csr_matrix(A)).nnz == 0 and (J2 != sp.csr_matrix(J)).nnz == 0
    assert np.array_equal(b2, b) and np.array_equal(c2, c) and np.array_equal(x2, x0)


def test_malformed_files_are_rejected(qp_dir):
    raw, *_ = tiny_qp()
    cases = {
        "missing.qp": None,
        "truncated.qp": raw[:-8],                                      # x0 short by one value
        "not_square.qp": sparse_bytes(6, 5, []) + raw[24 + 11 * 16:],
        "bad_index.qp": sparse_bytes(6, 6, [(6, 0, 1.0)]) + raw[24 + 11 * 16:],
        "j_columns.qp": raw[:24 + 11 * 16 + 48] + sparse_bytes(1, 5, []) + raw[24 + 11 * 16 + 48 + 24 + 48:],
    }
    for name, data in cases.items():
        path = os.path.join(qp_dir, name)
        if data is not None:
            with open(path, "wb") as f:
                f.write(data)
        with pytest.raises(M.MprgpError) as e:
            M.read_qp(path)
        assert e.value.code == -2, name  # MPRGP_EBADINPUT


def test_oversized_header_is_rejected_not_fatal(qp_dir):
    """A 24-byte file whose header claims 2^39 triplets: the counts are
    checked against the bytes left in the file before anything is sized by
    them, so the reader answers MPRGP_EBADINPUT.  Run in a child process, so a
    regression (an exception escaping the C ABI) fails this test rather than
    ending the test run."""
    import subprocess
    import sys
    path = os.path.join(qp_dir, "huge_nnz.qp")
    with open(path, "wb") as f:
        f.write(struct.pack("<QQQ", 3, 3, 1 << 39))
    code = ("import sys, mprgp_amd as M\n"
            "try:\n    M.qp_file_info(sys.argv[1])\nexcept M.MprgpError as e:\n    print('code', e.code)\n")
    r = subprocess.run([sys.executable, "-c", code, path], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip() == "code -2", r.stdout + r.stderr
