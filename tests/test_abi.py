"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/*.h, and the synthetic workloads have the stated sizes."""
import glob
import os
import re

import numpy as np

import mprgp_amd as M
from mprgp_amd import _lib, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(mprgp_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding table covers the whole header too
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_defaults_match_reference_facade():
    p = M.params()
    assert p.tol == 1e-3 and p.max_it == 1000          # MPRGPSolver.h:546
    assert p.variant == M.MONOTONIC and p.preconditioned == 1


def test_null_arguments_are_rejected_without_gpu():
    L = _lib.load()
    assert L.mprgp_qp_solve(None, None, None) == _lib.EBADINPUT
    assert L.mprgp_matrix_create_bsr3(None, 0, None, None, None, None) == _lib.EBADINPUT
    assert "NULL" in M.last_error() or "ctx" in M.last_error()


def test_workload_sizes():
    X, T = problems.box_mesh(20, 4, 4, 0.05)
    assert X.shape == (525, 3) and T.shape == (1600, 4)
    X, T = problems.box_mesh(100, 20, 20, 0.05)
    assert X.shape == (44541, 3) and T.shape == (200000, 4)
    assert np.all(problems.signed_volumes(X, T) > 0)
    # conforming 5-tet split: 18 edges per hex shared across cells
    e = np.sort(np.concatenate([T[:, [i, j]] for i in range(4) for j in range(i + 1, 4)]), 1)
    edges = np.unique(e[:, 0].astype(np.int64) * 10**6 + e[:, 1]).size
    assert edges == 253340 and 44541 + 2 * edges == 551221   # blocks of A


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror in mprgp_amd/_lib.py has the header's size and
    field offsets (gcc compiles a program that prints them)."""
    import ctypes
    import subprocess
    pairs = [("mprgp_params", _lib.Params), ("mprgp_stats", _lib.Stats),
             ("mprgp_step_params", _lib.StepParams), ("mprgp_step_stats", _lib.StepStats),
             ("mprgp_engine_config", _lib.EngineConfig), ("mprgp_engine_info", _lib.EngineInfo)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mprgp_amd.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.splitlines())
    for cname, py in pairs:
        assert int(got[f"{cname} size"]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname} {f}"]) == getattr(py, f).offset, (cname, f)
