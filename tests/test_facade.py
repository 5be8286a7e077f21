"""The C++ facade (include/mprgp_amd.hpp), the reference's operator API on
the C ABI.  CPU: the header compiles on its own with the reference's call
shapes.  GPU: tests/cpp/facade_check (built by build()) runs the three solve
overloads, the projector, the matrix methods and the ICA comparator on a QP
file; each result is checked against the oracle and the Python drop-in."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "facade_check")


def test_header_compiles_alone(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "mprgp_amd.hpp"\n'
                   '#include <vector>\n'
                   'int main() { std::vector<double> x; (void)x; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # and the check program compiles against it without warnings
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), CHECK + ".cpp"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_facade_overloads_match_oracle():
    import mprgp_amd as M
    import oracle
    from mprgp_amd import problems
    from tests.common import scene_qp

    assert os.path.exists(CHECK), "build() compiles tests/cpp/facade_check"
    sc = problems.config1_flat()
    A, b, node, normal, c, _ = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    x0 = sc.x_n.ravel().copy()
    x0[1::3] -= 0.015  # the bottom layer starts inside the ground
    d = os.path.join(ROOT, "tests", "_tmp", "facade")
    os.makedirs(d, exist_ok=True)
    f = os.path.join(d, "qp.bin")
    M.write_qp(f, A, b, J, c, x0)
    tol, max_it = 1e-8, 1000
    r = subprocess.run([CHECK, f, repr(tol), str(max_it)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        name, *rest = line.split()
        out[name] = rest
    vec = lambda k: (int(out[k][0]), np.array([float(v) for v in out[k][1:]]))
    n = A.shape[0]

    P = oracle.Projector(J, c)
    pc, y = vec("project")
    assert pc == 0
    np.testing.assert_allclose(y, P.project(x0), rtol=0, atol=1e-14 * np.abs(x0).max())
    assert out["feasible"] == ["0", "1"]

    code_r, xr, _, _ = oracle.decoupled_solve(A, b, J, c, x0.copy(), tol=tol, max_it=max_it)
    assert code_r == 0
    xs = x0.copy()
    assert M.MPRGPDecoupledCon.solve(A, b, J, c, xs, tol, max_it) == 0
    for k in ("projector_overload", "jc_overload", "file_overload"):
        code, x = vec(k)
        assert code == 0, k
        assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr), k
    # the J/c and file overloads run the drop-in's own pipeline: bitwise equal.
    # The projector overload projects in a separate kernel, whose FMA
    # contraction may differ from the step kernel's by an ulp; from there the
    # two solves agree to the tolerance only (the bar above).
    for k in ("jc_overload", "file_overload"):
        assert np.array_equal(vec(k)[1], xs), k
    assert int(out["ev_size"][0]) in (0, n)

    _, x = vec("jc_overload")
    _, ax = vec("multiply")
    assert np.all(np.abs(ax - A @ x) <= 1e-13 * (abs(A) @ np.abs(x)))
    fv = float(out["funcValue"][0])
    assert abs(fv - oracle.func_value(A, x, b)) <= 1e-13 * (np.abs(x) @ (abs(A) @ np.abs(x)) + np.abs(x) @ np.abs(b))
    _, dg = vec("diag")
    assert np.array_equal(dg, A.diagonal())

    # ICASolver (ICAFemSolver's comparator) against the oracle's ICA
    code_i, xi, li, si = oracle.ica_solve(A, b, J, c, x0.copy(), tol=tol, max_it=max_it)
    ci, x_ica = vec("ica")
    it_i, lam_ica = vec("ica_lambda")
    assert code_i == 0 and ci == 0 and it_i == si["iterations"]
    assert np.linalg.norm(x_ica - xi) <= 1e-10 * np.linalg.norm(xi)
    np.testing.assert_allclose(lam_ica, li, rtol=1e-10, atol=1e-14)
