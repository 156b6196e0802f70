"""The command-line tools over the C ABI: bcmgx_solve (driverSolve.cu:59-406)
and bcmgx_compare (compareSolutions.cu:20-71), run as the reference's
regression loop runs them (launchAndCompare.mk: solve A x = 1 from x = 0 with a
settings file, write the solution with -o, compare it with the reference's
solution file at 1e-5)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bootcmatchgx_amd", "bin")
SOLVE = os.path.join(BIN, "bcmgx_solve")
COMPARE = os.path.join(BIN, "bcmgx_compare")


def run(*args, timeout=120):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module", autouse=True)
def tools():
    if not (os.path.exists(SOLVE) and os.path.exists(COMPARE)):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "bootcmatchgx_amd", "csrc")])


# ------------------------------------------------------------- CPU checks
def test_compare_tool(tmp_path, golden):
    ref = os.path.join(golden, "lap3d_7p_10x10x10_1x1x1_sol.txt")
    r = run(COMPARE, ref, ref, "1e-5")
    assert r.returncode == 0 and r.stdout.strip() == "Relative error: 0"
    x = np.loadtxt(ref)
    y = x * (1 + 1e-3)
    f = tmp_path / "y.txt"
    np.savetxt(f, y, fmt="%g")
    r = run(COMPARE, ref, f, "1e-5")
    assert r.returncode == 1 and "exceeds threshold" in r.stdout
    got = float(r.stdout.split("Relative error:")[1].split()[0])
    assert abs(got - 1e-3) < 2e-5  # %g keeps 6 digits
    r = run(COMPARE, ref, f)  # no threshold: report only
    assert r.returncode == 0
    z = tmp_path / "z.txt"
    np.savetxt(z, np.zeros(5), fmt="%g")
    r = run(COMPARE, z, z)
    assert r.returncode == 1 and "sol1 must be != <0>" in r.stdout
    r = run(COMPARE, ref)
    assert r.returncode == 1 and "Usage" in r.stdout


def test_solver_rejects_bad_command_lines(tmp_path, golden):
    """Errors found before any device work: usage, missing files, settings format."""
    assert "Usage" in run(SOLVE).stdout
    assert run(SOLVE, "-h").returncode == 1
    props = os.path.join(golden, "FCG_BCMG.properties")
    assert "Usage" in run(SOLVE, "-g", "9p", "-a", "10", "-s", props).stdout
    r = run(SOLVE, "-l", tmp_path / "missing.cfg", "-s", props)
    assert r.returncode == 1 and "not found" in r.stderr
    bad = tmp_path / "bad.cfg"
    bad.write_text("nx = 10\nny\n")
    assert run(SOLVE, "-l", bad, "-s", props).returncode == 1
    r = run(SOLVE, "-a", "10", "-s", os.path.join(golden, "lap3d_10x10x10_1x1x1.cfg"))
    assert r.returncode == 1 and ".properties" in r.stderr


# ------------------------------------------------------------- GPU runs
REGRESSION = [
    # (matrix options, settings, reference solution)
    (["-g", "7p", "-l", "lap3d_10x10x10_1x1x1.cfg"], "FCG_BCMG", "lap3d_7p_10x10x10_1x1x1_sol.txt"),
    (["-g", "7p", "-l", "lap3d_10x10x10_2x1x1.cfg"], "FCG_BCMG", "lap3d_7p_10x10x10_2x1x1_sol.txt"),
    (["-g", "7p", "-l", "lap3d_10x10x10_2x2x1.cfg"], "FCG_BCMG", "lap3d_7p_10x10x10_2x2x1_sol.txt"),
    (["-g", "7p", "-l", "lap3d_10x10x10_2x2x2.cfg"], "FCG_BCMG", "lap3d_7p_10x10x10_2x2x2_sol.txt"),
    (["-g", "27p", "-l", "lap3d_10x10x10_1x1x1.cfg"], "FCG_BCMG", "lap3d_27p_10x10x10_1x1x1_sol.txt"),
    (["-g", "27p", "-l", "lap3d_10x10x10_2x1x1.cfg"], "FCG_BCMG", "lap3d_27p_10x10x10_2x1x1_sol.txt"),
    (["-g", "27p", "-l", "lap3d_10x10x10_2x2x1.cfg"], "FCG_BCMG", "lap3d_27p_10x10x10_2x2x1_sol.txt"),
    (["-g", "27p", "-l", "lap3d_10x10x10_2x2x2.cfg"], "FCG_BCMG", "lap3d_27p_10x10x10_2x2x2_sol.txt"),
    (["-a", "10"], "FCG_BCMG", "lap3d_7p_10x10x10_1x1x1_sol.txt"),
    (["-m", "bcsstk09.mtx"], "FCG_BCMG", "bcsstk09_sol.txt"),
] + [
    (["-g", "27p", "-l", "lap3d_10x10x10_2x2x1.cfg"], s, "lap3d_27p_10x10x10_2x2x1_sol.txt")
    for s in ["FCG_AFSAI", "FCG_l1Jacobi", "FCG_noPreconditioner", "CGHS_BCMG", "CGHS_AFSAI", "CGS2_BCMG",
              "CGS3_l1Jacobi", "CGS_CUBLAS2_BCMG", "PipelinedCGS3_BCMG", "PipelinedCGS2_AFSAI",
              "PipelinedCGS1_l1Jacobi"]
]


@pytest.mark.gpu
@pytest.mark.parametrize("mopts,settings,sol", REGRESSION)
def test_regression_loop(tmp_path, golden, mopts, settings, sol):
    args = [a if not a.endswith((".cfg", ".mtx")) else os.path.join(golden, a) for a in mopts]
    out, info = tmp_path / "x.txt", tmp_path / "info.txt"
    r = run(SOLVE, *args, "-s", os.path.join(golden, settings + ".properties"), "-o", out, "-i", info)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Solving" in r.stdout and "...done." in r.stdout
    c = run(COMPARE, os.path.join(golden, sol), out, "1e-5")
    print(settings, mopts[-1], c.stdout.strip())
    assert c.returncode == 0, c.stdout
    txt = info.read_text()
    assert "Total iterations" in txt and "*** Exceeded max iter ***" not in txt


@pytest.mark.gpu
def test_dump_matrix_and_log(tmp_path, golden):
    """-d writes the generated matrix (the reference's MATLAB-made fixture for the same grid), -e the log."""
    d, e = tmp_path / "A", tmp_path / "log"
    r = run(SOLVE, "-g", "7p", "-l", os.path.join(golden, "lap3d_10x10x10_2x1x1.cfg"), "-s",
            os.path.join(golden, "FCG_BCMG.properties"), "-d", d, "-e", e)
    assert r.returncode == 0, r.stderr
    import scipy.io

    A = scipy.io.mmread(str(d) + "0.mtx").tocsr()
    assert A.shape == (2000, 2000) and (A != A.T).nnz == 0 and np.all(A.diagonal() == 6)
    assert "Finalizing preconditioner" in (tmp_path / "log0").read_text()


@pytest.mark.gpu
def test_misspelt_solver_in_reference_settings_is_rejected(golden):
    """The reference's own PipelinedCGS2_BCMG.properties reads "solver = CGSPIPELINED_CGS": the
    key parses to INVALID and solve() stops with "Unhandled value for enum class SolverType"
    (solve.cu:41-42), here as an error status and exit code 1."""
    r = run(SOLVE, "-a", "10", "-s", os.path.join(golden, "PipelinedCGS2_BCMG.properties"))
    assert r.returncode == 1 and "Unhandled value for enum class SolverType" in r.stderr
