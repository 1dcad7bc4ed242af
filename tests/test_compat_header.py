"""The C++ drop-in header (include/mdtw_compat.hpp) compiles for a plain
host C++ caller -- g++, no HIP headers -- with the reference drivers' call
forms (GPU-dtw.cu:67-96, GPU-ed.cu:62-82, GPU-ldtw.cu:72-98), and every
symbol it forwards to is declared in include/mdtw.h."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CALLER = r'''
#include "mdtw_compat.hpp"
#define EUCLIDEAN (true)
#define MANHATTAN (false)
void caller(float* Subject, float* AvgS, float* StdS, float* out, int* Indices,
            float* zquery, float* zlower, float* zupper, double* q64, float* R,
            double* S64, double* A64, double* D64) {
    int M = 128, N = 1 << 20, W = M * (10 * 0.01);
    znormalize(zquery, M);
    znormalize(q64, M);
    envelope(zquery, W, zlower, zupper, M);
    mdtw_compat::set_query(zquery, zlower, zupper, M, W);
    avg_std<double>(Subject, AvgS, StdS, M, N);
    lb_kim<EUCLIDEAN>(Subject, AvgS, StdS, out, M, Indices, 100);
    lb_kim<MANHATTAN>(Subject, AvgS, StdS, out, M, N);
    lb_keogh<EUCLIDEAN>(Subject, AvgS, StdS, out, M, Indices, 100);
    lb_keogh<EUCLIDEAN>(Subject, AvgS, StdS, out, M, N, nullptr);
    gpu_cdtw<EUCLIDEAN>(Subject, AvgS, StdS, out, M, Indices, 100, W);
    mdtw_compat::match r = prune_cdtw<EUCLIDEAN>(Subject, AvgS, StdS, M, N, W);
    double bsf = calculate_ed(Subject, StdS, M, N);
    mdtw_compat::ztsh2_ldtw<true>(Subject, M, N, AvgS, StdS, R, W);
    mdtw_compat::tsh2_ldtw<false>(Subject, M, N, R, W);
    determine_min<true>(R, N - M + 1);
    // the ED driver's FP64 form (GPU-ed.cu:62-82)
    mdtw_compat::set_query(q64, M);
    avg_std<double>(S64, A64, D64, M, N);
    double bsf64 = calculate_ed(S64, D64, M, N);
    (void)r; (void)bsf; (void)bsf64;
}
'''


def test_header_compiles_with_plain_gxx(tmp_path):
    src = tmp_path / "caller.cpp"
    src.write_text(CALLER)
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_header_forwards_only_declared_abi():
    hdr = open(os.path.join(ROOT, "include", "mdtw_compat.hpp")).read()
    abi = open(os.path.join(ROOT, "include", "mdtw.h")).read()
    used = set(re.findall(r"\b(mdtw_[a-z0-9_]+)\(", hdr))
    declared = set(re.findall(r"\bint\s+(mdtw_[a-z0-9_]+)\(|char\*\s+(mdtw_[a-z0-9_]+)\(", abi))
    declared = {a or b for a, b in declared}
    assert used and used <= declared, used - declared


@pytest.mark.parametrize("prog", ["gpu_dtw", "gpu_ed", "gpu_ldtw"])
def test_cli_built(prog):
    path = os.path.join(ROOT, "cudadtw_amd", "bin", prog)
    if not os.path.exists(path):
        pytest.skip("drivers not built (make -C cudadtw_amd/cli)")
    r = subprocess.run([path], capture_output=True, text=True)
    assert r.returncode == 1 and "query.bin subject.bin M N" in r.stdout
