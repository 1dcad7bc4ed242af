"""The reference's CPU paths in the C++ drop-in header (include/mdtw_compat.hpp):
cpu_cdtw (cdtw.cuh:437-448, sequential_cdtw / openmp_cdtw), z_stream_ldtw /
omp_z_stream_ldtw (ldtw.cuh:151-174) and stream_ldtw / omp_stream_ldtw
(:130-149), compiled by g++ for a plain host caller and compared bit for bit
with the oracle's restatements (or_cdtw with host rounding, or_ldtw,
or_ldtw_raw) on seeded inputs.  No GPU: these functions are host loops."""
import os
import subprocess

import numpy as np
import pytest

from cudadtw_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include "mdtw_compat.hpp"
#include <cstdio>
#include <vector>
template <class T> std::vector<T> load(const char* p, long n) {
    std::vector<T> v(n); FILE* f = std::fopen(p, "rb");
    if (!f || std::fread(v.data(), sizeof(T), n, f) != (size_t)n) std::exit(2);
    std::fclose(f); return v;
}
template <class T> void save(const char* p, const std::vector<T>& v) {
    FILE* f = std::fopen(p, "wb"); std::fwrite(v.data(), sizeof(T), v.size(), f); std::fclose(f);
}
int main(int argc, char** argv) {
    // dir M N W metric
    std::string d = argv[1];
    int M = std::atoi(argv[2]), N = std::atoi(argv[3]), W = std::atoi(argv[4]);
    bool l2 = std::atoi(argv[5]) != 0;
    int L = N - M + 1;
    auto q = load<float>((d + "/q").c_str(), M);
    auto s = load<float>((d + "/s").c_str(), N);
    auto a = load<float>((d + "/a").c_str(), L);
    auto sd = load<float>((d + "/sd").c_str(), L);
    auto idx = load<int>((d + "/idx").c_str(), L);
    std::vector<float> r(L);
    for (int par = 0; par < 2; ++par) {
        if (l2) cpu_cdtw<true>(q.data(), s.data(), a.data(), sd.data(), r.data(), M, idx.data(), L, W, par != 0);
        else cpu_cdtw<false>(q.data(), s.data(), a.data(), sd.data(), r.data(), M, idx.data(), L, W, par != 0);
        save((d + "/cdtw" + std::to_string(par)).c_str(), r);
        if (par) { if (l2) omp_z_stream_ldtw<true>(q.data(), M, s.data(), N, a.data(), sd.data(), r.data(), W);
                   else omp_z_stream_ldtw<false>(q.data(), M, s.data(), N, a.data(), sd.data(), r.data(), W); }
        else { if (l2) z_stream_ldtw<true>(q.data(), M, s.data(), N, a.data(), sd.data(), r.data(), W);
               else z_stream_ldtw<false>(q.data(), M, s.data(), N, a.data(), sd.data(), r.data(), W); }
        save((d + "/ldtw" + std::to_string(par)).c_str(), r);
        if (par) { if (l2) omp_stream_ldtw<true>(q.data(), M, s.data(), N, r.data(), W);
                   else omp_stream_ldtw<false>(q.data(), M, s.data(), N, r.data(), W); }
        else { if (l2) stream_ldtw<true>(q.data(), M, s.data(), N, r.data(), W);
               else stream_ldtw<false>(q.data(), M, s.data(), N, r.data(), W); }
        save((d + "/raw" + std::to_string(par)).c_str(), r);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    d = tmp_path_factory.mktemp("compat_host")
    src = d / "host.cpp"
    src.write_text(PROGRAM)
    exe = d / "host"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off",
                        "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("m,w,metric", [(64, 6, 1), (128, 13, 1), (100, 10, 0), (37, 36, 1),
                                        (16, 0, 1)])
def test_host_paths_bit_exact_with_oracle(program, tmp_path, m, w, metric):
    n = 3000
    s = synth.random_walk(n, m + w)
    zq = O.znormalize_f32(synth.random_walk(m, 7))
    a, sd = O.stats_fp64(s, m)
    L = n - m + 1
    idx = np.random.default_rng(m).permutation(L).astype(np.int32)
    for name, arr in (("q", zq), ("s", s), ("a", a), ("sd", sd), ("idx", idx)):
        arr.tofile(tmp_path / name)
    r = subprocess.run([str(program), str(tmp_path), str(m), str(n), str(w), str(metric)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cd_ref = O.cdtw(zq, s, a, sd, w, idx=idx, metric=metric, fused=0)
    ld_ref = O.ldtw(zq, s, a, sd, w, metric=metric)
    raw_ref = O.ldtw_raw(zq, s, w, metric=metric)  # the raw walk of the query given
    for par in (0, 1):
        cd = np.fromfile(tmp_path / f"cdtw{par}", np.float32)
        ld = np.fromfile(tmp_path / f"ldtw{par}", np.float32)
        raw = np.fromfile(tmp_path / f"raw{par}", np.float32)
        assert np.array_equal(cd, cd_ref), par
        assert np.array_equal(ld, ld_ref), par
        assert np.array_equal(raw, raw_ref), par
