// gpu_ldtw -- command-line equivalent of the reference's GPU-LDTW driver
// (GPU-ldtw.cu): greedy LDTW of the z-normalised query against every window.
//   gpu_ldtw query.bin subject.bin M N P
// Raw FP32 files; W = M * (P * 0.01) truncated (GPU-ldtw.cu:19).  Prints the
// reference's determine_min line for the GPU result, then -- as GPU-ldtw.cu
// does (:108-141) -- the same search by the host paths on all cores and on
// one core over the GPU's window statistics, as a consistency check.
#include "common.hpp"

int main(int argc, char* argv[]) {
    if (argc != 6) {
        std::cout << "call" << argv[0] << " query.bin subject.bin M N P" << std::endl;
        return 1;
    }
    const int M = std::atoi(argv[3]);
    const int N = std::atoi(argv[4]);
    const int W = static_cast<int>(M * (std::atoi(argv[5]) * 0.01));
    std::cout << " M " << M << " N " << N << " W " << W << std::endl;
    try {
        std::cout << "\n= loading data =============================" << std::endl;
        Timer tl;
        std::vector<float> q = read_bin<float>(argv[1], M);
        std::vector<float> s = read_bin<float>(argv[2], N);
        znormalize(q.data(), M);
        std::cout << "Miliseconds to load data: " << tl.ms() << std::endl;

        const int L = N - M + 1;
        float *S = nullptr, *AvgS = nullptr, *StdS = nullptr, *R = nullptr;
        HIP_OK(hipMalloc(&S, sizeof(float) * N));
        HIP_OK(hipMalloc(&AvgS, sizeof(float) * L));
        HIP_OK(hipMalloc(&StdS, sizeof(float) * L));
        HIP_OK(hipMalloc(&R, sizeof(float) * L));
        HIP_OK(hipMemcpy(S, s.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        mdtw_compat::set_query(q.data(), nullptr, nullptr, M, W);
        avg_std<double>(S, AvgS, StdS, M, N);
        std::cout << "\n= aligning==================================" << std::endl;
        Timer ta;
        mdtw_compat::ztsh2_ldtw<true>(S, M, N, AvgS, StdS, R, W);
        mdtw_compat::check(mdtw_synchronize(mdtw_compat::engine()), "synchronize");
        std::cout << "Miliseconds to find best alignment on the GPU: " << ta.ms() << std::endl;
        std::vector<float> r(static_cast<size_t>(L));
        HIP_OK(hipMemcpy(r.data(), R, sizeof(float) * L, hipMemcpyDeviceToHost));
        determine_min<true>(r.data(), L);

        std::vector<float> avgs(static_cast<size_t>(L)), stds(static_cast<size_t>(L));
        HIP_OK(hipMemcpy(avgs.data(), AvgS, sizeof(float) * L, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(stds.data(), StdS, sizeof(float) * L, hipMemcpyDeviceToHost));
        std::fill(r.begin(), r.end(), 0.0f);
        Timer to;
        omp_z_stream_ldtw<true>(q.data(), M, s.data(), N, avgs.data(), stds.data(), r.data(), W);
        std::cout << "Miliseconds to find best alignment on all cores: " << to.ms() << std::endl;
        determine_min<true>(r.data(), L);
        std::fill(r.begin(), r.end(), 0.0f);
        Timer t1;
        z_stream_ldtw<true>(q.data(), M, s.data(), N, avgs.data(), stds.data(), r.data(), W);
        std::cout << "Miliseconds to find best alignment on a single core: " << t1.ms()
                  << std::endl;
        determine_min<true>(r.data(), L);
        HIP_OK(hipFree(S));
        HIP_OK(hipFree(AvgS));
        HIP_OK(hipFree(StdS));
        HIP_OK(hipFree(R));
    } catch (const mdtw_compat::error& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
