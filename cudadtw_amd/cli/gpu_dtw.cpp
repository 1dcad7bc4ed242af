// gpu_dtw -- command-line equivalent of the reference's GPU-DTW driver
// (GPU-dtw.cu): best z-normalised cDTW match of a query in a stream.
//   gpu_dtw query.bin subject.bin M N P
// Raw FP32 files; band W = M * (P * 0.01) truncated (GPU-dtw.cu:27).  Prints
// the reference's info, "Location:" and "Distance:" lines.
#include "common.hpp"

int main(int argc, char* argv[]) {
    if (argc != 6) {
        std::cout << "call" << argv[0] << " query.bin subject.bin M N P" << std::endl;
        return 1;
    }
    const int M = std::atoi(argv[3]);
    const int N = std::atoi(argv[4]);
    const int W = static_cast<int>(M * (std::atoi(argv[5]) * 0.01));
    std::cout << "\n= info =====================================" << std::endl;
    std::cout << "|Query| = " << M << "\t|Subject| = " << N << "\twindow = " << W << std::endl;
    try {
        std::cout << "\n= loading data =============================" << std::endl;
        Timer tl;
        std::vector<float> zquery = read_bin<float>(argv[1], M);
        std::vector<float> subject = read_bin<float>(argv[2], N);
        std::vector<float> zlower(M), zupper(M);
        znormalize(zquery.data(), M);
        envelope(zquery.data(), W, zlower.data(), zupper.data(), M);
        std::cout << "Miliseconds to load data: " << tl.ms() << std::endl;

        Timer ts;
        float *Subject = nullptr, *AvgS = nullptr, *StdS = nullptr;
        HIP_OK(hipMalloc(&Subject, sizeof(float) * N));
        HIP_OK(hipMalloc(&AvgS, sizeof(float) * (N - M + 1)));
        HIP_OK(hipMalloc(&StdS, sizeof(float) * (N - M + 1)));
        HIP_OK(hipMemcpy(Subject, subject.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        mdtw_compat::set_query(zquery.data(), zlower.data(), zupper.data(), M, W);
        avg_std<double>(Subject, AvgS, StdS, M, N);
        std::cout << "\n= pruning ==================================" << std::endl;
        prune_cdtw<true>(Subject, AvgS, StdS, M, N, W);
        std::cout << "Miliseconds to find best match: " << ts.ms() << std::endl;
        HIP_OK(hipFree(Subject));
        HIP_OK(hipFree(AvgS));
        HIP_OK(hipFree(StdS));
    } catch (const mdtw_compat::error& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
