// gpu_ed -- command-line equivalent of the reference's GPU-ED driver
// (GPU-ed.cu): best z-normalised Euclidean match.
//   gpu_ed query.bin subject.bin M N
// Raw FP64 files as the reference reads them (GPU-ed.cu:51-56), searched in
// FP64 as the reference does: znormalize<double> (GPU-ed.cu:62), the query
// set as FP64 (:72), avg_std<double> on the FP64 stream (:76), calculate_ed
// (:82).  Prints the reference's result lines.
#include "common.hpp"

int main(int argc, char* argv[]) {
    if (argc != 5) {
        std::cout << "call" << argv[0] << " query.bin subject.bin M N" << std::endl;
        return 1;
    }
    const int M = std::atoi(argv[3]);
    const int N = std::atoi(argv[4]);
    std::cout << "\n= info =====================================" << std::endl;
    std::cout << "|Query| = " << M << "\t|Subject| = " << N << "\t" << std::endl;
    try {
        std::cout << "\n= loading data =============================" << std::endl;
        Timer tl;
        std::vector<double> zquery = read_bin<double>(argv[1], M);
        std::vector<double> subject = read_bin<double>(argv[2], N);
        znormalize(zquery.data(), M);
        std::cout << "Miliseconds to load data: " << tl.ms() << std::endl;

        Timer ts;
        double *Subject = nullptr, *AvgS = nullptr, *StdS = nullptr;
        HIP_OK(hipMalloc(&Subject, sizeof(double) * N));
        HIP_OK(hipMalloc(&AvgS, sizeof(double) * (N - M + 1)));
        HIP_OK(hipMalloc(&StdS, sizeof(double) * (N - M + 1)));
        HIP_OK(hipMemcpy(Subject, subject.data(), sizeof(double) * N, hipMemcpyHostToDevice));
        mdtw_compat::set_query(zquery.data(), M);
        avg_std<double>(Subject, AvgS, StdS, M, N);
        HIP_OK(hipFree(AvgS));
        calculate_ed(Subject, StdS, M, N);
        std::cout << "Miliseconds to find best match: " << ts.ms() << std::endl;
        HIP_OK(hipFree(Subject));
        HIP_OK(hipFree(StdS));
    } catch (const mdtw_compat::error& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
