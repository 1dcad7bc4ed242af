// gpu_ed -- command-line equivalent of the reference's GPU-ED driver
// (GPU-ed.cu): best z-normalised Euclidean match.
//   gpu_ed query.bin subject.bin M N
// Raw FP64 files as the reference reads them (GPU-ed.cu:51-56); the query is
// z-normalised in FP64 (znormalize<double>, GPU-ed.cu:62), then query and
// stream are searched in FP32.  Prints the reference's result lines.
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
        std::vector<double> zq64 = read_bin<double>(argv[1], M);
        std::vector<double> s64 = read_bin<double>(argv[2], N);
        znormalize(zq64.data(), M);
        std::vector<float> zquery(zq64.begin(), zq64.end());
        std::vector<float> subject(s64.begin(), s64.end());
        std::cout << "Miliseconds to load data: " << tl.ms() << std::endl;

        Timer ts;
        float* Subject = nullptr;
        HIP_OK(hipMalloc(&Subject, sizeof(float) * N));
        HIP_OK(hipMemcpy(Subject, subject.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        mdtw_compat::set_query(zquery.data(), nullptr, nullptr, M, 0);
        calculate_ed(Subject, static_cast<const float*>(nullptr), M, N);
        std::cout << "Miliseconds to find best match: " << ts.ms() << std::endl;
        HIP_OK(hipFree(Subject));
    } catch (const mdtw_compat::error& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
