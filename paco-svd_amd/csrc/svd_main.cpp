// svd_main: host driver with the reference's file contract (main.cpp:15-203).
//
//   svd_main [args.txt] [A_gm.bin] [output_dir]
//     defaults: ../args.txt  ../input/A_gm.bin  ../output   (as main.cpp:20-24,64,169-174)
//   args.txt  : "M N"
//   A_gm.bin  : M*N FP32, row-major
//   writes    : sigma.bin (N FP32, descending), U.bin (M*M FP32), Vt.bin (N*N FP32)
//
// Same buffer sizes and zero-fill rules as the reference host (main.cpp:33-123), the two
// launches replaced by the C ABI of libpacosvd.so.  Unlike the reference's CHECK_ACL,
// every failing call aborts with a message and a non-zero exit status.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/paco_svd.h"

#define CHECK_HIP(x)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            std::cerr << "[ERROR] " #x " failed: " << hipGetErrorString(e_) << std::endl;               \
            std::exit(1);                                                                               \
        }                                                                                               \
    } while (0)
#define CHECK_PACO(x)                                                                                   \
    do {                                                                                                \
        int r_ = (x);                                                                                   \
        if (r_ != 0) {                                                                                  \
            std::cerr << "[ERROR] " #x " returned " << r_ << std::endl;                                 \
            std::exit(1);                                                                               \
        }                                                                                               \
    } while (0)

static bool read_file(const std::string &path, void *buf, size_t bytes)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.read(static_cast<char *>(buf), (std::streamsize)bytes);
    return (size_t)f.gcount() == bytes;
}

static bool write_file(const std::string &path, const void *buf, size_t bytes)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(static_cast<const char *>(buf), (std::streamsize)bytes);
    return (bool)f;
}

int main(int argc, char **argv)
{
    const std::string args_path = argc > 1 ? argv[1] : "../args.txt";
    const std::string a_path = argc > 2 ? argv[2] : "../input/A_gm.bin";
    const std::string out_dir = argc > 3 ? argv[3] : "../output";
    int M = 0, N = 0;
    {
        std::ifstream f(args_path);
        if (!(f >> M >> N)) {
            std::cerr << "[ERROR] cannot read M N from " << args_path << std::endl;
            return 1;
        }
    }
    if (M < N || N < 1) {
        std::cerr << "M must be greater than N" << std::endl;
        return 1;
    }
    const size_t szA = sizeof(float) * (size_t)M * N, szU = sizeof(float) * (size_t)M * M,
                 szVt = sizeof(float) * (size_t)N * N, szD = sizeof(float) * N,
                 szE = sizeof(float) * (size_t)(N > 1 ? N - 1 : 1), szQt = 2 * szVt,
                 szIdx = sizeof(uint32_t) * 16 * (size_t)N;
    const size_t szWs = paco_workspace_size(M, N);
    const size_t szTiling = (size_t)getTilingSize(N, 1);

    std::vector<float> hA((size_t)M * N), hU((size_t)M * M), hVt((size_t)N * N), hD(N);
    if (!read_file(a_path, hA.data(), szA)) {
        std::cerr << "[ERROR] cannot read " << szA << " bytes from " << a_path << std::endl;
        return 1;
    }
    std::vector<uint8_t> hTiling(szTiling);
    GenerateTiling(N, 1, hTiling.data());

    hipStream_t s;
    CHECK_HIP(hipStreamCreate(&s));
    float *A, *U, *Vt, *D, *E, *Tq, *Tp, *Qt, *Wt;
    uint32_t *Idx;
    void *Ws, *Tiling;
    CHECK_HIP(hipMalloc(&A, szA));
    CHECK_HIP(hipMalloc(&U, szU));
    CHECK_HIP(hipMalloc(&Vt, szVt));
    CHECK_HIP(hipMalloc(&D, szD));
    CHECK_HIP(hipMalloc(&E, szE));
    CHECK_HIP(hipMalloc(&Tq, szD));
    CHECK_HIP(hipMalloc(&Tp, szE));
    CHECK_HIP(hipMalloc(&Qt, szQt));
    CHECK_HIP(hipMalloc(&Wt, szQt));
    CHECK_HIP(hipMalloc(&Idx, szIdx));
    CHECK_HIP(hipMalloc(&Ws, szWs));
    CHECK_HIP(hipMalloc(&Tiling, szTiling));
    CHECK_HIP(hipMemset(U, 0, szU));
    CHECK_HIP(hipMemset(Vt, 0, szVt));
    CHECK_HIP(hipMemset(Qt, 0, szQt));
    CHECK_HIP(hipMemset(Wt, 0, szQt));
    CHECK_HIP(hipMemset(Idx, 0, szIdx));
    CHECK_HIP(hipMemcpy(A, hA.data(), szA, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(Tiling, hTiling.data(), szTiling, hipMemcpyHostToDevice));

    hipEvent_t e0, e1, e2;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_HIP(hipEventCreate(&e2));
    std::cout << "before bidiagonalization" << std::endl;
    CHECK_HIP(hipEventRecord(e0, s));
    CHECK_PACO(paco_upper_bidiagonalization(40, s, M, N, A, U, Vt, D, E, Tq, Tp, Ws));
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_PACO(paco_svd_DC(1, s, M, N, A, U, Vt, D, E, Qt, Wt, Idx, Ws, Tiling));
    CHECK_HIP(hipEventRecord(e2, s));
    CHECK_HIP(hipStreamSynchronize(s));
    float ms_a = 0.f, ms_b = 0.f;
    CHECK_HIP(hipEventElapsedTime(&ms_a, e0, e1));
    CHECK_HIP(hipEventElapsedTime(&ms_b, e1, e2));
    std::cout << "finish: bidiagonalization " << ms_a << " ms, divide and conquer " << ms_b << " ms" << std::endl;

    CHECK_HIP(hipMemcpy(hU.data(), U, szU, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(hVt.data(), Vt, szVt, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(hD.data(), D, szD, hipMemcpyDeviceToHost));
    bool ok = write_file(out_dir + "/sigma.bin", hD.data(), szD) && write_file(out_dir + "/U.bin", hU.data(), szU) &&
              write_file(out_dir + "/Vt.bin", hVt.data(), szVt);
    if (!ok) {
        std::cerr << "[ERROR] cannot write the outputs into " << out_dir << std::endl;
        return 1;
    }
    for (void *p : {(void *)A, (void *)U, (void *)Vt, (void *)D, (void *)E, (void *)Tq, (void *)Tp, (void *)Qt,
                    (void *)Wt, (void *)Idx, Ws, Tiling})
        CHECK_HIP(hipFree(p));
    CHECK_HIP(hipStreamDestroy(s));
    return 0;
}
