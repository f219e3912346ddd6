// Exposes the reference's own CPU GEMM oracle, cpu_gemm<float> and
// cpu_gemm<std::complex<float>> (reference tests/blas_funcs/test_gemm.hpp:83-201),
// compiled from the reference's headers where they lie, with a C ABI.
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ and used to pin the
// restatement in oracle.c; the library never links it.
#include "test_gemm.hpp"

extern "C" void ref_cpu_sgemm(int transa, int transb, int m, int n, int k, float alpha, const float* A, int lda,
                              const float* B, int ldb, float beta, float* C, int ldc)
{
    cpu_gemm<float>((iclblasOperation_t)transa, (iclblasOperation_t)transb, m, n, k, alpha, A, lda, B, ldb, beta, C,
                    ldc);
}

extern "C" void ref_cpu_cgemm(int transa, int transb, int m, int n, int k, const float* alpha, const float* A, int lda,
                              const float* B, int ldb, const float* beta, float* C, int ldc)
{
    using cf = std::complex<float>;
    cpu_gemm<cf>((iclblasOperation_t)transa, (iclblasOperation_t)transb, m, n, k, cf(alpha[0], alpha[1]),
                 reinterpret_cast<const cf*>(A), lda, reinterpret_cast<const cf*>(B), ldb, cf(beta[0], beta[1]),
                 reinterpret_cast<cf*>(C), ldc);
}
