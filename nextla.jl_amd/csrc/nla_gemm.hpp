// Internal GEMM entry shared by the QR and TRSM/TRMM code.
#pragma once
#include "nla_common.hpp"

namespace nla {
// C[M x N] = alpha * op(A) * op(B) + beta * C, column-major, op = transpose when ta/tb is 'T'/'C'.
template <typename T>
hipError_t gemm(char ta, char tb, i64 M, i64 N, i64 K, T alpha, const T *A, i64 lda, const T *B, i64 ldb, T beta,
                T *C, i64 ldc, hipStream_t s);
}  // namespace nla
