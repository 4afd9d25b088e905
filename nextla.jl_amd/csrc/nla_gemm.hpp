// Internal GEMM entry shared by the QR and TRSM/TRMM code.
#pragma once
#include "nla_common.hpp"

namespace nla {
// C[M x N] = alpha * op(A) * op(B) + beta * C, column-major, op = transpose when ta/tb is 'T'/'C'.
template <typename T>
hipError_t gemm(char ta, char tb, i64 M, i64 N, i64 K, T alpha, const T *A, i64 lda, const T *B, i64 ldb, T beta,
                T *C, i64 ldc, hipStream_t s);
// Operand precision of the Float32 GEMMs the calling host thread issues ('S' Float32, 'H' FP16,
// 'B' BF16 operands with Float32 accumulation); GemmLowp sets it for a scope.
char gemm_lowp_get();
void gemm_lowp_set(char p);
struct GemmLowp {
    char saved;
    explicit GemmLowp(char p) : saved(gemm_lowp_get()) { gemm_lowp_set(p); }
    ~GemmLowp() { gemm_lowp_set(saved); }
};
}  // namespace nla
