// Internal launchers of the tile-QR kernels, shared with the tile driver.
#pragma once
#include "nla_common.hpp"

namespace nla {
template <typename T>
hipError_t launch_ge_apply(bool trc, i64 m, i64 ncols, i64 k, i64 ib, const T *V, i64 ldv, const T *Tm, i64 ldt,
                           bool notrans, bool forward, T *C, i64 ldc, hipStream_t s);
template <typename T>
hipError_t launch_ts_apply(bool trc, i64 m2, i64 ncols, i64 k, i64 ib, const T *V, i64 ldv, const T *Tm, i64 ldt,
                           bool notrans, bool forward, T *A1, i64 lda1, T *A2, i64 lda2, hipStream_t s);
template <typename T>
hipError_t geqrt_async(i64 m, i64 n, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, T *tau, hipStream_t s);
template <typename T>
hipError_t tsqrt_async(i64 m, i64 n, i64 ib, i64 k, T *A1, i64 lda1, T *A2, i64 lda2, T *Tm, i64 ldt, T *tau,
                       hipStream_t s);
}  // namespace nla
