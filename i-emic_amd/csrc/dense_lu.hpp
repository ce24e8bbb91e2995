// Dense LU of the row-block preconditioner (dense_lu.hip), as used by
// rowlu.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstddef>

namespace iemic {

// Look-ahead resources of dense_getrf: a second stream with its own rocBLAS
// handle (bound to it, with its own workspace) and the two events of the
// hand-overs between the streams.
struct LuAhead {
    hipStream_t s2 = nullptr;
    rocblas_handle h2 = nullptr;
    hipEvent_t ev_panel = nullptr, ev_rest = nullptr;
};

template <typename T>
int dense_getrf(rocblas_handle h, hipStream_t s, int n, T *A, int *ipiv, int *info, void *work, int *perm,
                const LuAhead *la);
template <typename T>
int dense_inverse(rocblas_handle h, hipStream_t s, int n, T *LU, const int *perm, T *W);
size_t dense_getrf_work_bytes(int n);
void dense_getrf_size_query(rocblas_handle h, int n);

}  // namespace iemic
