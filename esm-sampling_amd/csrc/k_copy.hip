// k_copy.hip — copy one replica's columns between two replica sets (esm_copy_replica).
//
// Arrays are [row][replica] with the replica fastest; a replica's column is one element per
// row at stride R. One thread per row.
#include "esm_internal.h"

namespace esm {

template <typename E>
__global__ void k_copy_column(const E *__restrict__ src, int Rs, int js, E *__restrict__ dst, int Rd, int jd,
                              size_t nrows) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrows) return;
    dst[g * Rd + jd] = src[g * Rs + js];
}

template <typename E>
static int copy_column(hipStream_t st, const void *src, int Rs, int js, void *dst, int Rd, int jd, size_t nrows) {
    if (nrows == 0) return 0;
    const unsigned nb = (unsigned)((nrows + 255) / 256);
    hipLaunchKernelGGL((k_copy_column<E>), dim3(nb), dim3(256), 0, st, (const E *)src, Rs, js, (E *)dst, Rd, jd, nrows);
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_copy_replica(const esm_handle *hs, int js, esm_handle *hd, int jd, bool weather) {
    const size_t NT = (size_t)hs->N * hs->T;
    const bool f64 = hs->prec == ESM_FP64;
    int rc;
    auto col = [&](const void *s, void *d, size_t rows) {
        return f64 ? copy_column<double>(hd->stream, s, hs->R, js, d, hd->R, jd, rows)
                   : copy_column<float>(hd->stream, s, hs->R, js, d, hd->R, jd, rows);
    };
    if ((rc = col(hs->d_v, hd->d_v, ESM_NV * NT))) return rc;
    if ((rc = col(hs->d_x, hd->d_x, ESM_NREN * NT))) return rc;
    if (weather) {
        if ((rc = col(hs->d_z, hd->d_z, ESM_NREN * NT))) return rc;
        if ((rc = copy_column<int>(hd->stream, hs->d_m, hs->R, js, hd->d_m, hd->R, jd, (size_t)hs->T))) return rc;
    }
    return 0;
}

}  // namespace esm
