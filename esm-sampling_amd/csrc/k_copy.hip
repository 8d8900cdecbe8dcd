// k_copy.hip — state utilities: copy one replica's columns between two replica sets
// (esm_copy_replica), and the heuristic start state of every replica (esm_heuristic_state).
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

// The heuristic start of one (node, step, replica): zero angle, the gas unit covering the node's
// net load (load minus the renewables' output k·x, raw capacities and weather), clamped to
// [0, 1], both storage units half full. Computed in FP64, stored in the state precision.
template <typename S>
__global__ void __launch_bounds__(256) k_heuristic(DevModel<S> m, const double *__restrict__ k,
                                                   const S *__restrict__ x, S *__restrict__ v) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t TR = (size_t)m.T * m.R;
    if (g >= (size_t)m.N * TR) return;
    const int n = (int)(g / TR);
    const size_t o = g - (size_t)n * TR;
    const int t = (int)(o / m.R);
    double ren = 0.0;
    for (int c = 0; c < ESM_NREN; c++) ren += k[c * m.N + n] * (double)x[plane(c * m.N + n, m.T, m.R) + o];
    const double net = (double)m.load[(size_t)n * m.T + t] - ren;
    const double y = net / fmax(k[ESM_KGAS * m.N + n], 1e-9);
    v[plane(0, m.T, m.R) + g] = S(0);
    v[plane(m.N, m.T, m.R) + g] = (S)fmin(fmax(y, 0.0), 1.0);
    v[plane(2 * m.N, m.T, m.R) + g] = S(0.5);
    v[plane(3 * m.N, m.T, m.R) + g] = S(0.5);
}

int launch_heuristic(esm_handle *h) {
    const size_t n = (size_t)h->N * h->T * h->R;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (h->prec == ESM_FP64)
        hipLaunchKernelGGL((k_heuristic<double>), dim3(nb), dim3(256), 0, h->stream, dev_model<double>(h), h->d_k,
                           (const double *)h->d_x, (double *)h->d_v);
    else
        hipLaunchKernelGGL((k_heuristic<float>), dim3(nb), dim3(256), 0, h->stream, dev_model<float>(h), h->d_k,
                           (const float *)h->d_x, (float *)h->d_v);
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
