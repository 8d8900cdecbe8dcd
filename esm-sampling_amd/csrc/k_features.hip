// k_features.hip — the weather features of every replica (`replica_features`,
// src/replicas.jl:270-273, over `weather_features`, :250-267), recorded by the replica driver at
// every replica step.
//
// Per replica j, from the complete-level weather A = clamp(x, 0, 1) and the regimes m:
//   load_GW   mean_t D_t / 1e3, D_t = Σ_n max(load[n, t], 0)         (the load is not weather here)
//   cf_c      mean over (n, t) of A[c, n, t]                          (c = solar, onwind, offwind)
//   res_t     (D_t − Σ_c Σ_n kref[c, n] A[c, n, t]) / 1e3               (residual load, GW)
//   resload_mean_GW, resload_peak_GW   mean_t res_t, max_t res_t
//   resload_7step_GW   max_t mean(res_t … res_{t+6}), cyclic
//   regime_k  the fraction of steps in regime k; regime_switches  #{t : m_{t+1} ≠ m_t}
//
// MI355X layout: k_feat_hour, one thread per (t, j) with j fastest (every load of x coalesced
// over the replicas), sums the nodes of the three carriers and writes res and the three node sums
// as FP64 planes [4][T][R]; k_feat_final, one workgroup of 256 threads per replica, reduces the
// planes and the regimes over t (threads stride t; fixed-order tree in LDS, so the result does
// not depend on scheduling) and writes the replica's row.
#include "esm_internal.h"

namespace esm {

constexpr int FEAT_BLOCK = 256;
constexpr int FEAT_KMAX = 8;

template <typename S>
__global__ void __launch_bounds__(256) k_feat_hour(const S *__restrict__ x, const double *__restrict__ kref,
                                                   const double *__restrict__ dload, int N, int T, int R,
                                                   double *__restrict__ planes) {
    const size_t TR = (size_t)T * R;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= TR) return;
    const int t = (int)(g / R);
    double cf[ESM_NREN], rk = 0.0;
#pragma unroll
    for (int c = 0; c < ESM_NREN; c++) {
        double s = 0.0;
        for (int n = 0; n < N; n++) {
            const double a = (double)clampv(x[plane(c * N + n, T, R) + g], S(0), S(1));
            s += a;
            rk += kref[c * N + n] * a;
        }
        cf[c] = s;
    }
    planes[g] = (dload[t] - rk) / 1e3;
#pragma unroll
    for (int c = 0; c < ESM_NREN; c++) planes[(size_t)(c + 1) * TR + g] = cf[c];
}

// fixed-order tree reduction of FEAT_BLOCK values in LDS (sum or max); every thread gets it
__device__ double feat_reduce(double *sh, double v, bool mx) {
    const int i = threadIdx.x;
    __syncthreads();
    sh[i] = v;
    __syncthreads();
    for (int o = FEAT_BLOCK / 2; o > 0; o >>= 1) {
        if (i < o) sh[i] = mx ? fmax(sh[i], sh[i + o]) : sh[i] + sh[i + o];
        __syncthreads();
    }
    return sh[0];
}

__global__ void __launch_bounds__(FEAT_BLOCK) k_feat_final(const double *__restrict__ planes, const int *__restrict__ m,
                                                           int N, int T, int R, int K, double load_gw,
                                                           double *__restrict__ out) {
    __shared__ double sh[FEAT_BLOCK];
    const int j = blockIdx.x, i = threadIdx.x;
    const size_t TR = (size_t)T * R;
    const double *res = planes;
    double sres = 0.0, pk = -INFINITY, w7 = -INFINITY, scf[ESM_NREN] = {0.0, 0.0, 0.0}, sw = 0.0;
    double cnt[FEAT_KMAX];
#pragma unroll
    for (int k = 0; k < FEAT_KMAX; k++) cnt[k] = 0.0;
    for (int t = i; t < T; t += FEAT_BLOCK) {
        const size_t o = (size_t)t * R + j;
        const double r = res[o];
        sres += r;
        pk = fmax(pk, r);
        double s7 = 0.0;
        for (int q = 0; q < 7; q++) s7 += res[(size_t)((t + q) % T) * R + j];
        w7 = fmax(w7, s7 / 7.0);
#pragma unroll
        for (int c = 0; c < ESM_NREN; c++) scf[c] += planes[(size_t)(c + 1) * TR + o];
        const int mt = m[o];
#pragma unroll
        for (int k = 0; k < FEAT_KMAX; k++) cnt[k] += mt == k ? 1.0 : 0.0;
        if (t + 1 < T && m[o + R] != mt) sw += 1.0;
    }
    const int nf = ESM_NREN + 5 + K;
    double *row = out + (size_t)j * nf;
    const double v0 = feat_reduce(sh, sres, false);
    const double v1 = feat_reduce(sh, pk, true);
    const double v2 = feat_reduce(sh, w7, true);
    double vc[ESM_NREN];
#pragma unroll
    for (int c = 0; c < ESM_NREN; c++) vc[c] = feat_reduce(sh, scf[c], false);
    const double vs = feat_reduce(sh, sw, false);
    if (i == 0) {
        row[0] = load_gw;
        for (int c = 0; c < ESM_NREN; c++) row[1 + c] = vc[c] / ((double)N * T);
        row[1 + ESM_NREN] = v0 / T;
        row[2 + ESM_NREN] = v1;
        row[3 + ESM_NREN] = v2;
    }
#pragma unroll
    for (int k = 0; k < FEAT_KMAX; k++) {
        if (k >= K) break;   // K is the same in every thread: the barriers stay uniform
        const double vk = feat_reduce(sh, cnt[k], false);
        if (i == 0) row[4 + ESM_NREN + k] = vk / T;
    }
    if (i == 0) row[4 + ESM_NREN + K] = vs;
}

template <typename S>
int launch_features(esm_handle *h, const double *kref_host, double *out_host) {
    const int N = h->N, T = h->T, R = h->R, K = h->m.K;
    if (K < 1 || K > FEAT_KMAX) {
        set_error("esm_replica_features: the regime count must be 1 … 8");
        return ESM_EINVAL;
    }
    const int nf = ESM_NREN + 5 + K;
    // the replica-independent parts in FP64 on the host: D_t and the annual mean load
    std::vector<double> buf((size_t)ESM_NREN * N + T);
    for (int q = 0; q < ESM_NREN * N; q++) buf[q] = kref_host[q];
    double dsum = 0.0;
    for (int t = 0; t < T; t++) {
        double d = 0.0;
        for (int n = 0; n < N; n++) d += std::max(h->load[(size_t)n * T + t], 0.0);
        buf[(size_t)ESM_NREN * N + t] = d;
        dsum += d;
    }
    const size_t need = buf.size() + (size_t)R * nf;
    if (h->feat_len < need) {
        if (h->d_feat) ESM_HIP_CHECK(hipFree(h->d_feat));
        h->d_feat = nullptr;
        ESM_HIP_CHECK(hipMalloc((void **)&h->d_feat, need * sizeof(double)));
        h->feat_len = need;
    }
    double *d_kref = h->d_feat, *d_load = h->d_feat + (size_t)ESM_NREN * N, *d_out = h->d_feat + buf.size();
    ESM_HIP_CHECK(hipMemcpyAsync(h->d_feat, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    const size_t TR = (size_t)T * R;
    hipLaunchKernelGGL((k_feat_hour<S>), dim3((unsigned)((TR + 255) / 256)), dim3(256), 0, h->stream, (const S *)h->d_x,
                       (const double *)d_kref, (const double *)d_load, N, T, R, h->d_hparts);
    ESM_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_feat_final, dim3((unsigned)R), dim3(FEAT_BLOCK), 0, h->stream, (const double *)h->d_hparts,
                       (const int *)h->d_m, N, T, R, K, dsum / T / 1e3, d_out);
    ESM_HIP_CHECK(hipGetLastError());
    ESM_HIP_CHECK(hipMemcpyAsync(out_host, d_out, (size_t)R * nf * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    return ESM_OK;
}

template int launch_features<float>(esm_handle *, const double *, double *);
template int launch_features<double>(esm_handle *, const double *, double *);

}  // namespace esm
