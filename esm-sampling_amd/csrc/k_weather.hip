// k_weather.hip — the regime-switching VAR(1) weather process of every replica.
//
// Model (include/esm.h): per capacity-factor series i and replica j, regimes m_τ ∈ {0..K−1}, a
// Markov chain (π0, Ptr); anomalies z_1 = σ0 e_1, z_τ = φ z_{τ−1} + μ_{m_τ}(τ) + σ e_τ with
// μ_k(τ) = M_k[i] · (1, cos 2πτ/T, sin 2πτ/T) (τ 1-based, src/weather.jl:43 and :461-481);
// raw weather x = sigmoid(xbar(hour of day) + z).
//
// One weather step (src/replicas.jl:179-197) composes the pCN rotation of the innovations
// e ↦ cos θ e + sin θ ξ at fixed regimes with a regime window move that re-builds z from the
// (rotated) innovations at the new regimes. With Φ = φ I and Σ = σ² I both fuse into one affine
// recurrence per series,
//     z''_τ = φ z''_{τ−1} + b_τ,  b_τ = μ_{m'_τ}(τ) + σ (c e_τ + s ξ_τ),  b_1 = σ0 (c e_1 + s ξ_1),
// with e the innovations of the current z at the current regimes m and m' the new regimes. The
// recurrence runs as a chunked scan: (1) every chunk from zero, (2) the chunk carries in order,
// (3) the carries applied, with the new weather and the compensated re-dispatch
// (compensate_weather!, src/replicas.jl:207-236) fused into the same pass.
//
// Random numbers: Philox4x32-10, key = the call's key; normals ξ(i, τ, j) by Box–Muller from the
// counter (i, τ/2, j, call counter) (cos branch for even τ, sin for odd); regime and window
// uniforms from counters tagged per purpose.
#include "esm_internal.h"

namespace esm {

constexpr uint32_t TAG_REG = 0x52454731u;  // prior regime chain
constexpr uint32_t TAG_WIN = 0x57494E44u;  // window move
constexpr uint32_t TAG_FFBS = 0x46464253u; // centred FFBS of the regimes

__device__ __forceinline__ double normal_of(uint32_t k0, uint32_t k1, int i, int tau, int j, uint32_t ctr) {
    uint32_t w[4];
    philox4x32(k0, k1, (uint32_t)i, (uint32_t)(tau >> 1), (uint32_t)j, ctr, w);
    const double u1 = u01(w[0], w[1]), u2 = u01(w[2], w[3]);
    const double rad = sqrt(-2.0 * log(1.0 - u1));
    const double ang = 6.283185307179586 * u2;
    return (tau & 1) ? rad * sin(ang) : rad * cos(ang);
}

__device__ __forceinline__ int categorical(const double *p, int K, double u) {
    double tot = 0;
    for (int k = 0; k < K; k++) tot += p[k];
    const double t = u * tot;
    double c = 0;
    for (int k = 0; k < K; k++) {
        c += p[k];
        if (t <= c) return k;
    }
    return K - 1;
}

// Prior regime chain of every replica (one thread per replica; sequential in τ).
__global__ void k_regime_init(int *__restrict__ m, const double *__restrict__ Ptr, const double *__restrict__ pi0, int K,
                              int T, int R, uint32_t k0, uint32_t k1, uint32_t ctr) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= R) return;
    uint32_t w[4];
    int prev = 0;
    for (int t = 0; t < T; t++) {
        philox4x32(k0, k1, (uint32_t)j, (uint32_t)t, TAG_REG, ctr, w);
        const double u = u01(w[0], w[1]);
        prev = categorical(t == 0 ? pi0 : Ptr + prev * K, K, u);
        m[(size_t)t * R + j] = prev;
    }
}

// Per replica: the move flags and, with a window, the exact Gibbs redraw of m'[a..b]
// (window_bounds :167, regime_window_messages :131-141, regime_window :156-164).
// mnew holds a copy of m on entry.
__global__ void k_window(const int *__restrict__ m, int *__restrict__ mnew, const double *__restrict__ wpar,
                         int *__restrict__ moved, const double *__restrict__ Ptr, const double *__restrict__ pi0, int K,
                         int T, int R, int B, uint32_t k0, uint32_t k1, uint32_t ctr) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= R) return;
    const double s = wpar[R + j], wp = wpar[2 * R + j];
    int mv = s != 0.0 ? 1 : 0;
    uint32_t w[4];
    if (wp > 0 && B > 0) {
        philox4x32(k0, k1, (uint32_t)j, 0u, TAG_WIN, ctr, w);
        if (u01(w[0], w[1]) < wp) {
            mv = 2;
            const double us = u01(w[2], w[3]);
            int a = 1 + (int)floor(us * T);          // 1-based window [a, b]
            if (a > T) a = T;
            int b = a + B - 1;
            if (b > T) b = T;
            const int n = b - a + 1;
            double beta[ESM_MAXWIN * ESM_MAXK];
            for (int q = 0; q < n; q++)
                for (int i = 0; i < K; i++) beta[q * K + i] = 1.0;
            if (b < T) {
                const int mb1 = m[(size_t)b * R + j];   // m[b+1] (1-based) = index b
                for (int i = 0; i < K; i++) beta[(n - 1) * K + i] = Ptr[i * K + mb1];
                for (int q = n - 2; q >= 0; q--)
                    for (int i = 0; i < K; i++) {
                        double acc = 0;
                        for (int l = 0; l < K; l++) acc += Ptr[i * K + l] * beta[(q + 1) * K + l];
                        beta[q * K + i] = acc;
                    }
            }
            int prev = a >= 2 ? m[(size_t)(a - 2) * R + j] : 0;
            for (int q = 0; q < n; q++) {
                const int tau = a + q;                     // 1-based
                if ((q & 1) == 0) philox4x32(k0, k1, (uint32_t)j, (uint32_t)(1 + q / 2), TAG_WIN, ctr, w);
                const double u = (q & 1) ? u01(w[2], w[3]) : u01(w[0], w[1]);
                double p[ESM_MAXK];
                for (int i = 0; i < K; i++) p[i] = (tau == 1 ? pi0[i] : Ptr[prev * K + i]) * beta[q * K + i];
                prev = categorical(p, K, u);
                mnew[(size_t)(tau - 1) * R + j] = prev;
            }
        }
    }
    moved[j] = mv;
}

template <typename S>
struct WxArgs {
    int N, T, R, K, nchunk, chunk;
    S phi, sigma, sigma0;
    const S *mu;      // [K][3N][3]
    const S *xbar;    // [3N][24]
    const int *hod;   // [T]
    const double *harm;  // [T][2] cos, sin of 2π(τ+1)/T
    const double *wpar;  // [3][R]: c, s, window prob
    const int *moved;    // [R]
    const int *m, *mnew; // [T][R]
    uint32_t k0, k1, ctr;
};

// (1) chunk-local recurrence from zero: Y[i][τ][j] and carry[c][i][j] (thread per (chunk, node, replica)).
template <typename S>
__global__ void __launch_bounds__(256) k_wx_local(WxArgs<S> a, const S *__restrict__ z, S *__restrict__ Y,
                                                  S *__restrict__ carry) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t NR = (size_t)a.N * a.R;
    if (g >= (size_t)a.nchunk * NR) return;
    const int j = (int)(g % a.R);
    const int n = (int)((g / a.R) % a.N);
    const int c = (int)(g / NR);
    if (a.moved[j] == 0) return;
    const double cj = a.wpar[j], sj = a.wpar[a.R + j];
    const int t0 = c * a.chunk, t1 = min(a.T, t0 + a.chunk);
    const int nser = ESM_NREN * a.N;
    for (int q = 0; q < ESM_NREN; q++) {
        const int i = q * a.N + n;
        const S *zi = z + (size_t)i * a.T * a.R + j;
        S *Yi = Y + (size_t)i * a.T * a.R + j;
        S y = 0;
        S zp = t0 > 0 ? zi[(size_t)(t0 - 1) * a.R] : S(0);
        for (int t = t0; t < t1; t++) {
            const S zt = zi[(size_t)t * a.R];
            const int mo = a.m[(size_t)t * a.R + j], mn = a.mnew[(size_t)t * a.R + j];
            const S h1 = (S)a.harm[2 * t], h2 = (S)a.harm[2 * t + 1];
            const S *Mo = a.mu + ((size_t)mo * nser + i) * 3;
            const S *Mn = a.mu + ((size_t)mn * nser + i) * 3;
            const S muo = Mo[0] + h1 * Mo[1] + h2 * Mo[2];
            const S mun = Mn[0] + h1 * Mn[1] + h2 * Mn[2];
            const double xi = sj != 0.0 ? normal_of(a.k0, a.k1, i, t, j, a.ctr) : 0.0;
            S b;
            if (t == 0) {
                const S e = zt / a.sigma0;
                b = a.sigma0 * (S)(cj * (double)e + sj * xi);
            } else {
                const S e = (zt - a.phi * zp - muo) / a.sigma;
                b = mun + a.sigma * (S)(cj * (double)e + sj * xi);
            }
            y = a.phi * y + b;
            Yi[(size_t)t * a.R] = y;
            zp = zt;
        }
        carry[((size_t)c * nser + i) * a.R + j] = y;
    }
}

// (2) chunk entry values: zin[c][i][j] = z at the step before chunk c (thread per (series, replica)).
template <typename S>
__global__ void k_wx_carry(WxArgs<S> a, const S *__restrict__ carry, S *__restrict__ zin) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int nser = ESM_NREN * a.N;
    if (g >= nser * a.R) return;
    const int j = g % a.R, i = g / a.R;
    if (a.moved[j] == 0) return;
    S prev = 0;
    for (int c = 0; c < a.nchunk; c++) {
        zin[((size_t)c * nser + i) * a.R + j] = prev;
        const int len = min(a.T, (c + 1) * a.chunk) - c * a.chunk;
        S pw = 1;
        for (int q = 0; q < len; q++) pw *= a.phi;
        prev = carry[((size_t)c * nser + i) * a.R + j] + pw * prev;
    }
}

// (3) apply: z, x and (optionally) the compensated gas re-dispatch with the residual cache.
template <typename S>
__global__ void __launch_bounds__(256) k_wx_apply(WxArgs<S> a, DevModel<S> md, const S *__restrict__ Y,
                                                  const S *__restrict__ zin, S *__restrict__ z, S *__restrict__ x,
                                                  S *__restrict__ v, S *__restrict__ r, int compensate) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t NR = (size_t)a.N * a.R;
    if (g >= (size_t)a.nchunk * NR) return;
    const int j = (int)(g % a.R);
    const int n = (int)((g / a.R) % a.N);
    const int c = (int)(g / NR);
    if (a.moved[j] == 0) return;
    const int t0 = c * a.chunk, t1 = min(a.T, t0 + a.chunk);
    const int nser = ESM_NREN * a.N;
    const size_t TR = (size_t)a.T * a.R;
    S z0[ESM_NREN], pw = a.phi;
    for (int q = 0; q < ESM_NREN; q++) z0[q] = zin[((size_t)c * nser + q * a.N + n) * a.R + j];
    S *yg = v + plane(md.N, md.T, md.R);
    const S kg = md.kc[ESM_KGAS * md.N + n];
    for (int t = t0; t < t1; t++, pw *= a.phi) {
        const size_t o = (size_t)t * a.R + j;
        S dr = 0;
        const int hd = a.hod[t];
        for (int q = 0; q < ESM_NREN; q++) {
            const int i = q * a.N + n;
            const S zz = Y[(size_t)i * TR + o] + pw * z0[q];
            z[(size_t)i * TR + o] = zz;
            const S xn = S(1) / (S(1) + exp(-(a.xbar[i * 24 + hd] + zz)));
            const S xo = x[(size_t)i * TR + o];
            x[(size_t)i * TR + o] = xn;
            dr -= md.kc[q * md.N + n] * (clampv(xn, S(0), S(1)) - clampv(xo, S(0), S(1)));
        }
        const size_t pn = (size_t)n * TR + o;
        S rn = r[pn] + dr;
        if (compensate && fabs(dr) >= S(1e-6)) {
            // compensate_weather!: the gas unit takes up Δr, up to its headroom or its output
            const S y = yg[pn];
            const S kap = kg;
            const S p = kap * clampv(y, S(0), S(1));
            const S wq = dr > 0 ? fmax(kap - p, S(0)) : p;
            if (wq > S(1e-9) && fabs(kap) >= S(1e-9)) {
                const S aa = dr > 0 ? fmin(dr, wq) / wq : fmax(dr, -wq) / wq;
                const S dy = aa * wq / kap;
                if (dy != 0) {
                    const S y1 = y + dy;
                    yg[pn] = y1;
                    rn -= kap * (clampv(y1, S(0), S(1)) - clampv(y, S(0), S(1)));
                }
            }
        }
        r[pn] = rn;
    }
}

// ---------------------------------------------------------------------------------------------
// Centred FFBS of the regimes given the anomalies (replica_ffbs = true): fine_regime_step!
// (src/weather_fine.jl:290-300) at the single complete level redraws the regime path from
// p(m | z) by forward filtering, backward sampling (ffbs, src/weather.jl:160-174); z and the
// weather stay as they are. (1) the emission log-likelihoods (emission_loglik, :117-133) of every
// (step, replica) under every regime, in FP64 from the stored anomalies: step 1 under the
// stationary start N(0, σ0² I), later steps the VAR(1) residual under N(0, σ² I); (2) per replica
// the forward filter over the steps and the backward draws (sample_log, :176-185) with the
// uniforms Philox (replica, τ, TAG_FFBS, call counter) words 0–1.
// ---------------------------------------------------------------------------------------------
struct FfbsConst {
    double lP[ESM_MAXK * ESM_MAXK];   // log Ptr (row: from), from the host's log
    double lpi0[ESM_MAXK];            // log π0
    double c0, c1, cpi;               // 2·n·log σ0, 2·n·log σ, n·log 2π (n = 3N series)
};

template <typename S>
__global__ void __launch_bounds__(256) k_ffbs_ll(const S *__restrict__ z, const double *__restrict__ mu64,
                                                 const double *__restrict__ harm, const int *__restrict__ mask, int N,
                                                 int T, int R, int K, double phi, double sigma, double sigma0,
                                                 FfbsConst c, double *__restrict__ ll) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t TR = (size_t)T * R;
    if (g >= TR) return;
    const int j = (int)(g % R), t = (int)(g / R);
    if (!mask[j]) return;
    const int nser = ESM_NREN * N;
    const double sd = t == 0 ? sigma0 : sigma, h1 = harm[2 * t], h2 = harm[2 * t + 1];
    for (int k = 0; k < K; k++) {
        double q = 0.0;
        for (int i = 0; i < nser; i++) {
            const double zt = (double)z[(size_t)i * TR + g];
            double e;
            if (t == 0) {
                e = zt;
            } else {
                const double *M = mu64 + ((size_t)k * nser + i) * 3;
                e = zt - phi * (double)z[(size_t)i * TR + g - R] - (M[0] + h1 * M[1] + h2 * M[2]);
            }
            e /= sd;
            q += e * e;
        }
        ll[(size_t)k * TR + g] = -0.5 * (q + (t == 0 ? c.c0 : c.c1) + c.cpi);
    }
}

__device__ __forceinline__ double lse_k(const double *v, int K) {
    double mx = -INFINITY;
    for (int k = 0; k < K; k++) mx = fmax(mx, v[k]);
    if (!isfinite(mx)) return mx;
    double s = 0.0;
    for (int k = 0; k < K; k++) s += exp(v[k] - mx);
    return mx + log(s);
}

__device__ __forceinline__ int sample_log_u(const double *lw, int K, double u) {
    double mx = -INFINITY, w[ESM_MAXK], tot = 0.0;
    for (int k = 0; k < K; k++) mx = fmax(mx, lw[k]);
    for (int k = 0; k < K; k++) {
        w[k] = exp(lw[k] - mx);
        tot += w[k];
    }
    const double t = u * tot;
    double acc = 0.0;
    for (int k = 0; k < K; k++) {
        acc += w[k];
        if (t <= acc) return k;
    }
    return K - 1;
}

// one thread per replica: the forward filter in place over the emissions, then the draws
__global__ void k_ffbs_scan(double *__restrict__ la, int *__restrict__ m, const int *__restrict__ mask, int T, int R,
                            int K, FfbsConst c, uint32_t k0, uint32_t k1, uint32_t ctr) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= R || !mask[j]) return;
    const size_t TR = (size_t)T * R;
    double prev[ESM_MAXK], v[ESM_MAXK];
    for (int k = 0; k < K; k++) {
        prev[k] = c.lpi0[k] + la[(size_t)k * TR + j];
        la[(size_t)k * TR + j] = prev[k];
    }
    for (int t = 1; t < T; t++) {
        double cur[ESM_MAXK];
        for (int k = 0; k < K; k++) {
            for (int i = 0; i < K; i++) v[i] = prev[i] + c.lP[i * K + k];
            cur[k] = lse_k(v, K) + la[(size_t)k * TR + (size_t)t * R + j];
        }
        for (int k = 0; k < K; k++) {
            prev[k] = cur[k];
            la[(size_t)k * TR + (size_t)t * R + j] = cur[k];
        }
    }
    uint32_t w[4];
    philox4x32(k0, k1, (uint32_t)j, (uint32_t)(T - 1), TAG_FFBS, ctr, w);
    int next = sample_log_u(prev, K, u01(w[0], w[1]));
    m[(size_t)(T - 1) * R + j] = next;
    for (int t = T - 2; t >= 0; t--) {
        for (int i = 0; i < K; i++) v[i] = la[(size_t)i * TR + (size_t)t * R + j] + c.lP[i * K + next];
        philox4x32(k0, k1, (uint32_t)j, (uint32_t)t, TAG_FFBS, ctr, w);
        next = sample_log_u(v, K, u01(w[0], w[1]));
        m[(size_t)t * R + j] = next;
    }
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename S>
int launch_ffbs(esm_handle *h, const int *mask, uint32_t k0, uint32_t k1, uint32_t ctr) {
    const int N = h->N, T = h->T, R = h->R, K = h->m.K;
    const size_t TR = (size_t)T * R;
    if (!h->d_ffbs) {
        ESM_HIP_CHECK(hipMalloc((void **)&h->d_ffbs, (size_t)K * TR * sizeof(double)));
        ESM_HIP_CHECK(hipMalloc((void **)&h->d_ffbs_mask, (size_t)R * sizeof(int)));
    }
    FfbsConst c{};
    for (int q = 0; q < K * K; q++) c.lP[q] = log(h->Ptr_h[q]);
    for (int k = 0; k < K; k++) c.lpi0[k] = log(h->pi0_h[k]);
    const int nser = ESM_NREN * N;
    c.c0 = 2.0 * nser * log(h->m.sigma0);
    c.c1 = 2.0 * nser * log(h->m.sigma);
    c.cpi = nser * log(2 * 3.14159265358979323846);
    ESM_HIP_CHECK(hipMemcpyAsync(h->d_ffbs_mask, mask, (size_t)R * sizeof(int), hipMemcpyHostToDevice, h->stream));
    ProfScope ps(h, ESM_KERN_WEATHER);
    hipLaunchKernelGGL((k_ffbs_ll<S>), dim3(nblk(TR, 256)), dim3(256), 0, h->stream, (const S *)h->d_z, h->d_mu64,
                       h->d_harm, h->d_ffbs_mask, N, T, R, K, h->m.phi, h->m.sigma, h->m.sigma0, c, h->d_ffbs);
    hipLaunchKernelGGL(k_ffbs_scan, dim3(nblk(R, 64)), dim3(64), 0, h->stream, h->d_ffbs, h->d_m, h->d_ffbs_mask, T, R,
                       K, c, k0, k1, ctr);
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

template <typename S>
int launch_weather(esm_handle *h, bool init, int B, int compensate, uint32_t k0, uint32_t k1, uint32_t ctr) {
    const int N = h->N, T = h->T, R = h->R, K = h->m.K;
    const size_t TR = (size_t)T * R;
    ProfScope ps(h, ESM_KERN_WEATHER);
    if (init) {
        hipLaunchKernelGGL(k_regime_init, dim3(nblk(R, 64)), dim3(64), 0, h->stream, h->d_mnew, h->d_Ptr, h->d_pi0, K,
                           T, R, k0, k1, ctr);
        ESM_HIP_CHECK(hipMemsetAsync(h->d_moved, 0, R * sizeof(int), h->stream));
        // every replica moves (c = 0, s = 1 set by the caller in wpar)
        std::vector<int> ones(R, 1);
        ESM_HIP_CHECK(hipMemcpyAsync(h->d_moved, ones.data(), R * sizeof(int), hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    } else {
        ESM_HIP_CHECK(hipMemcpyAsync(h->d_mnew, h->d_m, TR * sizeof(int), hipMemcpyDeviceToDevice, h->stream));
        hipLaunchKernelGGL(k_window, dim3(nblk(R, 64)), dim3(64), 0, h->stream, h->d_m, h->d_mnew, h->d_wpar,
                           h->d_moved, h->d_Ptr, h->d_pi0, K, T, R, B, k0, k1, ctr);
    }
    WxArgs<S> a;
    a.N = N;
    a.T = T;
    a.R = R;
    a.K = K;
    a.chunk = h->wx_chunk;
    a.nchunk = (T + a.chunk - 1) / a.chunk;
    a.phi = (S)h->m.phi;
    a.sigma = (S)h->m.sigma;
    a.sigma0 = (S)h->m.sigma0;
    a.mu = (const S *)h->d_mu;
    a.xbar = (const S *)h->d_xbar;
    a.hod = h->d_hod;
    a.harm = h->d_harm;
    a.wpar = h->d_wpar;
    a.moved = h->d_moved;
    a.m = h->d_m;
    a.mnew = h->d_mnew;
    a.k0 = k0;
    a.k1 = k1;
    a.ctr = ctr;
    const size_t n1 = (size_t)a.nchunk * N * R;
    hipLaunchKernelGGL((k_wx_local<S>), dim3(nblk(n1, 256)), dim3(256), 0, h->stream, a, (const S *)h->d_z,
                       (S *)h->d_ytmp, (S *)h->d_carry);
    hipLaunchKernelGGL((k_wx_carry<S>), dim3(nblk((size_t)ESM_NREN * N * R, 256)), dim3(256), 0, h->stream, a,
                       (const S *)h->d_carry, (S *)h->d_zin);
    const DevModel<S> md = dev_model<S>(h);
    hipLaunchKernelGGL((k_wx_apply<S>), dim3(nblk(n1, 256)), dim3(256), 0, h->stream, a, md, (const S *)h->d_ytmp,
                       (const S *)h->d_zin, (S *)h->d_z, (S *)h->d_x, (S *)h->d_v, (S *)h->d_r, compensate);
    ESM_HIP_CHECK(hipGetLastError());
    std::swap(h->d_m, h->d_mnew);
    return 0;
}

template int launch_ffbs<float>(esm_handle *, const int *, uint32_t, uint32_t, uint32_t);
template int launch_ffbs<double>(esm_handle *, const int *, uint32_t, uint32_t, uint32_t);
template int launch_weather<float>(esm_handle *, bool, int, int, uint32_t, uint32_t, uint32_t);
template int launch_weather<double>(esm_handle *, bool, int, int, uint32_t, uint32_t, uint32_t);

}  // namespace esm
