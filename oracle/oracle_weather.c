/*
 * oracle_weather.c — FP64 CPU restatement of the replica weather process (TEST INFRASTRUCTURE).
 *
 * Follows src/replicas.jl:29-236 and src/weather.jl:43, :461-481 for the synthetic weather model
 * of include/esm.h (Φ = φ I, Σ_k = σ² I, Σ0 = σ0² I, harmonics (1, cos 2πτ/T, sin 2πτ/T)), step
 * by step as the reference does: innovations, pCN rotation, recursion, window messages and
 * forward draws, then the compensated re-dispatch. Random numbers: Philox4x32-10 addressed as
 * documented in the HIP library (normals ξ(i, τ, j) by Box–Muller from counter (i, τ/2, j, c)).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "esm_oracle.h"

#define TAG_REG 0x52454731u
#define TAG_WIN 0x57494E44u
#define ESM_MAXK_ORACLE 8   /* regimes of the FFBS restatement */

static double u01(uint32_t hi, uint32_t lo) {
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 11) * 0x1.0p-53;
}

static double normal_of(const uint32_t key[2], int i, int tau, int j, uint32_t c) {
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(tau >> 1), (uint32_t)j, c}, w[4];
    orc_philox(key, ctr, w);
    double u1 = u01(w[0], w[1]), u2 = u01(w[2], w[3]);
    double rad = sqrt(-2.0 * log(1.0 - u1));
    double ang = 6.283185307179586 * u2;
    return (tau & 1) ? rad * sin(ang) : rad * cos(ang);
}

/* categorical, src/replicas.jl:115-123 */
static int categorical(const double *p, int K, double u) {
    double tot = 0;
    for (int k = 0; k < K; k++) tot += p[k];
    double t = u * tot, c = 0;
    for (int k = 0; k < K; k++) {
        c += p[k];
        if (t <= c) return k;
    }
    return K - 1;
}

/* μ_k(τ) of series i, τ 0-based (harmonics of the 1-based step) */
static double var_mean_term(const esm_model *m, int k, int i, int tau) {
    int nser = ESM_NREN * m->N;
    const double *M = m->mu + ((size_t)k * nser + i) * 3;
    double ang = 2 * 3.14159265358979323846 * (tau + 1) / m->T;
    return M[0] + cos(ang) * M[1] + sin(ang) * M[2];
}

#define Z(i, t) z[((size_t)(i) * T + (t)) * R + j]
#define MR(t) mreg[(size_t)(t) * R + j]

/* z0_innovations, src/replicas.jl:49-58, of replica j into E [nser][T] */
static void innovations(const esm_model *m, int R, int j, const int *mreg, const double *z, double *E) {
    int T = m->T, nser = ESM_NREN * m->N;
    for (int i = 0; i < nser; i++) {
        E[(size_t)i * T] = Z(i, 0) / m->sigma0;
        for (int t = 1; t < T; t++)
            E[(size_t)i * T + t] = (Z(i, t) - (m->phi * Z(i, t - 1) + var_mean_term(m, MR(t), i, t))) / m->sigma;
    }
}

/* z0_from_innovations, src/replicas.jl:61-69 */
static void from_innovations(const esm_model *m, int R, int j, const int *mreg, const double *E, double *z) {
    int T = m->T, nser = ESM_NREN * m->N;
    for (int i = 0; i < nser; i++) {
        Z(i, 0) = m->sigma0 * E[(size_t)i * T];
        for (int t = 1; t < T; t++)
            Z(i, t) = m->phi * Z(i, t - 1) + var_mean_term(m, MR(t), i, t) + m->sigma * E[(size_t)i * T + t];
    }
}

static void set_x(const esm_model *m, int R, int j, const double *z, double *x) {
    int T = m->T, nser = ESM_NREN * m->N;
    for (int i = 0; i < nser; i++)
        for (int t = 0; t < T; t++) {
            int hod = (int)fmod(t * m->dt, 24.0);
            x[((size_t)i * T + t) * R + j] = 1.0 / (1.0 + exp(-(m->xbar[(size_t)i * 24 + hod] + Z(i, t))));
        }
}

/* sample_weather, src/weather.jl:461-481 (regimes by categorical inversion of one uniform each) */
int orc_weather_init(const esm_model *m, int R, const uint32_t key[2], uint32_t ctr, int *mreg, double *z, double *x) {
    int T = m->T, K = m->K, nser = ESM_NREN * m->N;
    double *E = (double *)malloc(sizeof(double) * (size_t)nser * T);
    if (!E) return -3;
    for (int j = 0; j < R; j++) {
        int prev = 0;
        for (int t = 0; t < T; t++) {
            uint32_t c4[4] = {(uint32_t)j, (uint32_t)t, TAG_REG, ctr}, w[4];
            orc_philox(key, c4, w);
            prev = categorical(t == 0 ? m->pi0 : m->Ptr + prev * K, K, u01(w[0], w[1]));
            MR(t) = prev;
        }
        for (int i = 0; i < nser; i++)
            for (int t = 0; t < T; t++) E[(size_t)i * T + t] = normal_of(key, i, t, j, ctr);
        from_innovations(m, R, j, mreg, E, z);
        set_x(m, R, j, z, x);
    }
    free(E);
    return 0;
}

/* regime_window_messages + regime_window, src/replicas.jl:131-164 (a, b 1-based) */
static void regime_window(const esm_model *m, int R, int j, int *mreg, int a, int b, const double *us) {
    int T = m->T, K = m->K, n = b - a + 1;
    double *beta = (double *)malloc(sizeof(double) * K * n);
    for (int q = 0; q < n * K; q++) beta[q] = 1.0;
    if (b < T) {
        int mb1 = MR(b);
        for (int i = 0; i < K; i++) beta[(n - 1) * K + i] = m->Ptr[i * K + mb1];
        for (int tau = b - 1; tau >= a; tau--) {
            int q = tau - a;
            for (int i = 0; i < K; i++) {
                double acc = 0;
                for (int l = 0; l < K; l++) acc += m->Ptr[i * K + l] * beta[(q + 1) * K + l];
                beta[q * K + i] = acc;
            }
        }
    }
    double p[8];
    for (int tau = a; tau <= b; tau++) {
        int q = tau - a;
        for (int i = 0; i < K; i++)
            p[i] = (tau == 1 ? m->pi0[i] : m->Ptr[MR(tau - 2) * K + i]) * beta[q * K + i];
        MR(tau - 1) = categorical(p, K, us[q]);
    }
    free(beta);
}

/* weather_process_step (src/replicas.jl:179-197) of every replica, then compensate_weather!
 * (:207-236) with the residuals of the state before and after the weather change. */
int orc_weather_step(const esm_model *m, int R, const double *theta, const double *win_prob, int B, int compensate,
                     const uint32_t key[2], uint32_t ctr, int *mreg, double *z, double *x, double *v, const double *k,
                     int *moved) {
    int N = m->N, T = m->T, nser = ESM_NREN * N;
    size_t TR = (size_t)T * R;
    double *E = (double *)malloc(sizeof(double) * (size_t)nser * T);
    double *r0 = (double *)malloc(sizeof(double) * (size_t)N * TR);
    double *r1 = (double *)malloc(sizeof(double) * (size_t)N * TR);
    double *us = (double *)malloc(sizeof(double) * (B > 0 ? B : 1));
    if (!E || !r0 || !r1 || !us) return -3;
    if (compensate) orc_energy(m, R, v, k, x, NULL, NULL, r0, NULL);
    for (int j = 0; j < R; j++) {
        int mv = 0;
        if (theta[j] > 0) {
            innovations(m, R, j, mreg, z, E);
            double c = cos(theta[j]), s = sin(theta[j]);
            for (int i = 0; i < nser; i++)
                for (int t = 0; t < T; t++)
                    E[(size_t)i * T + t] = c * E[(size_t)i * T + t] + s * normal_of(key, i, t, j, ctr);
            from_innovations(m, R, j, mreg, E, z);
            mv = 1;
        }
        if (win_prob[j] > 0 && B > 0) {
            uint32_t c4[4] = {(uint32_t)j, 0u, TAG_WIN, ctr}, w[4];
            orc_philox(key, c4, w);
            if (u01(w[0], w[1]) < win_prob[j]) {
                /* window_bounds, :167 */
                int a = 1 + (int)floor(u01(w[2], w[3]) * T);
                if (a > T) a = T;
                int b = a + B - 1;
                if (b > T) b = T;
                for (int q = 0; q < b - a + 1; q++) {
                    uint32_t cq[4] = {(uint32_t)j, (uint32_t)(1 + q / 2), TAG_WIN, ctr}, wq[4];
                    orc_philox(key, cq, wq);
                    us[q] = (q & 1) ? u01(wq[2], wq[3]) : u01(wq[0], wq[1]);
                }
                innovations(m, R, j, mreg, z, E);
                regime_window(m, R, j, mreg, a, b, us);
                from_innovations(m, R, j, mreg, E, z);
                mv = 2;
            }
        }
        if (mv) set_x(m, R, j, z, x);
        if (moved) moved[j] = mv;
    }
    if (compensate) {
        orc_energy(m, R, v, k, x, NULL, NULL, r1, NULL);
        double *yg = v + (size_t)N * TR;
        for (int n = 0; n < N; n++) {
            double kap = fmin(fmax(k[ESM_KGAS * N + n], m->klo[ESM_KGAS * N + n]), m->khi[ESM_KGAS * N + n]);
            for (size_t g = 0; g < TR; g++) {
                size_t o = (size_t)n * TR + g;
                double need = r1[o] - r0[o];
                if (fabs(need) < 1e-6) continue;
                double y = yg[o];
                double p = kap * fmin(fmax(y, 0.0), 1.0);
                double wq = need > 0 ? fmax(kap - p, 0.0) : p;
                double tot = wq;
                if (!(tot > 1e-9)) continue;
                double aa = need > 0 ? fmin(need, tot) / tot : fmax(need, -tot) / tot;
                if (fabs(kap) < 1e-9) continue;
                double dy = aa * wq / kap;
                if (dy == 0.0) continue;
                yg[o] = y + dy;
            }
        }
    }
    free(E); free(r0); free(r1); free(us);
    return 0;
}

int orc_categorical(const double *p, int K, double u) { return categorical(p, K, u); }

void orc_regime_window_path(const esm_model *m, int *path, int a, int b, const double *us) {
    regime_window(m, 1, 0, path, a, b, us);
}

/* ============================================================================================
 * Centred FFBS of the regimes given the anomalies (replica_ffbs = true): fine_regime_step!
 * (src/weather_fine.jl:290-300) at the single complete level — the regime path is redrawn from
 * p(m | z) by ffbs (src/weather.jl:160-174) with the emission log-likelihoods of
 * emission_loglik (:117-133) and sample_log (:176-185); z and x are unchanged (no detail
 * levels). Uniforms: Philox (replica, τ, TAG_FFBS, c) words 0–1 for the draw at step τ.
 * ============================================================================================ */
#define TAG_FFBS 0x46464253u

/* emission_loglik of replica j at step τ (0-based) under regime k: step 1 under the stationary
 * start N(0, σ0² I) (the same for every regime), later steps the VAR(1) residual
 * e = z_τ − φ z_{τ−1} − μ_k(τ) under N(0, σ² I) */
static double emission(const esm_model *m, int R, int j, const double *z, int k, int tau) {
    int T = m->T, nser = ESM_NREN * m->N;
    double sd = tau == 0 ? m->sigma0 : m->sigma, q = 0.0;
    for (int i = 0; i < nser; i++) {
        double e = tau == 0 ? Z(i, 0) : Z(i, tau) - m->phi * Z(i, tau - 1) - var_mean_term(m, k, i, tau);
        e /= sd;
        q += e * e;
    }
    return -0.5 * (q + 2.0 * nser * log(sd) + nser * log(2 * 3.14159265358979323846));
}

/* logsumexp (src/weather.jl:136) */
static double lse(const double *v, int K) {
    double mx = -INFINITY;
    for (int k = 0; k < K; k++) mx = fmax(mx, v[k]);
    if (!isfinite(mx)) return mx;
    double s = 0.0;
    for (int k = 0; k < K; k++) s += exp(v[k] - mx);
    return mx + log(s);
}

/* sample_log with a given uniform: weights exp(lw − max), inversion of u · Σ w */
static int sample_log_u(const double *lw, int K, double u) {
    double mx = -INFINITY, w[ESM_MAXK_ORACLE], tot = 0.0;
    for (int k = 0; k < K; k++) mx = fmax(mx, lw[k]);
    for (int k = 0; k < K; k++) {
        w[k] = exp(lw[k] - mx);
        tot += w[k];
    }
    double t = u * tot, c = 0.0;
    for (int k = 0; k < K; k++) {
        c += w[k];
        if (t <= c) return k;
    }
    return K - 1;
}

int orc_weather_ffbs(const esm_model *m, int R, const double *z, int *mreg, const int *mask, const uint32_t key[2],
                     uint32_t ctr, double *loglik) {
    int T = m->T, K = m->K;
    if (K < 1 || K > ESM_MAXK_ORACLE) return -1;
    double lP[ESM_MAXK_ORACLE * ESM_MAXK_ORACLE], v[ESM_MAXK_ORACLE];
    for (int q = 0; q < K * K; q++) lP[q] = log(m->Ptr[q]);
    double *la = (double *)malloc(sizeof(double) * (size_t)K * T);
    if (!la) return -3;
    for (int j = 0; j < R; j++) {
        if (mask && !mask[j]) continue;
        /* forward filter: la[k][τ] = lse_i(la[i][τ−1] + log Ptr[i][k]) + ll[k][τ] */
        for (int k = 0; k < K; k++) la[k] = log(m->pi0[k]) + emission(m, R, j, z, k, 0);
        for (int t = 1; t < T; t++)
            for (int k = 0; k < K; k++) {
                for (int i = 0; i < K; i++) v[i] = la[(size_t)(t - 1) * K + i] + lP[i * K + k];
                la[(size_t)t * K + k] = lse(v, K) + emission(m, R, j, z, k, t);
            }
        if (loglik) loglik[j] = lse(la + (size_t)(T - 1) * K, K);
        /* backward sampling: m[T] from la[:, T], then m[τ] from la[:, τ] + log Ptr[:, m[τ+1]] */
        uint32_t c4[4] = {(uint32_t)j, (uint32_t)(T - 1), TAG_FFBS, ctr}, w[4];
        orc_philox(key, c4, w);
        int next = sample_log_u(la + (size_t)(T - 1) * K, K, u01(w[0], w[1]));
        MR(T - 1) = next;
        for (int t = T - 2; t >= 0; t--) {
            for (int i = 0; i < K; i++) v[i] = la[(size_t)t * K + i] + lP[i * K + next];
            uint32_t c5[4] = {(uint32_t)j, (uint32_t)t, TAG_FFBS, ctr};
            orc_philox(key, c5, w);
            next = sample_log_u(v, K, u01(w[0], w[1]));
            MR(t) = next;
        }
    }
    free(la);
    return 0;
}
