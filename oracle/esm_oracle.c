/*
 * esm_oracle.c — FP64 CPU restatement of the reference algorithm (TEST INFRASTRUCTURE).
 *
 * The checker of the HIP path (tests, smoke check, CPU baseline of the benchmark only). Each
 * function names the reference code it restates. The restatement follows the reference's loops
 * for the synthetic system of include/esm.h: hard line limits without losses, no DC links, no
 * vintages, no unit commitment, balance cells = nodes (spatially complete level), renewables as
 * load (A k), storage in the state-of-charge form.
 */
#include "esm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(n, t, r, T, R) (((size_t)(n) * (T) + (t)) * (R) + (r))

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double signd(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

/* soc_power, src/energy.jl:437-441 */
static void soc_power(const esm_model *m, int s, int i, double dt, double ec, double ec1, double *pc, double *mcp) {
    int u = s * m->N + i;
    double d = m->stor_hours[u] * (ec1 - ec) / dt - m->stor_qt[u];
    if (d >= 0) {
        *pc = m->stor_eta_in[u] > 0 ? d / m->stor_eta_in[u] : 0.0;
        *mcp = 0.0;
        return;
    }
    *pc = 0.0;
    *mcp = fmin(-d * m->stor_eta_out[u], 1.0);
}

/* soc_cost, src/energy.jl:448-463: returns the cost, writes d and the injection k (p⁺ − p⁻) */
static double soc_cost(const esm_model *m, int s, int i, double k, double w, double dt, double cv,
                       double ye, double ye1, double *dout, double *inj) {
    int u = s * m->N + i;
    double ec = clampd(ye, 0.0, 1.0), ec1 = clampd(ye1, 0.0, 1.0);
    double h = m->stor_hours[u];
    double d = h * (ec1 - ec) / dt - m->stor_qt[u];
    double eta = m->stor_eta_in[u];
    double pc, mcp;
    if (d >= 0) {
        pc = eta > 0 ? d / eta : 0.0;
        mcp = 0.0;
    } else {
        pc = 0.0;
        mcp = fmin(-d * m->stor_eta_out[u], 1.0);
    }
    double c = dt * m->mc[(ESM_KSTOR + s) * m->N + i] * k * (pc + mcp) + cv * w * h * fabs(ye - ec);
    if (pc > 1) c += cv * dt * w * (pc - 1);
    if (d < 0 && m->stor_qt[u] <= 0) {
        double ex = -d * m->stor_eta_out[u] - 1.0;
        if (ex > 0) c += cv * dt * w * ex;
    }
    if (d >= 0 && eta <= 0) c += cv * dt * w * d;
    if (dout) *dout = d;
    if (inj) *inj = k * (pc - mcp);
    return c;
}

/* the clamped capacities kc and the capex, src/energy.jl:557-568 */
static double capex_and_kc(const esm_model *m, const double *k, double *kc) {
    int N = m->N;
    double capex = 0.0;
    for (int f = 0; f < ESM_NK; f++) {
        double dot = 0.0, pen = 0.0;
        for (int n = 0; n < N; n++) {
            int u = f * N + n;
            kc[u] = clampd(k[u], m->klo[u], m->khi[u]);
            dot += m->capex[u] * k[u];
            pen += m->capex[u] * fabs(k[u] - kc[u]);
        }
        capex += dot + m->gamma_k * pen;
    }
    return capex;
}

int orc_energy(const esm_model *m, int R, const double *v, const double *k, const double *x,
               double *parts, double *per_hour, double *rout, double *Fout) {
    int N = m->N, L = m->L, T = m->T;
    double dt = m->dt, cv = m->c_viol;
    double *kc = (double *)malloc(sizeof(double) * ESM_NK * N);
    double *r = (double *)malloc(sizeof(double) * (size_t)N * T);
    double *F = (double *)malloc(sizeof(double) * (size_t)L * T);
    double *ph = (double *)malloc(sizeof(double) * (size_t)T);
    if (!kc || !r || !F || !ph) return -3;
    double capex = capex_and_kc(m, k, kc);
    const double *th = v, *yg = v + (size_t)N * T * R, *ye[2] = {v + 2 * (size_t)N * T * R, v + 3 * (size_t)N * T * R};
    for (int j = 0; j < R; j++) {
        double pen = 0.0, cgen = 0.0, cst = 0.0, co2 = 0.0, shed = 0.0, dump = 0.0;
        for (int t = 0; t < T; t++) {
            ph[t] = 0.0;
            for (int n = 0; n < N; n++) r[(size_t)n * T + t] = fmax(m->load[(size_t)n * T + t], 0.0);
        }
        /* potentials and line flows, :589-598 */
        for (int t = 0; t < T; t++)
            for (int e = 0; e < L; e++) {
                int a = m->line_bus0[e], b = m->line_bus1[e];
                double fe = m->line_b[e] * (th[IDX(a, t, j, T, R)] - th[IDX(b, t, j, T, R)]);
                F[(size_t)e * T + t] = fe;
                r[(size_t)a * T + t] += fe;
                r[(size_t)b * T + t] += -fe;
                double ov = fmax(fabs(fe) - m->line_s[e], 0.0);   /* s·ψ(|f|/s), hard limit */
                pen += cv * dt * ov;
                ph[t] += cv * dt * ov;
            }
        /* renewables as load (vf = 0: Y = Yc = 1), :613-633 */
        for (int c = 0; c < ESM_NREN; c++)
            for (int t = 0; t < T; t++)
                for (int n = 0; n < N; n++) {
                    double A = clampd(x[IDX(c * N + n, t, j, T, R)], 0.0, 1.0);
                    double kv = kc[c * N + n];
                    double p = A * kv * 1.0;
                    double cc = m->mc[c * N + n] * p;
                    cgen += dt * cc;
                    ph[t] += dt * cc;
                    r[(size_t)n * T + t] -= p;
                }
        /* gas (utilization field, availability 1), :613-640 */
        for (int t = 0; t < T; t++)
            for (int n = 0; n < N; n++) {
                double Y = yg[IDX(n, t, j, T, R)], Yc = clampd(Y, 0.0, 1.0);
                double kv = kc[ESM_KGAS * N + n];
                double p = 1.0 * kv * Yc;
                double cc = m->mc[ESM_KGAS * N + n] * p;
                cgen += dt * cc;
                co2 += dt * m->ef[n] * p;
                r[(size_t)n * T + t] -= p;
                ph[t] += dt * cc;
            }
        for (int t = 0; t < T; t++)
            for (int n = 0; n < N; n++) {
                double Y = yg[IDX(n, t, j, T, R)], Yc = clampd(Y, 0.0, 1.0);
                double ex = fabs(Y - Yc);
                if (ex > 0) {
                    double c = cv * dt * (kc[ESM_KGAS * N + n] + m->kappa0) * ex;
                    pen += c;
                    ph[t] += c;
                }
            }
        /* storage, :643-654 */
        for (int s = 0; s < ESM_NSTOR; s++)
            for (int t = 0; t < T; t++) {
                int t1 = (t + 1) % T;
                for (int n = 0; n < N; n++) {
                    double kv = kc[(ESM_KSTOR + s) * N + n];
                    double y0 = ye[s][IDX(n, t, j, T, R)], y1 = ye[s][IDX(n, t1, j, T, R)];
                    double inj;
                    double c = soc_cost(m, s, n, kv, kv + m->kappa0, dt, cv, y0, y1, NULL, &inj);
                    double pc, mcp;
                    soc_power(m, s, n, dt, clampd(y0, 0.0, 1.0), clampd(y1, 0.0, 1.0), &pc, &mcp);
                    double run = dt * m->mc[(ESM_KSTOR + s) * N + n] * kv * (pc + mcp);
                    cst += run;
                    pen += c - run;
                    ph[t] += c;
                    r[(size_t)n * T + t] += inj;
                }
            }
        /* priced balance (complete level: the projection is the identity), :656-677 */
        for (int t = 0; t < T; t++)
            for (int n = 0; n < N; n++) {
                double xr = r[(size_t)n * T + t];
                if (xr > 0) {
                    shed += dt * xr;
                    ph[t] += dt * m->c_shed * xr;
                } else {
                    dump -= dt * xr;
                    ph[t] += -dt * m->c_dump * xr;
                }
            }
        double cbal = m->c_shed * shed + m->c_dump * dump;
        double op = cgen + cst + cbal + pen;
        if (parts) {
            double *p = parts + (size_t)j * ESM_NPARTS;
            p[0] = capex; p[1] = op; p[2] = cgen; p[3] = cst; p[4] = shed; p[5] = dump; p[6] = pen; p[7] = co2;
        }
        if (per_hour)
            for (int t = 0; t < T; t++) per_hour[(size_t)t * R + j] = ph[t];
        if (rout)
            for (int n = 0; n < N; n++)
                for (int t = 0; t < T; t++) rout[IDX(n, t, j, T, R)] = r[(size_t)n * T + t];
        if (Fout)
            for (int e = 0; e < L; e++)
                for (int t = 0; t < T; t++) Fout[IDX(e, t, j, T, R)] = F[(size_t)e * T + t];
    }
    free(kc); free(r); free(F); free(ph);
    return 0;
}

/* capacity_scores, src/replicas.jl:329-338, through energy_cost(grad = true),
 * src/energy.jl:680-757 (gen_grad :835-854, soc_grad :466-502) */
int orc_scores(const esm_model *m, int R, const double *v, const double *k, const double *x,
               double *sk, double *fk, double *sk_abs) {
    int N = m->N, T = m->T;
    double dt = m->dt, cv = m->c_viol;
    double *kc = (double *)malloc(sizeof(double) * ESM_NK * N);
    double *r = (double *)malloc(sizeof(double) * (size_t)N * T * R);
    double *dk = (double *)malloc(sizeof(double) * ESM_NK * N);
    double *ak = (double *)malloc(sizeof(double) * ESM_NK * N);
    if (!kc || !r || !dk || !ak) return -3;
    capex_and_kc(m, k, kc);
    orc_energy(m, R, v, k, x, NULL, NULL, r, NULL);
    const double *yg = v + (size_t)N * T * R, *ye[2] = {v + 2 * (size_t)N * T * R, v + 3 * (size_t)N * T * R};
#define GR(n, t) (dt * (r[IDX(n, t, j, T, R)] > 0 ? m->c_shed : (r[IDX(n, t, j, T, R)] < 0 ? -m->c_dump : 0.0)))
    for (int j = 0; j < R; j++) {
        memset(dk, 0, sizeof(double) * ESM_NK * N);
        memset(ak, 0, sizeof(double) * ESM_NK * N);
        /* renewables: comp (no utilization field), Yc = 1: dk += Δt mc A − A Gr */
        for (int c = 0; c < ESM_NREN; c++)
            for (int t = 0; t < T; t++)
                for (int n = 0; n < N; n++) {
                    double A = clampd(x[IDX(c * N + n, t, j, T, R)], 0.0, 1.0);
                    double dck = m->mc[c * N + n] * A * 1.0;
                    double a = dt * dck, b = A * 1.0 * GR(n, t);
                    dk[c * N + n] += a - b;
                    ak[c * N + n] += fabs(a) + fabs(b);
                }
        /* gas: box term outside (0, 1), then Δt dC/dk − A Yc Gr */
        for (int t = 0; t < T; t++)
            for (int n = 0; n < N; n++) {
                double Y = yg[IDX(n, t, j, T, R)], Yc = clampd(Y, 0.0, 1.0);
                int u = ESM_KGAS * N + n;
                if (!(0.0 < Y && Y < 1.0)) {
                    double a = cv * dt * fabs(Y - Yc);
                    dk[u] += a;
                    ak[u] += a;
                }
                double dck = m->mc[u] * 1.0 * Yc;
                double a = dt * dck, b = 1.0 * Yc * GR(n, t);
                dk[u] += a - b;
                ak[u] += fabs(a) + fabs(b);
            }
        /* storage, soc_grad */
        for (int s = 0; s < ESM_NSTOR; s++) {
            for (int t = 0; t < T; t++) {
                int t1 = (t + 1) % T;
                for (int n = 0; n < N; n++) {
                    int us = s * N + n, u = (ESM_KSTOR + s) * N + n;
                    double h = m->stor_hours[us];
                    double ec = clampd(ye[s][IDX(n, t, j, T, R)], 0.0, 1.0), ec1 = clampd(ye[s][IDX(n, t1, j, T, R)], 0.0, 1.0);
                    double d = h * (ec1 - ec) / dt - m->stor_qt[us];
                    double g = GR(n, t);
                    double mcs = m->mc[u];
                    double a;
                    if (d >= 0 && m->stor_eta_in[us] <= 0) {
                        a = cv * dt * d;
                        dk[u] += a; ak[u] += fabs(a);
                    } else if (d >= 0) {
                        double pc = d / m->stor_eta_in[us];
                        a = dt * mcs * pc + g * pc + (pc > 1 ? cv * dt * (pc - 1) : 0.0);
                        dk[u] += a;
                        ak[u] += fabs(dt * mcs * pc) + fabs(g * pc) + (pc > 1 ? cv * dt * (pc - 1) : 0.0);
                    } else {
                        double mc_ = -d * m->stor_eta_out[us];
                        double mcc = fmin(mc_, 1.0);
                        double pe = (mc_ > 1 && m->stor_qt[us] <= 0) ? cv * dt * (mc_ - 1) : 0.0;
                        a = dt * mcs * mcc - g * mcc + pe;
                        dk[u] += a;
                        ak[u] += fabs(dt * mcs * mcc) + fabs(g * mcc) + pe;
                    }
                }
            }
            for (int t = 0; t < T; t++)
                for (int n = 0; n < N; n++) {
                    int us = s * N + n, u = (ESM_KSTOR + s) * N + n;
                    double y = ye[s][IDX(n, t, j, T, R)];
                    if (!(0.0 < y && y < 1.0)) {
                        double a = cv * m->stor_hours[us] * fabs(y - clampd(y, 0.0, 1.0));
                        dk[u] += a;
                        ak[u] += a;
                    }
                }
        }
        /* gk_op = mask · dkc (mask: raw capacity strictly inside the box); ŝ = −β_v gk_op */
        for (int f = 0; f < ESM_NK; f++)
            for (int n = 0; n < N; n++) {
                int u = f * N + n;
                double mask = (m->klo[u] < k[u] && k[u] < m->khi[u]) ? 1.0 : 0.0;
                sk[((size_t)j * ESM_NK + f) * N + n] = -m->beta_v * (mask * dk[u]);
                if (sk_abs) sk_abs[((size_t)j * ESM_NK + f) * N + n] = m->beta_v * mask * ak[u];
            }
    }
#undef GR
    if (fk)
        for (int f = 0; f < ESM_NK; f++)
            for (int n = 0; n < N; n++) {
                int u = f * N + n;
                double ex = k[u] - kc[u];
                fk[u] = m->beta_k * (m->capex[u] + m->gamma_k * m->capex[u] * signd(ex));
            }
    free(kc); free(r); free(dk); free(ak);
    return 0;
}

/* log_measure, src/energy.jl:790-806: with the physical reference measure, n_u log(kc_u + κ0) per
 * dispatchable unit without weather availability (gas: n_u = T, no unit commitment) and per
 * storage unit (n_u = T), kc the capacity clamped to its box; renewables (weather-dependent
 * availability) keep the utilization measure. 0 for the utilization measure. */
double orc_log_measure(const esm_model *m, const double *k, int physical) {
    if (!physical) return 0.0;
    double val = 0.0;
    for (int f = ESM_KGAS; f < ESM_NK; f++)
        for (int n = 0; n < m->N; n++) {
            int u = f * m->N + n;
            val += (double)m->T * log(clampd(k[u], m->klo[u], m->khi[u]) + m->kappa0);
        }
    return val;
}

/* log_measure_grad, src/energy.jl:811-829: n_u/(k + κ0) where lo < k < hi, else 0, into g [NK][N]. */
void orc_log_measure_grad(const esm_model *m, const double *k, int physical, double *g) {
    for (int u = 0; u < ESM_NK * m->N; u++) g[u] = 0.0;
    if (!physical) return;
    for (int f = ESM_KGAS; f < ESM_NK; f++)
        for (int n = 0; n < m->N; n++) {
            int u = f * m->N + n;
            if (m->klo[u] < k[u] && k[u] < m->khi[u]) g[u] = (double)m->T / (k[u] + m->kappa0);
        }
}

/* langevin_update / langevin_capacities, src/replicas.jl:357-403 */
int orc_langevin(const esm_model *m, int R, const double *k, const double *sk, const double *fk,
                 const double *weights, const double *xi, const double *eta, const double *kappa_s,
                 double kmin, double clip, int box, double *knew, int *nclip) {
    int N = m->N;
    double rr = m->beta_k / m->beta_v;
    int nc = 0;
    for (int f = 0; f < ESM_NK; f++) {
        for (int n = 0; n < N; n++) {
            int u = f * N + n;
            knew[u] = k[u];
            if (eta[f] == 0) continue;
            double sbar = 0.0;
            for (int j = 0; j < R; j++) sbar += weights[j] * sk[((size_t)j * ESM_NK + f) * N + n];
            double g = -fk[u] + rr * sbar;
            double ks = kappa_s[f];
            double w = log(k[u] + ks);
            double dw0 = eta[f] * ((k[u] + ks) * g + 1);
            if (fabs(dw0) > clip) nc++;
            double w1 = w + clampd(dw0, -clip, clip) + sqrt(2 * eta[f]) * xi[u];
            double wmin = log(kmin + ks);
            if (w1 < wmin) w1 = 2 * wmin - w1;
            if (box) {
                double wm = log(fmax(m->khi[u], kmin + 1.0) + ks);
                if (w1 > wm && w <= wm) w1 = fmax(2 * wm - w1, wmin);
            }
            knew[u] = exp(w1) - ks;
        }
    }
    if (nclip) *nclip = nc;
    return 0;
}

/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11. Multipliers 0xD2511F53, 0xCD9E8D57; Weyl increments 0x9E3779B9, 0xBB67AE85. */
void orc_philox(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int i = 0; i < 10; i++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
