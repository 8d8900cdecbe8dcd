/*
 * oracle_follow.c — FP64 CPU restatement of the dispatch following of a capacity change
 * (TEST INFRASTRUCTURE): follow_capacities! (src/replicas.jl:448-472) with
 * set_dispatchable_capacity! (:412-435) and the kop curve of every renewable and storage unit
 * (kop_prepare / kop_term / kop_eval / kop_commit!, src/kop.jl:57-271), unit by unit in the
 * reference's unit order, on the local caches (the residual r), as the reference does.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "esm_oracle.h"

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

int orc_follow(const esm_model *m, int R, double *v, double *k, const double *x, const double *knew, double kmin) {
    int N = m->N, T = m->T;
    size_t TR = (size_t)T * R;
    double *yg = v + (size_t)N * TR;
    double *r = (double *)malloc(sizeof(double) * (size_t)N * TR);
    if (!r) return -3;
    /* dispatchable units at fixed output */
    for (int n = 0; n < N; n++) {
        int u = ESM_KGAS * N + n;
        double k0 = k[u], k1 = knew[u];
        k[u] += (k1 - k0);
        double kc0 = clampd(k0, m->klo[u], m->khi[u]), kc1 = clampd(k1, m->klo[u], m->khi[u]);
        double sc = kc0 == kc1 ? 1.0 : kc0 / kc1;
        if (sc == 1.0) continue;
        for (size_t g = 0; g < TR; g++) {
            double y = yg[(size_t)n * TR + g];
            yg[(size_t)n * TR + g] += y * (sc - 1);
        }
    }
    orc_energy(m, R, v, k, x, NULL, NULL, r, NULL);            /* LocalCost */
    /* renewables, then storage, in unit order */
    for (int f = 0; f < ESM_NK; f++) {
        if (f == ESM_KGAS) continue;
        for (int i = 0; i < N; i++) {
            int u = f * N + i;
            double k0 = k[u];
            if (!(k0 >= kmin)) {                                 /* the capacity alone */
                k[u] += knew[u] - k0;
                orc_energy(m, R, v, k, x, NULL, NULL, r, NULL);
                continue;
            }
            double uu = log(knew[u] / k0);
            double k1 = k0 * exp(uu);
            double kc0 = clampd(k0, m->klo[u], m->khi[u]), kc1 = clampd(k1, m->klo[u], m->khi[u]);
            double dkc = kc1 - kc0;
            int ug = ESM_KGAS * N + i;
            double kap = 1.0 * clampd(k[ug], m->klo[ug], m->khi[ug]);   /* κunit of the gas unit: A k */
            for (int t = 0; t < T; t++) {
                int t1 = (t + 1) % T;
                for (int j = 0; j < R; j++) {
                    size_t o = ((size_t)i * T + t) * R + j;
                    /* kop_own: injection per MW of clamped capacity */
                    double jinj;
                    if (f < ESM_NREN) {
                        jinj = clampd(x[((size_t)u * T + t) * R + j], 0.0, 1.0);
                    } else {
                        int s = f - ESM_KSTOR, us = s * N + i;
                        const double *ye = v + (size_t)(2 + s) * N * TR;
                        double ec = clampd(ye[o], 0.0, 1.0), ec1 = clampd(ye[((size_t)i * T + t1) * R + j], 0.0, 1.0);
                        double d = m->stor_hours[us] * (ec1 - ec) / m->dt - m->stor_qt[us];
                        double pc, mcp;
                        if (d >= 0) {
                            pc = m->stor_eta_in[us] > 0 ? d / m->stor_eta_in[us] : 0.0;
                            mcp = 0.0;
                        } else {
                            pc = 0.0;
                            mcp = fmin(-d * m->stor_eta_out[us], 1.0);
                        }
                        jinj = mcp - pc;
                    }
                    /* kop_term: the compensators of the step (the gas unit if |κ| ≥ κmin) */
                    double dinj = jinj * dkc, dS = -dinj;
                    if (fabs(kap) >= kmin) {
                        double w = yg[o];
                        double Pt = 0.0 + kap * w, Qs = 0.0 + kap * clampd(w, 0.0, 1.0);
                        if (fabs(Pt) >= 1e-6) {
                            double Q = r[o] + Qs;
                            double e = Pt - fmax(Q, 0.0);
                            double sig = kap * w / Pt;
                            double P1 = fmax(Q - dinj, 0.0) + e;
                            double y1 = sig * P1 / kap;
                            dS -= kap * (clampd(y1, 0.0, 1.0) - clampd(w, 0.0, 1.0));
                            yg[o] = y1;
                        }
                    }
                    r[o] += dS;
                }
            }
            k[u] += k1 - k0;
        }
    }
    free(r);
    return 0;
}
