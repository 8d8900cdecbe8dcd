/*
 * oracle_sweep.c — FP64 CPU restatement of the exact line-conditional operation sweep
 * (TEST INFRASTRUCTURE).
 *
 * Follows src/local.jl: the move workspace (MoveWS :451-535) with its entries, touch sets and
 * primary breakpoints (prep_node! :1579, prep_spair! :1708-1726, finalize! :656-757,
 * link_soc! :630-641), the recorded-term evaluation (eval_move! mode 1 / eval_cached!
 * :789-981) of a two-step block move evaluated later step first (mv_eval! :1219-1229), the
 * exact line sampler (walk! :1294-1443, piece_roots! :1267-1291, sample_line! :1501-1560,
 * logseg :1099, trexp_offset :1106) and the commit (commit_block! / apply_step! :2495-2527).
 * The synthetic system has hard line limits, no losses, links, vintages or commitment, and one
 * node per balance cell.
 *
 * Schedule (the rounds of the device port, device/src/rounds.jl; the HIP library runs the same):
 * gas atoms at every (n, τ); potential atoms colour by colour of a greedy distance-2 colouring;
 * SoC atoms and storage–gas pairs per storage carrier and time-parity class. Uniforms: Philox
 * counter (replica, n·T + τ, round id, sweep counter). Walk points are capped at PMAX per side
 * (t = 0 beyond, counted as a guard), as on the device.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "esm_oracle.h"

#define PMAX 4096
#define MAXE 8      /* entries per step   */
#define MAXL 16     /* lines per step     */
#define MAXC 24     /* balance cells      */
#define MAXSEC 64
#define MAXBP 64

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double u01(uint32_t hi, uint32_t lo) {
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 11) * 0x1.0p-53;
}

typedef struct {
    const esm_model *m;
    int R;
    double *th, *yg, *ye[2], *r, *F;     /* state and caches (FP64) */
    double kc[ESM_NK * 256 + 64];        /* clamped capacities */
    int *adj_ptr, *adj_line;              /* incident lines per node, line order */
} Sys;

#define AT(a, n, t) (a)[((size_t)(n) * S->m->T + (t)) * S->R + j]

/* one step of a move: MoveWS restricted to the fields of the synthetic system */
typedef struct {
    int tau;
    /* entries (f, i, dy): f 0 θ, 1 gas, 2 battery SoC, 3 H2 SoC */
    int ne, ef[MAXE], ei[MAXE];
    double edy[MAXE];
    /* touch sets */
    int nl, line[MAXL];
    double dfl[MAXL];
    int nn, node[MAXC];
    int nstu, stu_s[2], stu_i[2];
    double bps[MAXBP];
    int nbp;
    /* recorded terms (eval_move! mode 1) */
    int ng, g_node[MAXE];
    double g_y0[MAXE], g_k[MAXE], g_mc[MAXE], g_a[MAXE], g_c0[MAXE], g_p0[MAXE];
    int tl_b0[MAXL], tl_b1[MAXL];
    double tl_f0[MAXL], tl_df[MAXL], tl_se[MAXL], tl_c0[MAXL];
    double st_ye[2], st_ye1[2], st_dye[2], st_dye1[2], st_k[2], st_w[2], st_c0[2], st_j0[2];
    double c_S0[MAXC], c_phi0[MAXC];
    /* evaluation scratch */
    double dr[MAXC];
    double sec[MAXSEC];
    int nsec;
} Step;

static int node_pos(const Step *w, int n) {
    for (int q = 0; q < w->nn; q++)
        if (w->node[q] == n) return q;
    return -1;
}
static void touch_node(Step *w, int n) {
    if (node_pos(w, n) < 0) w->node[w->nn++] = n;
}
static void add_entry(Step *w, int f, int i, double a) {
    if (a == 0.0) return;
    for (int q = 0; q < w->ne; q++)
        if (w->ef[q] == f && w->ei[q] == i) {
            w->edy[q] += a;
            return;
        }
    w->ef[w->ne] = f;
    w->ei[w->ne] = i;
    w->edy[w->ne] = a;
    w->ne++;
}

static double phibal(const esm_model *m, double x) { return x > 0 ? m->c_shed * x : -m->c_dump * x; }

static void gen_cost(double mc, double A, double k, double w, double dt, double cv, double y, double *c, double *p) {
    double yc = clampd(y, 0.0, 1.0);
    *p = A * k * yc;
    *c = dt * mc * *p + cv * dt * w * fabs(y - yc);
}

/* soc_cost, src/energy.jl:448-463 (storage carrier s, unit at node i) */
static double soc_cost(const esm_model *m, int s, int i, double k, double w, double dt, double cv, double ye,
                       double ye1, double *dout, double *inj) {
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
    *dout = d;
    *inj = k * (pc - mcp);
    return c;
}

/* finalize! (src/local.jl:656-757) for the plain model: primary breakpoints, touch sets */
static void finalize(Sys *S, Step *w, int j) {
    const esm_model *m = S->m;
    int tau = w->tau;
    for (int q = 0; q < w->ne; q++) {
        double a = w->edy[q];
        if (fabs(a) < 1e-14) {
            w->edy[q] = 0.0;
            a = 0.0;
        }
        if (a == 0.0) continue;
        int f = w->ef[q], i = w->ei[q];
        if (f == 0) {            /* ROLE_THETA: its lines */
            for (int p = S->adj_ptr[i]; p < S->adj_ptr[i + 1]; p++) {
                int e = S->adj_line[p], seen = 0;
                for (int l = 0; l < w->nl; l++) seen |= w->line[l] == e;
                if (!seen) w->line[w->nl++] = e;
            }
            continue;
        }
        double y0 = f == 1 ? AT(S->yg, i, tau) : AT(S->ye[f - 2], i, tau);
        w->bps[w->nbp++] = (0.0 - y0) / a;
        w->bps[w->nbp++] = (1.0 - y0) / a;
        if (f == 1) {
            touch_node(w, i);
        } else {
            int s = f - 2, seen = 0;
            for (int q2 = 0; q2 < w->nstu; q2++) seen |= (w->stu_s[q2] == s && w->stu_i[q2] == i);
            if (!seen) {
                w->stu_s[w->nstu] = s;
                w->stu_i[w->nstu] = i;
                w->nstu++;
            }
            touch_node(w, i);
        }
    }
    /* lines: flow change per unit t; drop lines whose flow does not move */
    int keep = 0;
    for (int l = 0; l < w->nl; l++) {
        int e = w->line[l], a = m->line_bus0[e], b = m->line_bus1[e];
        double da = 0, db = 0;
        for (int q = 0; q < w->ne; q++)
            if (w->ef[q] == 0) {
                if (w->ei[q] == a) da += w->edy[q];
                if (w->ei[q] == b) db += w->edy[q];
            }
        double df = m->line_b[e] * (da - db);
        if (fabs(df) <= 1e-14) continue;
        w->line[keep] = e;
        w->dfl[keep] = df;
        keep++;
        touch_node(w, a);
        touch_node(w, b);
        double f0 = AT(S->F, e, tau), se = m->line_s[e];
        w->bps[w->nbp++] = (1.0 * se - f0) / df;   /* U = [0, 1]: kinks at ±s */
        w->bps[w->nbp++] = (-1.0 * se - f0) / df;
    }
    w->nl = keep;
}

/* link_soc! (src/local.jl:630-641): the storage term of wa sees wb's state-of-charge entries */
static void link_soc(Step *wa, const Step *wb) {
    for (int q = 0; q < wb->ne; q++) {
        int f = wb->ef[q];
        if (f < 2 || wb->edy[q] == 0.0) continue;
        int s = f - 2, i = wb->ei[q], seen = 0;
        for (int q2 = 0; q2 < wa->nstu; q2++) seen |= (wa->stu_s[q2] == s && wa->stu_i[q2] == i);
        if (!seen) {
            wa->stu_s[wa->nstu] = s;
            wa->stu_i[wa->nstu] = i;
            wa->nstu++;
        }
        touch_node(wa, i);
    }
}

static double entry_dy(const Step *w, int f, int i) {
    if (!w) return 0.0;
    for (int q = 0; q < w->ne; q++)
        if (w->ef[q] == f && w->ei[q] == i) return w->edy[q];
    return 0.0;
}

/* eval_move! mode 1 at t = 0: record the terms (EntTerm, LineTerm, StuTerm, CellTerm) */
static void record(Sys *S, Step *w, const Step *next, int j) {
    const esm_model *m = S->m;
    int tau = w->tau, T = m->T, N = m->N;
    double dt = m->dt, cv = m->c_viol, k0 = m->kappa0;
    w->ng = 0;
    for (int q = 0; q < w->ne; q++) {
        if (w->ef[q] != 1 || w->edy[q] == 0.0) continue;
        int i = w->ei[q], g = w->ng++;
        double k = S->kc[ESM_KGAS * N + i];
        w->g_node[g] = i;
        w->g_y0[g] = AT(S->yg, i, tau);
        w->g_k[g] = k;
        w->g_mc[g] = m->mc[ESM_KGAS * N + i];
        w->g_a[g] = w->edy[q];
        gen_cost(w->g_mc[g], 1.0, k, k + k0, dt, cv, w->g_y0[g], &w->g_c0[g], &w->g_p0[g]);
    }
    for (int l = 0; l < w->nl; l++) {
        int e = w->line[l];
        w->tl_b0[l] = m->line_bus0[e];
        w->tl_b1[l] = m->line_bus1[e];
        w->tl_f0[l] = AT(S->F, e, tau);
        w->tl_df[l] = w->dfl[l];
        w->tl_se[l] = m->line_s[e];
        w->tl_c0[l] = fmax(fabs(w->tl_f0[l]) - w->tl_se[l], 0.0);
    }
    int tau1 = (tau + 1) % T;
    for (int q = 0; q < w->nstu; q++) {
        int s = w->stu_s[q], i = w->stu_i[q];
        double k = S->kc[(ESM_KSTOR + s) * N + i], d;
        w->st_ye[q] = AT(S->ye[s], i, tau);
        w->st_ye1[q] = AT(S->ye[s], i, tau1);
        w->st_dye[q] = entry_dy(w, 2 + s, i);
        w->st_dye1[q] = entry_dy(next, 2 + s, i);
        w->st_k[q] = k;
        w->st_w[q] = k + k0;
        w->st_c0[q] = soc_cost(m, s, i, k, k + k0, dt, cv, w->st_ye[q], w->st_ye1[q], &d, &w->st_j0[q]);
    }
    for (int c = 0; c < w->nn; c++) {
        w->c_S0[c] = AT(S->r, w->node[c], tau);
        w->c_phi0[c] = phibal(m, w->c_S0[c] / 1.0);
    }
}

/* eval_cached! (src/local.jl:926-981) */
static double eval_step(Sys *S, Step *w, double t) {
    const esm_model *m = S->m;
    double dt = m->dt, cv = m->c_viol, k0 = m->kappa0;
    double D = 0.0;
    w->nsec = 0;
    for (int c = 0; c < w->nn; c++) w->dr[c] = 0.0;
    for (int g = 0; g < w->ng; g++) {
        double y1 = w->g_y0[g] + (t * w->g_a[g] + 0.0), c1, p1;
        gen_cost(w->g_mc[g], 1.0, w->g_k[g], w->g_k[g] + k0, dt, cv, y1, &c1, &p1);
        D += c1 - w->g_c0[g];
        w->dr[node_pos(w, w->g_node[g])] -= p1 - w->g_p0[g];
    }
    for (int l = 0; l < w->nl; l++) {
        double f1 = w->tl_f0[l] + t * w->tl_df[l];
        D += cv * dt * (fmax(fabs(f1) - w->tl_se[l], 0.0) - w->tl_c0[l]);
        w->dr[node_pos(w, w->tl_b0[l])] += f1 - w->tl_f0[l];
        w->dr[node_pos(w, w->tl_b1[l])] -= f1 - w->tl_f0[l];
    }
    for (int q = 0; q < w->nstu; q++) {
        int s = w->stu_s[q], i = w->stu_i[q], u = s * m->N + i;
        double d1, j1;
        double c1 = soc_cost(m, s, i, w->st_k[q], w->st_w[q], dt, cv, w->st_ye[q] + t * w->st_dye[q],
                             w->st_ye1[q] + t * w->st_dye1[q], &d1, &j1);
        D += c1 - w->st_c0[q];
        w->dr[node_pos(w, i)] += j1 - w->st_j0[q];
        w->sec[w->nsec++] = d1;
        w->sec[w->nsec++] = m->stor_eta_in[u] > 0 ? d1 - m->stor_eta_in[u] : 1.0;
        w->sec[w->nsec++] = -d1 * m->stor_eta_out[u] - 1.0;
    }
    for (int c = 0; c < w->nn; c++) {
        double S1 = w->c_S0[c] + (0.0 + w->dr[c]);
        D += dt * 1 * (phibal(m, S1 / 1) - w->c_phi0[c]);
        w->sec[w->nsec++] = S1;
    }
    return D;
}

typedef struct {
    Sys *S;
    Step *wa, *ws;     /* steps τm and τ */
    int j;
    double sec[2 * MAXSEC];
    int nsec;
    int gmask[2 * MAXSEC];
    int stor;
} Block;

static __thread long g_nev;   /* ΔC evaluations of the current thread (statistics) */

/* mv_eval! of the block move: later step first, secondaries appended in that order */
static double mv_eval(Block *b, double t) {
    double D = 0.0;
    g_nev++;
    b->nsec = 0;
    Step *order[2] = {b->ws, b->wa};
    for (int q = 0; q < 2; q++) {
        D += eval_step(b->S, order[q], t);
        for (int k = 0; k < order[q]->nsec; k++) b->sec[b->nsec++] = order[q]->sec[k];
    }
    return D;
}

/* storage_mask! (src/local.jl:1244-1265): the 3 storage secondaries of every storage term */
static void storage_mask(Block *b) {
    int p = 0;
    Step *order[2] = {b->ws, b->wa};
    for (int q = 0; q < 2; q++) {
        int nst = 3 * order[q]->nstu, tot = nst + order[q]->nn;
        for (int k = 0; k < tot; k++) b->gmask[p++] = k < nst;
    }
    b->stor = b->ws->nstu + b->wa->nstu > 0;
}

static void insort(double *v, int n, double s) {
    for (int i = 1; i < n; i++) {
        double x = v[i];
        int k = i - 1;
        while (k >= 0 && (s > 0 ? x < v[k] : x > v[k])) {
            v[k + 1] = v[k];
            k--;
        }
        v[k + 1] = x;
    }
}

typedef struct {
    double p[PMAX], c[PMAX];
    int n, full;
} Pts;

static void push_pt(Pts *P, double x, double c) {
    if (P->n >= PMAX) {
        P->full = 1;
        return;
    }
    P->p[P->n] = x;
    P->c[P->n] = c;
    P->n++;
}

static void piece_roots(Block *b, double beta, double *cmin, const double *A, const double *B, int ns, double a0,
                        double b0, double sgn, Pts *P) {
    double sr[2 * MAXSEC];
    int n = 0;
    for (int k = 0; k < ns; k++) {
        if (b->gmask[k]) continue;
        double u = A[k], v = B[k];
        if ((u < 0) != (v < 0) && u != v) {
            double x = a0 + (b0 - a0) * u / (u - v);
            if (sgn * (x - a0) > 0 && sgn * (b0 - x) > 0) sr[n++] = x;
        }
    }
    if (n == 0) return;
    insort(sr, n, sgn);
    for (int q = 0; q < n; q++) {
        double cx = beta * mv_eval(b, sr[q]);
        push_pt(P, sr[q], cx);
        *cmin = fmin(*cmin, cx);
    }
}

/* walk! without requirement-tracking kinks: 1 proper, 0 improper, -1 point cap */
static int walk(Block *b, double beta, const double *bp, int nb, double sgn, double maxnats, const double *sec0,
                int ns, Pts *P) {
    double secprev[2 * MAXSEC], sec[2 * MAXSEC], secB[2 * MAXSEC], secE[2 * MAXSEC];
    memcpy(secprev, sec0, sizeof(double) * ns);
    double cmin = 0.0, tp = 0.0, delta = 0.0;
    int idx = 0, nx = 0;
    P->n = 0;
    P->full = 0;
    for (;;) {
        double t;
        if (idx < nb) {
            t = bp[idx++];
            if (t == tp) continue;
        } else {
            double cprev = P->n == 0 ? 0.0 : P->c[P->n - 1];
            double hp = 1e-6 * fmax(1.0, fabs(tp));
            double sl = (beta * mv_eval(b, tp + sgn * hp) - cprev) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, 1e3 * fmax(fabs(tp), 1e-3));
            else delta = delta == 0.0 ? fmax(fabs(tp), 1e-6) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;   /* the ray left the floating-point range: improper */
        }
        double c = beta * mv_eval(b, t);
        memcpy(sec, b->sec, sizeof(double) * ns);
        double roots[2 * MAXSEC], groots[2 * MAXSEC];
        int nr = 0, ng = 0;
        for (int k = 0; k < ns; k++) {
            double a = secprev[k], bb = sec[k];
            if ((a < 0) != (bb < 0) && a != bb) {
                double x = tp + (t - tp) * a / (a - bb);
                if (sgn * (x - tp) > 0 && sgn * (t - x) > 0) {
                    if (b->stor && b->gmask[k]) groots[ng++] = x;
                    else roots[nr++] = x;
                }
            }
        }
        if (b->stor && ng > 0) {
            memcpy(secE, sec, sizeof(double) * ns);
            insort(groots, ng, sgn);
            double a0 = tp;
            for (int q = 0; q < ng; q++) {
                double x = groots[q];
                double cx = beta * mv_eval(b, x);
                memcpy(secB, b->sec, sizeof(double) * ns);
                piece_roots(b, beta, &cmin, secprev, secB, ns, a0, x, sgn, P);
                push_pt(P, x, cx);
                cmin = fmin(cmin, cx);
                memcpy(secprev, secB, sizeof(double) * ns);
                a0 = x;
            }
            piece_roots(b, beta, &cmin, secprev, secE, ns, a0, t, sgn, P);
            memcpy(secprev, secE, sizeof(double) * ns);
        } else {
            memcpy(secprev, sec, sizeof(double) * ns);
            if (nr > 0) {
                insort(roots, nr, sgn);
                for (int q = 0; q < nr; q++) {
                    double cx = beta * mv_eval(b, roots[q]);
                    push_pt(P, roots[q], cx);
                    cmin = fmin(cmin, cx);
                }
            }
        }
        push_pt(P, t, c);
        if (P->full) return -1;
        cmin = fmin(cmin, c);
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}

static double logseg(double L, double x) {
    if (fabs(x) < 1e-10) return log(L) - x / 2;
    if (x > 0) return log(L) + log(-expm1(-x)) - log(x);
    return log(L) - x + log(-expm1(x)) - log(-x);
}

static double trexp_offset(double g, double L, double u) {
    double x = g * L;
    if (fabs(x) < 1e-10) return u * L;
    if (g > 0) return -log1p(-u * (-expm1(-x))) / g;
    return L + log1p(-u * (-expm1(x))) / (-g);
}

/* segment log-masses of the walked points (left side reversed, t = 0, right side), the segment
 * by inversion of the first uniform, the offset (trexp_offset) with the second (sample_line!,
 * src/local.jl:1530-1560); status 0 drawn, 2 a non-finite draw */
static double draw_points(const Pts *PLp, const Pts *PRp, double useg, double uoff, int *status) {
    const Pts *PL = PLp, *PR = PRp;
    *status = 0;
    int np = PL->n + 1 + PR->n;
    static __thread double pts[2 * PMAX + 1], cs[2 * PMAX + 1], lm[2 * PMAX];
    for (int q = 0; q < PL->n; q++) {
        pts[q] = PL->p[PL->n - 1 - q];
        cs[q] = PL->c[PL->n - 1 - q];
    }
    pts[PL->n] = 0.0;
    cs[PL->n] = 0.0;
    for (int q = 0; q < PR->n; q++) {
        pts[PL->n + 1 + q] = PR->p[q];
        cs[PL->n + 1 + q] = PR->c[q];
    }
    int nseg = np - 1;
    double mx = -INFINITY;
    for (int q = 0; q < nseg; q++) {
        double L = pts[q + 1] - pts[q];
        lm[q] = L > 0 ? -fmin(cs[q], cs[q + 1]) + logseg(L, fabs(cs[q + 1] - cs[q])) : -INFINITY;
        mx = fmax(mx, lm[q]);
    }
    double tot = 0.0;
    for (int q = 0; q < nseg; q++) tot += exp(lm[q] - mx);
    double u = useg * tot, acc = 0.0;
    int jj = nseg - 1;
    for (int q = 0; q < nseg; q++) {
        acc += exp(lm[q] - mx);
        if (u <= acc) {
            jj = q;
            break;
        }
    }
    double L = pts[jj + 1] - pts[jj];
    double g = (cs[jj + 1] - cs[jj]) / L;
    double t = (g >= 0 || fabs(g * L) < 1e-10) ? pts[jj] + trexp_offset(g, L, uoff) : pts[jj + 1] - trexp_offset(-g, L, uoff);
    if (!isfinite(t)) {   /* a non-finite draw: the move is not made */
        *status = 2;
        return 0.0;
    }
    return t;
}

/* sample_line! with the uniforms given: status 0 drawn, 1 improper, 2 point cap */
static double sample_line(Block *b, double beta, double maxnats, double useg, double uoff, int *status) {
    double right[2 * MAXBP], left[2 * MAXBP];
    int nr = 0, nl = 0;
    Step *bs[2] = {b->wa, b->ws};          /* bm.bps = wa.bps ++ ws.bps */
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < bs[q]->nbp; k++) {
            double x = bs[q]->bps[k];
            if (!isfinite(x)) continue;
            if (x > 0) right[nr++] = x;
            else if (x < 0) left[nl++] = x;
        }
    insort(right, nr, 1.0);
    insort(left, nl, -1.0);
    storage_mask(b);
    mv_eval(b, 0.0);
    int ns = b->nsec;
    double sec0[2 * MAXSEC];
    memcpy(sec0, b->sec, sizeof(double) * ns);
    static __thread Pts PR, PL;
    int w = walk(b, beta, right, nr, 1.0, maxnats, sec0, ns, &PR);
    if (w != 1) {
        *status = w == 0 ? 1 : 2;
        return 0.0;
    }
    w = walk(b, beta, left, nl, -1.0, maxnats, sec0, ns, &PL);
    if (w != 1) {
        *status = w == 0 ? 1 : 2;
        return 0.0;
    }
    return draw_points(&PL, &PR, useg, uoff, status);
}

/* commit_block! / apply_step! *//* commit_block! / apply_step! */
static void commit(Block *b, double t) {
    Sys *S = b->S;
    int j = b->j;
    if (t == 0.0) return;
    mv_eval(b, t);
    Step *order[2] = {b->ws, b->wa};
    for (int q = 0; q < 2; q++) {
        Step *w = order[q];
        int tau = w->tau;
        for (int k = 0; k < w->ne; k++) {
            int f = w->ef[k], i = w->ei[k];
            double *arr = f == 0 ? S->th : f == 1 ? S->yg : S->ye[f - 2];
            AT(arr, i, tau) += t * w->edy[k];
        }
        for (int l = 0; l < w->nl; l++) AT(S->F, w->line[l], tau) += t * w->dfl[l];
        for (int c = 0; c < w->nn; c++) AT(S->r, w->node[c], tau) += w->dr[c];
    }
}

static void clear_step(Step *w, int tau) {
    w->tau = tau;
    w->ne = w->nl = w->nn = w->nstu = w->nbp = w->ng = w->nsec = 0;
}

static void stat_add(int64_t *st, int status, double t, long nev) {
    st[4] += nev;
    st[0]++;
    if (t != 0.0) st[1]++;
    if (status == 1) st[2]++;
    if (status == 2) st[3]++;
}

/* one move of family fam at (n, τ) of replica j with the uniforms of its line draw; s = storage
 * carrier. Moves that are not valid (a pair below κmin) are not made and not counted. */
static void do_move_u(Sys *S, int j, int fam, int s, int n, int tau, double useg, double uoff, double maxnats,
                      int64_t *stats) {
    const esm_model *m = S->m;
    int T = m->T, N = m->N;
    int taum = (tau - 1 + T) % T;
    Step wa, ws;
    clear_step(&wa, taum);
    clear_step(&ws, tau);
    if (fam == 0) {
        add_entry(&ws, 1, n, 1.0);
    } else if (fam == 1) {
        add_entry(&ws, 0, n, 1.0);
    } else if (fam == 2) {
        add_entry(&ws, 2 + s, n, 1.0);
    } else {
        /* prep_spair!: κs = k h/Δt; gas at τm (+κs/κ1) and at τ (−κs/κ2) */
        double ks = S->kc[(ESM_KSTOR + s) * N + n] * m->stor_hours[s * N + n] / m->dt;
        if (ks < 1.0) return;
        double k1 = 1.0 * S->kc[ESM_KGAS * N + n], k2 = k1;
        if (fabs(k1) < 1.0 || fabs(k2) < 1.0) return;
        add_entry(&wa, 1, n, ks / k1);
        add_entry(&ws, 2 + s, n, 1.0);
        add_entry(&ws, 1, n, -ks / k2);
    }
    finalize(S, &ws, j);
    if (ws.ne == 0) return;
    finalize(S, &wa, j);
    link_soc(&wa, &ws);
    record(S, &ws, NULL, j);
    record(S, &wa, &ws, j);
    Block b = {S, &wa, &ws, j};
    int status;
    g_nev = 0;
    double t = sample_line(&b, m->beta_v, maxnats, useg, uoff, &status);
    long nev = g_nev;
    commit(&b, t);
    stat_add(stats, status, t, nev);
}

/* an atom move: uniforms from counter (replica, n·T + τ, round, sweep counter) */
static void do_move(Sys *S, int j, int fam, int s, int n, int tau, const uint32_t key[2], uint32_t round,
                    uint32_t ctr, double maxnats, int64_t *stats) {
    uint32_t c4[4] = {(uint32_t)j, (uint32_t)(n * S->m->T + tau), round, ctr}, w[4];
    orc_philox(key, c4, w);
    do_move_u(S, j, fam, s, n, tau, u01(w[0], w[1]), u01(w[2], w[3]), maxnats, stats);
}


/* ============================================================================================
 * The random families (local_sweep! :1952-2044, plain_families :1912-1915) and the templates of
 * the table-driven ones.
 *
 * Templates (network only): xloc / xlocc after xloc_directions (:122-191) with one-hop partners
 * and ball radius 0, xpair after prep_reroute! / reroute! (:44-110, :592-601) grounded at the
 * first node of the component, cells of every partition level (E.parts, src/energy.jl:292;
 * LocalModel :265-283). A direction is finished as finalize! does (:656-757): angle entries
 * below 1e-14 are zero, lines of the angle entries with |Δf| > 1e-14 are kept, the touched nodes
 * are those of the gas entries, the kinks and the kept lines; lines and nodes are listed in
 * increasing order (the order of the evaluation sums).
 * ============================================================================================ */

typedef struct {
    int *a;
    int n, cap;
} IVec;
typedef struct {
    double *a;
    int n, cap;
} DVec;
static void ipush(IVec *v, int x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->a = (int *)realloc(v->a, sizeof(int) * v->cap);
    }
    v->a[v->n++] = x;
}
static void dpush(DVec *v, double x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->a = (double *)realloc(v->a, sizeof(double) * v->cap);
    }
    v->a[v->n++] = x;
}

typedef struct {
    IVec th_off, th_node, en_off, en_node, en_div, kk_off, kk_node, ln_off, ln_line, nd_off, nd_node, nd_ent, nd_kk,
        nd_l, nl_pos, nl_sgn, xp_off, lev_off;
    DVec th_val, en_c, kk_g, ln_df, qcum;
    int xp0, cl0, ntpl, nlev;
    /* storage pairs (sx_templates, storage_cross.jl:29-53): the slots sx_poff[n] … sx_poff[n+1]−1
     * of node n, the first the node itself, then its partners within one hop in increasing order;
     * per slot the partner, the angle entries of a unit dipole re-routed on the ball of radius one
     * around the pair (or on the whole component when the ball covers it), the leak nodes with
     * their requirement change per unit, the lines at the angle nodes with the angle change at
     * each end (unscaled: the move scales and filters them as `finalize!` does), the touched nodes
     * (line ends, leak nodes, the pair) with their lines and leak, and the footprint (sxfoot) */
    IVec sx_poff, sx_part, sx_th_off, sx_th_node, sx_kk_off, sx_kk_node, sx_ln_off, sx_ln_line, sx_nd_off,
        sx_nd_node, sx_nd_kk, sx_nd_l, sx_nl_pos, sx_nl_sgn, sx_ft_off, sx_ft_node;
    DVec sx_th_val, sx_kk_g, sx_ln_ta, sx_ln_tb;
} Tables;

static void tables_free(Tables *tb) {
    IVec *iv[] = {&tb->th_off, &tb->th_node, &tb->en_off, &tb->en_node, &tb->en_div, &tb->kk_off, &tb->kk_node,
                  &tb->ln_off, &tb->ln_line, &tb->nd_off, &tb->nd_node, &tb->nd_ent, &tb->nd_kk, &tb->nd_l,
                  &tb->nl_pos, &tb->nl_sgn, &tb->xp_off, &tb->lev_off, &tb->sx_poff, &tb->sx_part, &tb->sx_th_off,
                  &tb->sx_th_node, &tb->sx_kk_off, &tb->sx_kk_node, &tb->sx_ln_off, &tb->sx_ln_line, &tb->sx_nd_off,
                  &tb->sx_nd_node, &tb->sx_nd_kk, &tb->sx_nd_l, &tb->sx_nl_pos, &tb->sx_nl_sgn, &tb->sx_ft_off,
                  &tb->sx_ft_node};
    DVec *dv[] = {&tb->th_val, &tb->en_c, &tb->kk_g, &tb->ln_df, &tb->qcum, &tb->sx_th_val, &tb->sx_kk_g,
                  &tb->sx_ln_ta, &tb->sx_ln_tb};
    for (size_t q = 0; q < sizeof(iv) / sizeof(iv[0]); q++) free(iv[q]->a);
    for (size_t q = 0; q < sizeof(dv) / sizeof(dv[0]); q++) free(dv[q]->a);
}

/* a direction before finalize: angle entries, gas entries (slope c, /κ when div), kinks */
typedef struct {
    int nth, ne, nk;
    int th_node[256], en_node[256], en_div[256], kk_node[256];
    double th_val[256], en_c[256], kk_g[256];
} Dir;

typedef struct {
    const esm_model *m;
    int N, L;
    int *lptr, *lines;      /* lines at each node, line order */
    int *nptr, *nbr;        /* neighbours (one per line end, line order) */
    int *comp, ncomp;
    int *cptr, *cnodes;     /* nodes of each component, increasing */
    int *loc;               /* position among the non-ground nodes of the component (0: ground) */
    double **G;             /* grounded Laplacian factor per component */
} Net;

/* A = G Gᵀ in place (row-major lower); 0 if not positive definite */
static int chol(double *A, int n) {
    for (int i = 0; i < n; i++) {
        for (int j = 0; j <= i; j++) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            if (i == j) {
                if (!(s > 0)) return 0;
                A[(size_t)i * n + i] = sqrt(s);
            } else {
                A[(size_t)i * n + j] = s / A[(size_t)j * n + j];
            }
        }
        for (int j = i + 1; j < n; j++) A[(size_t)i * n + j] = 0.0;
    }
    return 1;
}
static void chol_solve(const double *G, int n, double *b) {
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= G[(size_t)i * n + k] * b[k];
        b[i] = s / G[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int k = i + 1; k < n; k++) s -= G[(size_t)k * n + i] * b[k];
        b[i] = s / G[(size_t)i * n + i];
    }
}

static int cmp_int(const void *a, const void *b) { return (*(const int *)a > *(const int *)b) - (*(const int *)a < *(const int *)b); }

static void emit(Tables *tb, const Net *net, const Dir *d) {
    const esm_model *m = net->m;
    int N = net->N;
    double *dth = (double *)calloc(N, sizeof(double));
    int *lmark = (int *)calloc(net->L, sizeof(int)), *nmark = (int *)calloc(N, sizeof(int));
    for (int q = 0; q < d->nth; q++) {
        if (fabs(d->th_val[q]) < 1e-14) continue;
        dth[d->th_node[q]] = d->th_val[q];
        ipush(&tb->th_node, d->th_node[q]);
        dpush(&tb->th_val, d->th_val[q]);
        for (int p = net->lptr[d->th_node[q]]; p < net->lptr[d->th_node[q] + 1]; p++) lmark[net->lines[p]] = 1;
    }
    int nl = 0, *ll = (int *)malloc(sizeof(int) * (net->L + 1));
    double *lf = (double *)malloc(sizeof(double) * (net->L + 1));
    for (int e = 0; e < net->L; e++) {
        if (!lmark[e]) continue;
        double df = m->line_b[e] * (dth[m->line_bus0[e]] - dth[m->line_bus1[e]]);
        if (fabs(df) <= 1e-14) continue;
        ll[nl] = e;
        lf[nl] = df;
        nl++;
        nmark[m->line_bus0[e]] = nmark[m->line_bus1[e]] = 1;
    }
    for (int q = 0; q < d->ne; q++) nmark[d->en_node[q]] = 1;
    for (int q = 0; q < d->nk; q++) nmark[d->kk_node[q]] = 1;
    for (int q = 0; q < d->ne; q++) {
        ipush(&tb->en_node, d->en_node[q]);
        dpush(&tb->en_c, d->en_c[q]);
        ipush(&tb->en_div, d->en_div[q]);
    }
    for (int q = 0; q < d->nk; q++) {
        ipush(&tb->kk_node, d->kk_node[q]);
        dpush(&tb->kk_g, d->kk_g[q]);
    }
    for (int l = 0; l < nl; l++) {
        ipush(&tb->ln_line, ll[l]);
        dpush(&tb->ln_df, lf[l]);
    }
    for (int n = 0; n < N; n++) {
        if (!nmark[n]) continue;
        int ent = -1, kk = -1;
        for (int q = 0; q < d->ne; q++)
            if (d->en_node[q] == n) ent = q;
        for (int q = 0; q < d->nk; q++)
            if (d->kk_node[q] == n) kk = q;
        ipush(&tb->nd_node, n);
        ipush(&tb->nd_ent, ent);
        ipush(&tb->nd_kk, kk);
        for (int l = 0; l < nl; l++) {
            int e = ll[l];
            if (m->line_bus0[e] == n || m->line_bus1[e] == n) {
                ipush(&tb->nl_pos, l);
                ipush(&tb->nl_sgn, m->line_bus0[e] == n ? 1 : -1);
            }
        }
        ipush(&tb->nd_l, tb->nl_pos.n);
    }
    ipush(&tb->th_off, tb->th_node.n);
    ipush(&tb->en_off, tb->en_node.n);
    ipush(&tb->kk_off, tb->kk_node.n);
    ipush(&tb->ln_off, tb->ln_line.n);
    ipush(&tb->nd_off, tb->nd_node.n);
    tb->ntpl++;
    free(dth); free(lmark); free(nmark); free(ll); free(lf);
}

/* Δθ with L Δθ = e_n − e_m on the component of n, grounded at its first node (reroute!) */
static void reroute(const Net *net, int n, int mm, double *x) {
    int c = net->comp[n], k = net->cptr[c + 1] - net->cptr[c] - 1;
    double *rhs = (double *)calloc(k > 0 ? k : 1, sizeof(double));
    if (net->loc[n] > 0) rhs[net->loc[n] - 1] += 1.0;
    if (net->loc[mm] > 0) rhs[net->loc[mm] - 1] -= 1.0;
    if (k > 0) chol_solve(net->G[c], k, rhs);
    for (int i = 0; i < net->N; i++) x[i] = 0.0;
    for (int q = net->cptr[c]; q < net->cptr[c + 1]; q++) {
        int i = net->cnodes[q];
        x[i] = net->loc[i] > 0 ? rhs[net->loc[i] - 1] : 0.0;
    }
    free(rhs);
}

/* the network of the templates: incident lines and neighbours per node, components (nodes
 * increasing) and the grounded Laplacian factor of each; -2 if one is not positive definite */
static int net_build(const esm_model *m, Net *netp) {
    int N = m->N, L = m->L;
    Net net;
    memset(&net, 0, sizeof(net));
    net.m = m;
    net.N = N;
    net.L = L;
    net.lptr = (int *)calloc(N + 1, sizeof(int));
    net.lines = (int *)malloc(sizeof(int) * 2 * L);
    net.nptr = (int *)calloc(N + 1, sizeof(int));
    net.nbr = (int *)malloc(sizeof(int) * 2 * L);
    for (int e = 0; e < L; e++) {
        net.lptr[m->line_bus0[e] + 1]++;
        net.lptr[m->line_bus1[e] + 1]++;
    }
    for (int n = 0; n < N; n++) net.lptr[n + 1] += net.lptr[n];
    memcpy(net.nptr, net.lptr, sizeof(int) * (N + 1));
    {
        int *f1 = (int *)malloc(sizeof(int) * N), *f2 = (int *)malloc(sizeof(int) * N);
        for (int n = 0; n < N; n++) f1[n] = f2[n] = net.lptr[n];
        for (int e = 0; e < L; e++) {
            int a = m->line_bus0[e], b = m->line_bus1[e];
            net.lines[f1[a]++] = e;
            net.lines[f1[b]++] = e;
            net.nbr[f2[a]++] = b;
            net.nbr[f2[b]++] = a;
        }
        free(f1); free(f2);
    }
    /* components, nodes increasing */
    net.comp = (int *)malloc(sizeof(int) * N);
    for (int n = 0; n < N; n++) net.comp[n] = -1;
    int *stack = (int *)malloc(sizeof(int) * N), *order = (int *)malloc(sizeof(int) * N), no = 0;
    net.cptr = (int *)calloc(N + 1, sizeof(int));
    for (int s = 0; s < N; s++) {
        if (net.comp[s] >= 0) continue;
        int c = net.ncomp++, sp = 0, c0 = no;
        stack[sp++] = s;
        net.comp[s] = c;
        while (sp) {
            int i = stack[--sp];
            order[no++] = i;
            for (int q = net.nptr[i]; q < net.nptr[i + 1]; q++) {
                int jn = net.nbr[q];
                if (net.comp[jn] < 0) {
                    net.comp[jn] = c;
                    stack[sp++] = jn;
                }
            }
        }
        qsort(order + c0, no - c0, sizeof(int), cmp_int);
        net.cptr[c + 1] = no;
    }
    net.cnodes = order;
    net.loc = (int *)calloc(N, sizeof(int));
    net.G = (double **)calloc(net.ncomp, sizeof(double *));
    for (int c = 0; c < net.ncomp; c++) {
        int c0 = net.cptr[c], k = net.cptr[c + 1] - c0 - 1;
        for (int q = 1; q <= k; q++) net.loc[net.cnodes[c0 + q]] = q;
        net.loc[net.cnodes[c0]] = 0;
        if (k == 0) continue;
        double *A = (double *)calloc((size_t)k * k, sizeof(double));
        for (int e = 0; e < L; e++) {
            int a = m->line_bus0[e], b = m->line_bus1[e];
            if (net.comp[a] != c) continue;
            int ia = net.loc[a], ib = net.loc[b];
            double w = m->line_b[e];
            if (ia > 0) A[(size_t)(ia - 1) * k + (ia - 1)] += w;
            if (ib > 0) A[(size_t)(ib - 1) * k + (ib - 1)] += w;
            if (ia > 0 && ib > 0) {
                A[(size_t)(ia - 1) * k + (ib - 1)] -= w;
                A[(size_t)(ib - 1) * k + (ia - 1)] -= w;
            }
        }
        if (!chol(A, k)) return -2;
        net.G[c] = A;
    }
    free(stack);
    *netp = net;
    return 0;
}

static void net_free(Net *net) {
    for (int c = 0; c < net->ncomp; c++) free(net->G[c]);
    free(net->G); free(net->lptr); free(net->lines); free(net->nptr); free(net->nbr); free(net->comp);
    free(net->cnodes); free(net->cptr); free(net->loc);
}

/* one storage-pair slot (pair n, mm): angle entries thn/thv, leak nodes lkn/lkg (sorted by node) */
static void sx_emit(Tables *tb, const Net *net, int n, int mm, const int *thn, const double *thv, int nth,
                    const int *lkn, const double *lkg, int nlk) {
    const esm_model *m = net->m;
    int N = net->N, L = net->L;
    double *dth = (double *)calloc(N, sizeof(double));
    int *lmark = (int *)calloc(L > 0 ? L : 1, sizeof(int)), *nmark = (int *)calloc(N, sizeof(int)),
        *fmark = (int *)calloc(N, sizeof(int));
    for (int q = 0; q < nth; q++) {
        int i = thn[q];
        ipush(&tb->sx_th_node, i);
        dpush(&tb->sx_th_val, thv[q]);
        dth[i] = thv[q];
        fmark[i] = 1;
        for (int p2 = net->lptr[i]; p2 < net->lptr[i + 1]; p2++) {
            int e = net->lines[p2];
            lmark[e] = 1;
            fmark[m->line_bus0[e]] = fmark[m->line_bus1[e]] = 1;
        }
    }
    int *ll = (int *)malloc(sizeof(int) * (L + 1)), nl = 0;
    for (int e = 0; e < L; e++) {
        if (!lmark[e]) continue;
        double ta = dth[m->line_bus0[e]], tb2 = dth[m->line_bus1[e]];
        if (ta == tb2) continue;                     /* no flow change at any scale */
        ipush(&tb->sx_ln_line, e);
        dpush(&tb->sx_ln_ta, ta);
        dpush(&tb->sx_ln_tb, tb2);
        ll[nl++] = e;
        nmark[m->line_bus0[e]] = nmark[m->line_bus1[e]] = 1;
    }
    for (int q = 0; q < nlk; q++) {
        ipush(&tb->sx_kk_node, lkn[q]);
        dpush(&tb->sx_kk_g, lkg[q]);
        nmark[lkn[q]] = fmark[lkn[q]] = 1;
    }
    nmark[n] = nmark[mm] = fmark[n] = fmark[mm] = 1;
    for (int i = 0; i < N; i++) {
        if (!nmark[i]) continue;
        int kk = -1;
        for (int q = 0; q < nlk; q++)
            if (lkn[q] == i) kk = q;
        ipush(&tb->sx_nd_node, i);
        ipush(&tb->sx_nd_kk, kk);
        for (int l = 0; l < nl; l++) {
            int e = ll[l];
            if (m->line_bus0[e] == i || m->line_bus1[e] == i) {
                ipush(&tb->sx_nl_pos, l);
                ipush(&tb->sx_nl_sgn, m->line_bus0[e] == i ? 1 : -1);
            }
        }
        ipush(&tb->sx_nd_l, tb->sx_nl_pos.n);
    }
    for (int i = 0; i < N; i++)
        if (fmark[i]) ipush(&tb->sx_ft_node, i);
    ipush(&tb->sx_part, mm);
    ipush(&tb->sx_th_off, tb->sx_th_node.n);
    ipush(&tb->sx_kk_off, tb->sx_kk_node.n);
    ipush(&tb->sx_ln_off, tb->sx_ln_line.n);
    ipush(&tb->sx_nd_off, tb->sx_nd_node.n);
    ipush(&tb->sx_ft_off, tb->sx_ft_node.n);
    free(dth); free(lmark); free(nmark); free(fmark); free(ll);
}

/* sx_templates (storage_cross.jl:29-53) with xloc_directions (src/local.jl:122-191) at partner
 * hops 1 and ball radius 1: per node its own slot, then per partner m (neighbours, increasing)
 * the ball B = ball(n) ∪ ball(m), L_BB Δθ = e_n − e_m with Δθ = 0 outside B (entries ≥ 1e-15
 * kept), the leak −Σ b Δθ_i into the first layer outside B (accumulated in B and line order,
 * sorted by node); the exact re-routing when B holds the whole component (no leak) */
static int build_sx(const Net *net, Tables *tb) {
    const esm_model *m = net->m;
    int N = net->N;
    int *inb = (int *)calloc(N, sizeof(int)), *B = (int *)malloc(sizeof(int) * N), *pos = (int *)malloc(sizeof(int) * N);
    int *thn = (int *)malloc(sizeof(int) * N), *lkn = (int *)malloc(sizeof(int) * N);
    double *thv = (double *)malloc(sizeof(double) * N), *lkg = (double *)malloc(sizeof(double) * N);
    double *x = (double *)malloc(sizeof(double) * N);
    ipush(&tb->sx_poff, 0);
    ipush(&tb->sx_th_off, 0);
    ipush(&tb->sx_kk_off, 0);
    ipush(&tb->sx_ln_off, 0);
    ipush(&tb->sx_nd_off, 0);
    ipush(&tb->sx_nd_l, 0);
    ipush(&tb->sx_ft_off, 0);
    int rc = 0;
    for (int n = 0; n < N && !rc; n++) {
        sx_emit(tb, net, n, n, NULL, NULL, 0, NULL, NULL, 0);
        int np = 0, part[256];
        for (int q = net->nptr[n]; q < net->nptr[n + 1]; q++) {
            int mm = net->nbr[q], seen = mm == n;
            for (int r = 0; r < np; r++) seen |= part[r] == mm;
            if (!seen) part[np++] = mm;
        }
        qsort(part, np, sizeof(int), cmp_int);
        for (int pi = 0; pi < np && !rc; pi++) {
            int mm = part[pi], nb = 0, nth = 0, nlk = 0;
            for (int i = 0; i < N; i++) inb[i] = 0;
            int ends[2] = {n, mm};
            for (int e2 = 0; e2 < 2; e2++) {
                int c = ends[e2];
                inb[c] = 1;
                for (int q = net->nptr[c]; q < net->nptr[c + 1]; q++) inb[net->nbr[q]] = 1;
            }
            for (int i = 0; i < N; i++) {
                pos[i] = -1;
                if (inb[i]) { pos[i] = nb; B[nb++] = i; }
            }
            int comp = net->comp[n], csize = net->cptr[comp + 1] - net->cptr[comp];
            if (nb >= csize) {
                reroute(net, n, mm, x);
                for (int q = net->cptr[comp]; q < net->cptr[comp + 1]; q++) {
                    int i = net->cnodes[q];
                    if (fabs(x[i]) >= 1e-15) { thn[nth] = i; thv[nth++] = x[i]; }
                }
            } else {
                double *A = (double *)calloc((size_t)nb * nb, sizeof(double)), *th = (double *)calloc(nb, sizeof(double));
                for (int q = 0; q < nb; q++)
                    for (int p2 = net->lptr[B[q]]; p2 < net->lptr[B[q] + 1]; p2++) {
                        int e = net->lines[p2], a = m->line_bus0[e], b = m->line_bus1[e];
                        int o = a == B[q] ? b : a;
                        A[(size_t)q * nb + q] += m->line_b[e];
                        if (pos[o] >= 0) A[(size_t)q * nb + pos[o]] -= m->line_b[e];
                    }
                if (!chol(A, nb)) rc = -2;
                if (!rc) {
                    th[pos[n]] += 1.0;
                    th[pos[mm]] -= 1.0;
                    chol_solve(A, nb, th);
                    for (int q = 0; q < nb; q++)
                        if (fabs(th[q]) >= 1e-15) { thn[nth] = B[q]; thv[nth++] = th[q]; }
                    for (int q = 0; q < nb; q++)
                        for (int p2 = net->lptr[B[q]]; p2 < net->lptr[B[q] + 1]; p2++) {
                            int e = net->lines[p2], a = m->line_bus0[e], b = m->line_bus1[e];
                            int o = a == B[q] ? b : a;
                            if (pos[o] >= 0) continue;
                            int jl = 0;
                            while (jl < nlk && lkn[jl] != o) jl++;
                            if (jl == nlk) { lkn[nlk] = o; lkg[nlk] = 0.0; nlk++; }
                            lkg[jl] -= m->line_b[e] * th[q];
                        }
                    for (int a = 1; a < nlk; a++)      /* insertion sort by node */
                        for (int b = a; b > 0 && lkn[b - 1] > lkn[b]; b--) {
                            int tn = lkn[b]; lkn[b] = lkn[b - 1]; lkn[b - 1] = tn;
                            double tg = lkg[b]; lkg[b] = lkg[b - 1]; lkg[b - 1] = tg;
                        }
                }
                free(A); free(th);
            }
            if (!rc) sx_emit(tb, net, n, mm, thn, thv, nth, lkn, lkg, nlk);
        }
        ipush(&tb->sx_poff, tb->sx_part.n);
    }
    free(inb); free(B); free(pos); free(thn); free(lkn); free(thv); free(lkg); free(x);
    return rc;
}

static int build_tables(const esm_model *m, Tables *tb) {
    int N = m->N;
    memset(tb, 0, sizeof(*tb));
    if (N > 256) return -1;
    Net net;
    if (net_build(m, &net)) return -2;
    ipush(&tb->th_off, 0);
    ipush(&tb->en_off, 0);
    ipush(&tb->kk_off, 0);
    ipush(&tb->ln_off, 0);
    ipush(&tb->nd_off, 0);
    ipush(&tb->nd_l, 0);
    ipush(&tb->xp_off, 0);
    double *x = (double *)malloc(sizeof(double) * N);
    static __thread Dir d;
    /* xloc / xlocc: partners = neighbours (one hop), ball B = {n, m} */
    for (int n = 0; n < N; n++) {
        int np = 0, part[256];
        for (int q = net.nptr[n]; q < net.nptr[n + 1]; q++) {
            int mm = net.nbr[q], seen = mm == n;
            for (int r = 0; r < np; r++) seen |= part[r] == mm;
            if (!seen) part[np++] = mm;
        }
        qsort(part, np, sizeof(int), cmp_int);
        for (int pi = 0; pi < np; pi++) {
            int mm = part[pi];
            d.nth = d.ne = d.nk = 0;
            d.en_node[d.ne] = n; d.en_c[d.ne] = 1.0; d.en_div[d.ne++] = 1;
            d.en_node[d.ne] = mm; d.en_c[d.ne] = -1.0; d.en_div[d.ne++] = 1;
            int B[2] = {n < mm ? n : mm, n < mm ? mm : n};
            int csize = net.cptr[net.comp[n] + 1] - net.cptr[net.comp[n]];
            if (2 >= csize) {
                reroute(&net, n, mm, x);
                for (int q = net.cptr[net.comp[n]]; q < net.cptr[net.comp[n] + 1]; q++) {
                    int i = net.cnodes[q];
                    if (fabs(x[i]) >= 1e-15) { d.th_node[d.nth] = i; d.th_val[d.nth++] = x[i]; }
                }
            } else {
                double A[4] = {0, 0, 0, 0};
                for (int q = 0; q < 2; q++)
                    for (int p2 = net.lptr[B[q]]; p2 < net.lptr[B[q] + 1]; p2++) {
                        int e = net.lines[p2], a = m->line_bus0[e], b = m->line_bus1[e];
                        int o = a == B[q] ? b : a;
                        A[q * 2 + q] += m->line_b[e];
                        int po = o == B[0] ? 0 : (o == B[1] ? 1 : -1);
                        if (po >= 0) A[q * 2 + po] -= m->line_b[e];
                    }
                if (!chol(A, 2)) return -2;
                double th[2] = {0, 0};
                th[n == B[0] ? 0 : 1] += 1.0;
                th[mm == B[0] ? 0 : 1] -= 1.0;
                chol_solve(A, 2, th);
                for (int q = 0; q < 2; q++)
                    if (fabs(th[q]) >= 1e-15) { d.th_node[d.nth] = B[q]; d.th_val[d.nth++] = th[q]; }
                /* leak −Σ b Δθ_i into the first layer outside B, sorted by node */
                int lkn[256], nlk = 0;
                double lkg[256];
                for (int q = 0; q < 2; q++)
                    for (int p2 = net.lptr[B[q]]; p2 < net.lptr[B[q] + 1]; p2++) {
                        int e = net.lines[p2], a = m->line_bus0[e], b = m->line_bus1[e];
                        int o = a == B[q] ? b : a;
                        if (o == B[0] || o == B[1]) continue;
                        int jl = 0;
                        while (jl < nlk && lkn[jl] != o) jl++;
                        if (jl == nlk) { lkn[nlk] = o; lkg[nlk] = 0.0; nlk++; }
                        lkg[jl] -= m->line_b[e] * th[q];
                    }
                for (int a = 1; a < nlk; a++)      /* insertion sort by node */
                    for (int b = a; b > 0 && lkn[b - 1] > lkn[b]; b--) {
                        int tn = lkn[b]; lkn[b] = lkn[b - 1]; lkn[b - 1] = tn;
                        double tg = lkg[b]; lkg[b] = lkg[b - 1]; lkg[b - 1] = tg;
                    }
                for (int q = 0; q < nlk; q++)
                    if (lkg[q] != 0.0) { d.kk_node[d.nk] = lkn[q]; d.kk_g[d.nk++] = lkg[q]; }
            }
            emit(tb, &net, &d);
        }
        ipush(&tb->xp_off, tb->ntpl);
    }
    /* xpair: every ordered pair of one component */
    tb->xp0 = tb->ntpl;
    for (int n = 0; n < N; n++)
        for (int mm = 0; mm < N; mm++) {
            d.nth = d.ne = d.nk = 0;
            if (mm != n && net.comp[mm] == net.comp[n]) {
                d.en_node[d.ne] = n; d.en_c[d.ne] = 1.0; d.en_div[d.ne++] = 1;
                d.en_node[d.ne] = mm; d.en_c[d.ne] = -1.0; d.en_div[d.ne++] = 1;
                reroute(&net, n, mm, x);
                for (int q = net.cptr[net.comp[n]]; q < net.cptr[net.comp[n] + 1]; q++) {
                    int i = net.cnodes[q];
                    if (fabs(x[i]) >= 1e-15) { d.th_node[d.nth] = i; d.th_val[d.nth++] = x[i]; }
                }
            }
            emit(tb, &net, &d);
        }
    /* cells: the whole system, the countries, the nodes (cells in order of first appearance) */
    tb->cl0 = tb->ntpl;
    ipush(&tb->lev_off, 0);
    int nlev = m->ncountry > 0 ? 3 : 2;
    double acc = 0.0;
    for (int lv = 0; lv < nlev; lv++) {
        int keys[256], nkey = 0;
        for (int n = 0; n < N; n++) {
            int lab = lv == 0 ? 0 : (lv == nlev - 1 ? n : m->country[n]), seen = 0;
            for (int q = 0; q < nkey; q++) seen |= keys[q] == lab;
            if (!seen) keys[nkey++] = lab;
        }
        for (int q = 0; q < nkey; q++) {
            d.nth = d.ne = d.nk = 0;
            for (int n = 0; n < N; n++) {
                int lab = lv == 0 ? 0 : (lv == nlev - 1 ? n : m->country[n]);
                if (lab == keys[q]) { d.en_node[d.ne] = n; d.en_c[d.ne] = 1.0; d.en_div[d.ne++] = 0; }
            }
            emit(tb, &net, &d);
        }
        ipush(&tb->lev_off, tb->lev_off.a[tb->lev_off.n - 1] + nkey);
        acc += (double)nkey;
        dpush(&tb->qcum, acc);
    }
    for (int q = 0; q < tb->qcum.n; q++) tb->qcum.a[q] /= acc;
    tb->nlev = nlev;
    free(x);
    int rc = build_sx(&net, tb);
    net_free(&net);
    return rc;
}

/* |S| of a pro-rata kink at node n (prorata!, src/local.jl:1784-1802): the node's dispatchable
 * units with |κ| ≥ κmin (κmin = 1), in this model its gas unit */
static int gas_units(const Sys *S, int n) { return fabs(S->kc[ESM_KGAS * S->m->N + n]) >= 1.0 ? 1 : 0; }

/* ---- a table-driven move of replica j at τ, evaluated on the fly from the caches ---- */
enum { SP_CELL = 0, SP_XLOC = 1, SP_XLOCC = 2, SP_XPAIR = 3 };

typedef struct {
    Sys *S;
    const Tables *tb;
    int j, tau, tid, comp;
    int e0, e1, q0, q1, l0, l1, n0, n1;
    double kmin;
} SMove;

static void sm_init(SMove *f, Sys *S, const Tables *tb, int j, int tau, int tid, int comp) {
    f->S = S; f->tb = tb; f->j = j; f->tau = tau; f->tid = tid; f->comp = comp; f->kmin = 1.0;
    f->e0 = tb->en_off.a[tid]; f->e1 = tb->en_off.a[tid + 1];
    f->q0 = tb->kk_off.a[tid]; f->q1 = tb->kk_off.a[tid + 1];
    f->l0 = tb->ln_off.a[tid]; f->l1 = tb->ln_off.a[tid + 1];
    f->n0 = tb->nd_off.a[tid]; f->n1 = tb->nd_off.a[tid + 1];
}
#define SMV(arr, n) (f->S->arr)[((size_t)(n) * f->S->m->T + f->tau) * f->S->R + f->j]
static double sm_kap(const SMove *f, int n) { return f->S->kc[ESM_KGAS * f->S->m->N + n]; }
static double sm_slope(const SMove *f, int e) {
    double c = f->tb->en_c.a[e];
    double a = f->tb->en_div.a[e] ? c / sm_kap(f, f->tb->en_node.a[e]) : c;
    return fabs(a) < 1e-14 ? 0.0 : a;
}
static void sm_gen(const SMove *f, int n, double y, double *c, double *p) {
    const esm_model *m = f->S->m;
    double k = sm_kap(f, n);
    gen_cost(m->mc[ESM_KGAS * m->N + n], 1.0, k, k + m->kappa0, m->dt, m->c_viol, y, c, p);
}
/* prorata! of the single unit of a leak node: 0 when no kink is made */
static int sm_kink(const SMove *f, int q, double *Q, double *g, double *ak, double *P, double *e, double *y0) {
    int n = f->tb->kk_node.a[q];
    double k = sm_kap(f, n);
    if (!(fabs(k) >= f->kmin)) return 0;
    double w = SMV(yg, n);
    *P = 0.0 + k * w;
    double Qs = 0.0 + k * clampd(w, 0.0, 1.0);
    if (fabs(*P) < 1e-6) return 0;
    *Q = SMV(r, n) + Qs;
    *e = *P - fmax(*Q, 0.0);
    *g = f->tb->kk_g.a[q];
    double sig = k * w / *P;
    *ak = sig == 0.0 ? 0.0 : sig / k;
    *y0 = w;
    return *g != 0.0;
}
static double sm_phi(double Q, double g, double t) { return fmax(Q + g * t, 0.0) - fmax(Q, 0.0); }

static double sm_node_S(const SMove *f, int p, double t) {
    const Tables *tb = f->tb;
    int n = tb->nd_node.a[p];
    double dr = 0.0;
    int ie = tb->nd_ent.a[p];
    if (ie >= 0) {
        double a = sm_slope(f, f->e0 + ie);
        if (a != 0.0) {
            double y0 = SMV(yg, n), c0, p0, c1, p1;
            sm_gen(f, n, y0, &c0, &p0);
            sm_gen(f, n, y0 + t * a, &c1, &p1);
            dr -= p1 - p0;
        }
    }
    int ik = tb->nd_kk.a[p];
    if (f->comp && ik >= 0) {
        double Q, g, ak, P, e, y0;
        if (sm_kink(f, f->q0 + ik, &Q, &g, &ak, &P, &e, &y0) && ak != 0.0) {
            double c0, p0, c1, p1;
            sm_gen(f, n, y0, &c0, &p0);
            sm_gen(f, n, y0 + ak * sm_phi(Q, g, t), &c1, &p1);
            dr -= p1 - p0;
        }
    }
    for (int q = tb->nd_l.a[p]; q < tb->nd_l.a[p + 1]; q++) {
        int pos = f->l0 + tb->nl_pos.a[q];
        double f0 = SMV(F, tb->ln_line.a[pos]);
        double dd = (f0 + t * tb->ln_df.a[pos]) - f0;
        dr += tb->nl_sgn.a[q] > 0 ? dd : -dd;
    }
    return SMV(r, n) + dr;
}

static double sm_eval(const SMove *f, double t) {
    const Tables *tb = f->tb;
    const esm_model *m = f->S->m;
    double D = 0.0;
    g_nev++;
    for (int e = f->e0; e < f->e1; e++) {
        double a = sm_slope(f, e);
        if (a == 0.0) continue;
        int n = tb->en_node.a[e];
        double y0 = SMV(yg, n), c0, p0, c1, p1;
        sm_gen(f, n, y0, &c0, &p0);
        sm_gen(f, n, y0 + t * a, &c1, &p1);
        D += c1 - c0;
    }
    if (f->comp)
        for (int q = f->q0; q < f->q1; q++) {
            double Q, g, ak, P, e, y0, c0, p0, c1, p1;
            if (!sm_kink(f, q, &Q, &g, &ak, &P, &e, &y0) || ak == 0.0) continue;
            int n = tb->kk_node.a[q];
            sm_gen(f, n, y0, &c0, &p0);
            sm_gen(f, n, y0 + ak * sm_phi(Q, g, t), &c1, &p1);
            D += c1 - c0;
        }
    for (int l = f->l0; l < f->l1; l++) {
        int e = tb->ln_line.a[l];
        double f0 = SMV(F, e), se = m->line_s[e];
        double f1 = f0 + t * tb->ln_df.a[l];
        D += m->c_viol * m->dt * (fmax(fabs(f1) - se, 0.0) - fmax(fabs(f0) - se, 0.0));
    }
    for (int p = f->n0; p < f->n1; p++) {
        double S0 = SMV(r, tb->nd_node.a[p]);
        double S1 = sm_node_S(f, p, t);
        D += m->dt * (phibal(m, S1) - phibal(m, S0));
    }
    return D;
}

static int sm_next_bp(const SMove *f, double tp, double sgn, double *t) {
    const Tables *tb = f->tb;
    const esm_model *m = f->S->m;
    int have = 0;
#define CAND(bv)                                                                   \
    do {                                                                           \
        double b_ = (bv);                                                          \
        if (isfinite(b_) && sgn * (b_ - tp) > 0 && (!have || sgn * (*t - b_) > 0)) { \
            *t = b_;                                                               \
            have = 1;                                                              \
        }                                                                          \
    } while (0)
    for (int e = f->e0; e < f->e1; e++) {
        double a = sm_slope(f, e);
        if (a == 0.0) continue;
        double y0 = SMV(yg, tb->en_node.a[e]);
        CAND((0.0 - y0) / a);
        CAND((1.0 - y0) / a);
    }
    if (f->comp)
        for (int q = f->q0; q < f->q1; q++) {
            double Q, g, ak, P, e, y0;
            if (!sm_kink(f, q, &Q, &g, &ak, &P, &e, &y0)) continue;
            CAND(-Q / g);
            if (ak != 0.0) {
                double sl = ak * g, off = ak * Q;
                CAND((0.0 - y0 - 0.0) / sl);
                CAND((1.0 - y0 - 0.0) / sl);
                CAND((0.0 - y0 - off) / sl);
                CAND((1.0 - y0 - off) / sl);
            }
        }
    for (int l = f->l0; l < f->l1; l++) {
        int e = tb->ln_line.a[l];
        double f0 = SMV(F, e), se = m->line_s[e], df = tb->ln_df.a[l];
        CAND((se - f0) / df);
        CAND((-se - f0) / df);
    }
#undef CAND
    return have;
}

static int sm_next_root(const SMove *f, double tp, double t, double x, double sgn, double *xr) {
    int have = 0;
    for (int p = f->n0; p < f->n1; p++) {
        double A = sm_node_S(f, p, tp), B = sm_node_S(f, p, t);
        if ((A < 0) != (B < 0) && A != B) {
            double rt = tp + (t - tp) * A / (A - B);
            if (sgn * (rt - tp) > 0 && sgn * (t - rt) > 0 && sgn * (rt - x) > 0 && (!have || sgn * (*xr - rt) > 0)) {
                *xr = rt;
                have = 1;
            }
        }
    }
    return have;
}

/* walk! of a table-driven move: 1 proper, 0 improper, -1 point cap */
static int sm_walk(const SMove *f, double beta, double sgn, double maxnats, Pts *P) {
    double cmin = 0.0, tp = 0.0, delta = 0.0;
    int nx = 0;
    P->n = 0;
    P->full = 0;
    for (;;) {
        double t;
        if (!sm_next_bp(f, tp, sgn, &t)) {
            double cprev = P->n == 0 ? 0.0 : P->c[P->n - 1];
            double hp = 1e-6 * fmax(1.0, fabs(tp));
            double sl = (beta * sm_eval(f, tp + sgn * hp) - cprev) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, 1e3 * fmax(fabs(tp), 1e-3));
            else delta = delta == 0.0 ? fmax(fabs(tp), 1e-6) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;
        }
        double c = beta * sm_eval(f, t);
        double x = tp, xr = 0.0;
        while (sm_next_root(f, tp, t, x, sgn, &xr)) {
            double cx = beta * sm_eval(f, xr);
            push_pt(P, xr, cx);
            cmin = fmin(cmin, cx);
            x = xr;
        }
        push_pt(P, t, c);
        if (P->full) return -1;
        cmin = fmin(cmin, c);
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}

static double sm_sample(const SMove *f, double beta, double maxnats, double useg, double uoff, int *status) {
    static __thread Pts PR, PL;
    int w = sm_walk(f, beta, 1.0, maxnats, &PR);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    w = sm_walk(f, beta, -1.0, maxnats, &PL);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    return draw_points(&PL, &PR, useg, uoff, status);
}

/* logjac (:1874): Σ (|S| − 1) log|P(t)/P| over the kinks made, |S| from the model */
static double sm_logjac(const SMove *f, double t, int *any) {
    double lj = 0.0;
    *any = 0;
    for (int q = f->q0; q < f->q1; q++) {
        double Q, g, ak, P, e, y0;
        if (!sm_kink(f, q, &Q, &g, &ak, &P, &e, &y0)) continue;
        *any = 1;
        double nm = (double)(gas_units(f->S, f->tb->kk_node.a[q]) - 1);
        lj += nm * log(fabs((fmax(Q + g * t, 0.0) + e) / P));
    }
    return lj;
}

/* commit_block! of a table-driven move: node by node, then flows, then angles */
static void sm_commit(const SMove *f, double t) {
    const Tables *tb = f->tb;
    if (t == 0.0) return;
    for (int p = f->n0; p < f->n1; p++) {
        int n = tb->nd_node.a[p];
        double S1 = sm_node_S(f, p, t), ynew = 0.0;
        int wy = 0, ie = tb->nd_ent.a[p], ik = tb->nd_kk.a[p];
        if (ie >= 0) {
            double a = sm_slope(f, f->e0 + ie);
            if (a != 0.0) { ynew = SMV(yg, n) + t * a; wy = 1; }
        }
        if (f->comp && ik >= 0) {
            double Q, g, ak, P, e, y0;
            if (sm_kink(f, f->q0 + ik, &Q, &g, &ak, &P, &e, &y0) && ak != 0.0) { ynew = y0 + ak * sm_phi(Q, g, t); wy = 1; }
        }
        SMV(r, n) = S1;
        if (wy) SMV(yg, n) = ynew;
    }
    for (int l = f->l0; l < f->l1; l++) {
        double *F1 = &SMV(F, tb->ln_line.a[l]);
        *F1 = *F1 + t * tb->ln_df.a[l];
    }
    for (int q = tb->th_off.a[f->tid]; q < tb->th_off.a[f->tid + 1]; q++) {
        double *th = &SMV(th, tb->th_node.a[q]);
        *th = *th + t * tb->th_val.a[q];
    }
}

/* whether the pair's units have |κ| ≥ κmin (cells have no such condition) */
static int sm_valid(const SMove *f, int kind) {
    if (kind == SP_CELL) return 1;
    if (f->e0 == f->e1) return 0;
    for (int e = f->e0; e < f->e1; e++)
        if (!(fabs(sm_kap(f, f->tb->en_node.a[e])) >= f->kmin)) return 0;
    return 1;
}

/* one table-driven move with given uniforms (umh: the Jacobian test's); returns t, counts it */
static double sm_move(Sys *S, const Tables *tb, int j, int tau, int tid, int kind, double useg, double uoff, double umh,
                      double maxnats, int64_t *stats) {
    SMove f;
    sm_init(&f, S, tb, j, tau, tid, kind == SP_XLOCC);
    if (!sm_valid(&f, kind)) return 0.0;
    int status;
    g_nev = 0;
    double t = sm_sample(&f, S->m->beta_v, maxnats, useg, uoff, &status);
    if (kind == SP_XLOCC && t != 0.0) {
        int any;
        double lj = sm_logjac(&f, t, &any);
        if (any && !(umh < exp(fmin(0.0, lj)))) {
            t = 0.0;
            status = 2;
        }
    }
    long nev = g_nev;
    sm_commit(&f, t);
    if (stats) stat_add(stats, status, t, nev);
    return t;
}

static uint32_t pick(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

/* the families of the draws: the plain ones, then the storage pair sxpair */
#define NW_ALL (ESM_NWEIGHT + 1)

/* family counts of every step: round(extra·xmul·N) draws over the cumulative weights; with sxpair
 * (weight px > 0, some storage unit taking part) its weight follows the plain ones and xmul =
 * Σ all / Σ plain keeps every plain family's rate (local_sweep! :1934-1941) */
static void move_counts(const esm_model *m, double extra, const double *wts, double px, const uint32_t key[2],
                        uint32_t ctr, int *counts) {
    int T = m->T;
    double wall[NW_ALL], cw[NW_ALL], acc = 0.0, tot = 0.0, tot0 = 0.0;
    for (int q = 0; q < ESM_NWEIGHT; q++) wall[q] = wts[q];
    wall[ESM_NWEIGHT] = px > 0 ? px : 0.0;
    for (int q = 0; q < ESM_NWEIGHT; q++) tot0 += wall[q];
    for (int q = 0; q < NW_ALL; q++) tot += wall[q];
    int nextra = tot > 0 ? (int)nearbyint(extra * (tot0 > 0 ? tot / tot0 : 1.0) * m->N) : 0;
    for (int q = 0; q < NW_ALL; q++) { acc += wall[q]; cw[q] = acc / tot; }
    for (int tau = 0; tau < T; tau++) {
        int cnt[NW_ALL] = {0};
        for (int d = 0; d < nextra; d++) {
            uint32_t c4[4] = {0x80000000u | (uint32_t)d, (uint32_t)tau, 9000u, ctr}, w[4];
            orc_philox(key, c4, w);
            double u = u01(w[0], w[1]);
            int f = NW_ALL - 1;
            for (int q = NW_ALL - 1; q >= 0; q--)
                if (cw[q] >= u) f = q;
            cnt[f]++;
        }
        for (int q = 0; q < NW_ALL; q++) counts[(size_t)q * T + tau] = nextra > 0 ? cnt[q] : 0;
    }
}

enum { W_UNIT = 0, W_PAIR = 1, W_XPAIR = 2, W_XLOCC = 3, W_XLOC = 4, W_POT = 5, W_CELLPAIR = 6, W_CELL = 7, W_SPAIR = 8,
       W_SXPAIR = 9 };

static void sxpair_move(Sys *S, const Tables *tb, int j, int tau, const uint32_t *s, double hmin, double useg,
                        double uoff, double umh, double maxnats, int64_t *stats);

/* the random moves of step τ for replica j, family by family (pair and cellpair are void); the
 * storage pairs last, with weight px and units of at least hmin hours */
static void random_step(Sys *S, const Tables *tb, int j, int tau, const int *counts, const double *wts, double px,
                        double hmin, const uint32_t key[2], uint32_t ctr, double maxnats, int64_t *st) {
    const esm_model *m = S->m;
    int N = m->N, T = m->T;
    static const int order[8] = {W_UNIT, W_POT, W_SPAIR, W_XPAIR, W_XLOCC, W_XLOC, W_CELL, W_SXPAIR};
    for (int oi = 0; oi < 8; oi++) {
        int w = order[oi];
        if (!((w == W_SXPAIR ? px : wts[w]) > 0)) continue;
        int cnt = counts[(size_t)w * T + tau];
        int nbucket = w == W_POT ? 3 : 1;
        for (int bk = 0; bk < nbucket; bk++)
            for (int k = 0; k < cnt; k++) {
                uint32_t c4[4] = {0x80000000u | ((uint32_t)w << 20) | (uint32_t)k, (uint32_t)tau, 9001u, ctr}, s[4];
                orc_philox(key, c4, s);
                uint32_t cl[4] = {(uint32_t)j, (uint32_t)tau, 0x40000000u | ((uint32_t)w << 20) | (uint32_t)k, ctr}, wl[4];
                orc_philox(key, cl, wl);
                double useg = u01(wl[0], wl[1]), uoff = u01(wl[2], wl[3]);
                int n = (int)pick(s[0], (uint32_t)N);
                if (w == W_SXPAIR) {
                    uint32_t cm[4] = {(uint32_t)j, (uint32_t)tau, 0x60000000u | ((uint32_t)w << 20) | (uint32_t)k, ctr}, wm[4];
                    orc_philox(key, cm, wm);
                    sxpair_move(S, tb, j, tau, s, hmin, useg, uoff, u01(wm[0], wm[1]), maxnats, st + 5 * ESM_FAM_SXPAIR);
                } else if (w == W_UNIT) {
                    do_move_u(S, j, 0, 0, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_UNIT);
                } else if (w == W_POT) {
                    int deg = S->adj_ptr[n + 1] - S->adj_ptr[n];
                    int lo = bk == 0 ? 0 : (bk == 1 ? 4 : 8), hi = bk == 0 ? 4 : (bk == 1 ? 8 : 16);
                    if (deg <= lo || deg > hi) continue;
                    do_move_u(S, j, 1, 0, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_POT);
                } else if (w == W_SPAIR) {
                    int sc = (int)pick(s[1], (uint32_t)ESM_NSTOR);
                    do_move_u(S, j, 3, sc, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_SPAIR);
                } else {
                    int kind, tid, fam;
                    if (w == W_CELL) {
                        double u = u01(s[0], s[1]);
                        int lv = tb->nlev - 1;
                        for (int q = tb->nlev - 1; q >= 0; q--)
                            if (tb->qcum.a[q] >= u) lv = q;
                        int nc = tb->lev_off.a[lv + 1] - tb->lev_off.a[lv];
                        tid = tb->cl0 + tb->lev_off.a[lv] + (int)pick(s[2], (uint32_t)nc);
                        kind = SP_CELL;
                        fam = ESM_FAM_CELL;
                    } else if (w == W_XPAIR) {
                        int mm = (int)pick(s[1], (uint32_t)N);
                        if (mm == n) continue;
                        tid = tb->xp0 + n * N + mm;
                        kind = SP_XPAIR;
                        fam = ESM_FAM_XPAIR;
                    } else {
                        int np = tb->xp_off.a[n + 1] - tb->xp_off.a[n];
                        if (np == 0) continue;
                        tid = tb->xp_off.a[n] + (int)pick(s[1], (uint32_t)np);
                        kind = w == W_XLOCC ? SP_XLOCC : SP_XLOC;
                        fam = w == W_XLOCC ? ESM_FAM_XLOCC : ESM_FAM_XLOC;
                    }
                    uint32_t cm[4] = {(uint32_t)j, (uint32_t)tau, 0x60000000u | ((uint32_t)w << 20) | (uint32_t)k, ctr}, wm[4];
                    orc_philox(key, cm, wm);
                    sm_move(S, tb, j, tau, tid, kind, useg, uoff, u01(wm[0], wm[1]), maxnats, st + 5 * fam);
                }
            }
    }
}

/* greedy distance-2 colouring in node order */
static int colouring(const esm_model *m, const int *adj_ptr, const int *adj_line, int *colour) {
    int N = m->N, nc = 0;
    int *used = (int *)malloc(sizeof(int) * (N + 1));
    for (int n = 0; n < N; n++) colour[n] = -1;
    for (int n = 0; n < N; n++) {
        memset(used, 0, sizeof(int) * (N + 1));
        for (int p = adj_ptr[n]; p < adj_ptr[n + 1]; p++) {
            int e = adj_line[p], a = m->line_bus0[e] == n ? m->line_bus1[e] : m->line_bus0[e];
            if (colour[a] >= 0) used[colour[a]] = 1;
            for (int p2 = adj_ptr[a]; p2 < adj_ptr[a + 1]; p2++) {
                int e2 = adj_line[p2], b2 = m->line_bus0[e2] == a ? m->line_bus1[e2] : m->line_bus0[e2];
                if (b2 != n && colour[b2] >= 0) used[colour[b2]] = 1;
            }
        }
        int c = 0;
        while (used[c]) c++;
        colour[n] = c;
        if (c + 1 > nc) nc = c + 1;
    }
    free(used);
    return nc;
}

int orc_colouring(const esm_model *m, int *colour) {
    int N = m->N, L = m->L;
    int *ptr = (int *)calloc(N + 1, sizeof(int)), *lines = (int *)malloc(sizeof(int) * 2 * L), *fill = (int *)malloc(sizeof(int) * N);
    for (int e = 0; e < L; e++) {
        ptr[m->line_bus0[e] + 1]++;
        ptr[m->line_bus1[e] + 1]++;
    }
    for (int n = 0; n < N; n++) ptr[n + 1] += ptr[n];
    for (int n = 0; n < N; n++) fill[n] = ptr[n];
    for (int e = 0; e < L; e++) {
        lines[fill[m->line_bus0[e]]++] = e;
        lines[fill[m->line_bus1[e]]++] = e;
    }
    int nc = colouring(m, ptr, lines, colour);
    free(ptr); free(lines); free(fill);
    return nc;
}

/* Explicit-uniform single move (no commit) and its cost along the line, for the pins of the
 * line sampler: the state arrays are read only. fam 0 gas, 1 potential, 2 SoC, 3 pair. */
static int prepare_single(Sys *S, int j, int fam, int s, int n, int tau, Step *wa, Step *ws) {
    const esm_model *m = S->m;
    int T = m->T, N = m->N;
    clear_step(wa, (tau - 1 + T) % T);
    clear_step(ws, tau);
    if (fam == 0) add_entry(ws, 1, n, 1.0);
    else if (fam == 1) add_entry(ws, 0, n, 1.0);
    else if (fam == 2) add_entry(ws, 2 + s, n, 1.0);
    else {
        double ks = S->kc[(ESM_KSTOR + s) * N + n] * m->stor_hours[s * N + n] / m->dt;
        double k1 = S->kc[ESM_KGAS * N + n];
        if (ks < 1.0 || fabs(k1) < 1.0) return -1;
        add_entry(wa, 1, n, ks / k1);
        add_entry(ws, 2 + s, n, 1.0);
        add_entry(ws, 1, n, -ks / k1);
    }
    finalize(S, ws, j);
    finalize(S, wa, j);
    link_soc(wa, ws);
    record(S, ws, NULL, j);
    record(S, wa, ws, j);
    return 0;
}

static void sys_view(Sys *S, const esm_model *m, int R, double *v, const double *k, double *r, double *F, int *ptr,
                     int *lines) {
    size_t NTR = (size_t)m->N * m->T * R;
    S->m = m;
    S->R = R;
    S->th = v;
    S->yg = v + NTR;
    S->ye[0] = v + 2 * NTR;
    S->ye[1] = v + 3 * NTR;
    S->r = r;
    S->F = F;
    for (int u = 0; u < ESM_NK * m->N; u++) S->kc[u] = clampd(k[u], m->klo[u], m->khi[u]);
    int N = m->N, L = m->L;
    int fill[256];
    memset(ptr, 0, sizeof(int) * (N + 1));
    for (int e = 0; e < L; e++) {
        ptr[m->line_bus0[e] + 1]++;
        ptr[m->line_bus1[e] + 1]++;
    }
    for (int q = 0; q < N; q++) ptr[q + 1] += ptr[q];
    for (int q = 0; q < N; q++) fill[q] = ptr[q];
    for (int e = 0; e < L; e++) {
        lines[fill[m->line_bus0[e]]++] = e;
        lines[fill[m->line_bus1[e]]++] = e;
    }
    S->adj_ptr = ptr;
    S->adj_line = lines;
}

int orc_move_draw(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int fam, int s,
                  int n, int tau, double maxnats, double useg, double uoff, double *t_out, int *status) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048) return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    Step wa, ws;
    if (prepare_single(&S, j, fam, s, n, tau, &wa, &ws)) return -2;
    Block b = {&S, &wa, &ws, j};
    *t_out = sample_line(&b, m->beta_v, maxnats, useg, uoff, status);
    return 0;
}

/* ΔC(t) (€) of the prepared move at each t[q] (recorded-term evaluation) */
int orc_move_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int fam, int s,
                  int n, int tau, const double *t, int nt, double *dc) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048) return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    Step wa, ws;
    if (prepare_single(&S, j, fam, s, n, tau, &wa, &ws)) return -2;
    Block b = {&S, &wa, &ws, j};
    for (int q = 0; q < nt; q++) dc[q] = mv_eval(&b, t[q]);
    return 0;
}

/* apply the prepared move at t (commit_block!) to v and the caches */
int orc_move_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int fam, int s,
                   int n, int tau, double t) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048) return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    Step wa, ws;
    if (prepare_single(&S, j, fam, s, n, tau, &wa, &ws)) return -2;
    Block b = {&S, &wa, &ws, j};
    commit(&b, t);
    return 0;
}

double orc_logseg(double L, double x) { return logseg(L, x); }
double orc_trexp_offset(double g, double L, double u) { return trexp_offset(g, L, u); }

/* ---- explicit-uniform table-driven moves and the template layout, for the pins ---- */

/* info [8]: templates, first xpair, first cell, levels, then the cells of levels 0, 1, 2 (0
 * beyond); xp_off [N + 1] (may be NULL): the xloc templates of each node */
int orc_tables_info(const esm_model *m, int *info, int *xp_off) {
    Tables tb;
    if (build_tables(m, &tb)) return -1;
    info[0] = tb.ntpl;
    info[1] = tb.xp0;
    info[2] = tb.cl0;
    info[3] = tb.nlev;
    for (int q = 0; q < 3; q++) info[4 + q] = q < tb.nlev ? tb.lev_off.a[q + 1] - tb.lev_off.a[q] : 0;
    info[7] = 0;
    if (xp_off) memcpy(xp_off, tb.xp_off.a, sizeof(int) * (m->N + 1));
    tables_free(&tb);
    return 0;
}

/* the touched nodes and leak nodes of template tid: nodes [≤ N] and kink nodes with their g */
int orc_template_nodes(const esm_model *m, int tid, int *nodes, int *nn, int *kk, double *kg, int *nk) {
    Tables tb;
    if (build_tables(m, &tb)) return -1;
    if (tid < 0 || tid >= tb.ntpl) { tables_free(&tb); return -2; }
    *nn = 0;
    for (int p = tb.nd_off.a[tid]; p < tb.nd_off.a[tid + 1]; p++) nodes[(*nn)++] = tb.nd_node.a[p];
    *nk = 0;
    for (int q = tb.kk_off.a[tid]; q < tb.kk_off.a[tid + 1]; q++) {
        kk[*nk] = tb.kk_node.a[q];
        kg[(*nk)++] = tb.kk_g.a[q];
    }
    tables_free(&tb);
    return 0;
}

static int sparse_setup(Sys *S, const esm_model *m, int R, double *v, const double *k, double *r, double *F, int *ptr,
                        int *lines, Tables *tb) {
    if (m->N > 256 || m->L > 2048) return -1;
    sys_view(S, m, R, v, k, r, F, ptr, lines);
    if (build_tables(m, tb)) return -2;
    return 0;
}

/* draw of template tid (kind: 0 cell, 1 xloc, 2 xlocc, 3 xpair) at (τ, replica j) with given
 * uniforms, no commit; t = 0 and status 3 when the move is not valid */
int orc_sparse_draw(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                    int tid, int tau, double maxnats, double useg, double uoff, double *t_out, int *status) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    Tables tb;
    int rc = sparse_setup(&S, m, R, v, k, r, F, ptr, lines, &tb);
    if (rc) return rc;
    SMove f;
    sm_init(&f, &S, &tb, j, tau, tid, kind == SP_XLOCC);
    if (!sm_valid(&f, kind)) {
        *t_out = 0.0;
        *status = 3;
    } else {
        *t_out = sm_sample(&f, m->beta_v, maxnats, useg, uoff, status);
    }
    tables_free(&tb);
    return 0;
}

int orc_sparse_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                    int tid, int tau, const double *t, int nt, double *dc) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    Tables tb;
    int rc = sparse_setup(&S, m, R, v, k, r, F, ptr, lines, &tb);
    if (rc) return rc;
    SMove f;
    sm_init(&f, &S, &tb, j, tau, tid, kind == SP_XLOCC);
    for (int q = 0; q < nt; q++) dc[q] = sm_eval(&f, t[q]);
    tables_free(&tb);
    return 0;
}

int orc_sparse_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                     int tid, int tau, double t) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    Tables tb;
    int rc = sparse_setup(&S, m, R, v, k, r, F, ptr, lines, &tb);
    if (rc) return rc;
    SMove f;
    sm_init(&f, &S, &tb, j, tau, tid, kind == SP_XLOCC);
    sm_commit(&f, t);
    tables_free(&tb);
    return 0;
}

/* nd draws of one table-driven move with given uniform pairs, the tables built once, no commit */
int orc_sparse_draw_many(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                         int tid, int tau, double maxnats, const double *useg, const double *uoff, int nd,
                         double *t_out, int *status) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    Tables tb;
    int rc = sparse_setup(&S, m, R, v, k, r, F, ptr, lines, &tb);
    if (rc) return rc;
    SMove f;
    sm_init(&f, &S, &tb, j, tau, tid, kind == SP_XLOCC);
    int ok = sm_valid(&f, kind);
    for (int q = 0; q < nd; q++) {
        if (!ok) {
            t_out[q] = 0.0;
            status[q] = 3;
            continue;
        }
        t_out[q] = sm_sample(&f, m->beta_v, maxnats, useg[q], uoff[q], &status[q]);
    }
    tables_free(&tb);
    return 0;
}

/* ============================================================================================
 * The storage multiscale sweep (storage_ms_sweep!, src/flowmoves.jl:255-324; prep_storage_ms!
 * :187-244), plateau moves (no flow-basis shapes in this basis), with the arithmetic of the
 * device's segment evaluation: block step q belongs to lane q mod W of a W-lane segment (W the
 * smallest power of two ≥ L + 1, at most 64); a lane sums its steps in increasing order, each
 * step its units in order (unit u: its node's kinked gas at the block ends, its storage term,
 * its node's balance); the lanes' sums are combined by the butterfly x_l + x_{l xor o},
 * o = W/2 … 1. The walk splits each primary segment at the storage secondaries' roots and finds
 * the balances' roots on the pieces (walk!, src/local.jl:1294-1443); equal roots are visited once.
 * ============================================================================================ */

#define SMS_NU 64   /* units per move: the largest country–carrier group */
#define SMS_PMAX 4096

typedef struct {
    double Q, g, ak, y0, P, e;
    int valid;
} SmsKink;

typedef struct {
    Sys *S;
    int j, L, b0, W, si, nu;
    int node[SMS_NU];
    double kst[SMS_NU], ks[SMS_NU];
    double *ye, *ye1, *S0;          /* [(L + 1)][SMS_NU] */
    SmsKink kk[2][SMS_NU];
} SmsMove;

static double sms_gen(const SmsMove *f, int n, double y, double *p) {
    const esm_model *m = f->S->m;
    double c, k = f->S->kc[ESM_KGAS * m->N + n];
    gen_cost(m->mc[ESM_KGAS * m->N + n], 1.0, k, k + m->kappa0, m->dt, m->c_viol, y, &c, p);
    return c;
}

static double sms_unit_step(const SmsMove *f, int q, int u, double t, double *sec3, double *S1) {
    const esm_model *m = f->S->m;
    int N = m->N, n = f->node[u], us = f->si * N + n;
    double D = 0.0, dr = 0.0;
    int kidx = q == 0 ? 0 : (q == f->L ? 1 : -1);
    if (kidx >= 0) {
        const SmsKink *kk = &f->kk[kidx][u];
        if (kk->valid && kk->ak != 0.0) {
            double p0, p1;
            double c0 = sms_gen(f, n, kk->y0, &p0);
            double c1 = sms_gen(f, n, kk->y0 + kk->ak * sm_phi(kk->Q, kk->g, t), &p1);
            D += c1 - c0;
            dr -= p1 - p0;
        }
    }
    double dye = q >= 1 ? 1.0 : 0.0, dye1 = q <= f->L - 1 ? 1.0 : 0.0;
    double ye = f->ye[q * SMS_NU + u], ye1 = f->ye1[q * SMS_NU + u];
    double w = f->kst[u] + m->kappa0, d0, j0, d1, j1;
    double c0 = soc_cost(m, f->si, n, f->kst[u], w, m->dt, m->c_viol, ye, ye1, &d0, &j0);
    double c1 = soc_cost(m, f->si, n, f->kst[u], w, m->dt, m->c_viol, ye + t * dye, ye1 + t * dye1, &d1, &j1);
    D += c1 - c0;
    dr += j1 - j0;
    if (sec3) {
        sec3[0] = d1;
        sec3[1] = m->stor_eta_in[us] > 0 ? d1 - m->stor_eta_in[us] : 1.0;
        sec3[2] = -d1 * m->stor_eta_out[us] - 1.0;
    }
    double S0 = f->S0[q * SMS_NU + u];
    *S1 = S0 + dr;
    D += m->dt * (phibal(m, *S1) - phibal(m, S0));
    return D;
}

static double sms_eval(const SmsMove *f, double t) {
    double part[64];
    g_nev++;
    for (int l = 0; l < f->W; l++) {
        double D = 0.0;
        for (int q = l; q <= f->L; q += f->W) {
            double Dq = 0.0, S1;
            for (int u = 0; u < f->nu; u++) Dq += sms_unit_step(f, q, u, t, NULL, &S1);
            D += Dq;
        }
        part[l] = D;
    }
    for (int off = f->W / 2; off > 0; off >>= 1) {
        double nx[64];
        for (int l = 0; l < f->W; l++) nx[l] = part[l] + part[l ^ off];
        memcpy(part, nx, sizeof(double) * f->W);
    }
    return part[0];
}

static int sms_next_bp(const SmsMove *f, double tp, double sgn, double *t) {
    int have = 0;
#define SCAND(bv)                                                                    \
    do {                                                                             \
        double b_ = (bv);                                                            \
        if (isfinite(b_) && sgn * (b_ - tp) > 0 && (!have || sgn * (*t - b_) > 0)) { \
            *t = b_;                                                                 \
            have = 1;                                                                \
        }                                                                            \
    } while (0)
    for (int q = 0; q <= f->L; q++) {
        int kidx = q == 0 ? 0 : (q == f->L ? 1 : -1);
        for (int u = 0; u < f->nu; u++) {
            if (kidx >= 0) {
                const SmsKink *kk = &f->kk[kidx][u];
                if (kk->valid) {
                    SCAND(-kk->Q / kk->g);
                    if (kk->ak != 0.0) {
                        double sl = kk->ak * kk->g, off = kk->ak * kk->Q;
                        SCAND((0.0 - kk->y0 - 0.0) / sl);
                        SCAND((1.0 - kk->y0 - 0.0) / sl);
                        SCAND((0.0 - kk->y0 - off) / sl);
                        SCAND((1.0 - kk->y0 - off) / sl);
                    }
                }
            }
            if (q >= 1) {
                double ye = f->ye[q * SMS_NU + u];
                SCAND((0.0 - ye) / 1.0);
                SCAND((1.0 - ye) / 1.0);
            }
        }
    }
#undef SCAND
    return have;
}

static int sms_next_root(const SmsMove *f, int stor, double a, double b, double x, double sgn, double *xr) {
    int have = 0;
    for (int q = 0; q <= f->L; q++)
        for (int u = 0; u < f->nu; u++) {
            double sa[3], sb[3], Ba, Bb;
            sms_unit_step(f, q, u, a, sa, &Ba);
            sms_unit_step(f, q, u, b, sb, &Bb);
            for (int k = 0; k < (stor ? 3 : 1); k++) {
                double A = stor ? sa[k] : Ba, B = stor ? sb[k] : Bb;
                if ((A < 0) != (B < 0) && A != B) {
                    double r = a + (b - a) * A / (A - B);
                    if (sgn * (r - a) > 0 && sgn * (b - r) > 0 && sgn * (r - x) > 0 && (!have || sgn * (*xr - r) > 0)) {
                        *xr = r;
                        have = 1;
                    }
                }
            }
        }
    return have;
}

typedef struct {
    double *p, *c;
    int n, full;
} BigPts;

static void bpush(BigPts *P, double x, double c) {
    if (P->n >= SMS_PMAX) {
        P->full = 1;
        return;
    }
    P->p[P->n] = x;
    P->c[P->n] = c;
    P->n++;
}

static int sms_walk(const SmsMove *f, double beta, double sgn, double maxnats, BigPts *P) {
    double cmin = 0.0, tp = 0.0, delta = 0.0;
    int nx = 0;
    P->n = 0;
    P->full = 0;
    for (;;) {
        double t;
        if (!sms_next_bp(f, tp, sgn, &t)) {
            double cprev = P->n == 0 ? 0.0 : P->c[P->n - 1];
            double hp = 1e-6 * fmax(1.0, fabs(tp));
            double sl = (beta * sms_eval(f, tp + sgn * hp) - cprev) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, 1e3 * fmax(fabs(tp), 1e-3));
            else delta = delta == 0.0 ? fmax(fabs(tp), 1e-6) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;
        }
        double c = beta * sms_eval(f, t);
        double a0 = tp, xs = tp, x = 0.0;
        while (sms_next_root(f, 1, tp, t, xs, sgn, &x)) {
            double cx = beta * sms_eval(f, x);
            double xb = a0, y = 0.0;
            while (sms_next_root(f, 0, a0, x, xb, sgn, &y)) {
                double cy = beta * sms_eval(f, y);
                bpush(P, y, cy);
                cmin = fmin(cmin, cy);
                xb = y;
            }
            bpush(P, x, cx);
            cmin = fmin(cmin, cx);
            a0 = x;
            xs = x;
        }
        double xb = a0, y = 0.0;
        while (sms_next_root(f, 0, a0, t, xb, sgn, &y)) {
            double cy = beta * sms_eval(f, y);
            bpush(P, y, cy);
            cmin = fmin(cmin, cy);
            xb = y;
        }
        bpush(P, t, c);
        if (P->full) return -1;
        cmin = fmin(cmin, c);
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}

/* the draw from two walked point lists of any length (the draw_points arithmetic) */
static double draw_points_big(const BigPts *PL, const BigPts *PR, double useg, double uoff, int *status) {
    int np = PL->n + 1 + PR->n;
    double *pts = (double *)malloc(sizeof(double) * np), *cs = (double *)malloc(sizeof(double) * np);
    double *lm = (double *)malloc(sizeof(double) * (np > 1 ? np - 1 : 1));
    for (int q = 0; q < PL->n; q++) {
        pts[q] = PL->p[PL->n - 1 - q];
        cs[q] = PL->c[PL->n - 1 - q];
    }
    pts[PL->n] = 0.0;
    cs[PL->n] = 0.0;
    for (int q = 0; q < PR->n; q++) {
        pts[PL->n + 1 + q] = PR->p[q];
        cs[PL->n + 1 + q] = PR->c[q];
    }
    int nseg = np - 1;
    double mx = -INFINITY;
    for (int q = 0; q < nseg; q++) {
        double L = pts[q + 1] - pts[q];
        lm[q] = L > 0 ? -fmin(cs[q], cs[q + 1]) + logseg(L, fabs(cs[q + 1] - cs[q])) : -INFINITY;
        mx = fmax(mx, lm[q]);
    }
    double tot = 0.0;
    for (int q = 0; q < nseg; q++) tot += exp(lm[q] - mx);
    double u = useg * tot, acc = 0.0;
    int jj = nseg - 1;
    for (int q = 0; q < nseg; q++) {
        acc += exp(lm[q] - mx);
        if (u <= acc) {
            jj = q;
            break;
        }
    }
    double L = pts[jj + 1] - pts[jj];
    double g = (cs[jj + 1] - cs[jj]) / L;
    double t = (g >= 0 || fabs(g * L) < 1e-10) ? pts[jj] + trexp_offset(g, L, uoff) : pts[jj + 1] - trexp_offset(-g, L, uoff);
    *status = 0;
    if (!isfinite(t)) {
        *status = 2;
        t = 0.0;
    }
    free(pts); free(cs); free(lm);
    return t;
}

static double sms_sample(const SmsMove *f, double beta, double maxnats, double useg, double uoff, int *status) {
    static __thread double bufs[4][SMS_PMAX];
    BigPts PR = {bufs[0], bufs[1], 0, 0}, PL = {bufs[2], bufs[3], 0, 0};
    int w = sms_walk(f, beta, 1.0, maxnats, &PR);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    w = sms_walk(f, beta, -1.0, maxnats, &PL);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    return draw_points_big(&PL, &PR, useg, uoff, status);
}

/* stage the block values of replica j and the kinks at the block's ends (prorata! with
 * soc_qadj, src/local.jl:1782-1840) */
static void sms_prepare(SmsMove *f, Sys *S, int j, int L, int b0, int si, const int *nodes, int nu) {
    const esm_model *m = S->m;
    int N = m->N, T = m->T;
    f->S = S; f->j = j; f->L = L; f->b0 = b0; f->si = si; f->nu = nu;
    f->W = 4;
    while (f->W < L + 1 && f->W < 64) f->W *= 2;
    for (int u = 0; u < nu; u++) {
        f->node[u] = nodes[u];
        f->kst[u] = S->kc[(ESM_KSTOR + si) * N + nodes[u]];
        f->ks[u] = f->kst[u] * m->stor_hours[si * N + nodes[u]] / m->dt;
    }
    double *ye_a = S->ye[si];
    for (int q = 0; q <= L; q++) {
        int tau = b0 + q, tau1 = tau + 1 == T ? 0 : tau + 1;
        for (int u = 0; u < nu; u++) {
            int n = nodes[u];
            f->ye[q * SMS_NU + u] = AT(ye_a, n, tau);
            f->ye1[q * SMS_NU + u] = AT(ye_a, n, tau1);
            f->S0[q * SMS_NU + u] = AT(S->r, n, tau);
        }
    }
    for (int kidx = 0; kidx < 2; kidx++) {
        int q = kidx == 0 ? 0 : L, tau = b0 + q;
        for (int u = 0; u < nu; u++) {
            SmsKink *kk = &f->kk[kidx][u];
            memset(kk, 0, sizeof(*kk));
            int n = nodes[u], us = si * N + n;
            double k = S->kc[ESM_KGAS * N + n];
            double g = kidx == 0 ? f->ks[u] * (1.0 - 0.0) : f->ks[u] * (0.0 - 1.0);
            if (!(fabs(k) >= 1.0) || g == 0.0) continue;
            double w = AT(S->yg, n, tau);
            double P = 0.0 + k * w, Qs = 0.0 + k * clampd(w, 0.0, 1.0);
            if (fabs(P) < 1e-6) continue;
            double Y = f->ye[q * SMS_NU + u], Y1 = f->ye1[q * SMS_NU + u];
            double ec = clampd(Y, 0.0, 1.0), ec1 = clampd(Y1, 0.0, 1.0);
            double d = m->stor_hours[us] * (ec1 - ec) / m->dt - m->stor_qt[us], pc, mcp;
            if (d >= 0) { pc = m->stor_eta_in[us] > 0 ? d / m->stor_eta_in[us] : 0.0; mcp = 0.0; }
            else { pc = 0.0; mcp = fmin(-d * m->stor_eta_out[us], 1.0); }
            double qadj = 0.0 + (f->kst[u] * m->stor_hours[us] * (Y1 - Y) / m->dt - f->kst[u] * (pc - mcp));
            kk->Q = AT(S->r, n, tau) + Qs + qadj;
            kk->g = g;
            kk->e = P - fmax(kk->Q, 0.0);
            kk->P = P;
            double sig = k * w / P;
            kk->ak = sig == 0.0 ? 0.0 : sig / k;
            kk->y0 = w;
            kk->valid = 1;
        }
    }
}

static void sms_commit(const SmsMove *f, double t) {
    Sys *S = f->S;
    int j = f->j;
    if (t == 0.0) return;
    for (int q = 0; q <= f->L; q++) {
        int tau = f->b0 + q, kidx = q == 0 ? 0 : (q == f->L ? 1 : -1);
        for (int u = 0; u < f->nu; u++) {
            int n = f->node[u];
            double S1;
            sms_unit_step(f, q, u, t, NULL, &S1);
            AT(S->r, n, tau) = S1;
            if (q >= 1) AT(S->ye[f->si], n, tau) = f->ye[q * SMS_NU + u] + t;
            if (kidx >= 0) {
                const SmsKink *kk = &f->kk[kidx][u];
                if (kk->valid && kk->ak != 0.0) AT(S->yg, n, tau) = kk->y0 + kk->ak * sm_phi(kk->Q, kk->g, t);
            }
        }
    }
}

static double sms_logjac(const SmsMove *f, double t, int *any) {
    double part[64] = {0};
    int cnt = 0;
    for (int kidx = 0; kidx < 2; kidx++) {
        int q = kidx == 0 ? 0 : f->L, lane = q % f->W;
        for (int u = 0; u < f->nu; u++) {
            const SmsKink *kk = &f->kk[kidx][u];
            if (!kk->valid) continue;
            cnt++;
            double nm = (double)(gas_units(f->S, f->node[u]) - 1);
            part[lane] += nm * log(fabs((fmax(kk->Q + kk->g * t, 0.0) + kk->e) / kk->P));
        }
    }
    *any = cnt > 0;
    for (int off = f->W / 2; off > 0; off >>= 1) {
        double nx[64];
        for (int l = 0; l < f->W; l++) nx[l] = part[l] + part[l ^ off];
        memcpy(part, nx, sizeof(double) * f->W);
    }
    return part[0];
}

/* units and country–carrier groups (cells sorted, then carrier; nodes increasing) */
typedef struct {
    int nunits, ngroups;
    int grp_off[2 * 256 + 1], grp_node[2 * 256], grp_si[2 * 256];
} SmsTables;

static int sms_tables(const esm_model *m, SmsTables *st) {
    int N = m->N;
    st->nunits = ESM_NSTOR * N;
    int cells[256], nc = 0;
    for (int n = 0; n < N; n++) {
        int c = m->ncountry > 0 ? m->country[n] : n, seen = 0;
        for (int q = 0; q < nc; q++) seen |= cells[q] == c;
        if (!seen) cells[nc++] = c;
    }
    qsort(cells, nc, sizeof(int), cmp_int);
    st->ngroups = 0;
    int off = 0;
    st->grp_off[0] = 0;
    for (int q = 0; q < nc; q++)
        for (int s = 0; s < ESM_NSTOR; s++) {
            int cnt = 0;
            for (int n = 0; n < N; n++)
                if ((m->ncountry > 0 ? m->country[n] : n) == cells[q]) { st->grp_node[off++] = n; cnt++; }
            if (cnt > SMS_NU) return -1;
            st->grp_si[st->ngroups] = s;
            st->grp_off[++st->ngroups] = off;
        }
    return 0;
}

/* one storage multiscale move of replica j; returns t and counts it */
static double sms_move(Sys *S, int j, int L, int b0, int si, const int *nodes, int nu, double useg, double uoff,
                       double umh, double maxnats, int64_t *stats) {
    static __thread double ye[129 * SMS_NU], ye1[129 * SMS_NU], S0[129 * SMS_NU];
    SmsMove f;
    f.ye = ye; f.ye1 = ye1; f.S0 = S0;
    sms_prepare(&f, S, j, L, b0, si, nodes, nu);
    int status;
    g_nev = 0;
    double t = sms_sample(&f, S->m->beta_v, maxnats, useg, uoff, &status);
    if (t != 0.0) {
        int any;
        double lj = sms_logjac(&f, t, &any);
        if (any && !(umh < exp(fmin(0.0, lj)))) { t = 0.0; status = 2; }
    }
    long nev = g_nev;
    sms_commit(&f, t);
    if (stats) stat_add(stats, status, t, nev);
    return t;
}

/* the schedule of one storage sweep: whether it runs, lengths in order, offsets, first class */
static int sms_schedule(int T, double p, const uint32_t key[2], uint32_t ctr, int *lengths, int *offs, int *first) {
    uint32_t c[4] = {0u, 0u, 9300u, ctr}, w[4];
    orc_philox(key, c, w);
    if (!(u01(w[0], w[1]) < p)) return 0;
    int n = 0;
    for (int L = 2; L <= 128; L *= 2)
        if (L < T / 2) lengths[n++] = L;
    for (int i = n - 1; i >= 1; i--) {
        uint32_t c2[4] = {(uint32_t)i, 0u, 9301u, ctr};
        orc_philox(key, c2, w);
        int jx = (int)(((uint64_t)w[0] * (uint32_t)(i + 1)) >> 32);
        int tmp = lengths[i]; lengths[i] = lengths[jx]; lengths[jx] = tmp;
    }
    for (int q = 0; q < n; q++) {
        uint32_t c3[4] = {(uint32_t)lengths[q], 0u, 9302u, ctr};
        orc_philox(key, c3, w);
        offs[q] = (int)(((uint64_t)w[0] * (uint32_t)lengths[q]) >> 32);
        first[q] = u01(w[2], w[3]) < 0.5 ? 0 : 1;
    }
    return n;
}

/* the storage sweep of replica j (nmoves per block, group probability p_agg) */
static void sms_sweep(Sys *S, const SmsTables *st, int j, double pms, const uint32_t key[2], uint32_t ctr,
                      double maxnats, int64_t *stats) {
    const esm_model *m = S->m;
    int T = m->T, N = m->N;
    int lengths[8], offs[8], first[8];
    int nlen = sms_schedule(T, pms, key, ctr, lengths, offs, first);
    for (int li = 0; li < nlen; li++) {
        int L = lengths[li], o = offs[li];
        if (o > T - L - 1) continue;
        int nb = (T - L - 1 - o) / L + 1;
        if (nb < 2) continue;
        for (int ci = 0; ci < 2; ci++) {
            int cls = ci == 0 ? first[li] : 1 - first[li];
            for (int bi = cls; bi < nb; bi += 2) {
                int b0 = o + L * bi;
                for (int mv = 0; mv < 8; mv++) {
                    uint32_t c4[4] = {(uint32_t)mv, (uint32_t)b0, 9200u + (uint32_t)L, ctr}, s[4];
                    orc_philox(key, c4, s);
                    int si, nu, nodes[SMS_NU];
                    if (u01(s[0], s[1]) < 0.5) {
                        int g = (int)(((uint64_t)s[2] * (uint32_t)st->ngroups) >> 32);
                        si = st->grp_si[g];
                        nu = st->grp_off[g + 1] - st->grp_off[g];
                        for (int u = 0; u < nu; u++) nodes[u] = st->grp_node[st->grp_off[g] + u];
                    } else {
                        int un = (int)(((uint64_t)s[3] * (uint32_t)st->nunits) >> 32);
                        si = un / N;
                        nu = 1;
                        nodes[0] = un % N;
                    }
                    uint32_t cl[4] = {(uint32_t)j, (uint32_t)b0, 0x50000000u + (uint32_t)(L * 16 + mv), ctr}, wl[4];
                    orc_philox(key, cl, wl);
                    uint32_t cm[4] = {(uint32_t)j, (uint32_t)b0, 0x70000000u + (uint32_t)(L * 16 + mv), ctr}, wm[4];
                    orc_philox(key, cm, wm);
                    sms_move(S, j, L, b0, si, nodes, nu, u01(wl[0], wl[1]), u01(wl[2], wl[3]), u01(wm[0], wm[1]),
                             maxnats, stats);
                }
            }
        }
    }
}

/* ---- explicit storage multiscale moves, for the pins ---- */
int orc_sms_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int L, int b0,
                 int si, const int *nodes, int nu, const double *t, int nt, double *dc) {
    static __thread Sys S;
    static __thread double ye[129 * SMS_NU], ye1[129 * SMS_NU], S0[129 * SMS_NU];
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048 || L > 128 || nu > SMS_NU || b0 + L >= m->T) return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    SmsMove f;
    f.ye = ye; f.ye1 = ye1; f.S0 = S0;
    sms_prepare(&f, &S, j, L, b0, si, nodes, nu);
    for (int q = 0; q < nt; q++) dc[q] = sms_eval(&f, t[q]);
    return 0;
}

int orc_sms_draw_many(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int L,
                      int b0, int si, const int *nodes, int nu, double maxnats, const double *useg, const double *uoff,
                      int nd, double *t_out, int *status) {
    static __thread Sys S;
    static __thread double ye[129 * SMS_NU], ye1[129 * SMS_NU], S0[129 * SMS_NU];
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048 || L > 128 || nu > SMS_NU || b0 + L >= m->T) return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    SmsMove f;
    f.ye = ye; f.ye1 = ye1; f.S0 = S0;
    sms_prepare(&f, &S, j, L, b0, si, nodes, nu);
    for (int q = 0; q < nd; q++) t_out[q] = sms_sample(&f, m->beta_v, maxnats, useg[q], uoff[q], &status[q]);
    return 0;
}

int orc_sms_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int L, int b0,
                  int si, const int *nodes, int nu, double t) {
    static __thread Sys S;
    static __thread double ye[129 * SMS_NU], ye1[129 * SMS_NU], S0[129 * SMS_NU];
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048 || L > 128 || nu > SMS_NU || b0 + L >= m->T) return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    SmsMove f;
    f.ye = ye; f.ye1 = ye1; f.S0 = S0;
    sms_prepare(&f, &S, j, L, b0, si, nodes, nu);
    sms_commit(&f, t);
    return 0;
}

/* ============================================================================================
 * The long-range storage sweep (storage_long_sweep!, src/storage_long.jl:245-298): per cell, for
 * every storage unit with at least `hmin` hours and κs = k h/Δt ≥ 1, a level move (L = T: every
 * instant rises) and `nmoves` long-range plateaus (the state of charge rises by t at the cyclic
 * instants a+1 … a+L, L uniform on 1 … T−1, a uniform; prep_slong! :77-146), in a shuffled
 * order; the node's gas unit compensates pro rata at both end steps (prorata! with soc_qadj),
 * with the Jacobian test. Interior steps whose two instants stay in [0, 1] change nothing: their
 * cost change is exactly zero and their secondaries keep their values at t = 0 (mv_eval!
 * :148-188). The arithmetic is the device's: step q of the move (q = 0 … L, or 1 … T for the
 * level move; step q is τ = a + q mod T) belongs to lane q mod 64; a lane sums its steps' terms
 * in increasing q, the lanes' sums are combined by the 64-lane butterfly. Selection (shared by
 * the replicas, Philox under the sweep's key and counter c): whether the sweep runs
 * (0, 0, 9410, c); per cell (its node n) and unit u of the cell: the level move (n, u, 9411, c);
 * plateau m: L and a from (n, 65536u + m, 9412, c); the cell's move list shuffled by
 * Fisher–Yates with (n, i, 9413, c); move i's line draw (replica, n, 0x54000000 + i, c) and
 * Jacobian test (replica, n, 0x74000000 + i, c).
 * ============================================================================================ */

typedef struct {
    Sys *S;
    int j, si, n, a0, L, level;
    double kst, ks;
    SmsKink kk[2];                 /* the kinks at step a (q = 0) and step b (q = L) */
} SlgMove;

static int slg_tau(const SlgMove *f, int q) { return (f->a0 + q) % f->S->m->T; }
static int slg_q0(const SlgMove *f) { return f->level ? 1 : 0; }

/* whether interior step q keeps both its instants in [0, 1] at t (its terms do not change) */
static int slg_inside(const SlgMove *f, int q, double t) {
    Sys *S = f->S;
    int j = f->j, T = S->m->T;
    if (!f->level && (q == 0 || q == f->L)) return 0;
    int tau = slg_tau(f, q), tau1 = tau + 1 == T ? 0 : tau + 1;
    double ye = AT(S->ye[f->si], f->n, tau), ye1 = AT(S->ye[f->si], f->n, tau1);
    return 0.0 <= ye && ye <= 1.0 && 0.0 <= ye1 && ye1 <= 1.0 && fmax(-ye, -ye1) <= t && t <= fmin(1.0 - ye, 1.0 - ye1);
}

/* term of step q at t; *d1 the storage quantity d, *S1 the balance, *moved 0 when an interior
 * step keeps both instants in [0, 1] (no change: d and the balance keep their values at t = 0) */
static double slg_step(const SlgMove *f, int q, double t, double *d1, double *S1, int *moved) {
    Sys *S = f->S;
    const esm_model *m = S->m;
    int j = f->j, N = m->N, T = m->T, n = f->n;
    int tau = slg_tau(f, q), tau1 = tau + 1 == T ? 0 : tau + 1;
    double ye = AT(S->ye[f->si], n, tau), ye1 = AT(S->ye[f->si], n, tau1), S0 = AT(S->r, n, tau);
    double w = f->kst + m->kappa0, d0, j0, j1;
    int end = !f->level && (q == 0 || q == f->L);
    double c0 = soc_cost(m, f->si, n, f->kst, w, m->dt, m->c_viol, ye, ye1, &d0, &j0);
    if (!end) {
        if (slg_inside(f, q, t)) {
            *d1 = d0;
            *S1 = S0;
            *moved = 0;
            return 0.0;
        }
        *moved = 1;
        double c1 = soc_cost(m, f->si, n, f->kst, w, m->dt, m->c_viol, ye + t, ye1 + t, d1, &j1);
        *S1 = S0 + (j1 - j0);
        return (c1 - c0) + m->dt * (phibal(m, *S1) - phibal(m, S0));
    }
    *moved = 1;
    double D = 0.0, dr = 0.0;
    const SmsKink *kk = &f->kk[q == 0 ? 0 : 1];
    if (kk->valid && kk->ak != 0.0) {
        double p0, p1, c = S->kc[ESM_KGAS * N + n], cg0, cg1;
        gen_cost(m->mc[ESM_KGAS * N + n], 1.0, c, c + m->kappa0, m->dt, m->c_viol, kk->y0, &cg0, &p0);
        gen_cost(m->mc[ESM_KGAS * N + n], 1.0, c, c + m->kappa0, m->dt, m->c_viol, kk->y0 + kk->ak * sm_phi(kk->Q, kk->g, t),
                 &cg1, &p1);
        D += cg1 - cg0;
        dr -= p1 - p0;
    }
    double dye = q >= 1 ? 1.0 : 0.0, dye1 = q <= f->L - 1 ? 1.0 : 0.0;
    double c1 = soc_cost(m, f->si, n, f->kst, w, m->dt, m->c_viol, ye + t * dye, ye1 + t * dye1, d1, &j1);
    D += c1 - c0;
    dr += j1 - j0;
    *S1 = S0 + dr;
    D += m->dt * (phibal(m, *S1) - phibal(m, S0));
    return D;
}

static double slg_eval(const SlgMove *f, double t) {
    double part[64] = {0};
    g_nev++;
    for (int q = slg_q0(f); q <= f->L; q++) {
        if (slg_inside(f, q, t)) continue;
        double d1, S1;
        int moved;
        double x = slg_step(f, q, t, &d1, &S1, &moved);
        part[q % 64] += x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        double nx[64];
        for (int l = 0; l < 64; l++) nx[l] = part[l] + part[l ^ off];
        memcpy(part, nx, sizeof(part));
    }
    return part[0];
}

static int slg_next_bp(const SlgMove *f, double tp, double sgn, double *t) {
    Sys *S = f->S;
    int j = f->j, have = 0;
#define SCAND(bv)                                                                    \
    do {                                                                             \
        double b_ = (bv);                                                            \
        if (isfinite(b_) && sgn * (b_ - tp) > 0 && (!have || sgn * (*t - b_) > 0)) { \
            *t = b_;                                                                 \
            have = 1;                                                                \
        }                                                                            \
    } while (0)
    if (!f->level)
        for (int kidx = 0; kidx < 2; kidx++) {
            const SmsKink *kk = &f->kk[kidx];
            if (!kk->valid) continue;
            SCAND(-kk->Q / kk->g);
            if (kk->ak != 0.0) {
                double sl = kk->ak * kk->g, off = kk->ak * kk->Q;
                SCAND((0.0 - kk->y0 - 0.0) / sl);
                SCAND((1.0 - kk->y0 - 0.0) / sl);
                SCAND((0.0 - kk->y0 - off) / sl);
                SCAND((1.0 - kk->y0 - off) / sl);
            }
        }
    for (int q = 1; q <= f->L; q++) {
        double ye = AT(S->ye[f->si], f->n, slg_tau(f, q));
        SCAND((0.0 - ye) / 1.0);
        SCAND((1.0 - ye) / 1.0);
    }
#undef SCAND
    return have;
}

static int slg_next_root(const SlgMove *f, int stor, double a, double b, double x, double sgn, double *xr) {
    const esm_model *m = f->S->m;
    int us = f->si * m->N + f->n, have = 0;
    double ein = m->stor_eta_in[us], eout = m->stor_eta_out[us];
    for (int q = slg_q0(f); q <= f->L; q++) {
        if (slg_inside(f, q, a) && slg_inside(f, q, b)) continue;   /* constant secondaries */
        double da, db, Ba, Bb;
        int mv;
        slg_step(f, q, a, &da, &Ba, &mv);
        slg_step(f, q, b, &db, &Bb, &mv);
        for (int k = 0; k < (stor ? 3 : 1); k++) {
            double A, B;
            if (!stor) { A = Ba; B = Bb; }
            else if (k == 0) { A = da; B = db; }
            else if (k == 1) { A = ein > 0 ? da - ein : 1.0; B = ein > 0 ? db - ein : 1.0; }
            else { A = -da * eout - 1.0; B = -db * eout - 1.0; }
            if ((A < 0) != (B < 0) && A != B) {
                double r = a + (b - a) * A / (A - B);
                if (sgn * (r - a) > 0 && sgn * (b - r) > 0 && sgn * (r - x) > 0 && (!have || sgn * (*xr - r) > 0)) {
                    *xr = r;
                    have = 1;
                }
            }
        }
    }
    return have;
}

static int slg_walk(const SlgMove *f, double beta, double sgn, double maxnats, BigPts *P) {
    double cmin = 0.0, tp = 0.0, delta = 0.0, cprev = 0.0;
    int nx = 0;
    while (1) {
        double t = 0.0;
        if (!slg_next_bp(f, tp, sgn, &t)) {
            double hp = 1e-6 * fmax(1.0, fabs(tp));
            double sl = (beta * slg_eval(f, tp + sgn * hp) - cprev) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, 1e3 * fmax(fabs(tp), 1e-3));
            else delta = delta == 0.0 ? fmax(fabs(tp), 1e-6) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;
        }
        double c = beta * slg_eval(f, t);
        double a0 = tp, xs = tp, x = 0.0;
        while (slg_next_root(f, 1, tp, t, xs, sgn, &x)) {
            double cx = beta * slg_eval(f, x);
            double xb = a0, y = 0.0;
            while (slg_next_root(f, 0, a0, x, xb, sgn, &y)) {
                double cy = beta * slg_eval(f, y);
                bpush(P, y, cy);
                cmin = fmin(cmin, cy);
                xb = y;
            }
            bpush(P, x, cx);
            cmin = fmin(cmin, cx);
            a0 = x;
            xs = x;
        }
        double xb = a0, y = 0.0;
        while (slg_next_root(f, 0, a0, t, xb, sgn, &y)) {
            double cy = beta * slg_eval(f, y);
            bpush(P, y, cy);
            cmin = fmin(cmin, cy);
            xb = y;
        }
        bpush(P, t, c);
        if (P->full) return -1;
        cmin = fmin(cmin, c);
        cprev = c;
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}

static double slg_sample(const SlgMove *f, double beta, double maxnats, double useg, double uoff, int *status) {
    static __thread double bufs[4][SMS_PMAX];
    BigPts PR = {bufs[0], bufs[1], 0, 0}, PL = {bufs[2], bufs[3], 0, 0};
    int w = slg_walk(f, beta, 1.0, maxnats, &PR);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    w = slg_walk(f, beta, -1.0, maxnats, &PL);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    return draw_points_big(&PL, &PR, useg, uoff, status);
}

/* a move of unit (si, n) of replica j: L < T a plateau from step a (0-based) to a + L, L = T the
 * level move; the kinks at both end steps (prorata! with soc_qadj, as the multiscale blocks) */
static void slg_prepare(SlgMove *f, Sys *S, int j, int si, int n, int a, int L) {
    const esm_model *m = S->m;
    int N = m->N, T = m->T, us = si * N + n;
    memset(f, 0, sizeof(*f));
    f->S = S; f->j = j; f->si = si; f->n = n; f->L = L;
    f->level = L >= T;
    f->a0 = f->level ? T - 1 : a;
    if (f->level) f->L = T;
    f->kst = S->kc[(ESM_KSTOR + si) * N + n];
    f->ks = f->kst * m->stor_hours[us] / m->dt;
    if (f->level) return;
    for (int kidx = 0; kidx < 2; kidx++) {
        SmsKink *kk = &f->kk[kidx];
        int tau = slg_tau(f, kidx == 0 ? 0 : f->L), tau1 = tau + 1 == T ? 0 : tau + 1;
        double k = S->kc[ESM_KGAS * N + n];
        double g = kidx == 0 ? f->ks * (1.0 - 0.0) : f->ks * (0.0 - 1.0);
        if (!(fabs(k) >= 1.0) || g == 0.0) continue;
        double w = AT(S->yg, n, tau);
        double P = 0.0 + k * w, Qs = 0.0 + k * clampd(w, 0.0, 1.0);
        if (fabs(P) < 1e-6) continue;
        double Y = AT(S->ye[si], n, tau), Y1 = AT(S->ye[si], n, tau1);
        double ec = clampd(Y, 0.0, 1.0), ec1 = clampd(Y1, 0.0, 1.0);
        double d = m->stor_hours[us] * (ec1 - ec) / m->dt - m->stor_qt[us], pc, mcp;
        if (d >= 0) { pc = m->stor_eta_in[us] > 0 ? d / m->stor_eta_in[us] : 0.0; mcp = 0.0; }
        else { pc = 0.0; mcp = fmin(-d * m->stor_eta_out[us], 1.0); }
        double qadj = 0.0 + (f->kst * m->stor_hours[us] * (Y1 - Y) / m->dt - f->kst * (pc - mcp));
        kk->Q = AT(S->r, n, tau) + Qs + qadj;
        kk->g = g;
        kk->e = P - fmax(kk->Q, 0.0);
        kk->P = P;
        double sig = k * w / P;
        kk->ak = sig == 0.0 ? 0.0 : sig / k;
        kk->y0 = w;
        kk->valid = 1;
    }
}

static void slg_commit(const SlgMove *f, double t) {
    Sys *S = f->S;
    int j = f->j, n = f->n;
    if (t == 0.0) return;
    double *nr = (double *)malloc(sizeof(double) * (f->L + 1));
    int *mv = (int *)malloc(sizeof(int) * (f->L + 1));
    for (int q = slg_q0(f); q <= f->L; q++) {   /* every term from the state before the move */
        double d1;
        slg_step(f, q, t, &d1, &nr[q], &mv[q]);
    }
    for (int q = slg_q0(f); q <= f->L; q++) {
        int tau = slg_tau(f, q);
        if (mv[q]) AT(S->r, n, tau) = nr[q];
        if (!f->level && (q == 0 || q == f->L)) {
            const SmsKink *kk = &f->kk[q == 0 ? 0 : 1];
            if (kk->valid && kk->ak != 0.0) AT(S->yg, n, tau) = kk->y0 + kk->ak * sm_phi(kk->Q, kk->g, t);
        }
    }
    for (int q = 1; q <= f->L; q++) AT(S->ye[f->si], n, slg_tau(f, q)) += t;
    free(nr); free(mv);
}

static double slg_logjac(const SlgMove *f, double t, int *any) {
    double part[64] = {0};
    int cnt = 0;
    if (!f->level)
        for (int kidx = 0; kidx < 2; kidx++) {
            const SmsKink *kk = &f->kk[kidx];
            if (!kk->valid) continue;
            cnt++;
            double nm = (double)(gas_units(f->S, f->n) - 1);
            part[(kidx == 0 ? 0 : f->L) % 64] += nm * log(fabs((fmax(kk->Q + kk->g * t, 0.0) + kk->e) / kk->P));
        }
    *any = cnt > 0;
    for (int off = 32; off > 0; off >>= 1) {
        double nx[64];
        for (int l = 0; l < 64; l++) nx[l] = part[l] + part[l ^ off];
        memcpy(part, nx, sizeof(part));
    }
    return part[0];
}

static double slg_move(Sys *S, int j, int si, int n, int a, int L, double useg, double uoff, double umh,
                       double maxnats, int64_t *stats) {
    SlgMove f;
    slg_prepare(&f, S, j, si, n, a, L);
    int status;
    g_nev = 0;
    double t = slg_sample(&f, S->m->beta_v, maxnats, useg, uoff, &status);
    if (t != 0.0) {
        int any;
        double lj = slg_logjac(&f, t, &any);
        if (any && !(umh < exp(fmin(0.0, lj)))) { t = 0.0; status = 2; }
    }
    long nev = g_nev;
    slg_commit(&f, t);
    if (stats) stat_add(stats, status, t, nev);
    return t;
}

/* the move list of the cell of node n (units of the node with hours ≥ hmin and κs ≥ 1, in
 * carrier order): level moves and plateaus, shuffled; entry = (unit carrier, a, L) */
static int slg_moves(const esm_model *m, const double *kc, int n, int nmoves, double hmin, const uint32_t key[2],
                     uint32_t ctr, int *msi, int *ma, int *mL) {
    int N = m->N, T = m->T, cnt = 0, u = 0;
    for (int si = 0; si < ESM_NSTOR; si++) {
        int us = si * N + n;
        double kst = kc[(ESM_KSTOR + si) * N + n];
        if (!(m->stor_hours[us] >= hmin && kst * m->stor_hours[us] / m->dt >= 1.0)) continue;
        uint32_t c1[4] = {(uint32_t)n, (uint32_t)u, 9411u, ctr}, w[4];
        orc_philox(key, c1, w);
        if (u01(w[0], w[1]) < 1.0) { msi[cnt] = si; ma[cnt] = 0; mL[cnt] = T; cnt++; }
        for (int mm = 0; mm < nmoves; mm++) {
            uint32_t c2[4] = {(uint32_t)n, (uint32_t)(65536u * (uint32_t)u + (uint32_t)mm), 9412u, ctr};
            orc_philox(key, c2, w);
            msi[cnt] = si;
            mL[cnt] = 1 + (int)(((uint64_t)w[0] * (uint32_t)(T - 1)) >> 32);
            ma[cnt] = (int)(((uint64_t)w[1] * (uint32_t)T) >> 32);
            cnt++;
        }
        u++;
    }
    for (int i = cnt - 1; i >= 1; i--) {
        uint32_t c3[4] = {(uint32_t)n, (uint32_t)i, 9413u, ctr}, w[4];
        orc_philox(key, c3, w);
        int jx = (int)(((uint64_t)w[0] * (uint32_t)(i + 1)) >> 32), tmp;
        tmp = msi[i]; msi[i] = msi[jx]; msi[jx] = tmp;
        tmp = ma[i]; ma[i] = ma[jx]; ma[jx] = tmp;
        tmp = mL[i]; mL[i] = mL[jx]; mL[jx] = tmp;
    }
    return cnt;
}

static int slg_sweep(Sys *S, int j, double plong, int nmoves, double hmin, const uint32_t key[2], uint32_t ctr,
                     double maxnats, int64_t *st_long, int64_t *st_level) {
    const esm_model *m = S->m;
    uint32_t c0[4] = {0u, 0u, 9410u, ctr}, w[4];
    orc_philox(key, c0, w);
    if (!(u01(w[0], w[1]) < plong)) return 0;
    int cap = ESM_NSTOR * (nmoves + 1);
    int *msi = (int *)malloc(sizeof(int) * cap), *ma = (int *)malloc(sizeof(int) * cap), *mL = (int *)malloc(sizeof(int) * cap);
    for (int n = 0; n < m->N; n++) {
        int cnt = slg_moves(m, S->kc, n, nmoves, hmin, key, ctr, msi, ma, mL);
        for (int i = 0; i < cnt; i++) {
            uint32_t cl[4] = {(uint32_t)j, (uint32_t)n, 0x54000000u + (uint32_t)i, ctr}, wl[4];
            orc_philox(key, cl, wl);
            uint32_t cm[4] = {(uint32_t)j, (uint32_t)n, 0x74000000u + (uint32_t)i, ctr}, wm[4];
            orc_philox(key, cm, wm);
            slg_move(S, j, msi[i], n, ma[i], mL[i], u01(wl[0], wl[1]), u01(wl[2], wl[3]), u01(wm[0], wm[1]), maxnats,
                     mL[i] >= m->T ? st_level : st_long);
        }
    }
    free(msi); free(ma); free(mL);
    return 1;
}

/* ---- explicit long-range moves, for the pins ---- */
int orc_slong_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int si, int n,
                   int a, int L, const double *t, int nt, double *dc) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048 || si < 0 || si >= ESM_NSTOR || n < 0 || n >= m->N || a < 0 || a >= m->T || L < 1 ||
        L > m->T)
        return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    SlgMove f;
    slg_prepare(&f, &S, j, si, n, a, L);
    for (int q = 0; q < nt; q++) dc[q] = slg_eval(&f, t[q]);
    return 0;
}

int orc_slong_draw_many(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int si,
                        int n, int a, int L, double maxnats, const double *useg, const double *uoff, int nd,
                        double *t_out, int *status) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048 || si < 0 || si >= ESM_NSTOR || n < 0 || n >= m->N || a < 0 || a >= m->T || L < 1 ||
        L > m->T)
        return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    SlgMove f;
    slg_prepare(&f, &S, j, si, n, a, L);
    for (int q = 0; q < nd; q++) t_out[q] = slg_sample(&f, m->beta_v, maxnats, useg[q], uoff[q], &status[q]);
    return 0;
}

int orc_slong_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int si, int n,
                    int a, int L, double t) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    if (m->N > 256 || m->L > 2048 || si < 0 || si >= ESM_NSTOR || n < 0 || n >= m->N || a < 0 || a >= m->T || L < 1 ||
        L > m->T)
        return -1;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    SlgMove f;
    slg_prepare(&f, &S, j, si, n, a, L);
    slg_commit(&f, t);
    return 0;
}

/* ============================================================================================
 * Cross-node storage pairs (storage_cross.jl): unit A's state of charge rises by t and unit B's
 * (at A's node or a partner node) falls by ρ t, ρ = E_A/E_B (equal nominal energy), at the cyclic
 * instants a+1 … a+L. At the end steps a (nominal injection −κ_A t at A's node, +κ_A t at B's) and
 * b = a+L (the opposite) the flow between them is re-routed on the pair's ball template and the
 * leak is compensated pro rata by the gas unit of each leak node (prorata!, Jacobian test); the
 * interior steps a+1 … a+L−1 hold both units' storage terms and the balances of their nodes.
 * L = 1 is `sxpair`, a family of the local sweep (prep_sxpair! :147-167, sxpair_move! :170-190);
 * L drawn on 1 … T−1 is `sxlong` (prep_sxlong! :252-350, mv_eval! :370-427, commit_sxlong!
 * :430-461, storage_cross_phase! :501-544). Every unit of this system can charge, so the loss
 * share λ of sx_loss (:108) is 0: the losses go into the residuals of the units' nodes.
 * The arithmetic is the device's, one wave per move and replica: the end steps' items on lanes by
 * kind (xm_ends), the interior steps on lanes q mod 64, the 64-lane butterfly (xm_eval).
 * ============================================================================================ */

typedef struct {
    Sys *S;
    const Tables *tb;
    int j, a0, L, W, slot;
    int si[2], nd[2];      /* units A and B: carrier, node */
    double cu[2];          /* their instants move by cu t: 1 and −ρ */
    double kst[2], kap;    /* clamped capacities; κ_A = k_A h_A/Δt */
} XMove;

typedef struct {
    double Q, g, ak, P, e, y0;
} XKink;

static int xm_tau(const XMove *f, int q) { return (f->a0 + q) % f->S->m->T; }
static double xm_y(const XMove *f, int u, int tau) {
    Sys *S = f->S;
    int j = f->j;
    return AT(S->ye[f->si[u]], f->nd[u], tau);
}
/* scale of the end step e (0: step a, 1: step b) */
static double xm_scale(const XMove *f, int e) { return e == 0 ? -f->kap : f->kap; }
/* an angle entry c θ as finalize! keeps it (below 1e-14 it is zero) */
static double xm_th(double c, double th) {
    double a = c * th;
    return fabs(a) < 1e-14 ? 0.0 : a;
}
/* flow change per unit t of template line l at scale c (0: finalize! drops the line) */
static double xm_df(const XMove *f, double c, int l) {
    const Tables *tb = f->tb;
    int e = tb->sx_ln_line.a[l];
    double df = f->S->m->line_b[e] * (xm_th(c, tb->sx_ln_ta.a[l]) - xm_th(c, tb->sx_ln_tb.a[l]));
    return fabs(df) <= 1e-14 ? 0.0 : df;
}
/* the box interval of an instant y + c t (sx_boxint :231) */
static void xm_box(double y, double c, double *lo, double *hi) {
    if (c > 0) { *lo = -y / c; *hi = (1.0 - y) / c; }
    else { *lo = (1.0 - y) / c; *hi = -y / c; }
}
/* unit u at interior step q keeps both its instants in [0, 1] at t (its terms do not change) */
static int xm_inside(const XMove *f, int u, int q, double t) {
    int T = f->S->m->T, tau = xm_tau(f, q), tau1 = tau + 1 == T ? 0 : tau + 1;
    double y = xm_y(f, u, tau), y1 = xm_y(f, u, tau1), l0, h0, l1, h1;
    if (!(0.0 <= y && y <= 1.0 && 0.0 <= y1 && y1 <= 1.0)) return 0;
    xm_box(y, f->cu[u], &l0, &h0);
    xm_box(y1, f->cu[u], &l1, &h1);
    return fmax(l0, l1) <= t && t <= fmin(h0, h1);
}
/* prorata! (qadj = 0) of the gas unit of leak kk at end step e: 0 when no kink is made */
static int xm_kink(const XMove *f, int e, int kk, XKink *k) {
    Sys *S = f->S;
    const esm_model *m = S->m;
    int j = f->j, n = f->tb->sx_kk_node.a[kk], tau = xm_tau(f, e == 0 ? 0 : f->L);
    double g = xm_scale(f, e) * f->tb->sx_kk_g.a[kk];
    if (g == 0.0) return 0;
    double kap = S->kc[ESM_KGAS * m->N + n];
    if (!(fabs(kap) >= 1.0)) return 0;
    double w = AT(S->yg, n, tau);
    double P = 0.0 + kap * w, Qs = 0.0 + kap * clampd(w, 0.0, 1.0);
    if (fabs(P) < 1e-6) return 0;
    k->Q = AT(S->r, n, tau) + Qs;
    k->g = g;
    k->e = P - fmax(k->Q, 0.0);
    k->P = P;
    double sig = kap * w / P;
    k->ak = sig == 0.0 ? 0.0 : sig / kap;
    k->y0 = w;
    return 1;
}
static void xm_gen(const XMove *f, int n, double y, double *c, double *p) {
    const esm_model *m = f->S->m;
    double k = f->S->kc[ESM_KGAS * m->N + n];
    gen_cost(m->mc[ESM_KGAS * m->N + n], 1.0, k, k + m->kappa0, m->dt, m->c_viol, y, c, p);
}
/* storage term of unit u at step q at t: cost change; d at t (*d1), injection change (*dj) */
static double xm_stor(const XMove *f, int u, int q, double t, double *d1, double *dj) {
    const esm_model *m = f->S->m;
    int T = m->T, tau = xm_tau(f, q), tau1 = tau + 1 == T ? 0 : tau + 1;
    double y = xm_y(f, u, tau), y1 = xm_y(f, u, tau1), c = f->cu[u];
    double dye = q >= 1 ? c : 0.0, dye1 = q <= f->L - 1 ? c : 0.0;
    double w = f->kst[u] + m->kappa0, d0, j0, j1;
    double c0 = soc_cost(m, f->si[u], f->nd[u], f->kst[u], w, m->dt, m->c_viol, y, y1, &d0, &j0);
    double c1 = soc_cost(m, f->si[u], f->nd[u], f->kst[u], w, m->dt, m->c_viol, y + t * dye, y1 + t * dye1, d1, &j1);
    *dj = j1 - j0;
    return c1 - c0;
}
/* the items of a move's end steps: per end the kinked gas entries of the leak nodes, the lines, storage
 * A and B, the balances of the template's nodes (eval_cached! order); end a then end b */
static int xm_nper(const XMove *f) {
    const Tables *tb = f->tb;
    int sl = f->slot;
    return (tb->sx_kk_off.a[sl + 1] - tb->sx_kk_off.a[sl]) + (tb->sx_ln_off.a[sl + 1] - tb->sx_ln_off.a[sl]) + 2 +
           (tb->sx_nd_off.a[sl + 1] - tb->sx_nd_off.a[sl]);
}
/* the change of node item p's inflow per unit t at scale c: Σ ± df over its template lines (line
 * order; + where the node is the line's bus0) */
static double xm_inflow(const XMove *f, double c, int p) {
    const Tables *tb = f->tb;
    int l0 = tb->sx_ln_off.a[f->slot];
    double cf = 0.0;
    for (int q = tb->sx_nd_l.a[p]; q < tb->sx_nd_l.a[p + 1]; q++) {
        double df = xm_df(f, c, l0 + tb->sx_nl_pos.a[q]);
        if (df == 0.0) continue;
        cf += tb->sx_nl_sgn.a[q] > 0 ? df : -df;
    }
    return cf;
}
/* end step e at t: the terms of its kinks tk, lines tl, storage terms ts and balances tn, each with a
 * flag (hk, hl, hn: whether it counts; the storage terms always do); d[2] the units' d, bal (may be
 * NULL) the balances */
static void xm_end(const XMove *f, int e, double t, double *tk, int *hk, double *tl, int *hl, double *ts, double *tn,
                   double *d, double *bal) {
    Sys *S = f->S;
    const esm_model *m = S->m;
    const Tables *tb = f->tb;
    int j = f->j, sl = f->slot, tau = xm_tau(f, e == 0 ? 0 : f->L);
    double c = xm_scale(f, e), dj[2];
    int k0 = tb->sx_kk_off.a[sl], k1 = tb->sx_kk_off.a[sl + 1], l0 = tb->sx_ln_off.a[sl], l1 = tb->sx_ln_off.a[sl + 1];
    int n0 = tb->sx_nd_off.a[sl], n1 = tb->sx_nd_off.a[sl + 1];
    for (int kq = k0; kq < k1; kq++) {
        XKink k;
        hk[kq - k0] = 0;
        if (!xm_kink(f, e, kq, &k) || k.ak == 0.0) continue;
        double c0, p0, c1, p1;
        int n = tb->sx_kk_node.a[kq];
        xm_gen(f, n, k.y0, &c0, &p0);
        xm_gen(f, n, k.y0 + k.ak * sm_phi(k.Q, k.g, t), &c1, &p1);
        tk[kq - k0] = c1 - c0;
        hk[kq - k0] = 1;
    }
    for (int l = l0; l < l1; l++) {
        double df = xm_df(f, c, l);
        hl[l - l0] = 0;
        if (df == 0.0) continue;
        int ln = tb->sx_ln_line.a[l];
        double f0 = AT(S->F, ln, tau), se = m->line_s[ln], f1 = f0 + t * df;
        tl[l - l0] = m->c_viol * m->dt * (fmax(fabs(f1) - se, 0.0) - fmax(fabs(f0) - se, 0.0));
        hl[l - l0] = 1;
    }
    for (int u = 0; u < 2; u++) ts[u] = xm_stor(f, u, e == 0 ? 0 : f->L, t, &d[u], &dj[u]);
    for (int p = n0; p < n1; p++) {
        int n = tb->sx_nd_node.a[p], ik = tb->sx_nd_kk.a[p];
        double S0 = AT(S->r, n, tau), dr = 0.0;
        if (ik >= 0) {
            XKink k;
            if (xm_kink(f, e, k0 + ik, &k) && k.ak != 0.0) {
                double c0, p0, c1, p1;
                xm_gen(f, n, k.y0, &c0, &p0);
                xm_gen(f, n, k.y0 + k.ak * sm_phi(k.Q, k.g, t), &c1, &p1);
                dr -= p1 - p0;
            }
        }
        dr += t * xm_inflow(f, c, p);
        if (n == f->nd[0]) dr += dj[0];
        if (n == f->nd[1]) dr += dj[1];
        double S1 = S0 + dr;
        tn[p - n0] = m->dt * (phibal(m, S1) - phibal(m, S0));
        if (bal) bal[p - n0] = S1;
    }
}
/* the end steps' terms on the W lanes of the device's segment: kinks (item e·K + k), lines
 * (e·Lw + l), storage terms (2e + u), balances (e·Nn + p), each kind over both ends on lanes item
 * mod W, a lane summing its kinks, lines, storage terms and balances in increasing item index */
static void xm_ends(const XMove *f, double t, double *part) {
    const Tables *tb = f->tb;
    int sl = f->slot;
    int K = tb->sx_kk_off.a[sl + 1] - tb->sx_kk_off.a[sl], Lw = tb->sx_ln_off.a[sl + 1] - tb->sx_ln_off.a[sl];
    int Nn = tb->sx_nd_off.a[sl + 1] - tb->sx_nd_off.a[sl];
    double *tk = (double *)malloc(sizeof(double) * 2 * (K + Lw + Nn + 3)), *tl = tk + 2 * (K + 1),
           *tn = tl + 2 * (Lw + 1), ts[4], d[2];
    int *hk = (int *)malloc(sizeof(int) * 2 * (K + Lw + 2)), *hl = hk + 2 * (K + 1);
    for (int e = 0; e < 2; e++)
        xm_end(f, e, t, tk + e * K, hk + e * K, tl + e * Lw, hl + e * Lw, ts + 2 * e, tn + e * Nn, d, NULL);
    int W = f->W;
    for (int i = 0; i < 2 * K; i++)
        if (hk[i]) part[i % W] += tk[i];
    for (int i = 0; i < 2 * Lw; i++)
        if (hl[i]) part[i % W] += tl[i];
    for (int i = 0; i < 4; i++) part[i % W] += ts[i];
    for (int i = 0; i < 2 * Nn; i++) part[i % W] += tn[i];
    free(tk);
    free(hk);
}
/* the units' d and the template's balances at end step e at t */
static void xm_end_bal(const XMove *f, int e, double t, double *d, double *bal) {
    const Tables *tb = f->tb;
    int sl = f->slot;
    int K = tb->sx_kk_off.a[sl + 1] - tb->sx_kk_off.a[sl], Lw = tb->sx_ln_off.a[sl + 1] - tb->sx_ln_off.a[sl];
    int Nn = tb->sx_nd_off.a[sl + 1] - tb->sx_nd_off.a[sl];
    double *tk = (double *)malloc(sizeof(double) * (K + Lw + Nn + 3)), *tl = tk + K + 1, *tn = tl + Lw + 1, ts[2];
    int *hk = (int *)malloc(sizeof(int) * (K + Lw + 2)), *hl = hk + K + 1;
    xm_end(f, e, t, tk, hk, tl, hl, ts, tn, d, bal);
    free(tk);
    free(hk);
}
/* interior step q at t: ΔC; d[2] (the value at t = 0 for a unit inside its box), the balances of
 * the units' nodes bal[2] (one when they share a node); *moved 0 when both units are inside */
static double xm_int(const XMove *f, int q, double t, double *d, double *bal, int *moved) {
    Sys *S = f->S;
    const esm_model *m = S->m;
    int j = f->j, T = m->T, tau = xm_tau(f, q), tau1 = tau + 1 == T ? 0 : tau + 1;
    double D = 0.0, dj[2] = {0.0, 0.0};
    int mv = 0;
    for (int u = 0; u < 2; u++) {
        if (xm_inside(f, u, q, t)) {
            double y = xm_y(f, u, tau), y1 = xm_y(f, u, tau1), j0;
            soc_cost(m, f->si[u], f->nd[u], f->kst[u], f->kst[u] + m->kappa0, m->dt, m->c_viol, y, y1, &d[u], &j0);
            continue;
        }
        D += xm_stor(f, u, q, t, &d[u], &dj[u]);
        mv = 1;
    }
    if (f->nd[0] == f->nd[1]) {
        double S0 = AT(S->r, f->nd[0], tau), S1 = S0 + (dj[0] + dj[1]);
        D += m->dt * (phibal(m, S1) - phibal(m, S0));
        bal[0] = bal[1] = S1;
    } else {
        for (int u = 0; u < 2; u++) {
            double S0 = AT(S->r, f->nd[u], tau), S1 = S0 + dj[u];
            D += m->dt * (phibal(m, S1) - phibal(m, S0));
            bal[u] = S1;
        }
    }
    *moved = mv;
    return D;
}
static void xm_butterfly(double *part, int W) {
    for (int off = W / 2; off > 0; off >>= 1) {
        double nx[64];
        for (int l = 0; l < W; l++) nx[l] = part[l] + part[l ^ off];
        memcpy(part, nx, sizeof(double) * W);
    }
}
/* ΔC(t) in the device's arithmetic: a segment of W lanes per (move, replica); the end steps' items
 * on lanes by kind (xm_ends), then the interior steps q = 1 … L−1 on lanes q mod W, each lane in
 * increasing order; the lanes' sums combined by the W-lane butterfly */
static double xm_eval(const XMove *f, double t) {
    double part[64] = {0}, d[2], bal[2];
    g_nev++;
    xm_ends(f, t, part);
    for (int q = 1; q <= f->L - 1; q++) {
        int mv;
        double x = xm_int(f, q, t, d, bal, &mv);
        if (!mv) continue;
        part[q % f->W] += x;
    }
    xm_butterfly(part, f->W);
    return part[0];
}
static int xm_next_bp(const XMove *f, double tp, double sgn, double *t) {
    Sys *S = f->S;
    const esm_model *m = S->m;
    const Tables *tb = f->tb;
    int j = f->j, sl = f->slot, have = 0;
#define XCAND(bv)                                                                    \
    do {                                                                             \
        double b_ = (bv);                                                            \
        if (isfinite(b_) && sgn * (b_ - tp) > 0 && (!have || sgn * (*t - b_) > 0)) { \
            *t = b_;                                                                 \
            have = 1;                                                                \
        }                                                                            \
    } while (0)
    for (int e = 0; e < 2; e++) {
        int tau = xm_tau(f, e == 0 ? 0 : f->L);
        double c = xm_scale(f, e);
        for (int kq = tb->sx_kk_off.a[sl]; kq < tb->sx_kk_off.a[sl + 1]; kq++) {
            XKink k;
            if (!xm_kink(f, e, kq, &k)) continue;
            XCAND(-k.Q / k.g);
            if (k.ak != 0.0) {
                double s1 = k.ak * k.g, off = k.ak * k.Q;
                XCAND((0.0 - k.y0 - 0.0) / s1);
                XCAND((1.0 - k.y0 - 0.0) / s1);
                XCAND((0.0 - k.y0 - off) / s1);
                XCAND((1.0 - k.y0 - off) / s1);
            }
        }
        for (int l = tb->sx_ln_off.a[sl]; l < tb->sx_ln_off.a[sl + 1]; l++) {
            double df = xm_df(f, c, l);
            if (df == 0.0) continue;
            int ln = tb->sx_ln_line.a[l];
            double f0 = AT(S->F, ln, tau), se = m->line_s[ln];
            XCAND((se - f0) / df);
            XCAND((-se - f0) / df);
        }
    }
    for (int q = 1; q <= f->L; q++)
        for (int u = 0; u < 2; u++) {
            double y = xm_y(f, u, xm_tau(f, q));
            XCAND((0.0 - y) / f->cu[u]);
            XCAND((1.0 - y) / f->cu[u]);
        }
#undef XCAND
    return have;
}
/* the storage secondaries of unit u (storage_mask!: 3 per storage term) */
static double xm_sec(const XMove *f, int u, int k, double d) {
    const esm_model *m = f->S->m;
    int us = f->si[u] * m->N + f->nd[u];
    if (k == 0) return d;
    if (k == 1) return m->stor_eta_in[us] > 0 ? d - m->stor_eta_in[us] : 1.0;
    return -d * m->stor_eta_out[us] - 1.0;
}
/* the nearest sign change beyond x, strictly inside (a, b), of the storage secondaries (stor) or
 * of the balances, from their values at a and b (linear in between) */
static int xm_next_root(const XMove *f, int stor, double a, double b, double x, double sgn, double *xr) {
    int have = 0, sl = f->slot, nn = f->tb->sx_nd_off.a[sl + 1] - f->tb->sx_nd_off.a[sl];
    double da[2], db[2], *ba = (double *)malloc(sizeof(double) * (nn + 2)), *bb = (double *)malloc(sizeof(double) * (nn + 2));
#define XTEST(A_, B_)                                                                                      \
    do {                                                                                                   \
        double A = (A_), B = (B_);                                                                         \
        if ((A < 0) != (B < 0) && A != B) {                                                                \
            double r_ = a + (b - a) * A / (A - B);                                                         \
            if (sgn * (r_ - a) > 0 && sgn * (b - r_) > 0 && sgn * (r_ - x) > 0 && (!have || sgn * (*xr - r_) > 0)) { \
                *xr = r_;                                                                                  \
                have = 1;                                                                                  \
            }                                                                                              \
        }                                                                                                  \
    } while (0)
    for (int e = 0; e < 2; e++) {
        xm_end_bal(f, e, a, da, ba);
        xm_end_bal(f, e, b, db, bb);
        if (stor) {
            for (int u = 0; u < 2; u++)
                for (int k = 0; k < 3; k++) XTEST(xm_sec(f, u, k, da[u]), xm_sec(f, u, k, db[u]));
        } else {
            for (int p = 0; p < nn; p++) XTEST(ba[p], bb[p]);
        }
    }
    for (int q = 1; q < f->L; q++) {
        if (xm_inside(f, 0, q, a) && xm_inside(f, 0, q, b) && xm_inside(f, 1, q, a) && xm_inside(f, 1, q, b)) continue;
        int mv;
        xm_int(f, q, a, da, ba, &mv);
        xm_int(f, q, b, db, bb, &mv);
        if (stor) {
            for (int u = 0; u < 2; u++)
                for (int k = 0; k < 3; k++) XTEST(xm_sec(f, u, k, da[u]), xm_sec(f, u, k, db[u]));
        } else {
            for (int u = 0; u < (f->nd[0] == f->nd[1] ? 1 : 2); u++) XTEST(ba[u], bb[u]);
        }
    }
#undef XTEST
    free(ba);
    free(bb);
    return have;
}
static int xm_walk(const XMove *f, double beta, double sgn, double maxnats, BigPts *P) {
    double cmin = 0.0, tp = 0.0, delta = 0.0, cprev = 0.0;
    int nx = 0;
    while (1) {
        double t = 0.0;
        if (!xm_next_bp(f, tp, sgn, &t)) {
            double hp = 1e-6 * fmax(1.0, fabs(tp));
            double sl = (beta * xm_eval(f, tp + sgn * hp) - cprev) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, 1e3 * fmax(fabs(tp), 1e-3));
            else delta = delta == 0.0 ? fmax(fabs(tp), 1e-6) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;
        }
        double c = beta * xm_eval(f, t);
        double a0 = tp, xs = tp, x = 0.0;
        while (xm_next_root(f, 1, tp, t, xs, sgn, &x)) {
            double cx = beta * xm_eval(f, x);
            double xb = a0, y = 0.0;
            while (xm_next_root(f, 0, a0, x, xb, sgn, &y)) {
                double cy = beta * xm_eval(f, y);
                bpush(P, y, cy);
                cmin = fmin(cmin, cy);
                xb = y;
            }
            bpush(P, x, cx);
            cmin = fmin(cmin, cx);
            a0 = x;
            xs = x;
        }
        double xb = a0, y = 0.0;
        while (xm_next_root(f, 0, a0, t, xb, sgn, &y)) {
            double cy = beta * xm_eval(f, y);
            bpush(P, y, cy);
            cmin = fmin(cmin, cy);
            xb = y;
        }
        bpush(P, t, c);
        if (P->full) return -1;
        cmin = fmin(cmin, c);
        cprev = c;
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}
static double xm_sample(const XMove *f, double beta, double maxnats, double useg, double uoff, int *status) {
    static __thread double bufs[4][SMS_PMAX];
    BigPts PR = {bufs[0], bufs[1], 0, 0}, PL = {bufs[2], bufs[3], 0, 0};
    int w = xm_walk(f, beta, 1.0, maxnats, &PR);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    w = xm_walk(f, beta, -1.0, maxnats, &PL);
    if (w != 1) { *status = w == 0 ? 1 : 2; return 0.0; }
    return draw_points_big(&PL, &PR, useg, uoff, status);
}
/* units of the node with |κ| ≥ 1 among its dispatchable units (this system: the gas unit) */
/* logjac (src/local.jl:1874): Σ (|S| − 1) log|P(t)/P| over the kinks made, |S| from the model */
static double xm_logjac(const XMove *f, double t, int *any) {
    double part[64] = {0};
    int cnt = 0, sl = f->slot;
    for (int e = 0; e < 2; e++)
        for (int kq = f->tb->sx_kk_off.a[sl]; kq < f->tb->sx_kk_off.a[sl + 1]; kq++) {
            XKink k;
            if (!xm_kink(f, e, kq, &k)) continue;
            cnt++;
            double nm = (double)(gas_units(f->S, f->tb->sx_kk_node.a[kq]) - 1);
            part[(e * (f->tb->sx_kk_off.a[sl + 1] - f->tb->sx_kk_off.a[sl]) + kq - f->tb->sx_kk_off.a[sl]) % f->W] +=
                nm * log(fabs((fmax(k.Q + k.g * t, 0.0) + k.e) / k.P));
        }
    *any = cnt > 0;
    xm_butterfly(part, f->W);
    return part[0];
}
/* commit_sxlong! / commit_block!: every new value from the state before the move, then the writes */
static void xm_commit(const XMove *f, double t) {
    if (t == 0.0) return;
    Sys *S = f->S;
    const esm_model *m = S->m;
    const Tables *tb = f->tb;
    int j = f->j, sl = f->slot, n0 = tb->sx_nd_off.a[sl], nn = tb->sx_nd_off.a[sl + 1] - n0;
    int k0 = tb->sx_kk_off.a[sl], nk = tb->sx_kk_off.a[sl + 1] - k0, L = f->L;
    double *bal = (double *)malloc(sizeof(double) * 2 * (nn + 1)), *gy = (double *)malloc(sizeof(double) * 2 * (nk + 1));
    int *gw = (int *)calloc(2 * (nk + 1), sizeof(int));
    double *ir = (double *)malloc(sizeof(double) * 2 * (L + 1));
    int *imv = (int *)calloc(L + 1, sizeof(int));
    double d[2];
    for (int e = 0; e < 2; e++) {
        xm_end_bal(f, e, t, d, bal + e * nn);
        for (int q = 0; q < nk; q++) {
            XKink k;
            if (xm_kink(f, e, k0 + q, &k) && k.ak != 0.0) {
                gy[e * nk + q] = k.y0 + k.ak * sm_phi(k.Q, k.g, t);
                gw[e * nk + q] = 1;
            }
        }
    }
    for (int q = 1; q < L; q++) xm_int(f, q, t, d, ir + 2 * q, &imv[q]);
    for (int e = 0; e < 2; e++) {
        int tau = xm_tau(f, e == 0 ? 0 : L);
        double c = xm_scale(f, e);
        for (int p = 0; p < nn; p++) AT(S->r, tb->sx_nd_node.a[n0 + p], tau) = bal[e * nn + p];
        for (int q = 0; q < nk; q++)
            if (gw[e * nk + q]) AT(S->yg, tb->sx_kk_node.a[k0 + q], tau) = gy[e * nk + q];
        for (int l = tb->sx_ln_off.a[sl]; l < tb->sx_ln_off.a[sl + 1]; l++) {
            double df = xm_df(f, c, l);
            if (df != 0.0) AT(S->F, tb->sx_ln_line.a[l], tau) += t * df;
        }
        for (int q = tb->sx_th_off.a[sl]; q < tb->sx_th_off.a[sl + 1]; q++) {
            double a = xm_th(c, tb->sx_th_val.a[q]);
            if (a != 0.0) AT(S->th, tb->sx_th_node.a[q], tau) += t * a;
        }
    }
    for (int q = 1; q < L; q++) {
        if (!imv[q]) continue;
        int tau = xm_tau(f, q);
        AT(S->r, f->nd[0], tau) = ir[2 * q];
        if (f->nd[1] != f->nd[0]) AT(S->r, f->nd[1], tau) = ir[2 * q + 1];
    }
    for (int q = 1; q <= L; q++) {
        int tau = xm_tau(f, q);
        for (int u = 0; u < 2; u++) AT(S->ye[f->si[u]], f->nd[u], tau) += f->cu[u] * t;
    }
    free(bal); free(gy); free(gw); free(ir); free(imv);
}
/* a pair move of replica j: units A = (sa, na) and B = (sb, nb) at slot `slot` of A's node,
 * instants a+1 … a+L (a 0-based); 0 when it is not made (κ_A or κ_B below 1) */
static int xm_prepare(XMove *f, Sys *S, const Tables *tb, int j, int sa, int na, int sb, int nb, int slot, int a, int L,
                      int W) {
    const esm_model *m = S->m;
    int N = m->N;
    memset(f, 0, sizeof(*f));
    f->S = S; f->tb = tb; f->j = j; f->a0 = a; f->L = L; f->W = W; f->slot = slot;
    f->si[0] = sa; f->nd[0] = na; f->si[1] = sb; f->nd[1] = nb;
    f->kst[0] = S->kc[(ESM_KSTOR + sa) * N + na];
    f->kst[1] = S->kc[(ESM_KSTOR + sb) * N + nb];
    double EA = f->kst[0] * m->stor_hours[sa * N + na], EB = f->kst[1] * m->stor_hours[sb * N + nb];
    f->kap = EA / m->dt;
    if (f->kap < 1.0 || EB / m->dt < 1.0) return 0;
    f->cu[0] = 1.0;
    f->cu[1] = -(EA / EB);
    return 1;
}
static double xm_move(XMove *f, double useg, double uoff, double umh, double maxnats, int64_t *stats) {
    int status;
    g_nev = 0;
    double t = xm_sample(f, f->S->m->beta_v, maxnats, useg, uoff, &status);
    if (t != 0.0) {
        int any;
        double lj = xm_logjac(f, t, &any);
        if (any && !(umh < exp(fmin(0.0, lj)))) { t = 0.0; status = 2; }
    }
    long nev = g_nev;
    xm_commit(f, t);
    if (stats) stat_add(stats, status, t, nev);
    return t;
}
/* sx_partner (:62-90): among the units with hours ≥ hmin and κ = k h/Δt ≥ 1 at the slots of A's
 * node (A itself excluded), one with probability ∝ min(k_A, k_B), from the uniform u; 0: none */
static int xm_partner(const Sys *S, const Tables *tb, double hmin, int sa, int na, double u, int *slot, int *sb,
                      int *nb) {
    const esm_model *m = S->m;
    int N = m->N, found = 0;
    double kA = S->kc[(ESM_KSTOR + sa) * N + na], W = 0.0;
    for (int pass = 0; pass < 2; pass++) {
        double x = u * W, acc = 0.0;
        for (int s = tb->sx_poff.a[na]; s < tb->sx_poff.a[na + 1]; s++) {
            int mm = tb->sx_part.a[s];
            for (int si = 0; si < ESM_NSTOR; si++) {
                if (si == sa && mm == na) continue;
                if (!(m->stor_hours[si * N + mm] >= hmin)) continue;
                double kB = S->kc[(ESM_KSTOR + si) * N + mm];
                if (!(kB * m->stor_hours[si * N + mm] / m->dt >= 1.0)) continue;
                if (pass == 0) {
                    W += fmin(kA, kB);
                    continue;
                }
                acc += fmin(kA, kB);
                *slot = s; *sb = si; *nb = mm;
                found = 1;
                if (x <= acc) return 1;
            }
        }
        if (pass == 0 && !(W > 0.0)) return 0;
    }
    return found;
}
/* the storage units of the pairs (LM.sxunits): hours ≥ hmin, node-major in carrier order */
static int sx_units(const esm_model *m, double hmin, int *us) {
    int n = 0;
    for (int i = 0; i < m->N; i++)
        for (int si = 0; si < ESM_NSTOR; si++)
            if (m->stor_hours[si * m->N + i] >= hmin) us[n++] = si * m->N + i;
    return n;
}
/* sxpair at step τ of replica j, selection words s (unit A pick(s[0]), partner u01(s[2], s[3])) */
static void sxpair_move(Sys *S, const Tables *tb, int j, int tau, const uint32_t *s, double hmin, double useg,
                        double uoff, double umh, double maxnats, int64_t *stats) {
    const esm_model *m = S->m;
    int N = m->N, T = m->T, us[2 * 256];
    int nsx = sx_units(m, hmin, us);
    if (nsx == 0) return;
    int ua = us[pick(s[0], (uint32_t)nsx)], sa = ua / N, na = ua % N, slot, sb, nb;
    if (S->kc[(ESM_KSTOR + sa) * N + na] * m->stor_hours[ua] / m->dt < 1.0) return;
    if (!xm_partner(S, tb, hmin, sa, na, u01(s[2], s[3]), &slot, &sb, &nb)) return;
    XMove f;
    /* the device runs each move on the narrowest segment its template fits: 16 lanes (kinks ≤ 8,
     * lines ≤ 24, balances ≤ 16 per end), 32 (twice that), else 64 */
    int K = tb->sx_kk_off.a[slot + 1] - tb->sx_kk_off.a[slot], Lw = tb->sx_ln_off.a[slot + 1] - tb->sx_ln_off.a[slot];
    int Nn = tb->sx_nd_off.a[slot + 1] - tb->sx_nd_off.a[slot];
    int W = 64;
    if (2 * K <= 32 && 2 * Lw <= 96 && 2 * Nn <= 64) W = 32;
    if (2 * K <= 16 && 2 * Lw <= 48 && 2 * Nn <= 32) W = 16;
    if (!xm_prepare(&f, S, tb, j, sa, na, sb, nb, slot, (tau - 1 + T) % T, 1, W)) return;
    xm_move(&f, useg, uoff, umh, maxnats, stats);
}
/* the sxlong moves of a sweep (the same for every replica): per unit of the pairs with κ ≥ 1 and
 * move m (unit index ui among them all) the partner, a and L from (ui, m, 9420, c); shuffled by
 * Fisher–Yates with (i, 0, 9421, c); rounds of disjoint footprints by first fit (sx_rounds
 * :470-492), listed round by round. mv[5·i …]: sa, na, sb, nb, slot, then ma[i], mL[i]; returns
 * the count, round offsets in roff (nround + 1 entries) */
static int sxlong_moves(const Sys *S, const Tables *tb, int nmoves, double hmin, const uint32_t key[2], uint32_t ctr,
                        int *mv, int *ma, int *mL, int *roff, int *nround) {
    const esm_model *m = S->m;
    int N = m->N, T = m->T, us[2 * 256], cnt = 0;
    int nsx = sx_units(m, hmin, us);
    for (int ui = 0; ui < nsx; ui++) {
        int sa = us[ui] / N, na = us[ui] % N;
        if (S->kc[(ESM_KSTOR + sa) * N + na] * m->stor_hours[us[ui]] / m->dt < 1.0) continue;
        for (int mm = 0; mm < nmoves; mm++) {
            uint32_t c4[4] = {(uint32_t)ui, (uint32_t)mm, 9420u, ctr}, w[4];
            orc_philox(key, c4, w);
            int slot, sb, nb;
            if (!xm_partner(S, tb, hmin, sa, na, u01(w[0], w[1]), &slot, &sb, &nb)) continue;
            int *o = mv + 5 * cnt;
            o[0] = sa; o[1] = na; o[2] = sb; o[3] = nb; o[4] = slot;
            ma[cnt] = (int)pick(w[2], (uint32_t)T);
            mL[cnt] = 1 + (int)pick(w[3], (uint32_t)(T - 1));
            cnt++;
        }
    }
    for (int i = cnt - 1; i >= 1; i--) {
        uint32_t c4[4] = {(uint32_t)i, 0u, 9421u, ctr}, w[4];
        orc_philox(key, c4, w);
        int jx = (int)pick(w[0], (uint32_t)(i + 1)), tmp[7];
        memcpy(tmp, mv + 5 * i, sizeof(int) * 5); tmp[5] = ma[i]; tmp[6] = mL[i];
        memcpy(mv + 5 * i, mv + 5 * jx, sizeof(int) * 5); ma[i] = ma[jx]; mL[i] = mL[jx];
        memcpy(mv + 5 * jx, tmp, sizeof(int) * 5); ma[jx] = tmp[5]; mL[jx] = tmp[6];
    }
    /* first fit into rounds of disjoint footprints, then the list round by round */
    int *rd = (int *)malloc(sizeof(int) * (cnt + 1)), nr = 0;
    unsigned char *mark = (unsigned char *)calloc((size_t)(cnt + 1) * N, 1);
    for (int i = 0; i < cnt; i++) {
        int sl = mv[5 * i + 4], f0 = tb->sx_ft_off.a[sl], f1 = tb->sx_ft_off.a[sl + 1], r = 0;
        for (; r < nr; r++) {
            int clash = 0;
            for (int q = f0; q < f1 && !clash; q++) clash = mark[(size_t)r * N + tb->sx_ft_node.a[q]];
            if (!clash) break;
        }
        if (r == nr) nr++;
        for (int q = f0; q < f1; q++) mark[(size_t)r * N + tb->sx_ft_node.a[q]] = 1;
        rd[i] = r;
    }
    int *mv2 = (int *)malloc(sizeof(int) * 5 * (cnt + 1)), *ma2 = (int *)malloc(sizeof(int) * (cnt + 1)),
        *mL2 = (int *)malloc(sizeof(int) * (cnt + 1)), k = 0;
    roff[0] = 0;
    for (int r = 0; r < nr; r++) {
        for (int i = 0; i < cnt; i++)
            if (rd[i] == r) {
                memcpy(mv2 + 5 * k, mv + 5 * i, sizeof(int) * 5);
                ma2[k] = ma[i]; mL2[k] = mL[i];
                k++;
            }
        roff[r + 1] = k;
    }
    memcpy(mv, mv2, sizeof(int) * 5 * cnt);
    memcpy(ma, ma2, sizeof(int) * cnt);
    memcpy(mL, mL2, sizeof(int) * cnt);
    *nround = nr;
    free(rd); free(mark); free(mv2); free(ma2); free(mL2);
    return cnt;
}
/* the sxlong phase of the long-range storage sweep for replica j: move i of the list (round by
 * round) with its line draw (replica, i, 0x5C000000, c) and Jacobian test (replica, i,
 * 0x7C000000, c) */
static void sxlong_sweep(Sys *S, const Tables *tb, int j, int nmoves, double hmin, const uint32_t key[2], uint32_t ctr,
                         double maxnats, int64_t *st) {
    const esm_model *m = S->m;
    int cap = 2 * m->N * (nmoves > 0 ? nmoves : 1) + 1;
    int *mv = (int *)malloc(sizeof(int) * 5 * cap), *ma = (int *)malloc(sizeof(int) * cap),
        *mL = (int *)malloc(sizeof(int) * cap), *roff = (int *)malloc(sizeof(int) * (cap + 1)), nr;
    int cnt = m->T >= 2 && nmoves > 0 ? sxlong_moves(S, tb, nmoves, hmin, key, ctr, mv, ma, mL, roff, &nr) : 0;
    for (int i = 0; i < cnt; i++) {
        const int *o = mv + 5 * i;
        XMove f;
        if (!xm_prepare(&f, S, tb, j, o[0], o[1], o[2], o[3], o[4], ma[i], mL[i], 64)) continue;
        uint32_t cl[4] = {(uint32_t)j, (uint32_t)i, 0x5C000000u, ctr}, wl[4];
        orc_philox(key, cl, wl);
        uint32_t cm[4] = {(uint32_t)j, (uint32_t)i, 0x7C000000u, ctr}, wm[4];
        orc_philox(key, cm, wm);
        xm_move(&f, u01(wl[0], wl[1]), u01(wl[2], wl[3]), u01(wm[0], wm[1]), maxnats, st);
    }
    free(mv); free(ma); free(mL); free(roff);
}

/* ---- explicit pair moves, for the pins ---- */
static int xm_args_ok(const esm_model *m, int sa, int na, int sb, int nb, int a, int L) {
    return m->N <= 256 && m->L <= 2048 && sa >= 0 && sa < ESM_NSTOR && sb >= 0 && sb < ESM_NSTOR && na >= 0 &&
           na < m->N && nb >= 0 && nb < m->N && a >= 0 && a < m->T && L >= 1 && L <= m->T - 1;
}
/* the slot of partner node nb among node na's slots (-1: none) */
static int xm_slot(const Tables *tb, int na, int nb) {
    for (int s = tb->sx_poff.a[na]; s < tb->sx_poff.a[na + 1]; s++)
        if (tb->sx_part.a[s] == nb) return s;
    return -1;
}

int orc_sx_slot(const esm_model *m, int n, int k, int *part, int *nth, int *thn, double *thv, int *nlk, int *lkn,
                double *lkg, int *nft, int *ft) {
    Tables tb;
    if (m->N > 256 || n < 0 || n >= m->N || build_tables(m, &tb)) return -1;
    int s = tb.sx_poff.a[n] + k, rc = 0;
    if (k < 0 || s >= tb.sx_poff.a[n + 1]) rc = -2;
    if (!rc) {
        *part = tb.sx_part.a[s];
        *nth = tb.sx_th_off.a[s + 1] - tb.sx_th_off.a[s];
        for (int q = 0; q < *nth; q++) {
            thn[q] = tb.sx_th_node.a[tb.sx_th_off.a[s] + q];
            thv[q] = tb.sx_th_val.a[tb.sx_th_off.a[s] + q];
        }
        *nlk = tb.sx_kk_off.a[s + 1] - tb.sx_kk_off.a[s];
        for (int q = 0; q < *nlk; q++) {
            lkn[q] = tb.sx_kk_node.a[tb.sx_kk_off.a[s] + q];
            lkg[q] = tb.sx_kk_g.a[tb.sx_kk_off.a[s] + q];
        }
        *nft = tb.sx_ft_off.a[s + 1] - tb.sx_ft_off.a[s];
        for (int q = 0; q < *nft; q++) ft[q] = tb.sx_ft_node.a[tb.sx_ft_off.a[s] + q];
    }
    tables_free(&tb);
    return rc;
}

/* per storage-pair slot s < *nslot (at most cap): out[4 s …] = kinks K, lines Lw, balances Nn and
 * node–line entries per end (the sizes that decide the device's lane form) */
int orc_sx_sizes(const esm_model *m, int cap, int *nslot, int *out) {
    Tables tb;
    if (m->N > 256 || build_tables(m, &tb)) return -1;
    int n = tb.sx_kk_off.n - 1;
    *nslot = n;
    for (int sl = 0; sl < n && sl < cap; sl++) {
        out[4 * sl + 0] = tb.sx_kk_off.a[sl + 1] - tb.sx_kk_off.a[sl];
        out[4 * sl + 1] = tb.sx_ln_off.a[sl + 1] - tb.sx_ln_off.a[sl];
        out[4 * sl + 2] = tb.sx_nd_off.a[sl + 1] - tb.sx_nd_off.a[sl];
        out[4 * sl + 3] = tb.sx_nd_l.a[tb.sx_nd_off.a[sl + 1]] - tb.sx_nd_l.a[tb.sx_nd_off.a[sl]];
    }
    tables_free(&tb);
    return 0;
}
int orc_sx_partner(const esm_model *m, const double *k, double hmin, int sa, int na, double u, int *sb, int *nb) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    Tables tb;
    if (m->N > 256 || m->L > 2048 || build_tables(m, &tb)) return -1;
    (void)ptr; (void)lines;
    S.m = m;
    S.R = 1;
    for (int u = 0; u < ESM_NK * m->N; u++) S.kc[u] = clampd(k[u], m->klo[u], m->khi[u]);
    int slot;
    int rc = xm_partner(&S, &tb, hmin, sa, na, u, &slot, sb, nb);
    tables_free(&tb);
    return rc;
}

/* op: 0 ΔC at t[0 … nt−1] (into out), 1 draws with useg/uoff (t into out, status), 2 commit t[0] */
int orc_sx_move(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int sa, int na,
                int sb, int nb, int a, int L, int W, int op, const double *t, int nt, double maxnats, const double *useg,
                const double *uoff, double *out, int *status) {
    static __thread Sys S;
    int ptr[257], lines[2 * 2048];
    Tables tb;
    if (!xm_args_ok(m, sa, na, sb, nb, a, L) || (W != 1 && W != 64) || build_tables(m, &tb)) return -1;
    int slot = xm_slot(&tb, na, nb), rc = 0;
    sys_view(&S, m, R, v, k, r, F, ptr, lines);
    XMove f;
    if (slot < 0 || (sa == sb && na == nb)) rc = -2;
    else if (!xm_prepare(&f, &S, &tb, j, sa, na, sb, nb, slot, a, L, W)) rc = -3;
    if (!rc) {
        if (op == 0) for (int q = 0; q < nt; q++) out[q] = xm_eval(&f, t[q]);
        else if (op == 1) for (int q = 0; q < nt; q++) out[q] = xm_sample(&f, m->beta_v, maxnats, useg[q], uoff[q], &status[q]);
        else xm_commit(&f, t[0]);
    }
    tables_free(&tb);
    return rc;
}

/* the view of a replica set for a sweep: state pointers, clamped capacities, incident lines */
static void sys_setup(Sys *S0, const esm_model *m, int R, double *v, const double *k, double *r, double *F) {
    int N = m->N, L = m->L;
    size_t NTR = (size_t)N * m->T * R;
    S0->m = m;
    S0->R = R;
    S0->th = v;
    S0->yg = v + NTR;
    S0->ye[0] = v + 2 * NTR;
    S0->ye[1] = v + 3 * NTR;
    S0->r = r;
    S0->F = F;
    for (int u = 0; u < ESM_NK * N; u++) S0->kc[u] = clampd(k[u], m->klo[u], m->khi[u]);
    S0->adj_ptr = (int *)calloc(N + 1, sizeof(int));
    S0->adj_line = (int *)malloc(sizeof(int) * 2 * (L > 0 ? L : 1));
    int *fill = (int *)malloc(sizeof(int) * N);
    for (int e = 0; e < L; e++) {
        S0->adj_ptr[m->line_bus0[e] + 1]++;
        S0->adj_ptr[m->line_bus1[e] + 1]++;
    }
    for (int n = 0; n < N; n++) S0->adj_ptr[n + 1] += S0->adj_ptr[n];
    for (int n = 0; n < N; n++) fill[n] = S0->adj_ptr[n];
    for (int e = 0; e < L; e++) {
        S0->adj_line[fill[m->line_bus0[e]]++] = e;
        S0->adj_line[fill[m->line_bus1[e]]++] = e;
    }
    free(fill);
}

static void sys_free(Sys *S0) {
    free(S0->adj_ptr);
    free(S0->adj_line);
}

int orc_sweeps(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int nsweeps,
               const uint32_t key[2], uint32_t ctr0, double maxnats, double extra, const double *weights, int atoms,
               double pms, double plong, int long_moves, double long_hours, double pxpair, int xlong_moves,
               int64_t *stats, int nthreads) {
    int N = m->N, L = m->L, T = m->T;
    if (N > 256) return -1;
    SmsTables smst;
    if (sms_tables(m, &smst)) return -3;
    static const double wdef[ESM_NWEIGHT] = {0.3, 0.25, 0.03, 0.08, 0.04, 0.1, 0.1, 0.1, 0.3};
    const double *wts = weights ? weights : wdef;
    Sys S0;
    sys_setup(&S0, m, R, v, k, r, F);
    int *colour = (int *)malloc(sizeof(int) * N);
    int nc = colouring(m, S0.adj_ptr, S0.adj_line, colour);
    Tables tb;
    if (build_tables(m, &tb)) return -2;
    /* time-parity classes */
    int cfirst[3], ccount[3], ncls;
    if (T % 2 == 0) {
        ncls = 2; cfirst[0] = 0; ccount[0] = T / 2; cfirst[1] = 1; ccount[1] = T / 2;
    } else {
        ncls = 3; cfirst[0] = 0; ccount[0] = (T - 1) / 2; cfirst[1] = 1; ccount[1] = (T - 1) / 2;
        cfirst[2] = T - 1; ccount[2] = 1;
    }
    int usx[2 * 256];
    double px = pxpair > 0 && sx_units(m, long_hours, usx) > 0 ? pxpair : 0.0;
    int *counts = (int *)malloc(sizeof(int) * NW_ALL * T * (nsweeps > 0 ? nsweeps : 1));
    for (int sw = 0; sw < nsweeps; sw++)
        move_counts(m, extra, wts, px, key, ctr0 + (uint32_t)sw, counts + (size_t)sw * NW_ALL * T);
    int64_t st[ESM_NFAM * ESM_NSTAT];
    memset(st, 0, sizeof(st));
    if (nthreads < 1) nthreads = 1;
    /* replicas are independent: threads over replicas, each with its own statistics */
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : st[:ESM_NFAM * ESM_NSTAT])
    for (int j = 0; j < R; j++) {
        Sys S = S0;
        for (int sw = 0; sw < nsweeps; sw++) {
            uint32_t ctr = ctr0 + (uint32_t)sw;
            if (atoms) {
                for (int n = 0; n < N; n++)
                    for (int t = 0; t < T; t++) do_move(&S, j, 0, 0, n, t, key, 0u, ctr, maxnats, st + 5 * ESM_FAM_GEN);
                for (int c = 0; c < nc; c++)
                    for (int n = 0; n < N; n++) {
                        if (colour[n] != c || S.adj_ptr[n + 1] == S.adj_ptr[n]) continue;
                        for (int t = 0; t < T; t++)
                            do_move(&S, j, 1, 0, n, t, key, 1000u + (uint32_t)c, ctr, maxnats, st + 5 * ESM_FAM_THETA);
                    }
                for (int s = 0; s < ESM_NSTOR; s++)
                    for (int c = 0; c < ncls; c++)
                        for (int n = 0; n < N; n++)
                            for (int q = 0; q < ccount[c]; q++) {
                                int t = cfirst[c] + 2 * q;
                                do_move(&S, j, 2, s, n, t, key, 2000u + (uint32_t)(s * 4 + c), ctr, maxnats,
                                        st + 5 * ESM_FAM_SOC);
                            }
            }
            const int *cnt = counts + (size_t)sw * NW_ALL * T;
            for (int c = 0; c < ncls; c++)
                for (int q = 0; q < ccount[c]; q++)
                    random_step(&S, &tb, j, cfirst[c] + 2 * q, cnt, wts, px, long_hours, key, ctr, maxnats, st);
            if (pms > 0) sms_sweep(&S, &smst, j, pms, key, ctr, maxnats, st + 5 * ESM_FAM_SMS);
            if (plong > 0 &&
                slg_sweep(&S, j, plong, long_moves, long_hours, key, ctr, maxnats, st + 5 * ESM_FAM_SLONG,
                          st + 5 * ESM_FAM_SLEVEL) &&
                xlong_moves > 0)
                sxlong_sweep(&S, &tb, j, xlong_moves, long_hours, key, ctr, maxnats, st + 5 * ESM_FAM_SXLONG);
        }
    }
    if (stats) memcpy(stats, st, sizeof(st));
    sys_free(&S0);
    free(colour); free(counts);
    tables_free(&tb);
    return 0;
}



/* ============================================================================================
 * The reference's own schedule of the operation sweep, for the chain-marginal exactness check
 * (test/test_exact.jl:71-106): `local_sweep!` (src/local.jl:1927-2062) with the plain families.
 * Per replica and sweep: the time-parity classes in a shuffled order (:1946-1949); per step τ one
 * list holding every atom (a node unit vector of each field: angle, gas, the two states of
 * charge; :1963-1966) and round(extra·N) draws over the family weights (:1968-1971), shuffled
 * (:1972), each move selected, drawn and committed in that order (:1973-2044) with the
 * selection of prep_random! / prep_xloc! / prep_spair! (:1608-1725): unit, potential, spair
 * at a uniform node (and storage carrier); xpair at a uniform ordered pair of distinct nodes;
 * xloc / xlocc at a uniform node and a uniform neighbour; cell at a partition level drawn
 * ∝ its number of cells and a uniform cell; pair and cellpair are void with one dispatchable
 * unit per node and one injecting carrier per cell. Every random number of a replica comes from
 * its own Philox stream (the reference: MersenneTwister per chain, Xoshiro per step), so the
 * replicas are independent chains. No storage sweeps, no learned moves.
 * ============================================================================================ */

typedef struct {
    const uint32_t *key;
    uint32_t c0, c1, c2, c3;
    uint32_t w[4];
    int used;
} RStream;

static void rs_init(RStream *s, const uint32_t key[2], uint32_t c0, uint32_t c1, uint32_t c3) {
    s->key = key;
    s->c0 = c0;
    s->c1 = c1;
    s->c2 = 0xB0000000u;
    s->c3 = c3;
    s->used = 4;
}
static uint32_t rs_u32(RStream *s) {
    if (s->used == 4) {
        uint32_t c[4] = {s->c0, s->c1, s->c2++, s->c3};
        orc_philox(s->key, c, s->w);
        s->used = 0;
    }
    return s->w[s->used++];
}
static double rs_u01(RStream *s) {
    uint32_t hi = rs_u32(s), lo = rs_u32(s);
    return u01(hi, lo);
}
static int rs_pick(RStream *s, int n) { return (int)pick(rs_u32(s), (uint32_t)n); }

int orc_ref_sweeps(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int nsweeps,
                   const uint32_t key[2], uint32_t ctr0, double maxnats, double extra, const double *weights,
                   int64_t *stats, int nthreads) {
    int N = m->N, T = m->T;
    if (N > 256) return -1;
    static const double wdef[ESM_NWEIGHT] = {0.3, 0.25, 0.03, 0.08, 0.04, 0.1, 0.1, 0.1, 0.3};
    const double *wts = weights ? weights : wdef;
    double cw[ESM_NWEIGHT], acc = 0.0, tot = 0.0;
    for (int q = 0; q < ESM_NWEIGHT; q++) tot += wts[q];
    for (int q = 0; q < ESM_NWEIGHT; q++) { acc += wts[q]; cw[q] = acc / tot; }
    int nextra = (int)nearbyint(extra * N);
    Sys S0;
    sys_setup(&S0, m, R, v, k, r, F);
    Tables tb;
    if (build_tables(m, &tb)) return -2;
    int cfirst[3], ccount[3], ncls;
    if (T % 2 == 0) {
        ncls = 2; cfirst[0] = 0; ccount[0] = T / 2; cfirst[1] = 1; ccount[1] = T / 2;
    } else {
        ncls = 3; cfirst[0] = 0; ccount[0] = (T - 1) / 2; cfirst[1] = 1; ccount[1] = (T - 1) / 2;
        cfirst[2] = T - 1; ccount[2] = 1;
    }
    int nlist = 4 * N + nextra;
    int64_t st[ESM_NFAM * ESM_NSTAT];
    memset(st, 0, sizeof(st));
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : st[:ESM_NFAM * ESM_NSTAT])
    for (int j = 0; j < R; j++) {
        Sys S = S0;
        int *list = (int *)malloc(sizeof(int) * (nlist > 0 ? nlist : 1));
        for (int sw = 0; sw < nsweeps; sw++) {
            uint32_t ctr = ctr0 + (uint32_t)sw;
            RStream rc;
            rs_init(&rc, key, (uint32_t)j, 0xFFFFFFFFu, ctr);
            int corder[3] = {0, 1, 2};
            for (int i = ncls - 1; i >= 1; i--) {
                int q = rs_pick(&rc, i + 1), tmp = corder[i];
                corder[i] = corder[q];
                corder[q] = tmp;
            }
            for (int ci = 0; ci < ncls; ci++) {
                int c = corder[ci];
                for (int qq = 0; qq < ccount[c]; qq++) {
                    int tau = cfirst[c] + 2 * qq;
                    RStream rs;
                    rs_init(&rs, key, (uint32_t)j, (uint32_t)tau, ctr);
                    /* atoms: code −1 − (f·N + n), f 0 θ, 1 gas, 2 battery SoC, 3 H2 SoC; draws: family q */
                    int nl = 0;
                    for (int f = 0; f < 4; f++)
                        for (int n = 0; n < N; n++) {
                            if (f == 0 && S.adj_ptr[n + 1] == S.adj_ptr[n]) continue;
                            list[nl++] = -1 - (f * N + n);
                        }
                    for (int d = 0; d < nextra; d++) {
                        double u = rs_u01(&rs);
                        int fq = ESM_NWEIGHT - 1;
                        for (int q = ESM_NWEIGHT - 1; q >= 0; q--)
                            if (cw[q] >= u) fq = q;
                        list[nl++] = fq;
                    }
                    for (int i = nl - 1; i >= 1; i--) {
                        int q = rs_pick(&rs, i + 1), tmp = list[i];
                        list[i] = list[q];
                        list[q] = tmp;
                    }
                    for (int i = 0; i < nl; i++) {
                        int code = list[i];
                        if (code < 0) {
                            int f = (-1 - code) / N, n = (-1 - code) % N;
                            double useg = rs_u01(&rs), uoff = rs_u01(&rs);
                            if (f == 0) do_move_u(&S, j, 1, 0, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_THETA);
                            else if (f == 1) do_move_u(&S, j, 0, 0, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_GEN);
                            else do_move_u(&S, j, 2, f - 2, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_SOC);
                            continue;
                        }
                        if (code == W_PAIR || code == W_CELLPAIR) continue;     /* void in this system */
                        int n = rs_pick(&rs, N);
                        if (code == W_UNIT) {
                            double useg = rs_u01(&rs), uoff = rs_u01(&rs);
                            do_move_u(&S, j, 0, 0, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_UNIT);
                        } else if (code == W_POT) {
                            if (S.adj_ptr[n + 1] == S.adj_ptr[n]) continue;
                            double useg = rs_u01(&rs), uoff = rs_u01(&rs);
                            do_move_u(&S, j, 1, 0, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_POT);
                        } else if (code == W_SPAIR) {
                            int sc = rs_pick(&rs, ESM_NSTOR);
                            double useg = rs_u01(&rs), uoff = rs_u01(&rs);
                            do_move_u(&S, j, 3, sc, n, tau, useg, uoff, maxnats, st + 5 * ESM_FAM_SPAIR);
                        } else {
                            int kind, tid, fam;
                            if (code == W_CELL) {
                                double u = rs_u01(&rs);
                                int lv = tb.nlev - 1;
                                for (int q = tb.nlev - 1; q >= 0; q--)
                                    if (tb.qcum.a[q] >= u) lv = q;
                                int ncell = tb.lev_off.a[lv + 1] - tb.lev_off.a[lv];
                                tid = tb.cl0 + tb.lev_off.a[lv] + rs_pick(&rs, ncell);
                                kind = SP_CELL;
                                fam = ESM_FAM_CELL;
                            } else if (code == W_XPAIR) {
                                int mm = rs_pick(&rs, N);
                                if (mm == n) continue;
                                tid = tb.xp0 + n * N + mm;
                                kind = SP_XPAIR;
                                fam = ESM_FAM_XPAIR;
                            } else {
                                int np = tb.xp_off.a[n + 1] - tb.xp_off.a[n];
                                if (np == 0) continue;
                                tid = tb.xp_off.a[n] + rs_pick(&rs, np);
                                kind = code == W_XLOCC ? SP_XLOCC : SP_XLOC;
                                fam = code == W_XLOCC ? ESM_FAM_XLOCC : ESM_FAM_XLOC;
                            }
                            double useg = rs_u01(&rs), uoff = rs_u01(&rs), umh = rs_u01(&rs);
                            sm_move(&S, &tb, j, tau, tid, kind, useg, uoff, umh, maxnats, st + 5 * fam);
                        }
                    }
                }
            }
        }
        free(list);
    }
    if (stats) memcpy(stats, st, sizeof(st));
    sys_free(&S0);
    tables_free(&tb);
    return 0;
}
