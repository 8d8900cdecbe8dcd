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

#define PMAX 48
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
    *status = 0;
    int np = PL.n + 1 + PR.n;
    double pts[2 * PMAX + 1], cs[2 * PMAX + 1], lm[2 * PMAX];
    for (int q = 0; q < PL.n; q++) {
        pts[q] = PL.p[PL.n - 1 - q];
        cs[q] = PL.c[PL.n - 1 - q];
    }
    pts[PL.n] = 0.0;
    cs[PL.n] = 0.0;
    for (int q = 0; q < PR.n; q++) {
        pts[PL.n + 1 + q] = PR.p[q];
        cs[PL.n + 1 + q] = PR.c[q];
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

/* commit_block! / apply_step! */
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

/* one move of family fam at (n, τ) of replica j; s = storage carrier */
static void do_move(Sys *S, int j, int fam, int s, int n, int tau, const uint32_t key[2], uint32_t round,
                    uint32_t ctr, double maxnats, int64_t *stats) {
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
    uint32_t c4[4] = {(uint32_t)j, (uint32_t)(n * T + tau), round, ctr}, w[4];
    orc_philox(key, c4, w);
    int status;
    g_nev = 0;
    double t = sample_line(&b, m->beta_v, maxnats, u01(w[0], w[1]), u01(w[2], w[3]), &status);
    long nev = g_nev;
    commit(&b, t);
    stat_add(stats, status, t, nev);
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

int orc_sweeps(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int nsweeps,
               const uint32_t key[2], uint32_t ctr0, double maxnats, int64_t *stats, int nthreads) {
    int N = m->N, L = m->L, T = m->T;
    if (N > 256) return -1;
    size_t NTR = (size_t)N * T * R;
    Sys S0;
    S0.m = m;
    S0.R = R;
    S0.th = v;
    S0.yg = v + NTR;
    S0.ye[0] = v + 2 * NTR;
    S0.ye[1] = v + 3 * NTR;
    S0.r = r;
    S0.F = F;
    for (int u = 0; u < ESM_NK * N; u++) S0.kc[u] = clampd(k[u], m->klo[u], m->khi[u]);
    S0.adj_ptr = (int *)calloc(N + 1, sizeof(int));
    S0.adj_line = (int *)malloc(sizeof(int) * 2 * L);
    int *fill = (int *)malloc(sizeof(int) * N), *colour = (int *)malloc(sizeof(int) * N);
    for (int e = 0; e < L; e++) {
        S0.adj_ptr[m->line_bus0[e] + 1]++;
        S0.adj_ptr[m->line_bus1[e] + 1]++;
    }
    for (int n = 0; n < N; n++) S0.adj_ptr[n + 1] += S0.adj_ptr[n];
    for (int n = 0; n < N; n++) fill[n] = S0.adj_ptr[n];
    for (int e = 0; e < L; e++) {
        S0.adj_line[fill[m->line_bus0[e]]++] = e;
        S0.adj_line[fill[m->line_bus1[e]]++] = e;
    }
    int nc = colouring(m, S0.adj_ptr, S0.adj_line, colour);
    /* time-parity classes */
    int cfirst[3], ccount[3], ncls;
    if (T % 2 == 0) {
        ncls = 2; cfirst[0] = 0; ccount[0] = T / 2; cfirst[1] = 1; ccount[1] = T / 2;
    } else {
        ncls = 3; cfirst[0] = 0; ccount[0] = (T - 1) / 2; cfirst[1] = 1; ccount[1] = (T - 1) / 2;
        cfirst[2] = T - 1; ccount[2] = 1;
    }
    int64_t st[20];
    memset(st, 0, sizeof(st));
    if (nthreads < 1) nthreads = 1;
    /* replicas are independent: threads over replicas, each with its own statistics */
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : st[:20])
    for (int j = 0; j < R; j++) {
        Sys S = S0;
        for (int sw = 0; sw < nsweeps; sw++) {
            uint32_t ctr = ctr0 + (uint32_t)sw;
            for (int n = 0; n < N; n++)
                for (int t = 0; t < T; t++) do_move(&S, j, 0, 0, n, t, key, 0u, ctr, maxnats, st + 0);
            for (int c = 0; c < nc; c++)
                for (int n = 0; n < N; n++) {
                    if (colour[n] != c || S.adj_ptr[n + 1] == S.adj_ptr[n]) continue;
                    for (int t = 0; t < T; t++) do_move(&S, j, 1, 0, n, t, key, 1000u + (uint32_t)c, ctr, maxnats, st + 5);
                }
            for (int fam = 2; fam <= 3; fam++)
                for (int s = 0; s < ESM_NSTOR; s++)
                    for (int c = 0; c < ncls; c++)
                        for (int n = 0; n < N; n++)
                            for (int q = 0; q < ccount[c]; q++) {
                                int t = cfirst[c] + 2 * q;
                                do_move(&S, j, fam, s, n, t, key, (fam == 2 ? 2000u : 3000u) + (uint32_t)(s * 4 + c),
                                        ctr, maxnats, st + 5 * fam);
                            }
        }
    }
    if (stats) memcpy(stats, st, sizeof(st));
    free(S0.adj_ptr); free(S0.adj_line); free(fill); free(colour);
    return 0;
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
