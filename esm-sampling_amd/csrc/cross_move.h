// cross_move.h — the cross-node storage pair move (src/storage_cross.jl) on the device, shared by
// the single-instant form `sxpair` (a family of the local sweep, k_random.hip) and the interval
// form `sxlong` (the long-range storage sweep, k_cross.hip).
//
// Unit A's state of charge rises by t and unit B's (at A's node or at a neighbour) falls by ρ t,
// ρ = E_A/E_B (equal nominal energy), at the cyclic instants a+1 … a+L. At the end steps a
// (nominal injection −κ_A t at A's node, +κ_A t at B's) and b = a+L (the opposite) the flow
// between them is re-routed on the pair's ball template (angles c θ with c = ∓κ_A, entries and
// lines dropped below 1e-14 as `finalize!` does) and the leak is compensated pro rata by the gas
// unit of each leak node (`prorata!` with its kink, the Jacobian test); the interior steps hold
// both units' storage terms and the balances of their nodes, which do not change while both
// instants of a unit stay in [0, 1] (`mv_eval!`, :370-427). Every unit of this system can
// charge, so the losses go into the residuals of the units' nodes (sx_loss = 0, :108).
//
// Lanes: a segment of W lanes (W = 1 for sxpair, one thread per replica; W = 64 for sxlong, a
// wave per replica) owns the move; lane q mod W owns step q (0 … L: step a+q), lane 0 end a and
// lane L mod W end b, and each lane sums its steps' terms in increasing q; ΔC(t) is the segment's
// butterfly sum. The walk is `walk_stor` (storage_walk.h) without recording: the secondaries at
// the ends of an interval are evaluated again by the root searches.
#pragma once
#include "storage_walk.h"

namespace esm {

// the storage-pair templates (MoveTablesDev sx_*)
template <typename S>
struct SxTab {
    const int *poff, *part, *th_off, *th_node, *kk_off, *kk_node, *ln_off, *ln_line, *nd_off, *nd_node, *nd_kk, *nd_l,
        *nl_pos, *nl_sgn, *ft_off, *ft_node;
    const S *th_val, *kk_g, *ln_ta, *ln_tb;
};

// the units of the pairs (storage units with at least hmin hours) and the templates
template <typename S>
struct SxArgs {
    SxTab<S> tb;
    const int *units;   // [nsx] carrier·N + node, node-major in carrier order
    int nsx;
    S hmin;
};

// the pair arguments of a handle (host)
template <typename S>
inline SxArgs<S> sx_args(const esm_handle *h) {
    const MoveTablesDev &d = h->dmt;
    SxArgs<S> a;
    a.tb = {d.sx_poff,   d.sx_part,  d.sx_th_off, d.sx_th_node, d.sx_kk_off, d.sx_kk_node, d.sx_ln_off, d.sx_ln_line,
            d.sx_nd_off, d.sx_nd_node, d.sx_nd_kk, d.sx_nd_l,   d.sx_nl_pos, d.sx_nl_sgn,  d.sx_ft_off, d.sx_ft_node,
            (const S *)d.sx_th_val, (const S *)d.sx_kk_g, (const S *)d.sx_ln_ta, (const S *)d.sx_ln_tb};
    a.units = h->d_sxu;
    a.nsx = h->nsx;
    a.hmin = (S)h->sweep_long_hours;
    return a;
}

// `sx_partner` (:62-90): among the units with hours ≥ hmin and κ = k h/Δt ≥ 1 at the slots of A's
// node (A itself excluded), one with probability ∝ min(k_A, k_B) from the uniform u; false: none.
// A function of the capacities only (the same in every replica).
template <typename S>
__device__ bool sx_partner(const DevModel<S> &m, const SxArgs<S> &sx, int sa, int na, double u, int &slot, int &sb,
                           int &nb) {
    const int N = m.N;
    const S kA = m.kc[(ESM_KSTOR + sa) * N + na];
    S W = 0;
    bool found = false;
    for (int pass = 0; pass < 2; pass++) {
        const S x = (S)u * W;
        S acc = 0;
        for (int s = sx.tb.poff[na]; s < sx.tb.poff[na + 1]; s++) {
            const int mm = sx.tb.part[s];
            for (int si = 0; si < ESM_NSTOR; si++) {
                if (si == sa && mm == na) continue;
                const S h = m.hours[si * N + mm];
                if (!(h >= sx.hmin)) continue;
                const S kB = m.kc[(ESM_KSTOR + si) * N + mm];
                if (!(kB * h / m.dt >= S(1))) continue;
                if (pass == 0) {
                    W += fmin(kA, kB);
                    continue;
                }
                acc += fmin(kA, kB);
                slot = s;
                sb = si;
                nb = mm;
                found = true;
                if (x <= acc) return true;
            }
        }
        if (pass == 0 && !(W > S(0))) return false;
    }
    return found;
}

template <typename S>
struct XKinkD {
    S Q, g, ak, P, e, y0;
};

template <typename S, int W>
struct FXPair {
    enum { REC_NONE, REC_B, REC_X, REC_A, REC_T };
    static constexpr bool CACHED = false;   // the root searches evaluate the interval ends again
    const DevModel<S> *m;
    SxTab<S> tb;
    const S *ye0, *ye1;    // the units' state-of-charge planes at (node, ·, replica), stride R
    const S *rp, *Fp, *gp; // residual, flow and gas planes at (·, ·, replica)
    int T, R, lane, L, a0, slot;
    int nd0, nd1;
    S cu0, cu1, kst0, kst1, w0, w1, kap, kmin;
    StoPar<S> par0, par1;
    S llo, lhi;            // t in [llo, lhi]: the lane's interior steps keep every instant in [0, 1]
    mutable int nev = 0;
    // the lane's nearest clamp point of its rising instants beyond the walk's last point (the
    // points are fixed during a move and a walk only moves forward)
    mutable S bp_tp = 0, bp_sgn = 0, bp_next = 0;
    mutable bool bp_have = false;

    __device__ int tau(int q) const {
        const int x = a0 + q;
        return x >= T ? x - T : x;
    }
    __device__ int next(int t) const { return t + 1 == T ? 0 : t + 1; }
    __device__ S y(int u, int t) const { return (u == 0 ? ye0 : ye1)[(size_t)t * R]; }
    __device__ S rv(int n, int t) const { return rp[((size_t)n * T + t) * R]; }
    __device__ S Fv(int e, int t) const { return Fp[((size_t)e * T + t) * R]; }
    __device__ S gv(int n, int t) const { return gp[((size_t)n * T + t) * R]; }
    __device__ int first(int q0) const { return lane >= q0 ? lane : lane + W * ((q0 - lane + W - 1) / W); }
    __device__ bool owns_a() const { return lane == 0; }
    __device__ bool owns_b() const { return lane == L % W; }
    __device__ S sc(int e) const { return e == 0 ? -kap : kap; }
    __device__ static S thv(S c, S v) {
        const S a = c * v;
        return fabs(a) < S(1e-14) ? S(0) : a;
    }
    // flow change per unit t of template line l at scale c (0: dropped)
    __device__ S dfl(S c, int l) const {
        const S df = m->lb[tb.ln_line[l]] * (thv(c, tb.ln_ta[l]) - thv(c, tb.ln_tb[l]));
        return fabs(df) <= S(1e-14) ? S(0) : df;
    }
    // `prorata!` (qadj = 0) of the gas unit of leak kq at end step e
    __device__ bool kink(int e, int kq, XKinkD<S> &k) const {
        const int n = tb.kk_node[kq], t = tau(e == 0 ? 0 : L);
        const S g = sc(e) * tb.kk_g[kq];
        if (g == S(0)) return false;
        const S kp = m->kc[ESM_KGAS * m->N + n];
        if (!(fabs(kp) >= kmin)) return false;
        const S w = gv(n, t);
        const S P = S(0) + kp * w, Qs = S(0) + kp * clampv(w, S(0), S(1));
        if (fabs(P) < S(1e-6)) return false;
        k.Q = rv(n, t) + Qs;
        k.g = g;
        k.e = P - fmax(k.Q, S(0));
        k.P = P;
        const S sig = kp * w / P;
        k.ak = sig == S(0) ? S(0) : sig / kp;
        k.y0 = w;
        return true;
    }
    __device__ static S kphi(S Q, S g, S t) { return fmax(Q + g * t, S(0)) - fmax(Q, S(0)); }
    __device__ void gen(int n, S yy, S &c, S &p) const {
        const S k = m->kc[ESM_KGAS * m->N + n];
        gen_cost(m->mc[ESM_KGAS * m->N + n], k, k + m->kappa0, m->dt, m->cv, yy, c, p);
    }
    // storage term of unit u at step q at t: cost change; d at t, injection change
    __device__ S sterm(int u, int q, S t, S &d1, S &dj) const {
        const int ta = tau(q);
        const S y0 = y(u, ta), y1 = y(u, next(ta)), c = u == 0 ? cu0 : cu1;
        const S dye = q >= 1 ? c : S(0), dye1 = q <= L - 1 ? c : S(0);
        const StoPar<S> &p = u == 0 ? par0 : par1;
        const S k = u == 0 ? kst0 : kst1, w = u == 0 ? w0 : w1;
        S d0, j0, j1;
        const S c0 = soc_cost_p(p, m->dt, m->cv, k, w, y0, y1, d0, j0);
        const S c1 = soc_cost_p(p, m->dt, m->cv, k, w, y0 + t * dye, y1 + t * dye1, d1, j1);
        dj = j1 - j0;
        return c1 - c0;
    }
    __device__ S sec(int u, int k, S d) const {
        const StoPar<S> &p = u == 0 ? par0 : par1;
        if (k == 0) return d;
        if (k == 1) return p.ein > 0 ? d - p.ein : S(1);
        return -d * p.eout - S(1);
    }
    // the balance of template node p at end step e at t, given the units' injection changes
    __device__ S node_S(int e, int p, S t, S djA, S djB) const {
        const int n = tb.nd_node[p], ta = tau(e == 0 ? 0 : L);
        const S c = sc(e);
        S dr = 0;
        const int ik = tb.nd_kk[p];
        if (ik >= 0) {
            XKinkD<S> k;
            if (kink(e, tb.kk_off[slot] + ik, k) && k.ak != S(0)) {
                S c0, p0, c1, p1;
                gen(n, k.y0, c0, p0);
                gen(n, k.y0 + k.ak * kphi(k.Q, k.g, t), c1, p1);
                dr -= p1 - p0;
            }
        }
        const int l0 = tb.ln_off[slot];
        for (int q = tb.nd_l[p]; q < tb.nd_l[p + 1]; q++) {
            const int pos = l0 + tb.nl_pos[q];
            const S df = dfl(c, pos);
            if (df == S(0)) continue;
            const S f0 = Fv(tb.ln_line[pos], ta), dd = (f0 + t * df) - f0;
            dr += tb.nl_sgn[q] > 0 ? dd : -dd;
        }
        if (n == nd0) dr += djA;
        if (n == nd1) dr += djB;
        return rv(n, ta) + dr;
    }
    // end step e at t (eval_cached! order: kinked gas, lines, storage A and B, balances)
    __device__ S end_term(int e, S t) const {
        const int ta = tau(e == 0 ? 0 : L), q = e == 0 ? 0 : L;
        const S c = sc(e);
        S D = 0;
        for (int kq = tb.kk_off[slot]; kq < tb.kk_off[slot + 1]; kq++) {
            XKinkD<S> k;
            if (!kink(e, kq, k) || k.ak == S(0)) continue;
            S c0, p0, c1, p1;
            const int n = tb.kk_node[kq];
            gen(n, k.y0, c0, p0);
            gen(n, k.y0 + k.ak * kphi(k.Q, k.g, t), c1, p1);
            D += c1 - c0;
        }
        for (int l = tb.ln_off[slot]; l < tb.ln_off[slot + 1]; l++) {
            const S df = dfl(c, l);
            if (df == S(0)) continue;
            const int ln = tb.ln_line[l];
            const S f0 = Fv(ln, ta), se = m->ls[ln], f1 = f0 + t * df;
            D += m->cv * m->dt * (fmax(fabs(f1) - se, S(0)) - fmax(fabs(f0) - se, S(0)));
        }
        S d, djA, djB;
        D += sterm(0, q, t, d, djA);
        D += sterm(1, q, t, d, djB);
        for (int p = tb.nd_off[slot]; p < tb.nd_off[slot + 1]; p++) {
            const S S0 = rv(tb.nd_node[p], ta), S1 = node_S(e, p, t, djA, djB);
            D += m->dt * (phibal(S1, m->c_shed, m->c_dump) - phibal(S0, m->c_shed, m->c_dump));
        }
        return D;
    }
    // the box interval of an instant y + c t
    __device__ static void box(S yy, S c, S &lo, S &hi) {
        if (c > 0) {
            lo = -yy / c;
            hi = (S(1) - yy) / c;
        } else {
            lo = (S(1) - yy) / c;
            hi = -yy / c;
        }
    }
    __device__ bool inside(int u, int q, S t) const {
        const int ta = tau(q);
        const S y0 = y(u, ta), y1 = y(u, next(ta)), c = u == 0 ? cu0 : cu1;
        if (!(S(0) <= y0 && y0 <= S(1) && S(0) <= y1 && y1 <= S(1))) return false;
        S l0, h0, l1, h1;
        box(y0, c, l0, h0);
        box(y1, c, l1, h1);
        return fmax(l0, l1) <= t && t <= fmin(h0, h1);
    }
    // interior step q at t: false when both units stay inside (nothing changes); else its term x,
    // the units' d (the value at t = 0 for a unit inside) and the balances bA, bB of their nodes
    __device__ bool int_term(int q, S t, S &x, S &dA, S &dB, S &bA, S &bB) const {
        const int ta = tau(q);
        S D = 0, djA = 0, djB = 0;
        bool mv = false;
        for (int u = 0; u < 2; u++) {
            S &d = u == 0 ? dA : dB;
            S &dj = u == 0 ? djA : djB;
            if (inside(u, q, t)) {
                S j0;
                soc_cost_p(u == 0 ? par0 : par1, m->dt, m->cv, u == 0 ? kst0 : kst1, u == 0 ? w0 : w1, y(u, ta),
                           y(u, next(ta)), d, j0);
                continue;
            }
            D += sterm(u, q, t, d, dj);
            mv = true;
        }
        if (nd0 == nd1) {
            const S S0 = rv(nd0, ta), S1 = S0 + (djA + djB);
            D += m->dt * (phibal(S1, m->c_shed, m->c_dump) - phibal(S0, m->c_shed, m->c_dump));
            bA = bB = S1;
        } else {
            const S S0A = rv(nd0, ta), S1A = S0A + djA;
            D += m->dt * (phibal(S1A, m->c_shed, m->c_dump) - phibal(S0A, m->c_shed, m->c_dump));
            const S S0B = rv(nd1, ta), S1B = S0B + djB;
            D += m->dt * (phibal(S1B, m->c_shed, m->c_dump) - phibal(S0B, m->c_shed, m->c_dump));
            bA = S1A;
            bB = S1B;
        }
        x = D;
        return mv;
    }

    // ΔC(t): the lane's steps in increasing q, summed over the segment
    __device__ S eval(S t, int) const {
        nev++;
        const bool oa = owns_a(), ob = owns_b();
        S xa = 0, xb = 0;
        if (oa) xa = end_term(0, t);
        if (ob) xb = end_term(1, t);
        S D = 0;
        if (oa) D += xa;
        if (!(llo <= t && t <= lhi))
            for (int q = first(1); q <= L - 1; q += W) {
                S x, d0, d1, b0, b1;
                if (int_term(q, t, x, d0, d1, b0, b1)) D += x;
            }
        if (ob) D += xb;
        return seg_sum<S, W>(D);
    }
    __device__ void shift(bool) const {}

    __device__ bool next_bp(S tp, S sgn, S &t) const {
        bool have = false;
        S best = 0;
        auto cand = [&](S b) {
            if (isfinite(b) && sgn * (b - tp) > 0 && (!have || sgn * (best - b) > 0)) {
                best = b;
                have = true;
            }
        };
        for (int e = 0; e < 2; e++) {
            if (!(e == 0 ? owns_a() : owns_b())) continue;
            const int ta = tau(e == 0 ? 0 : L);
            const S c = sc(e);
            for (int kq = tb.kk_off[slot]; kq < tb.kk_off[slot + 1]; kq++) {
                XKinkD<S> k;
                if (!kink(e, kq, k)) continue;
                cand(-k.Q / k.g);
                if (k.ak != S(0)) {
                    const S s1 = k.ak * k.g, off = k.ak * k.Q;
                    cand((S(0) - k.y0 - S(0)) / s1);
                    cand((S(1) - k.y0 - S(0)) / s1);
                    cand((S(0) - k.y0 - off) / s1);
                    cand((S(1) - k.y0 - off) / s1);
                }
            }
            for (int l = tb.ln_off[slot]; l < tb.ln_off[slot + 1]; l++) {
                const S df = dfl(c, l);
                if (df == S(0)) continue;
                const int ln = tb.ln_line[l];
                const S f0 = Fv(ln, ta), se = m->ls[ln];
                cand((se - f0) / df);
                cand((-se - f0) / df);
            }
        }
        // the clamp points of the rising instants a+1 … a+L (cached while the walk moves forward)
        if (!(bp_sgn == sgn && sgn * (tp - bp_tp) >= 0 && (!bp_have || sgn * (bp_next - tp) > 0))) {
            bool h2 = false;
            S b2 = 0;
            auto cand2 = [&](S b) {
                if (isfinite(b) && sgn * (b - tp) > 0 && (!h2 || sgn * (b2 - b) > 0)) {
                    b2 = b;
                    h2 = true;
                }
            };
            for (int q = first(1); q <= L; q += W) {
                const int ta = tau(q);
                const S yA = y(0, ta), yB = y(1, ta);
                cand2((S(0) - yA) / cu0);
                cand2((S(1) - yA) / cu0);
                cand2((S(0) - yB) / cu1);
                cand2((S(1) - yB) / cu1);
            }
            bp_have = h2;
            bp_next = b2;
            bp_sgn = sgn;
        }
        bp_tp = tp;
        if (bp_have) cand(bp_next);
        return seg_best<S, W>(have, best, sgn, t);
    }

    // the nearest sign change beyond x, strictly inside (a, b), of the storage secondaries (stor)
    // or of the balances, from their values at a and b (linear in between)
    __device__ bool next_root(bool stor, S a, S b, S x, S sgn, S &xr) const {
        bool have = false;
        S best = 0;
        auto test = [&](S A, S B) {
            if ((A < 0) != (B < 0) && A != B) {
                const S r = a + (b - a) * A / (A - B);
                if (sgn * (r - a) > 0 && sgn * (b - r) > 0 && sgn * (r - x) > 0 && (!have || sgn * (best - r) > 0)) {
                    best = r;
                    have = true;
                }
            }
        };
        for (int e = 0; e < 2; e++) {
            if (!(e == 0 ? owns_a() : owns_b())) continue;
            const int q = e == 0 ? 0 : L;
            S dAa, dBa, dAb, dBb, jAa, jBa, jAb, jBb;
            sterm(0, q, a, dAa, jAa);
            sterm(1, q, a, dBa, jBa);
            sterm(0, q, b, dAb, jAb);
            sterm(1, q, b, dBb, jBb);
            if (stor) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    test(sec(0, k, dAa), sec(0, k, dAb));
                    test(sec(1, k, dBa), sec(1, k, dBb));
                }
            } else {
                for (int p = tb.nd_off[slot]; p < tb.nd_off[slot + 1]; p++)
                    test(node_S(e, p, a, jAa, jBa), node_S(e, p, b, jAb, jBb));
            }
        }
        if (!(llo <= a && a <= lhi && llo <= b && b <= lhi))
            for (int q = first(1); q <= L - 1; q += W) {
                if (inside(0, q, a) && inside(0, q, b) && inside(1, q, a) && inside(1, q, b)) continue;
                S xx, dAa, dBa, bAa, bBa, dAb, dBb, bAb, bBb;
                int_term(q, a, xx, dAa, dBa, bAa, bBa);
                int_term(q, b, xx, dAb, dBb, bAb, bBb);
                if (stor) {
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        test(sec(0, k, dAa), sec(0, k, dAb));
                        test(sec(1, k, dBa), sec(1, k, dBb));
                    }
                } else {
                    test(bAa, bAb);
                    if (nd0 != nd1) test(bBa, bBb);
                }
            }
        return seg_best<S, W>(have, best, sgn, xr);
    }

    // log Jacobian Σ (|S| − 1) log|P(t)/P| of the kinks made (FP64), |S| the node's dispatchable
    // units with |κ| ≥ κmin: the gas unit, so every term is 0 in this system
    __device__ double logjac(S t, bool &any) const {
        double lj = 0.0;
        int cnt = 0;
        for (int e = 0; e < 2; e++) {
            if (!(e == 0 ? owns_a() : owns_b())) continue;
            for (int kq = tb.kk_off[slot]; kq < tb.kk_off[slot + 1]; kq++) {
                XKinkD<S> k;
                if (!kink(e, kq, k)) continue;
                cnt++;
                const double nm = kink_nm(*m, tb.kk_node[kq], kmin);
                lj += nm * log(fabs((fmax((double)k.Q + (double)k.g * (double)t, 0.0) + (double)k.e) / (double)k.P));
            }
        }
        any = seg_sum<double, W>((double)cnt) > 0;
        return seg_sum<double, W>(lj);
    }

    // commit at t ≠ 0: the end steps (node by node: the balance and the leak's gas read only that
    // node's values and the flows, which are written after all nodes; then flows and angles), the
    // interior balances, then — after every lane has read the instants — the rising instants
    __device__ void commit(S t, S *r_w, S *F_w, S *g_w, S *th_w, S *yA_w, S *yB_w) const {
        for (int e = 0; e < 2; e++) {
            if (!(e == 0 ? owns_a() : owns_b())) continue;
            const int q = e == 0 ? 0 : L, ta = tau(q);
            const S c = sc(e);
            S d, djA, djB;
            sterm(0, q, t, d, djA);
            sterm(1, q, t, d, djB);
            for (int p = tb.nd_off[slot]; p < tb.nd_off[slot + 1]; p++) {
                const int n = tb.nd_node[p], ik = tb.nd_kk[p];
                const S S1 = node_S(e, p, t, djA, djB);
                bool wy = false;
                S ynew = 0;
                if (ik >= 0) {
                    XKinkD<S> k;
                    if (kink(e, tb.kk_off[slot] + ik, k) && k.ak != S(0)) {
                        ynew = k.y0 + k.ak * kphi(k.Q, k.g, t);
                        wy = true;
                    }
                }
                r_w[((size_t)n * T + ta) * R] = S1;
                if (wy) g_w[((size_t)n * T + ta) * R] = ynew;
            }
            for (int l = tb.ln_off[slot]; l < tb.ln_off[slot + 1]; l++) {
                const S df = dfl(c, l);
                if (df == S(0)) continue;
                S &F1 = F_w[((size_t)tb.ln_line[l] * T + ta) * R];
                F1 = F1 + t * df;
            }
            for (int k = tb.th_off[slot]; k < tb.th_off[slot + 1]; k++) {
                const S a = thv(c, tb.th_val[k]);
                if (a == S(0)) continue;
                S &th = th_w[((size_t)tb.th_node[k] * T + ta) * R];
                th = th + t * a;
            }
        }
        for (int q = first(1); q <= L - 1; q += W) {
            S x, dA, dB, bA, bB;
            if (!int_term(q, t, x, dA, dB, bA, bB)) continue;
            const int ta = tau(q);
            r_w[((size_t)nd0 * T + ta) * R] = bA;
            if (nd1 != nd0) r_w[((size_t)nd1 * T + ta) * R] = bB;
        }
        if (W > 1) {
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        for (int q = first(1); q <= L; q += W) {
            const int ta = tau(q);
            yA_w[(size_t)ta * R] = y(0, ta) + cu0 * t;
            yB_w[(size_t)ta * R] = y(1, ta) + cu1 * t;
        }
    }
};

// Set up the pair move of units A = (sa, na) and B = (sb, nb) at template slot `slot`, instants
// a+1 … a+L, for replica j seen by `lane` of its segment; false when it is not made (κ_A or κ_B
// below 1: the same in every lane and replica).
template <typename S, int W>
__device__ bool fx_prepare(FXPair<S, W> &f, const DevModel<S> &m, const SweepState<S> &st, const SxTab<S> &tb, int sa,
                           int na, int sb, int nb, int slot, int a, int L, int j, int lane, S kmin) {
    const int N = m.N, T = m.T, R = m.R;
    const size_t TR = (size_t)T * R;
    f.m = &m;
    f.tb = tb;
    f.T = T;
    f.R = R;
    f.lane = lane;
    f.L = L;
    f.a0 = a;
    f.slot = slot;
    f.nd0 = na;
    f.nd1 = nb;
    f.kmin = kmin;
    const int ua = sa * N + na, ub = sb * N + nb;
    f.kst0 = m.kc[(ESM_KSTOR + sa) * N + na];
    f.kst1 = m.kc[(ESM_KSTOR + sb) * N + nb];
    f.w0 = f.kst0 + m.kappa0;
    f.w1 = f.kst1 + m.kappa0;
    f.par0 = {m.hours[ua], m.eta_in[ua], m.eta_out[ua], m.qt[ua], m.mc[(ESM_KSTOR + sa) * N + na]};
    f.par1 = {m.hours[ub], m.eta_in[ub], m.eta_out[ub], m.qt[ub], m.mc[(ESM_KSTOR + sb) * N + nb]};
    const S EA = f.kst0 * m.hours[ua], EB = f.kst1 * m.hours[ub];
    f.kap = EA / m.dt;
    if (f.kap < S(1) || EB / m.dt < S(1)) return false;
    f.cu0 = S(1);
    f.cu1 = -(EA / EB);
    f.ye0 = st.v + plane((2 + sa) * N, T, R) + (size_t)na * TR + j;
    f.ye1 = st.v + plane((2 + sb) * N, T, R) + (size_t)nb * TR + j;
    f.rp = st.r + j;
    f.Fp = st.F + j;
    f.gp = st.v + plane(N, T, R) + j;
    // the lane's window over its interior steps
    S lo = -INFINITY, hi = INFINITY;
    for (int q = f.first(1); q <= L - 1; q += W)
        for (int u = 0; u < 2; u++) {
            const int ta = f.tau(q);
            const S y0 = f.y(u, ta), y1 = f.y(u, f.next(ta)), c = u == 0 ? f.cu0 : f.cu1;
            if (S(0) <= y0 && y0 <= S(1) && S(0) <= y1 && y1 <= S(1)) {
                S l0, h0, l1, h1;
                FXPair<S, W>::box(y0, c, l0, h0);
                FXPair<S, W>::box(y1, c, l1, h1);
                lo = fmax(lo, fmax(l0, l1));
                hi = fmin(hi, fmin(h0, h1));
            } else {
                lo = INFINITY;
                hi = -INFINITY;
            }
        }
    f.llo = lo;
    f.lhi = hi;
    return true;
}

}  // namespace esm
