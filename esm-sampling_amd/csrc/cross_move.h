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
// Lanes: one wave per (move, replica) (FXLane below): the lanes hold the items of the end
// steps' template and the interior steps q mod 64; the walk is `walk_stor` (storage_walk.h) with
// recording.
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


// ---------------------------------------------------------------------------------------------
// FXLane: the pair move of one replica on one wave. The end steps' terms are items of four kinds
// (`eval_cached!`, src/local.jl): the kinked gas entries of the leak nodes, the template's lines,
// the storage terms of A and B, the balances of the template's nodes. Items of one kind over both
// ends (end a first) are spread over the lanes, item k of a kind on lane k mod 64, so that a pass
// runs each kind's code once on all lanes at the same time. Each lane reads its items' values at
// t = 0 once (prepare) and keeps them in registers; a pass at t computes the kinks', lines' and
// storage terms, which publish their change of the node balances in LDS, then the balances; the
// storage d and the balances are recorded at the walk's interval and piece ends (the storage
// walk's CACHED protocol), so the root searches evaluate nothing. The interior steps a+1 … a+L−1
// (sxlong) belong to lane q mod 64. ΔC(t): a lane sums its kinks, lines, storage terms, balances
// (each kind in increasing item index), then its interior steps in increasing q; the lanes' sums
// are combined by the 64-lane butterfly.
// ---------------------------------------------------------------------------------------------

// A move is owned by a segment of W lanes (W = 32 or 64: the segments of a wave are replicas).
// kink slots per lane KS (a template parameter): at most KS·W kinks of both ends; 1, or 2 for the
// largest templates (hub nodes of 256-node networks: up to 46 leak nodes per end), which also take
// a third balance slot (up to 66 balances per end)
constexpr int XL_LS = 3;                 // line slots: at most 3 W lines of both ends
__host__ __device__ constexpr int xl_ns(int KS) { return KS + 1; }   // node slots: (KS + 1)·W balances
__host__ __device__ constexpr int xl_nx(int W, int KS) { return KS * W + 4; }

// the LDS of one segment's pair move
constexpr int XL_NP = 32;   // walk points kept per side for the selection
constexpr int XL_MAXCH = 16;   // window chunks per lane of the interior steps (sxlong)

// interior steps of a lane per window chunk (at least 8) over a horizon of T, and the chunks
__host__ __device__ constexpr int xl_ch(int T) {
    return ((T + 63) / 64 + XL_MAXCH - 1) / XL_MAXCH > 8 ? ((T + 63) / 64 + XL_MAXCH - 1) / XL_MAXCH : 8;
}
// float bounds of a window rounded inward (a conservative filter of the exact test)
__device__ __forceinline__ float xl_inward_lo(double x) {
    float f = (float)x;
    return (double)f < x ? nextafterf(f, INFINITY) : f;
}
__device__ __forceinline__ float xl_inward_hi(double x) {
    float f = (float)x;
    return (double)f > x ? nextafterf(f, -INFINITY) : f;
}

template <typename S, int W, int KS = 1>
struct XlLds {
    S xv[xl_nx(W, KS)];          // at the pass's t: each kink's and storage term's change of the balances
    S pts[2 * XL_NP * 2];    // the mass walks' points (sample_stor_pts)
};

// whether a template of K kinks, Lw lines and Nn balances per end fits a segment of W lanes
// (checked on the host for every slot)
__host__ __device__ __forceinline__ bool xl_fits_counts(int W, int KS, int K, int Lw, int Nn, int NLp) {
    (void)NLp;
    return 2 * K <= W * KS && 2 * Lw <= W * XL_LS && 2 * Nn <= W * xl_ns(KS);
}
template <typename S, int W, int KS>
__device__ __forceinline__ bool xl_fits(const SxTab<S> &tb, int slot) {
    return xl_fits_counts(W, KS, tb.kk_off[slot + 1] - tb.kk_off[slot], tb.ln_off[slot + 1] - tb.ln_off[slot],
                          tb.nd_off[slot + 1] - tb.nd_off[slot],
                          tb.nd_l[tb.nd_off[slot + 1]] - tb.nd_l[tb.nd_off[slot]]);
}

// IN: the move has interior steps (sxlong); sxpair (L = 1) compiles without them. WW: the
// segment's lanes.
template <typename S, bool IN, int WW, int KSS = 1>
struct FXLane {
    enum { REC_NONE, REC_B, REC_X, REC_A, REC_T };
    static constexpr bool CACHED = true;
    static constexpr int W = WW;
    static constexpr int XL_KS = KSS, XL_NS = xl_ns(KSS);
    static constexpr int XK = 0, XS = XL_KS * W;
    const DevModel<S> *m;
    SxTab<S> tb;
    XlLds<S, W, XL_KS> *sh;
    const S *ye0, *ye1;    // the units' state-of-charge planes at (node, ·, replica), stride R
    const S *rp, *Fp, *gp; // residual, flow and gas planes at (·, ·, replica)
    int T, R, lane, L, a0, slot;   // lane: the lane within the segment
    int nd0, nd1;
    int K, Lw, Nn, NLp;
    S cu0, cu1, kst0, kst1, w0, w1, kap, kmin;
    StoPar<S> par0, par1;
    S llo, lhi;            // t in [llo, lhi]: the lane's interior steps keep every instant in [0, 1]
    // the same per chunk of ch consecutive interior steps of the lane (LDS [chunk][64], IN only)
    const float *cwl = nullptr, *cwh = nullptr;
    int ch = 8;
    // kinks (item e·K + k): node, made (bit 0), ak ≠ 0 (bit 1); Q, g, ak, y0, c0, p0, P, e
    int kn[XL_KS], kok[XL_KS];
    S kQ[XL_KS], kg[XL_KS], ka[XL_KS], ky[XL_KS], kc0[XL_KS], kp0[XL_KS], kP[XL_KS], ke[XL_KS];
    S kb[XL_KS][5];       // its primary breakpoints (the kink, the clamp points of the gas; NaN: none)
    // lines (item e·Lw + l): line; f0, df (0: dropped), limit, max(|f0| − limit, 0); the breakpoints
    // where |f| reaches the limit
    int ll[XL_LS];
    S lf[XL_LS], ld[XL_LS], ls_[XL_LS], lp0[XL_LS], lb0[XL_LS], lb1[XL_LS];
    // storage terms (item 2e + u on lane 2e + u): y0, y1, c0, j0, dye, dye1; d at the interval's ends
    int su, sq;
    S sy0, sy1, sc0, sj0, sdy0, sdy1;
    mutable S sdA, sdB;
    // balances (item e·Nn + p): node, xv index of its kink (−1: none made); S0, φ(S0), the change of
    // the node's inflow per unit t over its template lines (Σ ± df in line order); the balance at
    // the piece's ends
    int nn_[XL_NS], nk[XL_NS];
    S nS0[XL_NS], nph[XL_NS], ncf[XL_NS];
    mutable S nA[XL_NS], nB[XL_NS];
    mutable int nev = 0;
    // the lane's nearest clamp point of its rising instants beyond the walk's last point (the
    // points are fixed during a move and a walk only moves forward)
    mutable S bp_tp = 0, bp_sgn = 0, bp_next = 0;
    mutable bool bp_have = false;

    // unit u's parameters, selected field by field as values (a selected member, an lvalue, would
    // keep the walker in scratch memory)
    __device__ __forceinline__ StoPar<S> par(int u) const {
        StoPar<S> p;
        p.h = u == 0 ? +par0.h : +par1.h;
        p.ein = u == 0 ? +par0.ein : +par1.ein;
        p.eout = u == 0 ? +par0.eout : +par1.eout;
        p.qt = u == 0 ? +par0.qt : +par1.qt;
        p.mcs = u == 0 ? +par0.mcs : +par1.mcs;
        return p;
    }
    __device__ __forceinline__ int tau(int q) const {
        const int x = a0 + q;
        return x >= T ? x - T : x;
    }
    __device__ __forceinline__ int next(int t) const { return t + 1 == T ? 0 : t + 1; }
    __device__ __forceinline__ S y(int u, int t) const { return (u == 0 ? +ye0 : +ye1)[(size_t)t * R]; }
    __device__ __forceinline__ S rv(int n, int t) const { return rp[((size_t)n * T + t) * R]; }
    __device__ __forceinline__ S Fv(int e, int t) const { return Fp[((size_t)e * T + t) * R]; }
    __device__ __forceinline__ S gv(int n, int t) const { return gp[((size_t)n * T + t) * R]; }
    __device__ __forceinline__ int first(int q0) const {
        return lane >= q0 ? lane : lane + W * ((q0 - lane + W - 1) / W);
    }
    __device__ __forceinline__ S sc(int e) const { return e == 0 ? -kap : kap; }
    __device__ __forceinline__ static S thv(S c, S v) {
        const S a = c * v;
        return fabs(a) < S(1e-14) ? S(0) : a;
    }
    // flow change per unit t of template line l at scale c (0: dropped)
    __device__ __forceinline__ S dfl(S c, int l) const {
        const S df = m->lb[tb.ln_line[l]] * (thv(c, tb.ln_ta[l]) - thv(c, tb.ln_tb[l]));
        return fabs(df) <= S(1e-14) ? S(0) : df;
    }
    __device__ __forceinline__ static S kphi(S Q, S g, S t) { return fmax(Q + g * t, S(0)) - fmax(Q, S(0)); }
    __device__ __forceinline__ void gen(int n, S yy, S &c, S &p) const {
        const S k = m->kc[ESM_KGAS * m->N + n];
        gen_cost(m->mc[ESM_KGAS * m->N + n], k, k + m->kappa0, m->dt, m->cv, yy, c, p);
    }
    __device__ __forceinline__ void sync() const {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }

    // the lane's end-step terms at t summed in its item order (kinks, lines, storage, balances);
    // sd: the storage term's d at t, sb[s]: the balances at t
    __device__ __forceinline__ S items(S t, S &sd, S *sb) const {
        S D = 0;
#pragma unroll
        for (int s = 0; s < XL_KS; s++) {
            S x = 0;
            if (kok[s] & 2) {
                S c1, p1;
                gen(kn[s], ky[s] + ka[s] * kphi(kQ[s], kg[s], t), c1, p1);
                D += c1 - kc0[s];
                x = p1 - kp0[s];
            }
            sh->xv[XK + lane + W * s] = x;
        }
#pragma unroll
        for (int s = 0; s < XL_LS; s++) {
            if (ld[s] != S(0)) {
                const S f1 = lf[s] + t * ld[s];
                D += m->cv * m->dt * (fmax(fabs(f1) - ls_[s], S(0)) - lp0[s]);
            }
        }
        sd = 0;
        if (su >= 0) {
            S j1;
            const S c1 = soc_cost_p(par(su), m->dt, m->cv, su == 0 ? +kst0 : +kst1, su == 0 ? +w0 : +w1,
                                    sy0 + t * sdy0, sy1 + t * sdy1, sd, j1);
            D += c1 - sc0;
            sh->xv[XS + lane] = j1 - sj0;
        }
        sync();
#pragma unroll
        for (int s = 0; s < XL_NS; s++) {
            sb[s] = 0;
            if (nn_[s] < 0) continue;
            const int e = lane + W * s >= Nn ? 1 : 0;
            S dr = 0;
            if (nk[s] >= 0) dr -= sh->xv[nk[s]];
            dr += t * ncf[s];
            if (nn_[s] == nd0) dr += sh->xv[XS + 2 * e];
            if (nn_[s] == nd1) dr += sh->xv[XS + 2 * e + 1];
            const S S1 = nS0[s] + dr;
            D += m->dt * (phibal(S1, m->c_shed, m->c_dump) - nph[s]);
            sb[s] = S1;
        }
        sync();   // the next pass writes xv again
        return D;
    }

    // storage term of unit u at interior step q at t: cost change; d at t, injection change
    __device__ __forceinline__ S sterm(int u, int q, S t, S &d1, S &dj) const {
        const int ta = tau(q);
        const S y0 = y(u, ta), y1 = y(u, next(ta)), c = u == 0 ? +cu0 : +cu1;
        const S dye = q >= 1 ? c : S(0), dye1 = q <= L - 1 ? c : S(0);
        const StoPar<S> p = par(u);
        const S k = u == 0 ? +kst0 : +kst1, w = u == 0 ? +w0 : +w1;
        S d0, j0, j1;
        const S c0 = soc_cost_p(p, m->dt, m->cv, k, w, y0, y1, d0, j0);
        const S c1 = soc_cost_p(p, m->dt, m->cv, k, w, y0 + t * dye, y1 + t * dye1, d1, j1);
        dj = j1 - j0;
        return c1 - c0;
    }
    __device__ __forceinline__ S sec(int u, int k, S d) const {
        const StoPar<S> p = par(u);
        if (k == 0) return d;
        if (k == 1) return p.ein > 0 ? d - p.ein : S(1);
        return -d * p.eout - S(1);
    }
    // the box interval of an instant y + c t
    __device__ __forceinline__ static void box(S yy, S c, S &lo, S &hi) {
        if (c > 0) {
            lo = -yy / c;
            hi = (S(1) - yy) / c;
        } else {
            lo = (S(1) - yy) / c;
            hi = -yy / c;
        }
    }
    __device__ __forceinline__ bool inside(int u, int q, S t) const {
        const int ta = tau(q);
        const S y0 = y(u, ta), y1 = y(u, next(ta)), c = u == 0 ? +cu0 : +cu1;
        if (!(S(0) <= y0 && y0 <= S(1) && S(0) <= y1 && y1 <= S(1))) return false;
        S l0, h0, l1, h1;
        box(y0, c, l0, h0);
        box(y1, c, l1, h1);
        return fmax(l0, l1) <= t && t <= fmin(h0, h1);
    }
    // interior step q at t: false when both units stay inside (nothing changes); else its term x,
    // the units' d (the value at t = 0 for a unit inside) and the balances bA, bB of their nodes
    __device__ __forceinline__ bool int_term(int q, S t, S &x, S &dA, S &dB, S &bA, S &bB) const {
        const int ta = tau(q);
        S D = 0, djA = 0, djB = 0;
        bool mv = false;
        for (int u = 0; u < 2; u++) {
            S d = 0, dj = 0;
            if (inside(u, q, t)) {
                S j0;
                soc_cost_p(par(u), m->dt, m->cv, u == 0 ? +kst0 : +kst1, u == 0 ? +w0 : +w1, y(u, ta), y(u, next(ta)),
                           d, j0);
            } else {
                D += sterm(u, q, t, d, dj);
                mv = true;
            }
            if (u == 0) {
                dA = d;
                djA = dj;
            } else {
                dB = d;
                djB = dj;
            }
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

    // ΔC(t); recording (storage walk protocol): REC_B d and balances at t (an interval's end),
    // REC_X balances (a storage root), REC_A d and balances at the walk's start, REC_T balances at
    // an interval's end after a split. Counted as evaluations: REC_NONE, REC_B, REC_X. (Every
    // recorded value is written as a selected value: stores under branches would be merged into
    // stores through selected addresses, which keep the walker in scratch memory.)
    __device__ __forceinline__ S eval(S t, int rec) const {
        if (rec <= REC_X) nev++;
        S sd, sb[XL_NS];
        S D = items(t, sd, sb);
        sdA = rec == REC_A ? sd : sdA;
        sdB = rec == REC_B ? sd : sdB;
        const bool wa = rec == REC_A, wb = rec != REC_NONE && rec != REC_A;
#pragma unroll
        for (int s = 0; s < XL_NS; s++) {
            nA[s] = wa ? sb[s] : nA[s];
            nB[s] = wb ? sb[s] : nB[s];
        }
        if (IN && !(llo <= t && t <= lhi))
            for (int q0 = first(1), c = 0; q0 <= L - 1; q0 += W * ch, c++) {
                if (cwl[c * 64 + lane] <= t && t <= cwh[c * 64 + lane]) continue;
                for (int q = q0; q <= L - 1 && q < q0 + W * ch; q += W) {
                    S x, d0, d1, b0, b1;
                    if (int_term(q, t, x, d0, d1, b0, b1)) D += x;
                }
            }
        return seg_sum<S, W>(D);
    }
    // the piece's end becomes the next piece's start (balances; with d: the interval's end)
    __device__ __forceinline__ void shift(bool d) const {
        sdA = d ? sdB : sdA;
#pragma unroll
        for (int s = 0; s < XL_NS; s++) nA[s] = nB[s];
    }

    __device__ __forceinline__ bool next_bp(S tp, S sgn, S &t) const {
        bool have = false;
        S best = 0;
        auto cand = [&](S b) {
            if (isfinite(b) && sgn * (b - tp) > 0 && (!have || sgn * (best - b) > 0)) {
                best = b;
                have = true;
            }
        };
#pragma unroll
        for (int s = 0; s < XL_KS; s++)
#pragma unroll
            for (int c = 0; c < 5; c++) cand(kb[s][c]);   // NaN (none) fails isfinite
#pragma unroll
        for (int s = 0; s < XL_LS; s++) {
            cand(lb0[s]);
            cand(lb1[s]);
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

    // the nearest sign change beyond x, strictly inside (a, b), of the storage secondaries (stor,
    // from d recorded at a and b) or of the balances (recorded at the piece's ends a and b); the
    // interior steps evaluate theirs at a and b
    __device__ __forceinline__ bool next_root(bool stor, S a, S b, S x, S sgn, S &xr) const {
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
        if (stor) {
            if (su >= 0) {
#pragma unroll
                for (int k = 0; k < 3; k++) test(sec(su, k, sdA), sec(su, k, sdB));
            }
        } else {
#pragma unroll
            for (int s = 0; s < XL_NS; s++)
                if (nn_[s] >= 0) test(nA[s], nB[s]);
        }
        if (IN && !(llo <= a && a <= lhi && llo <= b && b <= lhi))
            for (int q = first(1), c = 0; q <= L - 1; q += W) {
                if ((q - first(1)) / W == (c + 1) * ch) c++;
                const float wl = cwl[c * 64 + lane], wh = cwh[c * 64 + lane];
                if (wl <= a && a <= wh && wl <= b && b <= wh) continue;
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
    __device__ __forceinline__ double logjac(S t, bool &any) const {
        double lj = 0.0;
        int cnt = 0;
#pragma unroll
        for (int s = 0; s < XL_KS; s++) {
            if (!(kok[s] & 1)) continue;
            cnt++;
            const double nm = kink_nm(*m, kn[s], kmin);
            lj += nm * log(fabs((fmax((double)kQ[s] + (double)kg[s] * (double)t, 0.0) + (double)ke[s]) / (double)kP[s]));
        }
        any = seg_sum<double, W>((double)cnt) > 0;
        return seg_sum<double, W>(lj);
    }

    // commit at t ≠ 0: the end steps' balances, leak gas, flows and angles, the interior balances,
    // then — after every lane has read the instants — the rising instants
    __device__ __forceinline__ void commit(S t, S *r_w, S *F_w, S *g_w, S *th_w, S *yA_w, S *yB_w) const {
        S sd, sb[XL_NS];
        items(t, sd, sb);
        const int tA = tau(0), tB = tau(L);
#pragma unroll
        for (int s = 0; s < XL_NS; s++)
            if (nn_[s] >= 0) r_w[((size_t)nn_[s] * T + (lane + W * s >= Nn ? tB : tA)) * R] = sb[s];
#pragma unroll
        for (int s = 0; s < XL_KS; s++)
            if (kok[s] & 2)
                g_w[((size_t)kn[s] * T + (lane + W * s >= K ? tB : tA)) * R] = ky[s] + ka[s] * kphi(kQ[s], kg[s], t);
#pragma unroll
        for (int s = 0; s < XL_LS; s++)
            if (ld[s] != S(0)) F_w[((size_t)ll[s] * T + (lane + W * s >= Lw ? tB : tA)) * R] = lf[s] + t * ld[s];
        const int nth = tb.th_off[slot + 1] - tb.th_off[slot];
        for (int i = lane; i < 2 * nth; i += W) {
            const int e = i >= nth ? 1 : 0, k = tb.th_off[slot] + (i - e * nth);
            const S a = thv(sc(e), tb.th_val[k]);
            if (a == S(0)) continue;
            S &th = th_w[((size_t)tb.th_node[k] * T + (e == 0 ? tA : tB)) * R];
            th = th + t * a;
        }
        for (int q = first(1), c = 0; IN && q <= L - 1; q += W) {
            if ((q - first(1)) / W == (c + 1) * ch) c++;
            if (cwl[c * 64 + lane] <= t && t <= cwh[c * 64 + lane]) continue;
            S x, dA, dB, bA, bB;
            if (!int_term(q, t, x, dA, dB, bA, bB)) continue;
            const int ta = tau(q);
            r_w[((size_t)nd0 * T + ta) * R] = bA;
            if (nd1 != nd0) r_w[((size_t)nd1 * T + ta) * R] = bB;
        }
        sync();
        for (int q = first(1); q <= L; q += W) {
            const int ta = tau(q);
            yA_w[(size_t)ta * R] = y(0, ta) + cu0 * t;
            yB_w[(size_t)ta * R] = y(1, ta) + cu1 * t;
        }
        sync();
    }
};

// Set up the pair move of units A = (sa, na) and B = (sb, nb) at template slot `slot`, instants
// a+1 … a+L, for replica j on the calling wave (lane `lane`); false when it is not made (κ_A or κ_B
// below 1: the same in every lane and replica). Every lane reads its items' values at t = 0.
// cw (IN): the wave's LDS of 2 XL_MAXCH·64 floats for the chunk windows.
template <typename S, bool IN, int W, int KS>
__device__ __forceinline__ bool xl_prepare(FXLane<S, IN, W, KS> &f, const DevModel<S> &m, const SweepState<S> &st,
                                           const SxTab<S> &tb, XlLds<S, W, KS> *sh, int sa, int na, int sb, int nb,
                                           int slot, int a, int L, int j, int lane, S kmin, float *cw = nullptr) {
    const int N = m.N, T = m.T, R = m.R;
    const size_t TR = (size_t)T * R;
    f.m = &m;
    f.tb = tb;
    f.sh = sh;
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
    // the lane's window over its interior steps, and per chunk of ch of them
    S lo = -INFINITY, hi = INFINITY;
    if (IN) {
        f.ch = xl_ch(T);
        f.cwl = cw;
        f.cwh = cw + XL_MAXCH * 64;
        for (int q0 = f.first(1), c = 0; q0 <= L - 1; q0 += W * f.ch, c++) {
            S clo = -INFINITY, chi = INFINITY;
            for (int q = q0; q <= L - 1 && q < q0 + W * f.ch; q += W)
                for (int u = 0; u < 2; u++) {
                    const int ta = f.tau(q);
                    const S y0 = f.y(u, ta), y1 = f.y(u, f.next(ta)), cu = u == 0 ? +f.cu0 : +f.cu1;
                    if (S(0) <= y0 && y0 <= S(1) && S(0) <= y1 && y1 <= S(1)) {
                        S l0, h0, l1, h1;
                        FXLane<S, IN, W, KS>::box(y0, cu, l0, h0);
                        FXLane<S, IN, W, KS>::box(y1, cu, l1, h1);
                        clo = fmax(clo, fmax(l0, l1));
                        chi = fmin(chi, fmin(h0, h1));
                    } else {
                        clo = INFINITY;
                        chi = -INFINITY;
                    }
                }
            cw[c * 64 + lane] = xl_inward_lo((double)clo);
            cw[XL_MAXCH * 64 + c * 64 + lane] = xl_inward_hi((double)chi);
            lo = fmax(lo, clo);
            hi = fmin(hi, chi);
        }
    }
    f.llo = lo;
    f.lhi = hi;
    const int k0 = tb.kk_off[slot], l0 = tb.ln_off[slot], n0 = tb.nd_off[slot];
    f.K = tb.kk_off[slot + 1] - k0;
    f.Lw = tb.ln_off[slot + 1] - l0;
    f.Nn = tb.nd_off[slot + 1] - n0;
    f.NLp = tb.nd_l[n0 + f.Nn] - tb.nd_l[n0];
    const int tA = f.tau(0), tB = f.tau(L);
    // kinks: `prorata!` (qadj = 0) of the gas unit of a leak node. (Each slot's values are set as
    // locals and stored once: stores to different fields under branches would be merged into stores
    // through selected addresses.)
#pragma unroll
    for (int s = 0; s < KS; s++) {
        const int i = lane + W * s;
        int n = 0, okv = 0;
        S Q = 0, g = 0, ak = 0, w = 0, c0 = 0, p0 = 0, P = 0, e0 = 0;
        if (i < 2 * f.K) {
            const int e = i >= f.K ? 1 : 0, kq = k0 + (i - e * f.K), ta = e == 0 ? tA : tB;
            n = tb.kk_node[kq];
            g = f.sc(e) * tb.kk_g[kq];
            const S kp = m.kc[ESM_KGAS * N + n];
            if (g != S(0) && fabs(kp) >= kmin) {
                w = f.gv(n, ta);
                P = S(0) + kp * w;
                const S Qs = S(0) + kp * clampv(w, S(0), S(1));
                if (!(fabs(P) < S(1e-6))) {
                    Q = f.rv(n, ta) + Qs;
                    e0 = P - fmax(Q, S(0));
                    const S sig = kp * w / P;
                    ak = sig == S(0) ? S(0) : sig / kp;
                    okv = ak != S(0) ? 3 : 1;
                    if (okv & 2) f.gen(n, w, c0, p0);
                }
            }
        }
        S b0 = S(NAN), b1 = S(NAN), b2 = S(NAN), b3 = S(NAN), b4 = S(NAN);
        if (okv & 1) {
            b0 = -Q / g;
            if (okv & 2) {
                const S s1 = ak * g, off = ak * Q;
                b1 = (S(0) - w - S(0)) / s1;
                b2 = (S(1) - w - S(0)) / s1;
                b3 = (S(0) - w - off) / s1;
                b4 = (S(1) - w - off) / s1;
            }
        }
        f.kb[s][0] = b0;
        f.kb[s][1] = b1;
        f.kb[s][2] = b2;
        f.kb[s][3] = b3;
        f.kb[s][4] = b4;
        f.kn[s] = n;
        f.kok[s] = okv;
        f.kQ[s] = Q;
        f.kg[s] = g;
        f.ka[s] = ak;
        f.ky[s] = w;
        f.kc0[s] = c0;
        f.kp0[s] = p0;
        f.kP[s] = P;
        f.ke[s] = e0;
    }
#pragma unroll
    for (int s = 0; s < XL_LS; s++) {
        const int i = lane + W * s;
        int ln = 0;
        S f0 = 0, df = 0, se = 0, p0 = 0;
        if (i < 2 * f.Lw) {
            const int e = i >= f.Lw ? 1 : 0, l = l0 + (i - e * f.Lw);
            ln = tb.ln_line[l];
            df = f.dfl(f.sc(e), l);
            if (df != S(0)) {
                f0 = f.Fv(ln, e == 0 ? tA : tB);
                se = m.ls[ln];
                p0 = fmax(fabs(f0) - se, S(0));
            }
        }
        f.ll[s] = ln;
        f.lf[s] = f0;
        f.ld[s] = df;
        f.ls_[s] = se;
        f.lp0[s] = p0;
        f.lb0[s] = df != S(0) ? (se - f0) / df : S(NAN);
        f.lb1[s] = df != S(0) ? (-se - f0) / df : S(NAN);
    }
    {   // storage: lane 2e + u holds unit u's term at end e
        int u = -1, q = 0;
        S y0 = 0, y1 = 0, c0 = 0, j0 = 0, dy0 = 0, dy1 = 0;
        if (lane < 4) {
            const int e = lane >> 1;
            u = lane & 1;
            q = e == 0 ? 0 : L;
            const int ta = e == 0 ? tA : tB;
            const S cu = u == 0 ? +f.cu0 : +f.cu1;
            y0 = f.y(u, ta);
            y1 = f.y(u, f.next(ta));
            dy0 = q >= 1 ? cu : S(0);
            dy1 = q <= L - 1 ? cu : S(0);
            S d0;
            c0 = soc_cost_p(f.par(u), m.dt, m.cv, u == 0 ? +f.kst0 : +f.kst1, u == 0 ? +f.w0 : +f.w1, y0, y1, d0, j0);
        }
        f.su = u;
        f.sq = q;
        f.sy0 = y0;
        f.sy1 = y1;
        f.sc0 = c0;
        f.sj0 = j0;
        f.sdy0 = dy0;
        f.sdy1 = dy1;
        f.sdA = 0;
        f.sdB = 0;
    }
#pragma unroll
    for (int s = 0; s < xl_ns(KS); s++) {
        const int i = lane + W * s;
        int n = -1, kx = -1;
        S S0 = 0, ph = 0, cf = 0;
        if (i < 2 * f.Nn) {
            const int e = i >= f.Nn ? 1 : 0, p = n0 + (i - e * f.Nn);
            n = tb.nd_node[p];
            kx = tb.nd_kk[p] >= 0 ? FXLane<S, IN, W, KS>::XK + e * f.K + tb.nd_kk[p] : -1;
            S0 = f.rv(n, e == 0 ? tA : tB);
            ph = phibal(S0, m.c_shed, m.c_dump);
            const S c = f.sc(e);
            for (int q2 = tb.nd_l[p]; q2 < tb.nd_l[p + 1]; q2++) {
                const S df = f.dfl(c, l0 + tb.nl_pos[q2]);
                if (df == S(0)) continue;
                cf += tb.nl_sgn[q2] > 0 ? df : -df;
            }
        }
        f.nn_[s] = n;
        f.nk[s] = kx;
        f.nS0[s] = S0;
        f.nph[s] = ph;
        f.ncf[s] = cf;
        f.nA[s] = 0;
        f.nB[s] = 0;
    }
    // which kinks are made with a nonzero share (the balances take their gas change)
#pragma unroll
    for (int s = 0; s < KS; s++) sh->xv[FXLane<S, IN, W, KS>::XK + lane + W * s] = (f.kok[s] & 2) ? S(1) : S(0);
    f.sync();
#pragma unroll
    for (int s = 0; s < xl_ns(KS); s++) f.nk[s] = (f.nk[s] >= 0 && sh->xv[f.nk[s]] != S(0)) ? f.nk[s] : -1;
    f.sync();
    return true;
}

}  // namespace esm
