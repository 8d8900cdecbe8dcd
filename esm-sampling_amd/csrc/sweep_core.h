// sweep_core.h — device building blocks of the exact line-conditional moves: the families'
// recorded terms and ΔC(t), the register-resident line sampler, the per-move random numbers and
// the wave-aggregated statistics. Included by k_sweep.hip (atom rounds) and k_random.hip (the
// random families).
#pragma once
#include "esm_internal.h"

namespace esm {

constexpr int PMAX = 4096;    // walk points per side (guard: t = 0 beyond)

template <typename S>
__device__ __forceinline__ S phibal(S x, S cs, S cd) { return x > 0 ? cs * x : -cd * x; }

template <typename S>
__device__ __forceinline__ void gen_cost(S mc, S kap, S w, S dt, S cv, S y, S &c, S &p) {
    const S yc = clampv(y, S(0), S(1));
    p = kap * yc;
    c = dt * mc * p + cv * dt * w * fabs(y - yc);
}

// log ∫₀ᴸ e^{−g s} ds at x = gL (src/local.jl:1099-1103)
template <typename S>
__device__ __forceinline__ S logseg(S L, S x) {
    if (fabs(x) < S(1e-10)) return log(L) - x / 2;
    if (x > 0) return log(L) + log(-expm1(-x)) - log(x);
    return log(L) - x + log(-expm1(x)) - log(-x);
}

// offset in [0, L] with density ∝ e^{−g s} from the uniform u (src/local.jl:1106-1115)
template <typename S>
__device__ __forceinline__ S trexp_offset(S g, S L, S u) {
    const S x = g * L;
    if (fabs(x) < S(1e-10)) return u * L;
    if (g > 0) return -log1p(-u * (-expm1(-x))) / g;
    return L + log1p(-u * (-expm1(x))) / (-g);
}

template <typename S>
__device__ __forceinline__ void insort(S *v, int n, S sgn) {
    for (int i = 1; i < n; i++) {
        const S x = v[i];
        int j = i - 1;
        while (j >= 0 && (sgn > 0 ? x < v[j] : x > v[j])) {
            v[j + 1] = v[j];
            j--;
        }
        v[j + 1] = x;
    }
}

// ---------------------------------------------------------------------------------------------
// families: recorded terms and ΔC(t) with the secondary quantities
// ---------------------------------------------------------------------------------------------

template <typename S>
struct Ctx {
    S dt, cv, cs, cd;
};

// gas atom: one generator term and the balance of its node at τ
template <typename S>
struct FGen {
    static constexpr int NSEC = 1, NBP = 2;
    static constexpr bool STOR = false;
    __device__ static constexpr bool gmask(int) { return false; }
    Ctx<S> k;
    S y0, kap, w, mc, a, c0, p0, S0, phi0;
    mutable int nev = 0;   // ΔC evaluations (statistics)
    __device__ S eval(S t, S *sec, S &dr) const {
        nev++;
        S c1, p1;
        gen_cost(mc, kap, w, k.dt, k.cv, y0 + t * a, c1, p1);
        S d = c1 - c0;
        dr = -(p1 - p0);
        const S S1 = S0 + dr;
        d += k.dt * (phibal(S1, k.cs, k.cd) - phi0);
        sec[0] = S1;
        return d;
    }
    __device__ S eval(S t, S *sec) const {
        S dr;
        return eval(t, sec, dr);
    }
};

// potential of node n: every incident line (flow change ±b per unit t) and the balances of n
// and of its neighbours, in the reference's touch order (line order, bus0 before bus1)
template <typename S, int D>
struct FTheta {
    static constexpr int NSEC = D + 1, NBP = 2 * D;
    static constexpr bool STOR = false;
    __device__ static constexpr bool gmask(int) { return false; }
    Ctx<S> k;
    int nl, npos;          // lines; position of node n among the balances (0 or 1)
    S f0[D], df[D], se[D], c0[D], sgn_n[D];   // sgn_n: +1 if n is the line's bus0
    S S0n, phin, S0o[D], phio[D];
    mutable int nev = 0;   // ΔC evaluations (statistics)
    __device__ S eval(S t, S *sec, S &drn, S *dro) const {
        nev++;
        S d = 0;
        drn = 0;
#pragma unroll
        for (int l = 0; l < D; l++) {
            if (l < nl) {
                const S f1 = f0[l] + t * df[l];
                d += k.cv * k.dt * (fmax(fabs(f1) - se[l], S(0)) - c0[l]);
                const S dfl = f1 - f0[l];
                drn += sgn_n[l] > 0 ? dfl : -dfl;
                dro[l] = sgn_n[l] > 0 ? -dfl : dfl;
            }
        }
        const S S1n = S0n + drn;
        S S1o[D];
#pragma unroll
        for (int l = 0; l < D; l++) S1o[l] = l < nl ? S0o[l] + dro[l] : S(1);
        // balances in touch order: neighbours in line order with n at position npos
#pragma unroll
        for (int p = 0; p <= D; p++) {
            if (p <= nl) {
                // neighbour of line p (before n) or p − 1 (after n): static register indices
                const S s_lo = p < D ? S1o[p < D ? p : 0] : S(0), f_lo = p < D ? phio[p < D ? p : 0] : S(0);
                const S s_hi = p > 0 ? S1o[p > 0 ? p - 1 : 0] : S(0), f_hi = p > 0 ? phio[p > 0 ? p - 1 : 0] : S(0);
                const S S1 = p == npos ? S1n : (p < npos ? s_lo : s_hi);
                const S ph0 = p == npos ? phin : (p < npos ? f_lo : f_hi);
                d += k.dt * (phibal(S1, k.cs, k.cd) - ph0);
                sec[p] = S1;
            } else {
                sec[p] = S(1);
            }
        }
        return d;
    }
    __device__ S eval(S t, S *sec) const {
        S drn, dro[D];
        return eval(t, sec, drn, dro);
    }
};

// per-unit storage parameters (the SoC form of src/energy.jl:437-463)
template <typename S>
struct StoPar {
    S h, ein, eout, qt, mcs;
};

// soc_cost with the unit's parameters in registers: returns the cost, writes d and the injection
template <typename S>
__device__ __forceinline__ S soc_cost_p(const StoPar<S> &p, S dt, S cv, S k, S w, S ye, S ye1, S &d, S &inj) {
    const S ec = clampv(ye, S(0), S(1)), ec1 = clampv(ye1, S(0), S(1));
    d = p.h * (ec1 - ec) / dt - p.qt;
    S pc, mcp;
    if (d >= 0) {
        pc = p.ein > 0 ? d / p.ein : S(0);
        mcp = 0;
    } else {
        pc = 0;
        mcp = fmin(-d * p.eout, S(1));
    }
    S c = dt * p.mcs * k * (pc + mcp) + cv * w * p.h * fabs(ye - ec);
    if (pc > 1) c += cv * dt * w * (pc - 1);
    if (d < 0 && p.qt <= 0) {
        const S ex = -d * p.eout - S(1);
        if (ex > 0) c += cv * dt * w * ex;
    }
    if (d >= 0 && p.ein <= 0) c += cv * dt * w * d;
    inj = k * (pc - mcp);
    return c;
}

// storage term of a unit at one step, with both ends moving linearly
template <typename S>
struct StoTerm {
    S ye, ye1, dye, dye1, kk, w, c0, j0;
    __device__ S eval(const StoPar<S> &p, S dt, S cv, S t, S *sec3, S &dj) const {
        S d1, j1;
        const S c1 = soc_cost_p(p, dt, cv, kk, w, ye + t * dye, ye1 + t * dye1, d1, j1);
        dj = j1 - j0;
        sec3[0] = d1;
        sec3[1] = p.ein > 0 ? d1 - p.ein : S(1);
        sec3[2] = -d1 * p.eout - S(1);
        return c1 - c0;
    }
};

// SoC atom (s, n) at instant τ: storage terms of τ (moving start) and τm (moving end), the
// balances of n at τ and τm; secondaries [τ: 3 storage, balance; τm: 3 storage, balance].
// With GEN: the storage–gas pair, gas at τm (slope aa) and at τ (slope ab).
template <typename S, bool GEN>
struct FSoc {
    static constexpr int NSEC = 8, NBP = GEN ? 6 : 2;
    static constexpr bool STOR = true;
    __device__ static constexpr bool gmask(int j) { return j != 3 && j != 7; }
    Ctx<S> k;
    StoPar<S> p;
    StoTerm<S> st_t, st_m;     // the storage term at τ and at τm
    S S0t, phit, S0m, phim;
    // gas terms (GEN): at τ (b) and τm (a)
    S ya, yb, kap, w, mc, aa, ab, c0a, p0a, c0b, p0b;
    mutable int nev = 0;   // ΔC evaluations (statistics)
    __device__ S eval(S t, S *sec, S &drt, S &drm) const {
        nev++;
        // step τ first (the reference evaluates the later step first)
        S d = 0;
        drt = 0;
        if (GEN) {
            S c1, p1;
            gen_cost(mc, kap, w, k.dt, k.cv, yb + t * ab, c1, p1);
            d += c1 - c0b;
            drt -= p1 - p0b;
        }
        S dj;
        d += st_t.eval(p, k.dt, k.cv, t, sec, dj);
        drt += dj;
        {
            const S S1 = S0t + drt;
            d += k.dt * (phibal(S1, k.cs, k.cd) - phit);
            sec[3] = S1;
        }
        S e = 0;
        drm = 0;
        if (GEN) {
            S c1, p1;
            gen_cost(mc, kap, w, k.dt, k.cv, ya + t * aa, c1, p1);
            e += c1 - c0a;
            drm -= p1 - p0a;
        }
        e += st_m.eval(p, k.dt, k.cv, t, sec + 4, dj);
        drm += dj;
        {
            const S S1 = S0m + drm;
            e += k.dt * (phibal(S1, k.cs, k.cd) - phim);
            sec[7] = S1;
        }
        return d + e;
    }
    __device__ S eval(S t, S *sec) const {
        S a, b;
        return eval(t, sec, a, b);
    }
};

// ---------------------------------------------------------------------------------------------
// the exact line sampler, register-resident
//
// walk! (src/local.jl:1294-1443) visits the points of one side in walking order: the next
// primary breakpoint beyond the current point is found by a scan of the breakpoint registers
// (the same sequence as the sorted list with duplicates skipped), the secondary sign changes of
// a segment by a scan of the secondaries, storage roots first and the other roots on the pieces
// between them (piece_roots!, :1267-1291), ray steps beyond the last breakpoint. The points are
// not stored: a visitor consumes them. sample_line! (:1501-1560) walks both sides accumulating
// the segment log-masses (logseg) online, draws the side and the segment with the first uniform,
// re-walks that side to the segment, and draws the offset (trexp_offset) with the second.
// ---------------------------------------------------------------------------------------------

// next root in walking order among the flagged entries; consumes it; false if none
template <typename S, int NS>
__device__ __forceinline__ bool next_root(const S *rt, bool *hr, S sgn, S &x) {
    int jb = -1;
    S xb = 0;
#pragma unroll
    for (int j = 0; j < NS; j++)
        if (hr[j] && (jb < 0 || sgn * (xb - rt[j]) > 0)) {
            jb = j;
            xb = rt[j];
        }
    if (jb < 0) return false;
#pragma unroll
    for (int j = 0; j < NS; j++)
        if (j == jb) hr[j] = false;
    x = xb;
    return true;
}

// visit one point: 0 go on, 2 the visitor stops the walk, -1 the side has more than PMAX points
template <typename S, class V>
__device__ __forceinline__ int visit_point(V &vis, S x, S cx, S &cmin, int &np) {
    if (++np > PMAX) return -1;
    if (vis(x, cx)) return 2;
    cmin = fmin(cmin, cx);
    return 0;
}

// sign changes of the non-storage secondaries on a piece (a, b) where they are linear
template <typename S, class F, class V>
__device__ int piece_roots(const F &f, S beta, S &cmin, const S *A, const S *B, S a, S b, S sgn, int &np, V &vis) {
    constexpr int NS = F::NSEC;
    S rt[NS];
    bool hr[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) {
        hr[j] = false;
        rt[j] = 0;
        if (F::gmask(j)) continue;
        const S u = A[j], v = B[j];
        if ((u < 0) != (v < 0) && u != v) {
            const S x = a + (b - a) * u / (u - v);
            if (sgn * (x - a) > 0 && sgn * (b - x) > 0) {
                rt[j] = x;
                hr[j] = true;
            }
        }
    }
    S x, tmp[NS];
    while (next_root<S, NS>(rt, hr, sgn, x)) {
        const int w = visit_point(vis, x, beta * f.eval(x, tmp), cmin, np);
        if (w) return w;
    }
    return 0;
}

// 1 proper (stopped maxnats above the minimum), 0 improper, -1 point cap, 2 stopped by the visitor
template <typename S, class F, class V>
__device__ int walk(const F &f, S beta, S maxnats, const S *bp, int nbp, S sgn, const S *sec0, V &vis) {
    constexpr int NS = F::NSEC, NB = F::NBP;
    S secprev[NS], sec[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) secprev[j] = sec0[j];
    S cmin = 0, tp = 0, delta = 0, clast = 0;
    int nx = 0, np = 0;
    while (true) {
        S t = 0;
        bool have = false;
#pragma unroll
        for (int q = 0; q < NB; q++)
            if (q < nbp) {
                const S b = bp[q];
                if (sgn * (b - tp) > 0 && (!have || sgn * (t - b) > 0)) {
                    t = b;
                    have = true;
                }
            }
        if (!have) {
            // beyond the last breakpoint: the step from the outgoing slope at tp, capped
            const S hp = S(1e-6) * fmax(S(1), fabs(tp));
            S tmp[NS];
            const S sl = (beta * f.eval(tp + sgn * hp, tmp) - clast) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, S(1e3) * fmax(fabs(tp), S(1e-3)));
            else delta = delta == 0 ? fmax(fabs(tp), S(1e-6)) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;   // the ray left the floating-point range: improper
        }
        const S c = beta * f.eval(t, sec);
        S rt[NS];
        bool hr[NS], hg[NS];
        bool anyg = false;
#pragma unroll
        for (int j = 0; j < NS; j++) {
            hr[j] = hg[j] = false;
            rt[j] = 0;
            const S a = secprev[j], b = sec[j];
            if ((a < 0) != (b < 0) && a != b) {
                const S x = tp + (t - tp) * a / (a - b);
                if (sgn * (x - tp) > 0 && sgn * (t - x) > 0) {
                    rt[j] = x;
                    if (F::STOR && F::gmask(j)) {
                        hg[j] = true;
                        anyg = true;
                    } else {
                        hr[j] = true;
                    }
                }
            }
        }
        int w;
        if (F::STOR && anyg) {
            // storage roots split the segment; the other roots are found on the pieces
            S a0 = tp, x;
            while (next_root<S, NS>(rt, hg, sgn, x)) {
                S secB[NS];
                const S cx = beta * f.eval(x, secB);
                if ((w = piece_roots(f, beta, cmin, secprev, secB, a0, x, sgn, np, vis))) return w;
                if ((w = visit_point(vis, x, cx, cmin, np))) return w;
#pragma unroll
                for (int j = 0; j < NS; j++) secprev[j] = secB[j];
                a0 = x;
            }
            if ((w = piece_roots(f, beta, cmin, secprev, sec, a0, t, sgn, np, vis))) return w;
#pragma unroll
            for (int j = 0; j < NS; j++) secprev[j] = sec[j];
        } else {
#pragma unroll
            for (int j = 0; j < NS; j++) secprev[j] = sec[j];
            S x, tmp[NS];
            while (next_root<S, NS>(rt, hr, sgn, x))
                if ((w = visit_point(vis, x, beta * f.eval(x, tmp), cmin, np))) return w;
        }
        if ((w = visit_point(vis, t, c, cmin, np))) return w;
        clast = c;
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}

template <typename S>
__device__ __forceinline__ S seg_logmass(S ta, S ca, S tb, S cb) {
    const S L = fabs(tb - ta);
    return L > 0 ? -fmin(ca, cb) + logseg<S>(L, fabs(cb - ca)) : S(-INFINITY);
}

// pass 1: online log-sum-exp of the segment masses of one side
template <typename S>
struct MassAcc {
    S tp = 0, cp = 0, mx = -INFINITY, s = 0;
    __device__ bool operator()(S t, S c) {
        const S lm = seg_logmass(tp, cp, t, c);
        if (lm > mx) {
            s = s * exp(mx - lm) + S(1);
            mx = lm;
        } else {
            s += exp(lm - mx);
        }
        tp = t;
        cp = c;
        return false;
    }
};

// pass 2: the segment where the cumulative mass from t = 0 outward crosses the target
template <typename S>
struct SegSel {
    S tp = 0, cp = 0, c0 = 0, target = 0, MX = 0;
    bool left = false;
    S ta = 0, ca = 0, tb = 0, cb = 0;
    __device__ bool operator()(S t, S c) {
        const S lm = seg_logmass(tp, cp, t, c);
        const S m = exp(lm - MX);
        c0 += m;
        if (left) {
            ta = t; ca = c; tb = tp; cb = cp;
        } else {
            ta = tp; ca = cp; tb = t; cb = c;
        }
        tp = t;
        cp = c;
        return lm > S(-INFINITY) && (left ? c0 > target : target <= c0);
    }
};

// sample_line!: status 0 drawn, 1 improper (t = 0), 2 guard (t = 0)
template <typename S, class F>
__device__ S sample_line(const F &f, const S *bps, int nbp, S beta, S maxnats, S useg, S uoff, int &status) {
    constexpr int NB = F::NBP;
    S bp[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const S x = q < nbp ? bps[q] : S(0);
        bp[q] = isfinite(x) ? x : S(0);   // 0 is never beyond the current point of a walk
    }
    S sec0[F::NSEC];
    f.eval(S(0), sec0);
    MassAcc<S> mr, ml;
    int w = walk(f, beta, maxnats, bp, NB, S(1), sec0, mr);
    if (w != 1) {
        status = w == 0 ? 1 : 2;
        return 0;
    }
    w = walk(f, beta, maxnats, bp, NB, S(-1), sec0, ml);
    if (w != 1) {
        status = w == 0 ? 1 : 2;
        return 0;
    }
    const S MX = fmax(ml.mx, mr.mx);
    const S Ml = ml.s * exp(ml.mx - MX), Mr = mr.s * exp(mr.mx - MX);
    const S u = useg * (Ml + Mr);
    SegSel<S> sel;
    sel.MX = MX;
    sel.left = u <= Ml;
    sel.target = sel.left ? Ml - u : u - Ml;
    walk(f, beta, maxnats, bp, NB, sel.left ? S(-1) : S(1), sec0, sel);
    status = 0;
    const S L = sel.tb - sel.ta;
    const S g = (sel.cb - sel.ca) / L;
    const S t = (g >= 0 || fabs(g * L) < S(1e-10)) ? sel.ta + trexp_offset(g, L, uoff)
                                                     : sel.tb - trexp_offset(-g, L, uoff);
    if (!isfinite(t)) {   // a non-finite draw (round-off at the range limits): the move is not made
        status = 2;
        return 0;
    }
    return t;
}

// ---------------------------------------------------------------------------------------------
// kernels (one per family): thread g = site·R + replica
// ---------------------------------------------------------------------------------------------

struct SweepArgs {
    uint32_t k0, k1, ctr, round;
    int tau0, tstride, ntau;      // steps of the round: tau0 + q·tstride, q < ntau
    int s;                         // storage carrier (SOC, SPAIR)
    const int *nodes;              // THETA: node list of the round
    int nnodes;
    unsigned long long *stats;     // [5][STAT_SLOTS]: attempted, accepted, improper, guard, ΔC evaluations
};

template <typename S>
struct SweepState {
    S *v, *r, *F;
    const S *x;
};

__device__ __forceinline__ void uniforms2(uint32_t k0, uint32_t k1, int j, uint32_t site, uint32_t round, uint32_t ctr,
                                          double &u1, double &u2) {
    uint32_t w[4];
    philox4x32(k0, k1, (uint32_t)j, site, round, ctr, w);
    u1 = u01(w[0], w[1]);
    u2 = u01(w[2], w[3]);
}

// A uniform of [0, 1) in the state precision: the 53-bit value rounded to S stays below 1 (in
// FP32 a value within 2^-25 of 1 would round to 1, and log1p(−u(1 − e^{−x})) would be −∞).
template <typename S>
__device__ __forceinline__ S unit_open(double u) {
    const S v = (S)u;
    return v < S(1) ? v : (S)0x1.fffffep-1;   // the largest float below 1 (a double u is always < 1)
}




// ---------------------------------------------------------------------------------------------
// one move at a site of one replica: prepare the family's terms from the state at (τ, j), draw
// t with the two uniforms, commit. Shared by the atom rounds and the random families.
// ---------------------------------------------------------------------------------------------

struct MoveOut {
    int att = 0, acc = 0, improper = 0, guard = 0, nev = 0;
    __device__ void add(int status, bool accepted, int ev) {
        att++;
        acc += accepted;
        improper += status == 1;
        guard += status == 2;
        nev += ev;
    }
};

// wave-reduced counts of a thread's moves, one atomic per statistic per wave (every lane of the
// wave must call it)
__device__ __forceinline__ void add_counts(unsigned long long *st, MoveOut o) {
    int v[5] = {o.att, o.acc, o.improper, o.guard, o.nev};
#pragma unroll
    for (int q = 0; q < 5; q++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
    if ((threadIdx.x & 63) == 0) {
        const unsigned slot = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) & (STAT_SLOTS - 1);
#pragma unroll
        for (int q = 0; q < 5; q++)
            if (v[q]) atomicAdd(st + q * STAT_SLOTS + slot, (unsigned long long)v[q]);
    }
}

// gas utilization of node n moves by t (`prep_atom!` / `unit`)
template <typename S>
__device__ void site_gen(const DevModel<S> &m, SweepState<S> &st, int n, int tau, int j, double u1, double u2, S beta,
                         S maxnats, MoveOut &out) {
    const size_t o = ((size_t)n * m.T + tau) * m.R + j;
    S *yg = st.v + plane(m.N, m.T, m.R);
    FGen<S> f;
    f.k = {m.dt, m.cv, m.c_shed, m.c_dump};
    f.kap = m.kc[ESM_KGAS * m.N + n];
    f.w = f.kap + m.kappa0;
    f.mc = m.mc[ESM_KGAS * m.N + n];
    f.a = 1;
    f.y0 = yg[o];
    gen_cost(f.mc, f.kap, f.w, m.dt, m.cv, f.y0, f.c0, f.p0);
    f.S0 = st.r[o];
    f.phi0 = phibal(f.S0, m.c_shed, m.c_dump);
    const S bps[2] = {(S(0) - f.y0) / f.a, (S(1) - f.y0) / f.a};
    int status = 0;
    const S t = sample_line(f, bps, 2, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
    out.add(status, t != 0, f.nev);
    if (t != 0) {
        S sec[1], dr;
        f.eval(t, sec, dr);
        yg[o] = f.y0 + t * f.a;
        st.r[o] += dr;
    }
}

// voltage angle of node n moves by t (`prep_atom!` of the angle field / `potential`); D bounds
// the node's degree
template <typename S, int D>
__device__ void site_theta(const DevModel<S> &m, SweepState<S> &st, int n, int tau, int j, double u1, double u2, S beta,
                           S maxnats, MoveOut &out) {
    const size_t TR = (size_t)m.T * m.R;
    const size_t o = (size_t)tau * m.R + j;
    FTheta<S, D> f;
    f.k = {m.dt, m.cv, m.c_shed, m.c_dump};
    const int q0 = m.adj_ptr[n];
    f.nl = m.adj_ptr[n + 1] - q0;
    int other[D], line[D];
    f.npos = 0;
    S bps[2 * D];
#pragma unroll
    for (int l = 0; l < D; l++) {
        if (l < f.nl) {
            const int e = m.adj_line[q0 + l];
            line[l] = e;
            other[l] = m.adj_other[q0 + l];
            const bool isb0 = m.bus0[e] == n;
            if (l == 0) f.npos = isb0 ? 0 : 1;
            f.sgn_n[l] = isb0 ? S(1) : S(-1);
            f.df[l] = isb0 ? m.lb[e] : -m.lb[e];
            f.se[l] = m.ls[e];
            f.f0[l] = st.F[(size_t)e * TR + o];
            f.c0[l] = fmax(fabs(f.f0[l]) - f.se[l], S(0));
            f.S0o[l] = st.r[(size_t)other[l] * TR + o];
            f.phio[l] = phibal(f.S0o[l], m.c_shed, m.c_dump);
            bps[2 * l] = (f.se[l] - f.f0[l]) / f.df[l];
            bps[2 * l + 1] = (-f.se[l] - f.f0[l]) / f.df[l];
        }
    }
    f.S0n = st.r[(size_t)n * TR + o];
    f.phin = phibal(f.S0n, m.c_shed, m.c_dump);
    int status = 0;
    const S t = sample_line(f, bps, 2 * f.nl, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
    out.add(status, t != 0, f.nev);
    if (t != 0) {
        S sec[D + 1], drn, dro[D];
        f.eval(t, sec, drn, dro);
        st.v[(size_t)n * TR + o] += t;
        st.r[(size_t)n * TR + o] += drn;
#pragma unroll
        for (int l = 0; l < D; l++) {
            if (l < f.nl) {
                st.F[(size_t)line[l] * TR + o] += t * f.df[l];
                st.r[(size_t)other[l] * TR + o] += dro[l];
            }
        }
    }
}

// state of charge of storage s at node n and instant τ moves by t; with GEN the storage–gas pair
// (`prep_spair!`, src/local.jl:1708-1726): the gas unit compensates at τm (+κs/κ) and τ (−κs/κ).
// Returns without a move when the pair is not valid (κs or κ below kmin).
template <typename S, bool GEN>
__device__ void site_soc(const DevModel<S> &m, SweepState<S> &st, int s, int n, int tau, int j, double u1, double u2,
                         S beta, S maxnats, S kmin, MoveOut &out) {
    const int taum = tau == 0 ? m.T - 1 : tau - 1;
    const int tau1 = tau + 1 == m.T ? 0 : tau + 1;
    const size_t TR = (size_t)m.T * m.R;
    S *ye = st.v + plane((2 + s) * m.N, m.T, m.R) + (size_t)n * TR;
    S *yg = st.v + plane(m.N, m.T, m.R) + (size_t)n * TR;
    S *rr = st.r + (size_t)n * TR;
    const size_t ot = (size_t)tau * m.R + j, om = (size_t)taum * m.R + j, o1 = (size_t)tau1 * m.R + j;
    FSoc<S, GEN> f;
    f.k = {m.dt, m.cv, m.c_shed, m.c_dump};
    {
        const int us = s * m.N + n;
        f.p = {m.hours[us], m.eta_in[us], m.eta_out[us], m.qt[us], m.mc[(ESM_KSTOR + s) * m.N + n]};
    }
    const S kst = m.kc[(ESM_KSTOR + s) * m.N + n];
    if (GEN) {
        const int u = s * m.N + n;
        const S ks = kst * m.hours[u] / m.dt;
        f.kap = m.kc[ESM_KGAS * m.N + n];
        if (!(ks >= kmin && fabs(f.kap) >= kmin)) return;
        f.w = f.kap + m.kappa0;
        f.mc = m.mc[ESM_KGAS * m.N + n];
        f.aa = ks / f.kap;
        f.ab = -ks / f.kap;
        f.ya = yg[om];
        f.yb = yg[ot];
        gen_cost(f.mc, f.kap, f.w, m.dt, m.cv, f.ya, f.c0a, f.p0a);
        gen_cost(f.mc, f.kap, f.w, m.dt, m.cv, f.yb, f.c0b, f.p0b);
    }
    const S e_t = ye[ot], e_1 = ye[o1], e_m = ye[om];
    S d;
    f.st_t = {e_t, e_1, S(1), S(0), kst, kst + m.kappa0, S(0), S(0)};
    f.st_t.c0 = soc_cost_p(f.p, m.dt, m.cv, kst, kst + m.kappa0, e_t, e_1, d, f.st_t.j0);
    f.st_m = {e_m, e_t, S(0), S(1), kst, kst + m.kappa0, S(0), S(0)};
    f.st_m.c0 = soc_cost_p(f.p, m.dt, m.cv, kst, kst + m.kappa0, e_m, e_t, d, f.st_m.j0);
    f.S0t = rr[ot];
    f.phit = phibal(f.S0t, m.c_shed, m.c_dump);
    f.S0m = rr[om];
    f.phim = phibal(f.S0m, m.c_shed, m.c_dump);
    // primary breakpoints: step τm's entries, then step τ's (the SoC entry before the gas)
    S bps[6];
    int nb = 0;
    if (GEN) {
        bps[nb++] = (S(0) - f.ya) / f.aa;
        bps[nb++] = (S(1) - f.ya) / f.aa;
    }
    bps[nb++] = (S(0) - e_t) / S(1);
    bps[nb++] = (S(1) - e_t) / S(1);
    if (GEN) {
        bps[nb++] = (S(0) - f.yb) / f.ab;
        bps[nb++] = (S(1) - f.yb) / f.ab;
    }
    int status = 0;
    const S t = sample_line(f, bps, nb, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
    out.add(status, t != 0, f.nev);
    if (t != 0) {
        S sec[8], drt, drm;
        f.eval(t, sec, drt, drm);
        ye[ot] = e_t + t;
        rr[ot] += drt;
        rr[om] += drm;
        if (GEN) {
            yg[om] = f.ya + t * f.aa;
            yg[ot] = f.yb + t * f.ab;
        }
    }
}

}  // namespace esm
