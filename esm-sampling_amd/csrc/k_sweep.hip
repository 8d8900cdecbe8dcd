// k_sweep.hip — one operation sweep of exact line-conditional moves on every replica.
//
// A move is a direction d at step τ (and the linked step τm = τ − 1, cyclic, for state-of-
// charge moves); the conditional of t along it is ∝ exp(−β_v ΔC(t)) with ΔC piecewise linear.
// Every thread draws one site (node, step) of one replica exactly, as the reference's line
// sampler does (src/local.jl:1294-1560): primary breakpoints sorted outward from t = 0, the
// sign changes of the secondary quantities (storage dynamics, node balances) between them, the
// storage roots first where they bend the balances (`piece_roots!`), the walk stopping
// `maxnats` above the running minimum (ray steps beyond the last breakpoint, at most 80, else
// the line is improper and t = 0), then segment log-masses (`logseg`) and inversion with one
// uniform, and the truncated-exponential offset (`trexp_offset`) with a second.
// ΔC is evaluated from the terms recorded at t = 0, in the arithmetic of `eval_cached!`
// (src/local.jl:926-981): generator terms, line overloads, storage terms, node balances.
//
// Families (the complete level of the synthetic system, one dispatchable unit per node):
//   GEN    gas utilization atom at (n, τ)                      (atoms / `unit`)
//   THETA  voltage angle of node n at τ                        (`potential`)
//   SOC    state of charge of storage s at instant τ           (atoms of the SoC fields)
//   SPAIR  storage s at τ with the gas unit compensating at τm and τ (`prep_spair!`, :1708-1726)
// A round is a set of sites whose terms are disjoint, so every site of a round reads the state
// before the round and the order of sites inside a round does not matter: GEN all (n, τ);
// THETA the nodes of one colour of a distance-2 colouring, every τ; SOC and SPAIR one storage
// carrier and one time-parity class (τ and τ ± 1 share a storage term).
//
// Random numbers: per (site, replica) two uniforms from Philox4x32-10 with counter
// (replica, n·T + τ, round id, sweep counter).
#include "esm_internal.h"

namespace esm {

constexpr int PMAX = 48;      // walk points kept per side (guard: t = 0 beyond)

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
    __device__ S eval(S t, S *sec, S &dr) const {
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
    __device__ S eval(S t, S *sec, S &drn, S *dro) const {
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
                S S1, ph0;
                if (p == npos) {
                    S1 = S1n;
                    ph0 = phin;
                } else {
                    const int l = p < npos ? p : p - 1;
                    S1 = S1o[l];
                    ph0 = phio[l];
                }
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

// storage term of a unit at one step, with both ends moving linearly
template <typename S>
struct StoTerm {
    S ye, ye1, dye, dye1, kk, w, c0, j0;
    __device__ S eval(const DevModel<S> &m, int s, int n, S t, S *sec3, S &dj) const {
        S d1, j1, run;
        const S c1 = soc_cost(m, s, n, kk, w, ye + t * dye, ye1 + t * dye1, d1, j1, run);
        dj = j1 - j0;
        const int u = s * m.N + n;
        const S ein = m.eta_in[u];
        sec3[0] = d1;
        sec3[1] = ein > 0 ? d1 - ein : S(1);
        sec3[2] = -d1 * m.eta_out[u] - S(1);
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
    const DevModel<S> *m;
    int s, n;
    StoTerm<S> st_t, st_m;     // the storage term at τ and at τm
    S S0t, phit, S0m, phim;
    // gas terms (GEN): at τ (b) and τm (a)
    S ya, yb, kap, w, mc, aa, ab, c0a, p0a, c0b, p0b;
    __device__ S eval(S t, S *sec, S &drt, S &drm) const {
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
        d += st_t.eval(*m, s, n, t, sec, dj);
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
        e += st_m.eval(*m, s, n, t, sec + 4, dj);
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
// the exact line sampler
// ---------------------------------------------------------------------------------------------

// Sign changes of the non-storage secondaries on a piece (a, b) where they are linear, appended
// in walking order (piece_roots!, src/local.jl:1267-1291). Returns false when the walk buffer
// is full.
template <typename S, class F>
__device__ bool piece_roots(const F &f, S beta, S &cmin, const S *A, const S *B, S a, S b, S sgn, S *pts, S *cs,
                            int &np) {
    S sr[F::NSEC];
    int ns = 0;
#pragma unroll
    for (int j = 0; j < F::NSEC; j++) {
        if (F::gmask(j)) continue;
        const S u = A[j], v = B[j];
        if ((u < 0) != (v < 0) && u != v) {
            const S x = a + (b - a) * u / (u - v);
            if (sgn * (x - a) > 0 && sgn * (b - x) > 0) sr[ns++] = x;
        }
    }
    if (ns == 0) return true;
    insort(sr, ns, sgn);
    S tmp[F::NSEC];
    for (int q = 0; q < ns; q++) {
        const S cx = beta * f.eval(sr[q], tmp);
        if (np >= PMAX) return false;
        pts[np] = sr[q];
        cs[np] = cx;
        np++;
        cmin = fmin(cmin, cx);
    }
    return true;
}

// walk! (src/local.jl:1294-1443) without requirement-tracking kinks. Returns 1 (proper),
// 0 (improper: no point maxnats above the minimum within 80 ray steps) or −1 (buffer full).
template <typename S, class F>
__device__ int walk(const F &f, S beta, S maxnats, const S *bp, int nb, S sgn, const S *sec0, S *pts, S *cs, int &np) {
    constexpr int NS = F::NSEC;
    S secprev[NS], sec[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) secprev[j] = sec0[j];
    S cmin = 0, tp = 0, delta = 0;
    int idx = 0, nx = 0;
    np = 0;
    while (true) {
        S t;
        if (idx < nb) {
            t = bp[idx++];
            if (t == tp) continue;
        } else {
            const S cprev = np == 0 ? S(0) : cs[np - 1];
            const S hp = S(1e-6) * fmax(S(1), fabs(tp));
            S tmp[NS];
            const S sl = (beta * f.eval(tp + sgn * hp, tmp) - cprev) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, S(1e3) * fmax(fabs(tp), S(1e-3)));
            else delta = delta == 0 ? fmax(fabs(tp), S(1e-6)) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;   // the ray left the floating-point range: improper
        }
        const S c = beta * f.eval(t, sec);
        S roots[NS], groots[NS];
        int nr = 0, ng = 0;
#pragma unroll
        for (int j = 0; j < NS; j++) {
            const S a = secprev[j], b = sec[j];
            if ((a < 0) != (b < 0) && a != b) {
                const S x = tp + (t - tp) * a / (a - b);
                if (sgn * (x - tp) > 0 && sgn * (t - x) > 0) {
                    if (F::STOR && F::gmask(j)) groots[ng++] = x;
                    else roots[nr++] = x;
                }
            }
        }
        if (F::STOR && ng > 0) {
            S secB[NS];
            insort(groots, ng, sgn);
            S a0 = tp;
            for (int q = 0; q < ng; q++) {
                const S x = groots[q];
                const S cx = beta * f.eval(x, secB);
                if (!piece_roots(f, beta, cmin, secprev, secB, a0, x, sgn, pts, cs, np)) return -1;
                if (np >= PMAX) return -1;
                pts[np] = x;
                cs[np] = cx;
                np++;
                cmin = fmin(cmin, cx);
#pragma unroll
                for (int j = 0; j < NS; j++) secprev[j] = secB[j];
                a0 = x;
            }
            if (!piece_roots(f, beta, cmin, secprev, sec, a0, t, sgn, pts, cs, np)) return -1;
#pragma unroll
            for (int j = 0; j < NS; j++) secprev[j] = sec[j];
        } else {
#pragma unroll
            for (int j = 0; j < NS; j++) secprev[j] = sec[j];
            if (nr > 0) {
                insort(roots, nr, sgn);
                S tmp[NS];
                for (int q = 0; q < nr; q++) {
                    const S cx = beta * f.eval(roots[q], tmp);
                    if (np >= PMAX) return -1;
                    pts[np] = roots[q];
                    cs[np] = cx;
                    np++;
                    cmin = fmin(cmin, cx);
                }
            }
        }
        if (np >= PMAX) return -1;
        pts[np] = t;
        cs[np] = c;
        np++;
        cmin = fmin(cmin, c);
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}

// sample_line! (src/local.jl:1501-1560): status 0 drawn, 1 improper (t = 0), 2 guard (t = 0)
template <typename S, class F>
__device__ S sample_line(const F &f, const S *bps, int nbp, S beta, S maxnats, S useg, S uoff, int &status) {
    S right[F::NBP], left[F::NBP];
    int nr = 0, nl = 0;
    for (int q = 0; q < nbp; q++) {
        const S x = bps[q];
        if (!isfinite(x)) continue;
        if (x > 0) right[nr++] = x;
        else if (x < 0) left[nl++] = x;
    }
    insort(right, nr, S(1));
    insort(left, nl, S(-1));
    S sec0[F::NSEC];
    f.eval(S(0), sec0);
    S pr[PMAX], cr[PMAX], pl[PMAX], cl[PMAX];
    int npr, npl;
    int w = walk(f, beta, maxnats, right, nr, S(1), sec0, pr, cr, npr);
    if (w != 1) {
        status = w == 0 ? 1 : 2;
        return 0;
    }
    w = walk(f, beta, maxnats, left, nl, S(-1), sec0, pl, cl, npl);
    if (w != 1) {
        status = w == 0 ? 1 : 2;
        return 0;
    }
    status = 0;
    // points in increasing order: the left walk reversed, 0, the right walk
    const int P = npl + 1 + npr;
    auto pt = [&](int q) -> S { return q < npl ? pl[npl - 1 - q] : (q == npl ? S(0) : pr[q - npl - 1]); };
    auto cv = [&](int q) -> S { return q < npl ? cl[npl - 1 - q] : (q == npl ? S(0) : cr[q - npl - 1]); };
    auto lmass = [&](int q) -> S {
        const S L = pt(q + 1) - pt(q);
        const S c0 = cv(q), c1 = cv(q + 1);
        return L > 0 ? -fmin(c0, c1) + logseg<S>(L, fabs(c1 - c0)) : S(-INFINITY);
    };
    S mx = -INFINITY;
    for (int q = 0; q < P - 1; q++) mx = fmax(mx, lmass(q));
    S tot = 0;
    for (int q = 0; q < P - 1; q++) tot += exp(lmass(q) - mx);
    const S u = useg * tot;
    S acc = 0;
    int j = P - 2;
    for (int q = 0; q < P - 1; q++) {
        acc += exp(lmass(q) - mx);
        if (u <= acc) {
            j = q;
            break;
        }
    }
    const S L = pt(j + 1) - pt(j);
    const S g = (cv(j + 1) - cv(j)) / L;
    const S t = (g >= 0 || fabs(g * L) < S(1e-10)) ? pt(j) + trexp_offset(g, L, uoff) : pt(j + 1) - trexp_offset(-g, L, uoff);
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
    unsigned long long *stats;     // [4]: attempted, accepted, improper, guard
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

// wave-aggregated statistics
__device__ __forceinline__ void add_stats(unsigned long long *st, bool att, bool acc, int status) {
    const unsigned long long ma = __ballot(att), mc = __ballot(att && acc), mi = __ballot(att && status == 1),
                             mg = __ballot(att && status == 2);
    if ((threadIdx.x & 63) == 0) {
        if (ma) atomicAdd(st + 0, (unsigned long long)__popcll(ma));
        if (mc) atomicAdd(st + 1, (unsigned long long)__popcll(mc));
        if (mi) atomicAdd(st + 2, (unsigned long long)__popcll(mi));
        if (mg) atomicAdd(st + 3, (unsigned long long)__popcll(mg));
    }
}

template <typename S>
__global__ void __launch_bounds__(256) k_gen(DevModel<S> m, SweepState<S> st, SweepArgs a, S beta, S maxnats) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nsite = (size_t)m.N * a.ntau;
    bool att = false, acc = false;
    int status = 0;
    if (g < nsite * m.R) {
        const int j = (int)(g % m.R);
        const size_t site = g / m.R;
        const int n = (int)(site / a.ntau);
        const int tau = a.tau0 + (int)(site % a.ntau) * a.tstride;
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
        double u1, u2;
        uniforms2(a.k0, a.k1, j, (uint32_t)(n * m.T + tau), a.round, a.ctr, u1, u2);
        att = true;
        const S t = sample_line(f, bps, 2, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
        if (t != 0) {
            acc = true;
            S sec[1], dr;
            f.eval(t, sec, dr);
            yg[o] = f.y0 + t * f.a;
            st.r[o] += dr;
        }
    }
    add_stats(a.stats, att, acc, status);
}

template <typename S, int D>
__global__ void __launch_bounds__(256) k_theta(DevModel<S> m, SweepState<S> st, SweepArgs a, S beta, S maxnats) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nsite = (size_t)a.nnodes * a.ntau;
    bool att = false, acc = false;
    int status = 0;
    if (g < nsite * m.R) {
        const int j = (int)(g % m.R);
        const size_t site = g / m.R;
        const int n = a.nodes[site / a.ntau];
        const int tau = a.tau0 + (int)(site % a.ntau) * a.tstride;
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
        double u1, u2;
        uniforms2(a.k0, a.k1, j, (uint32_t)(n * m.T + tau), a.round, a.ctr, u1, u2);
        att = true;
        const S t = sample_line(f, bps, 2 * f.nl, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
        if (t != 0) {
            acc = true;
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
    add_stats(a.stats, att, acc, status);
}

template <typename S, bool GEN>
__global__ void __launch_bounds__(256) k_soc(DevModel<S> m, SweepState<S> st, SweepArgs a, S beta, S maxnats,
                                             S kmin) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nsite = (size_t)m.N * a.ntau;
    bool att = false, acc = false;
    int status = 0;
    if (g < nsite * m.R) {
        const int j = (int)(g % m.R);
        const size_t site = g / m.R;
        const int n = (int)(site / a.ntau);
        const int tau = a.tau0 + (int)(site % a.ntau) * a.tstride;
        const int taum = tau == 0 ? m.T - 1 : tau - 1;
        const int tau1 = tau + 1 == m.T ? 0 : tau + 1;
        const size_t TR = (size_t)m.T * m.R;
        const int s = a.s;
        S *ye = st.v + plane((2 + s) * m.N, m.T, m.R) + (size_t)n * TR;
        S *yg = st.v + plane(m.N, m.T, m.R) + (size_t)n * TR;
        S *rr = st.r + (size_t)n * TR;
        const size_t ot = (size_t)tau * m.R + j, om = (size_t)taum * m.R + j, o1 = (size_t)tau1 * m.R + j;
        FSoc<S, GEN> f;
        f.k = {m.dt, m.cv, m.c_shed, m.c_dump};
        f.m = &m;
        f.s = s;
        f.n = n;
        const S kst = m.kc[(ESM_KSTOR + s) * m.N + n];
        bool ok = true;
        if (GEN) {
            const int u = s * m.N + n;
            const S ks = kst * m.hours[u] / m.dt;
            f.kap = m.kc[ESM_KGAS * m.N + n];
            ok = ks >= kmin && fabs(f.kap) >= kmin;
            f.w = f.kap + m.kappa0;
            f.mc = m.mc[ESM_KGAS * m.N + n];
            f.aa = ks / f.kap;
            f.ab = -ks / f.kap;
            f.ya = yg[om];
            f.yb = yg[ot];
            gen_cost(f.mc, f.kap, f.w, m.dt, m.cv, f.ya, f.c0a, f.p0a);
            gen_cost(f.mc, f.kap, f.w, m.dt, m.cv, f.yb, f.c0b, f.p0b);
        }
        if (ok) {
            const S e_t = ye[ot], e_1 = ye[o1], e_m = ye[om];
            S d, run;
            f.st_t = {e_t, e_1, S(1), S(0), kst, kst + m.kappa0, S(0), S(0)};
            f.st_t.c0 = soc_cost(m, s, n, kst, kst + m.kappa0, e_t, e_1, d, f.st_t.j0, run);
            f.st_m = {e_m, e_t, S(0), S(1), kst, kst + m.kappa0, S(0), S(0)};
            f.st_m.c0 = soc_cost(m, s, n, kst, kst + m.kappa0, e_m, e_t, d, f.st_m.j0, run);
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
            double u1, u2;
            uniforms2(a.k0, a.k1, j, (uint32_t)(n * m.T + tau), a.round, a.ctr, u1, u2);
            att = true;
            const S t = sample_line(f, bps, nb, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
            if (t != 0) {
                acc = true;
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
    }
    add_stats(a.stats, att, acc, status);
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename S>
int launch_sweeps(esm_handle *h, int nsweeps, uint32_t k0, uint32_t k1, uint32_t ctr0, double maxnats) {
    const DevModel<S> m = dev_model<S>(h);
    SweepState<S> st{(S *)h->d_v, (S *)h->d_r, (S *)h->d_F, (const S *)h->d_x};
    const S beta = (S)h->m.beta_v, mxn = (S)maxnats, kmin = S(1);
    const int N = h->N, T = h->T, R = h->R;
    const unsigned B = 256;
    ProfScope ps(h, ESM_KERN_SWEEP);
    for (int sw = 0; sw < nsweeps; sw++) {
        SweepArgs a{};
        a.k0 = k0;
        a.k1 = k1;
        a.ctr = ctr0 + (uint32_t)sw;
        // GEN: every (n, τ)
        a.round = 0;
        a.tau0 = 0;
        a.tstride = 1;
        a.ntau = T;
        a.stats = h->d_stats + 0;
        {
            ProfScope pg(h, ESM_KERN_GEN);
            hipLaunchKernelGGL((k_gen<S>), dim3(nblk((size_t)N * T * R, B)), dim3(B), 0, h->stream, m, st, a, beta, mxn);
        }
        // THETA: colour by colour, degree buckets
        for (size_t rnd = 0; rnd < h->theta_rounds.size(); rnd++) {
            const auto &tr = h->theta_rounds[rnd];
            a.round = 1000u + (uint32_t)tr.colour;
            a.nodes = h->d_theta_nodes + tr.off;
            a.nnodes = tr.count;
            a.stats = h->d_stats + 4;
            const unsigned nb = nblk((size_t)tr.count * T * R, B);
            ProfScope pt(h, ESM_KERN_THETA);
            switch (tr.dmax) {
                case 4: hipLaunchKernelGGL((k_theta<S, 4>), dim3(nb), dim3(B), 0, h->stream, m, st, a, beta, mxn); break;
                case 8: hipLaunchKernelGGL((k_theta<S, 8>), dim3(nb), dim3(B), 0, h->stream, m, st, a, beta, mxn); break;
                case 16: hipLaunchKernelGGL((k_theta<S, 16>), dim3(nb), dim3(B), 0, h->stream, m, st, a, beta, mxn); break;
                default: set_error("theta degree bucket"); return ESM_EINVAL;
            }
        }
        a.nodes = nullptr;
        // SOC then SPAIR: per storage carrier and time-parity class
        for (int fam = 0; fam < 2; fam++) {
            for (int s = 0; s < ESM_NSTOR; s++) {
                for (size_t c = 0; c < h->tclass.size(); c++) {
                    a.s = s;
                    a.tau0 = h->tclass[c].first;
                    a.tstride = 2;
                    a.ntau = h->tclass[c].second;
                    a.round = (fam == 0 ? 2000u : 3000u) + (uint32_t)(s * 4 + (int)c);
                    a.stats = h->d_stats + (fam == 0 ? 8 : 12);
                    const unsigned nb = nblk((size_t)N * a.ntau * R, B);
                    ProfScope pf(h, fam == 0 ? ESM_KERN_SOC : ESM_KERN_SPAIR);
                    if (fam == 0)
                        hipLaunchKernelGGL((k_soc<S, false>), dim3(nb), dim3(B), 0, h->stream, m, st, a, beta, mxn, kmin);
                    else
                        hipLaunchKernelGGL((k_soc<S, true>), dim3(nb), dim3(B), 0, h->stream, m, st, a, beta, mxn, kmin);
                }
            }
        }
    }
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

template int launch_sweeps<float>(esm_handle *, int, uint32_t, uint32_t, uint32_t, double);
template int launch_sweeps<double>(esm_handle *, int, uint32_t, uint32_t, uint32_t, double);

}  // namespace esm
