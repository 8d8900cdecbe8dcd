// k_random.hip — the random moves of an operation sweep (`local_sweep!`, src/local.jl:1927-2062).
//
// Per step τ the reference draws round(extra·N) moves over the plain families
// (plain_families, :1912-1915) and runs them in a shuffled order together with the atoms; steps
// of one time-parity class run in parallel. Here, per step and sweep, k_movecounts makes the same
// draws (one uniform each, inverted on the cumulative weights) and counts them per family; per
// parity class the families then run in a fixed order (unit, potential, spair, xpair, xlocc,
// xloc, cell), every thread (step τ, replica j) running a family's moves of step τ in sequence:
// one kernel per atom-site family and one for the four table-driven families together (or one
// per family, esm_sweep_launch). A move's selection (node, partner, cell) comes from a
// counter of (family, index, step, sweep) and is the same for every replica, so the 64 lanes of a
// wave (64 replicas of one step) run one move shape on their own states; the line draw and the
// Jacobian test use counters of the replica.
//
// Families and their directions (the complete level of the synthetic system: one dispatchable
// unit per node, so `pair` and `cellpair` are void and are drawn but not run):
//   unit       the gas utilization of a node (`prep_random!` :unit)           site_gen
//   potential  the angle of a node (:potential)                                site_theta
//   spair      storage–gas pair at a node (`prep_spair!`)                      site_soc<GEN>
//   cell       +1 on every gas utilization of a cell of a partition level      FSparse
//   xloc       gas +1/κ_n at n, −1/κ_m at a neighbour m, the flow re-routed on the ball {n, m},
//              the leak into the residuals of the first layer (`prep_xloc!`)   FSparse
//   xlocc      xloc with the leak compensated pro rata by the gas unit of each leak node: a
//              kinked entry (σ/κ)·φ(t), φ(t) = (Q + g t)₊ − Q₊, and the Jacobian test
//              (`prorata!` :1782-1815, `logjac` :1874)                         FSparse
//   xpair      gas ±1/κ at any two nodes of one component, the flow re-routed globally,
//              Δθ = L⁺(e_n − e_m) (`prep_reroute!`)                            FSparse
// FSparse evaluates ΔC(t) of a table-driven move on the fly from the caches of (τ, j) (the terms
// at t = 0 are the cache values; nothing is written before the commit), in the arithmetic of
// `eval_cached!` (:926-981): gas entries, kinked entries, line overloads, node balances, each
// group summed in its listed order. The walk visits the primary breakpoints (clamp points of
// the entries, kinks, line limits) and the sign changes of the node balances between them in
// walking order, as `walk!` does; equal roots of several nodes are visited once.
#include <algorithm>

#include "cross_move.h"

namespace esm {

// positions of the plain families in the weight list
enum { W_UNIT = 0, W_PAIR = 1, W_XPAIR = 2, W_XLOCC = 3, W_XLOC = 4, W_POT = 5, W_CELLPAIR = 6, W_CELL = 7, W_SPAIR = 8,
       W_SXPAIR = 9 };
constexpr int NW_ALL = ESM_NWEIGHT + 1;   // the plain families, then the storage pair sxpair

struct CumW {
    double c[NW_ALL];
};

struct RandArgs {
    uint32_t k0, k1, ctr;
    int tau0, tstride, ntau;
    int w;                       // the family's position in the weight list
    const int *counts;           // [NW_ALL][T]
    unsigned long long *stats;   // [ESM_NSTAT][STAT_SLOTS] of the family
    const int *perm;             // the class's steps in launch order (null: in order)
    int rp;                      // threads per step: R, or R rounded up to a multiple of 64 (padded)
};


__device__ __forceinline__ uint32_t pick(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

// With UNI (the replica count a multiple of 64, so every wave holds the replicas of one step) the
// step, the move counts and the selection words are the same in all lanes of a wave; saying so
// lets the template and model-table reads become scalar loads.
template <bool UNI>
__device__ __forceinline__ int uniform(int x) { return UNI ? __builtin_amdgcn_readfirstlane(x) : x; }
template <bool UNI>
__device__ __forceinline__ void uniform4(uint32_t *s) {
    if (UNI)
#pragma unroll
        for (int i = 0; i < 4; i++) s[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)s[i]);
}

// the step of the wave that holds element g (replica fastest)
template <bool UNI>
__device__ __forceinline__ int step_of(const RandArgs &a, size_t g, int R) {
    const int i = (int)(g / R);
    return uniform<UNI>(a.perm ? a.perm[i] : a.tau0 + i * a.tstride);
}

// thread g of a one-wave launch over a class's steps: its replica j (threads j ≥ R of a padded
// step are idle) and whether it has work
__device__ __forceinline__ bool slot_of(const RandArgs &a, size_t g, int R, int &j) {
    j = (int)(g % (size_t)a.rp);
    return g < (size_t)a.ntau * a.rp && j < R;
}

// the selection words of move `k` of family `w` at step τ
__device__ __forceinline__ void selection(const RandArgs &a, int k, int tau, uint32_t *s) {
    philox4x32(a.k0, a.k1, 0x80000000u | ((uint32_t)a.w << 20) | (uint32_t)k, (uint32_t)tau, 9001u, a.ctr, s);
}

__device__ __forceinline__ void line_uniforms(const RandArgs &a, int j, int k, int tau, double &u1, double &u2) {
    uniforms2(a.k0, a.k1, j, (uint32_t)tau, 0x40000000u | ((uint32_t)a.w << 20) | (uint32_t)k, a.ctr, u1, u2);
}

// per step: round(extra·xmul·N) family draws, counted per family (xmul = 1 without sxpair)
__global__ void k_movecounts(int T, int nextra, CumW cw, uint32_t k0, uint32_t k1, uint32_t ctr, int *counts) {
    const int tau = blockIdx.x * blockDim.x + threadIdx.x;
    if (tau >= T) return;
    int cnt[NW_ALL];
#pragma unroll
    for (int q = 0; q < NW_ALL; q++) cnt[q] = 0;
    for (int d = 0; d < nextra; d++) {
        uint32_t w[4];
        philox4x32(k0, k1, 0x80000000u | (uint32_t)d, (uint32_t)tau, 9000u, ctr, w);
        const double u = u01(w[0], w[1]);
        int f = NW_ALL - 1;
#pragma unroll
        for (int q = NW_ALL - 1; q >= 0; q--)
            if (cw.c[q] >= u) f = q;
#pragma unroll
        for (int q = 0; q < NW_ALL; q++) cnt[q] += q == f;
    }
#pragma unroll
    for (int q = 0; q < NW_ALL; q++) counts[(size_t)q * T + tau] = cnt[q];
}

// ---------------------------------------------------------------------------------------------
// unit, potential, spair: the atom sites at a drawn node
// ---------------------------------------------------------------------------------------------

template <typename S, bool UNI>
__device__ __forceinline__ void run_unit(const DevModel<S> &m, SweepState<S> &st, const RandArgs &a, int tau, int j,
                                         S beta, S maxnats, MoveOut &out) {
    const int cnt = uniform<UNI>(a.counts[(size_t)a.w * m.T + tau]);
    for (int k = 0; k < cnt; k++) {
        uint32_t s[4];
        selection(a, k, tau, s);
        uniform4<UNI>(s);
        const int n = (int)pick(s[0], (uint32_t)m.N);
        double u1, u2;
        line_uniforms(a, j, k, tau, u1, u2);
        site_gen(m, st, n, tau, j, u1, u2, beta, maxnats, out);
    }
}

template <typename S, bool UNI>
__global__ void __launch_bounds__(256) k_rand_unit(DevModel<S> m, SweepState<S> st, RandArgs a, S beta, S maxnats) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    MoveOut out;
    int j;
    if (slot_of(a, g, m.R, j)) run_unit<S, UNI>(m, st, a, step_of<UNI>(a, g, a.rp), j, beta, maxnats, out);
    add_counts(a.stats, out);
}

// potential moves whose node degree falls in the bucket (DLO, D]: one kernel per bucket keeps
// the register footprint of its degree; a step's potential moves run bucket by bucket, each
// bucket in draw order
template <typename S, int DLO, int D, bool UNI>
__device__ __forceinline__ void run_pot(const DevModel<S> &m, SweepState<S> &st, const RandArgs &a, int tau, int j,
                                        S beta, S maxnats, MoveOut &out) {
    const int cnt = uniform<UNI>(a.counts[(size_t)a.w * m.T + tau]);
    for (int k = 0; k < cnt; k++) {
        uint32_t s[4];
        selection(a, k, tau, s);
        uniform4<UNI>(s);
        const int n = (int)pick(s[0], (uint32_t)m.N);
        const int deg = m.adj_ptr[n + 1] - m.adj_ptr[n];
        if (deg <= DLO || deg > D) continue;   // another bucket (degree 0: no angle direction)
        double u1, u2;
        line_uniforms(a, j, k, tau, u1, u2);
        site_theta<S, D>(m, st, n, tau, j, u1, u2, beta, maxnats, out);
    }
}

template <typename S, int DLO, int D, bool UNI>
__global__ void __launch_bounds__(256) k_rand_pot(DevModel<S> m, SweepState<S> st, RandArgs a, S beta, S maxnats) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    MoveOut out;
    int j;
    if (slot_of(a, g, m.R, j)) run_pot<S, DLO, D, UNI>(m, st, a, step_of<UNI>(a, g, a.rp), j, beta, maxnats, out);
    add_counts(a.stats, out);
}

template <typename S, bool UNI>
__device__ __forceinline__ void run_spair(const DevModel<S> &m, SweepState<S> &st, const RandArgs &a, int tau, int j,
                                          S beta, S maxnats, S kmin, MoveOut &out) {
    const int cnt = uniform<UNI>(a.counts[(size_t)a.w * m.T + tau]);
    for (int k = 0; k < cnt; k++) {
        uint32_t s[4];
        selection(a, k, tau, s);
        uniform4<UNI>(s);
        const int n = (int)pick(s[0], (uint32_t)m.N);
        const int sc = (int)pick(s[1], (uint32_t)ESM_NSTOR);   // the node's storage units: battery, H2
        double u1, u2;
        line_uniforms(a, j, k, tau, u1, u2);
        site_soc<S, true>(m, st, sc, n, tau, j, u1, u2, beta, maxnats, kmin, out);
    }
}

template <typename S, bool UNI>
__global__ void __launch_bounds__(256, min_waves<S>(3)) k_rand_spair(DevModel<S> m, SweepState<S> st, RandArgs a, S beta, S maxnats,
                                                    S kmin) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    MoveOut out;
    int j;
    if (slot_of(a, g, m.R, j)) run_spair<S, UNI>(m, st, a, step_of<UNI>(a, g, a.rp), j, beta, maxnats, kmin, out);
    add_counts(a.stats, out);
}

// ---------------------------------------------------------------------------------------------
// sxpair: the cross-node storage pair at one instant (`prep_sxpair!`, src/storage_cross.jl:147-167):
// unit A uniform among the units of the pairs (selection word 0), the partner by `sx_partner`
// (words 2–3), the instant τ (steps τm and τ). One wave per (step, replica), grid order replica
// fastest (the waves of one step read the same cache lines); its lanes hold the items of the
// pair's template (FXLane), and it runs the step's sxpair moves in draw order.
// ---------------------------------------------------------------------------------------------

constexpr int SXP_WAVES = 4;   // waves per workgroup of the sxpair launch

// one sxpair move of replica j on its segment of W lanes (lane `sub` of the segment)
template <typename S, int W, int KS>
__device__ __forceinline__ void sx_pair_move(const DevModel<S> &m, SweepState<S> &st, const RandArgs &a,
                                             const SxArgs<S> &sx, XlLds<S, W, KS> *sh, int sa, int na, int sb,
                                             int nb, int slot, int tau, int k, int j, int sub, S beta, S maxnats,
                                             S kmin, MoveOut &out) {
    const int N = m.N, T = m.T;
    FXLane<S, false, W, KS> f;
    if (!xl_prepare(f, m, st, sx.tb, sh, sa, na, sb, nb, slot, tau == 0 ? T - 1 : tau - 1, 1, j, sub, kmin)) return;
    double u1, u2;
    line_uniforms(a, j, k, tau, u1, u2);
    int status = 0;
    S t = sample_stor_pts<XL_NP>(f, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status, sh->pts, sub);
    if (t != 0) {
        bool any;
        const double lj = f.logjac(t, any);
        if (any) {
            uint32_t w[4];
            philox4x32(a.k0, a.k1, (uint32_t)j, (uint32_t)tau, 0x60000000u | ((uint32_t)a.w << 20) | (uint32_t)k, a.ctr,
                       w);
            if (!(u01(w[0], w[1]) < exp(fmin(0.0, lj)))) {
                t = 0;
                status = 2;
            }
        }
    }
    if (sub == 0) out.add(status, t != 0, f.nev);
    if (t != 0) {
        const size_t TR = (size_t)T * m.R;
        f.commit(t, st.r + j, st.F + j, st.v + plane(N, T, m.R) + j, st.v + j,
                 st.v + plane((2 + sa) * N, T, m.R) + (size_t)na * TR + j,
                 st.v + plane((2 + sb) * N, T, m.R) + (size_t)nb * TR + j);
    }
}

// the move's selection (unit A, partner, slot), the same in every lane; false: no move
template <typename S>
__device__ __forceinline__ bool sx_select(const DevModel<S> &m, const SxArgs<S> &sx, const RandArgs &a, int k, int tau,
                                          int &sa, int &na, int &sb, int &nb, int &slot) {
    uint32_t s[4];
    selection(a, k, tau, s);
    uniform4<true>(s);
    const int N = m.N;
    const int ua = sx.units[pick(s[0], (uint32_t)sx.nsx)];
    sa = ua / N;
    na = ua % N;
    if (m.kc[(ESM_KSTOR + sa) * N + na] * m.hours[ua] / m.dt < S(1)) return false;
    return sx_partner(m, sx, sa, na, u01(s[2], s[3]), slot, sb, nb);
}

// Four replicas per wave: each move runs on the narrowest segments its template fits — four
// 16-lane segments at once, two 32-lane segments twice, or (K64 > 0: a system with templates
// beyond 32 lanes) one 64-lane segment per replica, four times, with K64 kink slots per lane. The
// moves of a replica keep their draw order; the oracle sums each move over the same lanes.
template <typename S, int K64>
__global__ void __launch_bounds__(64 * SXP_WAVES, min_waves<S>(3, 1)) k_rand_sxpair_mix(DevModel<S> m,
                                                                                         SweepState<S> st, RandArgs a,
                                                                                         SxArgs<S> sx, S beta,
                                                                                         S maxnats, S kmin) {
    __shared__ XlLds<S, 16, 1> sh16[SXP_WAVES][4];
    __shared__ XlLds<S, 32, 1> sh32[SXP_WAVES][2];
    __shared__ XlLds<S, 64, (K64 > 0 ? K64 : 1)> sh64[SXP_WAVES];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ngrp = (m.R + 3) / 4;
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x, ngrp % SXP_WAVES == 0 ? ngrp / SXP_WAVES : 0);
    const size_t g = (size_t)blk * SXP_WAVES + wv;
    MoveOut out;
    if (g < (size_t)a.ntau * ngrp && sx.nsx > 0) {
        const int i = (int)(g / ngrp), j0 = (int)(g % ngrp) * 4;
        const int tau = __builtin_amdgcn_readfirstlane(a.perm ? a.perm[i] : a.tau0 + i * a.tstride);
        const int cnt = __builtin_amdgcn_readfirstlane(a.counts[(size_t)a.w * m.T + tau]);
        for (int k = 0; k < cnt; k++) {
            int sa, na, sb, nb, slot;
            if (!sx_select(m, sx, a, k, tau, sa, na, sb, nb, slot)) continue;
            if (xl_fits<S, 16, 1>(sx.tb, slot)) {
                const int j = j0 + lane / 16;
                if (j < m.R)
                    sx_pair_move<S, 16, 1>(m, st, a, sx, &sh16[wv][lane / 16], sa, na, sb, nb, slot, tau, k, j,
                                           lane % 16, beta, maxnats, kmin, out);
            } else if (xl_fits<S, 32, 1>(sx.tb, slot)) {
                for (int h = 0; h < 2; h++) {
                    const int j = j0 + 2 * h + lane / 32;
                    if (j < m.R)
                        sx_pair_move<S, 32, 1>(m, st, a, sx, &sh32[wv][lane / 32], sa, na, sb, nb, slot, tau, k, j,
                                               lane % 32, beta, maxnats, kmin, out);
                }
            } else if (K64 > 0 && xl_fits<S, 64, (K64 > 0 ? K64 : 1)>(sx.tb, slot)) {
                for (int h = 0; h < 4; h++) {
                    const int j = j0 + h;
                    if (j < m.R)
                        sx_pair_move<S, 64, (K64 > 0 ? K64 : 1)>(m, st, a, sx, &sh64[wv], sa, na, sb, nb, slot, tau, k,
                                                                  j, lane, beta, maxnats, kmin, out);
                }
            } else if ((lane & 15) == 0 && j0 + lane / 16 < m.R) {   // refused on the host; a guard if it ever came
                out.add(2, false, 0);
            }
        }
    }
    add_counts(a.stats, out);
}

// the lane form of the storage pairs: the narrowest of (32 lanes, 1 kink slot), (64, 1), (64, 2)
// with at least wmin lanes that every pair template of the system fits; 0 when none does (ks: the
// kink slots)
int sx_lane_form(const esm_handle *h, int wmin, int &ks) {
    const MoveTablesHost &t = h->mt;
    static const int forms[3][2] = {{32, 1}, {64, 1}, {64, 2}};
    for (const auto &f : forms) {
        if (f[0] < wmin) continue;
        bool ok = true;
        for (size_t sl = 0; ok && sl + 1 < t.sx_kk_off.size(); sl++) {
            const int K = t.sx_kk_off[sl + 1] - t.sx_kk_off[sl], Lw = t.sx_ln_off[sl + 1] - t.sx_ln_off[sl];
            const int Nn = t.sx_nd_off[sl + 1] - t.sx_nd_off[sl];
            const int NLp = t.sx_nd_l[t.sx_nd_off[sl + 1]] - t.sx_nd_l[t.sx_nd_off[sl]];
            ok = xl_fits_counts(f[0], f[1], K, Lw, Nn, NLp);
        }
        if (ok) {
            ks = f[1];
            return f[0];
        }
    }
    ks = 0;
    return 0;
}

// ---------------------------------------------------------------------------------------------
// table-driven moves (cell, xloc, xlocc, xpair)
// ---------------------------------------------------------------------------------------------

template <typename S>
struct Tables {
    const int *th_off, *th_node, *en_off, *en_node, *en_div, *kk_off, *kk_node, *ln_off, *ln_line, *nd_off, *nd_node,
        *nd_ent, *nd_kk, *nd_l, *nl_pos, *nl_sgn, *xp_off, *lev_off;
    const S *th_val, *en_c, *kk_g, *ln_df;
    const double *qcum;
    int xp0, cl0, nlev;
};

template <typename S>
struct FSparse {
    const DevModel<S> *m;
    const Tables<S> *tb;
    const S *yg, *r, *F;    // gas utilization, residual and flow planes
    size_t o, TR;           // (τ, j) offset inside a [·][T][R] plane, plane stride
    int e0, e1, q0, q1, l0, l1, n0, n1;
    bool comp;              // the kinks of the template are active (xlocc)
    S kmin;
    mutable int nev = 0;
    // split over the nw waves of a workgroup (wave wv takes every nw-th entry, node and line; the
    // partial results are combined through `red`, LDS [nw][RED][64], in wave order); nw = 1: none
    int wv = 0, nw = 1;
    S *red = nullptr;
    static constexpr int RED = 9;   // c, hb, nb, nr, more, rt[KR]

    __device__ S kap(int n) const { return m->kc[ESM_KGAS * m->N + n]; }
    // slope of gas entry e (entries below 1e-14 are zero, as `finalize!` makes them)
    __device__ S slope(int e) const {
        const S c = tb->en_c[e];
        const S a = tb->en_div[e] ? c / kap(tb->en_node[e]) : c;
        return fabs(a) < S(1e-14) ? S(0) : a;
    }
    __device__ S ygv(int n) const { return yg[(size_t)n * TR + o]; }
    __device__ S rv(int n) const { return r[(size_t)n * TR + o]; }
    __device__ S Fv(int e) const { return F[(size_t)e * TR + o]; }
    __device__ void gen(int n, S y, S &c, S &p) const {
        const S k = kap(n);
        gen_cost(m->mc[ESM_KGAS * m->N + n], k, k + m->kappa0, m->dt, m->cv, y, c, p);
    }
    // kink q (`prorata!` of one unit): false when it is not made; Q, e, P at t = 0, ak = σ/κ
    __device__ bool kink(int q, S &Q, S &g, S &ak, S &P, S &e, S &y0) const {
        const int n = tb->kk_node[q];
        const S k = kap(n);
        if (!(fabs(k) >= kmin)) return false;
        const S w = ygv(n);
        P = S(0) + k * w;
        const S Qs = S(0) + k * clampv(w, S(0), S(1));
        if (fabs(P) < S(1e-6)) return false;
        Q = rv(n) + Qs;
        e = P - fmax(Q, S(0));
        g = tb->kk_g[q];
        const S sig = k * w / P;
        ak = sig == S(0) ? S(0) : sig / k;
        y0 = w;
        return g != S(0);
    }
    __device__ static S kphi(S Q, S g, S t) { return fmax(Q + g * t, S(0)) - fmax(Q, S(0)); }

    // balance of touched node p at t
    __device__ S node_S(int p, S t) const {
        const int n = tb->nd_node[p];
        S dr = 0;
        const int ie = tb->nd_ent[p];
        if (ie >= 0) {
            const S a = slope(e0 + ie);
            if (a != S(0)) {
                const S y0 = ygv(n);
                S c0, p0, c1, p1;
                gen(n, y0, c0, p0);
                gen(n, y0 + t * a, c1, p1);
                dr -= p1 - p0;
            }
        }
        const int ik = tb->nd_kk[p];
        if (comp && ik >= 0) {
            S Q, g, ak, P, e, y0;
            if (kink(q0 + ik, Q, g, ak, P, e, y0) && ak != S(0)) {
                S c0, p0, c1, p1;
                gen(n, y0, c0, p0);
                gen(n, y0 + ak * kphi(Q, g, t), c1, p1);
                dr -= p1 - p0;
            }
        }
        for (int q = tb->nd_l[p]; q < tb->nd_l[p + 1]; q++) {
            const int pos = l0 + tb->nl_pos[q];
            const S f0 = Fv(tb->ln_line[pos]);
            const S d = (f0 + t * tb->ln_df[pos]) - f0;
            dr += tb->nl_sgn[q] > 0 ? d : -d;
        }
        return rv(n) + dr;
    }

    // balance of touched node p at ta and at tb (one read of the node's values)
    __device__ void node_S2(int p, S ta, S tb, S &A, S &B) const {
        const int n = tb_()->nd_node[p];
        S da = 0, db = 0;
        const int ie = tb_()->nd_ent[p];
        if (ie >= 0) {
            const S a = slope(e0 + ie);
            if (a != S(0)) {
                const S y0 = ygv(n);
                S c0, p0, c1, p1, c2, p2;
                gen(n, y0, c0, p0);
                gen(n, y0 + ta * a, c1, p1);
                gen(n, y0 + tb * a, c2, p2);
                da -= p1 - p0;
                db -= p2 - p0;
            }
        }
        const int ik = tb_()->nd_kk[p];
        if (comp && ik >= 0) {
            S Q, g, ak, P, e, y0;
            if (kink(q0 + ik, Q, g, ak, P, e, y0) && ak != S(0)) {
                S c0, p0, c1, p1, c2, p2;
                gen(n, y0, c0, p0);
                gen(n, y0 + ak * kphi(Q, g, ta), c1, p1);
                gen(n, y0 + ak * kphi(Q, g, tb), c2, p2);
                da -= p1 - p0;
                db -= p2 - p0;
            }
        }
        for (int q = tb_()->nd_l[p]; q < tb_()->nd_l[p + 1]; q++) {
            const int pos = l0 + tb_()->nl_pos[q];
            const S f0 = Fv(tb_()->ln_line[pos]), df = tb_()->ln_df[pos];
            const S d1 = (f0 + ta * df) - f0, d2 = (f0 + tb * df) - f0;
            const bool b0 = tb_()->nl_sgn[q] > 0;
            da += b0 ? d1 : -d1;
            db += b0 ? d2 : -d2;
        }
        const S S0 = rv(n);
        A = S0 + da;
        B = S0 + db;
    }
    __device__ const Tables<S> *tb_() const { return tb; }

    // One pass at a primary point t reached from tp: ΔC(t); the sign changes of the node balances
    // strictly inside (tp, t) in walking order (at most KR kept; `more` when some were left out);
    // the nearest primary breakpoint beyond t. With nodes = false only the breakpoint.
    static constexpr int KR = 4;
    struct Pass {
        S c = 0;
        int nr = 0;
        S rt[KR];
        bool more = false, hb = false;
        S nb = 0;
        // keep root x among the KR nearest distinct roots in walking order (more: one was dropped)
        __device__ void insert(S x, S sgn) {
            bool dup = false;
#pragma unroll
            for (int i = 0; i < KR; i++) dup = dup || (i < nr && rt[i] == x);
            if (dup) return;
            bool carry = true;
#pragma unroll
            for (int i = 0; i < KR; i++) {
                if (i < nr) {
                    if (sgn * (x - rt[i]) < 0) {
                        const S tmp = rt[i];
                        rt[i] = x;
                        x = tmp;
                    }
                } else if (i == nr && carry) {
                    rt[i] = x;
                    carry = false;
                }
            }
            if (carry) more = true;
            else nr++;
        }
    };
    __device__ void pass(S tp, S t, S sgn, bool nodes, Pass &P) const {
        if (nodes) nev++;
        S D = 0;
        auto cand = [&](S b) {
            if (isfinite(b) && sgn * (b - t) > 0 && (!P.hb || sgn * (P.nb - b) > 0)) {
                P.nb = b;
                P.hb = true;
            }
        };
        for (int e = e0 + wv; e < e1; e += nw) {
            const S a = slope(e);
            if (a == S(0)) continue;
            const int n = tb->en_node[e];
            const S y0 = ygv(n);
            if (nodes) {
                S c0, p0, c1, p1;
                gen(n, y0, c0, p0);
                gen(n, y0 + t * a, c1, p1);
                D += c1 - c0;
            }
            cand((S(0) - y0) / a);
            cand((S(1) - y0) / a);
        }
        if (comp)
            for (int q = q0 + wv; q < q1; q += nw) {
                S Q, g, ak, P_, e, y0;
                if (!kink(q, Q, g, ak, P_, e, y0)) continue;
                cand(-Q / g);
                if (ak != S(0)) {
                    if (nodes) {
                        const int n = tb->kk_node[q];
                        S c0, p0, c1, p1;
                        gen(n, y0, c0, p0);
                        gen(n, y0 + ak * kphi(Q, g, t), c1, p1);
                        D += c1 - c0;
                    }
                    const S sl = ak * g, off = ak * Q;
                    cand((S(0) - y0 - S(0)) / sl);
                    cand((S(1) - y0 - S(0)) / sl);
                    cand((S(0) - y0 - off) / sl);
                    cand((S(1) - y0 - off) / sl);
                }
            }
        for (int l = l0 + wv; l < l1; l += nw) {
            const int e = tb->ln_line[l];
            const S f0 = Fv(e), se = m->ls[e], df = tb->ln_df[l];
            if (nodes) {
                const S f1 = f0 + t * df;
                D += m->cv * m->dt * (fmax(fabs(f1) - se, S(0)) - fmax(fabs(f0) - se, S(0)));
            }
            cand((se - f0) / df);
            cand((-se - f0) / df);
        }
        if (nodes)
            for (int p = n0 + wv; p < n1; p += nw) {
                S A, B;
                node_S2(p, tp, t, A, B);
                const S S0 = rv(tb->nd_node[p]);
                D += m->dt * (phibal(B, m->c_shed, m->c_dump) - phibal(S0, m->c_shed, m->c_dump));
                if ((A < 0) != (B < 0) && A != B) {
                    const S x = tp + (t - tp) * A / (A - B);
                    if (sgn * (x - tp) > 0 && sgn * (t - x) > 0) P.insert(x, sgn);
                }
            }
        P.c = D;
        combine(P, sgn, nodes);
    }

    // ΔC(t)
    __device__ S eval(S t) const {
        nev++;
        S D = 0;
        for (int e = e0 + wv; e < e1; e += nw) {
            const S a = slope(e);
            if (a == S(0)) continue;
            const int n = tb->en_node[e];
            const S y0 = ygv(n);
            S c0, p0, c1, p1;
            gen(n, y0, c0, p0);
            gen(n, y0 + t * a, c1, p1);
            D += c1 - c0;
        }
        if (comp)
            for (int q = q0 + wv; q < q1; q += nw) {
                S Q, g, ak, P, e, y0;
                if (!kink(q, Q, g, ak, P, e, y0) || ak == S(0)) continue;
                const int n = tb->kk_node[q];
                S c0, p0, c1, p1;
                gen(n, y0, c0, p0);
                gen(n, y0 + ak * kphi(Q, g, t), c1, p1);
                D += c1 - c0;
            }
        for (int l = l0 + wv; l < l1; l += nw) {
            const int e = tb->ln_line[l];
            const S f0 = Fv(e), se = m->ls[e];
            const S f1 = f0 + t * tb->ln_df[l];
            D += m->cv * m->dt * (fmax(fabs(f1) - se, S(0)) - fmax(fabs(f0) - se, S(0)));
        }
        for (int p = n0 + wv; p < n1; p += nw) {
            const S S0 = rv(tb->nd_node[p]);
            const S S1 = node_S(p, t);
            D += m->dt * (phibal(S1, m->c_shed, m->c_dump) - phibal(S0, m->c_shed, m->c_dump));
        }
        return sum_waves(D);
    }

    // the nearest primary breakpoint beyond tp in walking direction sgn
    __device__ bool next_bp(S tp, S sgn, S &t) const {
        bool have = false;
        auto cand = [&](S b) {
            if (isfinite(b) && sgn * (b - tp) > 0 && (!have || sgn * (t - b) > 0)) {
                t = b;
                have = true;
            }
        };
        for (int e = e0 + wv; e < e1; e += nw) {
            const S a = slope(e);
            if (a == S(0)) continue;
            const S y0 = ygv(tb->en_node[e]);
            cand((S(0) - y0) / a);
            cand((S(1) - y0) / a);
        }
        if (comp)
            for (int q = q0; q < q1; q++) {
                S Q, g, ak, P, e, y0;
                if (!kink(q, Q, g, ak, P, e, y0)) continue;
                cand(-Q / g);
                if (ak != S(0)) {
                    const S sl = ak * g, off = ak * Q;
                    cand((S(0) - y0 - S(0)) / sl);
                    cand((S(1) - y0 - S(0)) / sl);
                    cand((S(0) - y0 - off) / sl);
                    cand((S(1) - y0 - off) / sl);
                }
            }
        for (int l = l0 + wv; l < l1; l += nw) {
            const int e = tb->ln_line[l];
            const S f0 = Fv(e), se = m->ls[e], df = tb->ln_df[l];
            cand((se - f0) / df);
            cand((-se - f0) / df);
        }
        return have;
    }

    // the next KR distinct sign changes of the node balances strictly between x and t (walking
    // order) on the segment (tp, t), where the balances are linear; more: others remain
    __device__ void roots_beyond(S tp, S t, S x, S sgn, Pass &P) const {
        for (int p = n0 + wv; p < n1; p += nw) {
            S A, B;
            node_S2(p, tp, t, A, B);
            if ((A < 0) != (B < 0) && A != B) {
                const S rt = tp + (t - tp) * A / (A - B);
                if (sgn * (rt - tp) > 0 && sgn * (t - rt) > 0 && sgn * (rt - x) > 0) P.insert(rt, sgn);
            }
        }
        combine(P, sgn, true);
    }

    // log Jacobian Σ (|S| − 1) log|P(t)/P| of the compensated kinks (FP64)
    __device__ double logjac(S t, bool &any) const {
        double lj = 0.0;
        any = false;
        for (int q = q0 + wv; q < q1; q += nw) {
            S Q, g, ak, P, e, y0;
            if (!kink(q, Q, g, ak, P, e, y0)) continue;
            any = true;
            const double nm = kink_nm(*m, tb->kk_node[q], kmin);
            lj += nm * log(fabs((fmax((double)Q + (double)g * (double)t, 0.0) + (double)e) / (double)P));
        }
        if (nw > 1) {
            any = sum_waves(any ? S(1) : S(0)) > S(0);
            lj = (double)sum_waves((S)lj);
        }
        return lj;
    }

    // the sum over the waves of a per-lane partial value (in wave order)
    __device__ S sum_waves(S x) const {
        if (nw == 1) return x;
        const int lane = threadIdx.x & 63;
        red[(size_t)wv * RED * 64 + lane] = x;
        __syncthreads();
        S s = 0;
        for (int w = 0; w < nw; w++) s += red[(size_t)w * RED * 64 + lane];
        __syncthreads();
        return s;
    }

    // the waves' partial passes combined: the costs summed in wave order, the nearest breakpoint,
    // the KR nearest distinct roots (more: a wave had more, or the merge dropped one)
    __device__ void combine(Pass &P, S sgn, bool roots) const {
        if (nw == 1) return;
        const int lane = threadIdx.x & 63;
        S *o = red + (size_t)wv * RED * 64 + lane;
        o[0] = P.c;
        o[64] = P.hb ? S(1) : S(0);
        o[128] = P.nb;
        o[192] = (S)P.nr;
        o[256] = P.more ? S(1) : S(0);
#pragma unroll
        for (int k = 0; k < KR; k++) o[(5 + k) * 64] = P.rt[k];
        __syncthreads();
        Pass Q;
        for (int w = 0; w < nw; w++) {
            const S *x = red + (size_t)w * RED * 64 + lane;
            Q.c += x[0];
            if (x[64] != S(0) && (!Q.hb || sgn * (Q.nb - x[128]) > 0)) {
                Q.nb = x[128];
                Q.hb = true;
            }
            if (roots) {
                const int nr = (int)x[192];
                for (int k = 0; k < nr; k++) Q.insert(x[(5 + k) * 64], sgn);
                if (x[256] != S(0)) Q.more = true;
            }
        }
        __syncthreads();
        P = Q;
    }
};

// walk! of a table-driven move: 1 proper, 0 improper, -1 point cap, 2 stopped by the visitor.
// One pass per primary point (cost, roots of the segment, next breakpoint); roots beyond the
// KR kept by a pass are found by scans.
template <typename S, class V>
__device__ int walk_sparse(const FSparse<S> &f, S beta, S maxnats, S sgn, V &vis) {
    S cmin = 0, tp = 0, delta = 0, clast = 0;
    int nx = 0, np = 0;
    typename FSparse<S>::Pass P0;
    f.pass(S(0), S(0), sgn, false, P0);
    bool hb = P0.hb;
    S nb = P0.nb;
    while (true) {
        S t = 0;
        if (hb) {
            t = nb;
        } else {
            const S hp = S(1e-6) * fmax(S(1), fabs(tp));
            const S sl = (beta * f.eval(tp + sgn * hp) - clast) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, S(1e3) * fmax(fabs(tp), S(1e-3)));
            else delta = delta == 0 ? fmax(fabs(tp), S(1e-6)) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;
        }
        typename FSparse<S>::Pass P;
        f.pass(tp, t, sgn, true, P);
        const S c = beta * P.c;
        hb = P.hb;
        nb = P.nb;
        int w;
        // the roots of the segment in walking order: the ones the pass kept, then KR per scan
        int nr = P.nr;
        bool more = P.more;
        S rt[FSparse<S>::KR];
#pragma unroll
        for (int k = 0; k < FSparse<S>::KR; k++) rt[k] = P.rt[k];
        while (true) {
#pragma nounroll
            for (int i = 0; i < nr; i++) {
                S xr = rt[0];
#pragma unroll
                for (int k = 1; k < FSparse<S>::KR; k++)
                    if (k == i) xr = rt[k];
                if ((w = visit_point(vis, xr, beta * f.eval(xr), cmin, np))) return w;
            }
            if (!more) break;
            typename FSparse<S>::Pass Q;
            f.roots_beyond(tp, t, rt[FSparse<S>::KR - 1], sgn, Q);
            nr = Q.nr;
            more = Q.more;
#pragma unroll
            for (int k = 0; k < FSparse<S>::KR; k++) rt[k] = Q.rt[k];
        }
        if ((w = visit_point(vis, t, c, cmin, np))) return w;
        clast = c;
        if (c - cmin > maxnats) return 1;
        tp = t;
    }
}

// the two visitors of the sampler in one type (mode 0: segment masses of one side; mode 1: the
// segment of the target mass), so that the three walks share one call site
template <typename S>
struct Vis2 {
    int mode = 0;
    MassAcc<S> ma;
    SegSel<S> ss;
    __device__ bool operator()(S t, S c) { return mode == 0 ? ma(t, c) : ss(t, c); }
};

template <typename S>
__device__ S sample_sparse(const FSparse<S> &f, S beta, S maxnats, S useg, S uoff, int &status) {
    Vis2<S> vis;
    S Mx[2] = {S(-INFINITY), S(-INFINITY)}, Ms[2] = {S(0), S(0)};
    S sgn = 1;
#pragma nounroll
    for (int k = 0; k < 3; k++) {
        if (k == 2) {
            const S MX = fmax(Mx[0], Mx[1]);
            const S Mr = Ms[0] * exp(Mx[0] - MX), Ml = Ms[1] * exp(Mx[1] - MX);
            const S u = useg * (Ml + Mr);
            vis.mode = 1;
            vis.ss = SegSel<S>();
            vis.ss.MX = MX;
            vis.ss.left = u <= Ml;
            vis.ss.target = vis.ss.left ? Ml - u : u - Ml;
            sgn = vis.ss.left ? S(-1) : S(1);
        } else {
            vis.mode = 0;
            vis.ma = MassAcc<S>();
            sgn = k == 0 ? S(1) : S(-1);
        }
        const int w = walk_sparse(f, beta, maxnats, sgn, vis);
        if (k < 2) {
            if (w != 1) {
                status = w == 0 ? 1 : 2;
                return 0;
            }
            Mx[k] = vis.ma.mx;
            Ms[k] = vis.ma.s;
        }
    }
    status = 0;
    const SegSel<S> &sel = vis.ss;
    const S L = sel.tb - sel.ta;
    const S g = (sel.cb - sel.ca) / L;
    const S t = (g >= 0 || fabs(g * L) < S(1e-10)) ? sel.ta + trexp_offset(g, L, uoff) : sel.tb - trexp_offset(-g, L, uoff);
    if (!isfinite(t)) {
        status = 2;
        return 0;
    }
    return t;
}

enum { SP_CELL = 0, SP_XLOC = 1, SP_XLOCC = 2, SP_XPAIR = 3 };

template <typename S, int KIND, bool UNI>
__device__ __forceinline__ void run_sparse(const DevModel<S> &m, SweepState<S> &st, const RandArgs &a, const Tables<S> &tb,
                                           int tau, int j, S beta, S maxnats, S kmin, MoveOut &out, int wv = 0,
                                           int nw = 1, S *red = nullptr) {
    {
        const int cnt = uniform<UNI>(a.counts[(size_t)a.w * m.T + tau]);
        const size_t TR = (size_t)m.T * m.R, o = (size_t)tau * m.R + j;
        S *yg = st.v + plane(m.N, m.T, m.R);
        for (int k = 0; k < cnt; k++) {
            uint32_t s[4];
            selection(a, k, tau, s);
            uniform4<UNI>(s);
            int tid;
            if (KIND == SP_CELL) {
                const double u = u01(s[0], s[1]);
                int lv = tb.nlev - 1;
                for (int q = tb.nlev - 1; q >= 0; q--)
                    if (tb.qcum[q] >= u) lv = q;
                const int nc = tb.lev_off[lv + 1] - tb.lev_off[lv];
                tid = tb.cl0 + tb.lev_off[lv] + (int)pick(s[2], (uint32_t)nc);
            } else if (KIND == SP_XPAIR) {
                const int n = (int)pick(s[0], (uint32_t)m.N), mm = (int)pick(s[1], (uint32_t)m.N);
                if (mm == n) continue;
                tid = tb.xp0 + n * m.N + mm;
                if (tb.en_off[tid + 1] == tb.en_off[tid]) continue;   // another component
            } else {
                const int n = (int)pick(s[0], (uint32_t)m.N);
                const int np = tb.xp_off[n + 1] - tb.xp_off[n];
                if (np == 0) continue;
                tid = tb.xp_off[n] + (int)pick(s[1], (uint32_t)np);
            }
            FSparse<S> f;
            f.m = &m;
            f.tb = &tb;
            f.yg = yg;
            f.r = st.r;
            f.F = st.F;
            f.o = o;
            f.TR = TR;
            f.e0 = tb.en_off[tid];
            f.e1 = tb.en_off[tid + 1];
            f.q0 = tb.kk_off[tid];
            f.q1 = tb.kk_off[tid + 1];
            f.l0 = tb.ln_off[tid];
            f.l1 = tb.ln_off[tid + 1];
            f.n0 = tb.nd_off[tid];
            f.n1 = tb.nd_off[tid + 1];
            f.comp = KIND == SP_XLOCC;
            f.kmin = kmin;
            f.wv = wv;
            f.nw = nw;
            f.red = red;
            if (KIND != SP_CELL) {   // the pair's units need |κ| ≥ κmin
                bool ok = true;
                for (int e = f.e0; e < f.e1; e++) ok = ok && fabs(f.kap(tb.en_node[e])) >= kmin;
                if (!ok) continue;
            }
            double u1, u2;
            line_uniforms(a, j, k, tau, u1, u2);
            int status = 0;
            S t = sample_sparse(f, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
            if (KIND == SP_XLOCC && t != 0) {
                bool any;
                const double lj = f.logjac(t, any);
                if (any) {
                    uint32_t w[4];
                    philox4x32(a.k0, a.k1, (uint32_t)j, (uint32_t)tau, 0x60000000u | ((uint32_t)a.w << 20) | (uint32_t)k,
                               a.ctr, w);
                    if (!(u01(w[0], w[1]) < exp(fmin(0.0, lj)))) {
                        t = 0;
                        status = 2;
                    }
                }
            }
            if (wv == 0) out.add(status, t != 0, f.nev);
            if (t != 0) {
                // node by node: the balance and the node's entry read only that node's values (and
                // the flows, written after all nodes; with nw waves, each writes every nw-th node)
                for (int p = f.n0 + wv; p < f.n1; p += nw) {
                    const int n = tb.nd_node[p];
                    const S S1 = f.node_S(p, t);
                    S ynew = 0;
                    bool wy = false;
                    const int ie = tb.nd_ent[p];
                    if (ie >= 0) {
                        const S av = f.slope(f.e0 + ie);
                        if (av != S(0)) {
                            ynew = f.ygv(n) + t * av;
                            wy = true;
                        }
                    }
                    const int ik = tb.nd_kk[p];
                    if (f.comp && ik >= 0) {
                        S Q, gk, ak, P, e, y0;
                        if (f.kink(f.q0 + ik, Q, gk, ak, P, e, y0) && ak != S(0)) {
                            ynew = y0 + ak * FSparse<S>::kphi(Q, gk, t);
                            wy = true;
                        }
                    }
                    st.r[(size_t)n * TR + o] = S1;
                    if (wy) yg[(size_t)n * TR + o] = ynew;
                }
            }
            if (nw > 1) __syncthreads();   // every wave's balances read the flows before the update
            if (t != 0) {
                for (int l = f.l0 + wv; l < f.l1; l += nw) {
                    S &F1 = st.F[(size_t)tb.ln_line[l] * TR + o];
                    F1 = F1 + t * tb.ln_df[l];
                }
                for (int q = tb.th_off[tid] + wv; q < tb.th_off[tid + 1]; q += nw) {
                    S &th = st.v[(size_t)tb.th_node[q] * TR + o];
                    th = th + t * tb.th_val[q];
                }
            }
            if (nw > 1) __syncthreads();   // the next move reads the new state
        }
    }
}

template <typename S, int KIND, bool UNI>
__global__ void __launch_bounds__(256) k_sparse(DevModel<S> m, SweepState<S> st, RandArgs a, Tables<S> tb, S beta,
                                                S maxnats, S kmin) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    MoveOut out;
    int j;
    if (slot_of(a, g, m.R, j))
        run_sparse<S, KIND, UNI>(m, st, a, tb, step_of<UNI>(a, g, a.rp), j, beta, maxnats, kmin, out);
    add_counts(a.stats, out);
}

// The table-driven families of a step (xpair, xlocc, xloc, cell: the last four of the family
// order) in one thread (step τ, replica j), in that order. Steps of one parity class share no
// values, so states and statistics equal those of one launch per family; a launch then lasts as
// long as its slowest step's whole list instead of the sum over the families of each family's
// slowest step. (unit, potential and spair keep their own launches: their register footprints,
// up to 256 VGPRs for the potential moves of degree > 8, would cut the merged kernel's occupancy.)
constexpr int NSPARSE = 4;
struct SparseAll {
    unsigned long long *stats[NSPARSE];   // xpair, xlocc, xloc, cell
    uint32_t mask;                        // families with a positive weight
};

// The steps of a parity class, heaviest first by Σ_f wk[f] · (moves of family f at the step): a
// launch hands its first waves the longest lists, so the waves that do not fit in the first round
// are the shortest. Which wave runs a step does not change any result. One block per class: a
// counting sort over SORT_BINS keys.
constexpr int SORT_BINS = 2048;
struct StepWeights {
    int wk[ESM_NWEIGHT];
};
__global__ void __launch_bounds__(1024) k_step_order(const int *counts, int T, int tau0, int tstride, int ntau,
                                                     StepWeights sw, int *perm) {
    __shared__ int hist[SORT_BINS];
    for (int b = threadIdx.x; b < SORT_BINS; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    auto key = [&](int tau) {
        int w = 0;
#pragma unroll
        for (int q = 0; q < ESM_NWEIGHT; q++) w += sw.wk[q] * counts[(size_t)q * T + tau];
        return SORT_BINS - 1 - min(w, SORT_BINS - 1);   // descending weight
    };
    for (int i = threadIdx.x; i < ntau; i += blockDim.x) atomicAdd(&hist[key(tau0 + i * tstride)], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = 0; b < SORT_BINS; b++) {
            const int c = hist[b];
            hist[b] = acc;
            acc += c;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ntau; i += blockDim.x) {
        const int tau = tau0 + i * tstride;
        perm[atomicAdd(&hist[key(tau)], 1)] = tau;
    }
}

// nw = 1: a thread per (step, replica), four steps per workgroup. nw > 1 (R a multiple of 64): a
// workgroup of nw waves per (step, 64 replicas), lane = replica, the waves splitting each move's
// template entries (FSparse's wave split); every wave runs the same walk on the combined values.
template <typename S, bool UNI, bool SPLIT>
__global__ void __launch_bounds__(SPLIT ? 1024 : 256, SPLIT ? 1 : min_waves<S>(4))
    k_sparse_all(DevModel<S> m, SweepState<S> st, RandArgs a, SparseAll sa, Tables<S> tb, S beta, S maxnats, S kmin,
                 int nw_) {
    extern __shared__ unsigned char sp_lds[];
    const int nw = SPLIT ? nw_ : 1;   // a compile-time 1 keeps the combining code out of the one-wave form
    size_t g;
    int wv = 0;
    if (!SPLIT) {
        g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    } else {
        wv = threadIdx.x >> 6;
        g = (size_t)blockIdx.x * 64 + (threadIdx.x & 63);
    }
    S *red = nw > 1 ? (S *)sp_lds : nullptr;
    MoveOut o[NSPARSE];
    int j;
    if (slot_of(a, g, m.R, j)) {   // the wave-split form runs only with R a multiple of 64 (rp = R, no idle lanes)
        const int tau = step_of<UNI>(a, g, a.rp);
        RandArgs b = a;
        if (sa.mask & 1u) {
            b.w = W_XPAIR;
            run_sparse<S, SP_XPAIR, UNI>(m, st, b, tb, tau, j, beta, maxnats, kmin, o[0], wv, nw, red);
        }
        if (sa.mask & 2u) {
            b.w = W_XLOCC;
            run_sparse<S, SP_XLOCC, UNI>(m, st, b, tb, tau, j, beta, maxnats, kmin, o[1], wv, nw, red);
        }
        if (sa.mask & 4u) {
            b.w = W_XLOC;
            run_sparse<S, SP_XLOC, UNI>(m, st, b, tb, tau, j, beta, maxnats, kmin, o[2], wv, nw, red);
        }
        if (sa.mask & 8u) {
            b.w = W_CELL;
            run_sparse<S, SP_CELL, UNI>(m, st, b, tb, tau, j, beta, maxnats, kmin, o[3], wv, nw, red);
        }
    }
#pragma unroll
    for (int q = 0; q < NSPARSE; q++)
        if (sa.mask >> q & 1u) add_counts(sa.stats[q], o[q]);
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// waves per (step, 64 replicas) of the merged table-driven launch: set by esm_sweep_launch, else
// one. Measured (FP32, storage sweeps off, ms per sweep of the merged launch for 1 / 2 / 4 / 8 / 16
// waves): 40 nodes × 8760 h × 64 replicas 57 / 59 / 55 / 75 / 152; 256 nodes × 2920 h × 64
// replicas 753 / 1041 / 853 / 780 / 797. The launch is VALU-bound over the whole GPU at 256
// nodes (every walk pass rescans all ~700 template entries), so more waves add barriers, not
// throughput.
static int sparse_waves(const esm_handle *h) { return h->sweep_waves > 0 ? h->sweep_waves : 1; }

template <typename S>
int launch_random(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats) {
    const int N = h->N, T = h->T, R = h->R;
    // with sxpair (a positive weight and some unit taking part) its weight follows the plain
    // families', and the draws are scaled by Σ all / Σ plain so that every plain family keeps its
    // rate (`local_sweep!`, src/local.jl:1934-1941)
    const bool usx = h->sweep_pxpair > 0 && h->nsx > 0;
    double wall[NW_ALL], tot0 = 0.0, tot = 0.0, acc = 0.0;
    for (int q = 0; q < ESM_NWEIGHT; q++) wall[q] = h->sweep_w[q];
    wall[ESM_NWEIGHT] = usx ? h->sweep_pxpair : 0.0;
    for (int q = 0; q < ESM_NWEIGHT; q++) tot0 += wall[q];
    for (int q = 0; q < NW_ALL; q++) tot += wall[q];
    const int nextra = (int)nearbyint(h->sweep_extra * (tot0 > 0 ? tot / tot0 : 1.0) * N);
    if (nextra <= 0 || !(tot > 0)) return 0;
    CumW cw;
    for (int q = 0; q < NW_ALL; q++) {
        acc += wall[q];
        cw.c[q] = acc / tot;
    }
    hipLaunchKernelGGL(k_movecounts, dim3(nblk((size_t)T, 256)), dim3(256), 0, h->stream, T, nextra, cw, k0, k1, ctr,
                       h->d_counts);
    ESM_HIP_CHECK(hipGetLastError());
    const DevModel<S> m = dev_model<S>(h);
    SweepState<S> st{(S *)h->d_v, (S *)h->d_r, (S *)h->d_F, (const S *)h->d_x};
    const S beta = (S)h->m.beta_v, mxn = (S)maxnats, kmin = S(1);
    const MoveTablesDev &d = h->dmt;
    Tables<S> tb{d.th_off, d.th_node, d.en_off, d.en_node, d.en_div, d.kk_off, d.kk_node, d.ln_off, d.ln_line,
                 d.nd_off, d.nd_node, d.nd_ent, d.nd_kk, d.nd_l, d.nl_pos, d.nl_sgn, d.xp_off, d.lev_off,
                 (const S *)d.th_val, (const S *)d.en_c, (const S *)d.kk_g, (const S *)d.ln_df, d.qcum,
                 d.xp0, d.cl0, d.nlev};
    // the family order (pair and cellpair are void here): the atom-site families, then the
    // table-driven ones
    const int order[7] = {W_UNIT, W_POT, W_SPAIR, W_XPAIR, W_XLOCC, W_XLOC, W_CELL};
    // padded: every step gets whole waves (R rounded up to a multiple of 64, the extra lanes
    // idle), so that a wave holds one step and its selection is wave-uniform
    const int rp = h->rand_pad && R % 64 ? (R + 63) / 64 * 64 : R;
    const bool uni = rp % 64 == 0;
    const int rb = h->rand_block;   // every kernel of RLAUNCH is bounded to 256 threads and counts per wave
    int dmax = 0;
    for (int n = 0; n < N; n++) dmax = std::max(dmax, h->adj_ptr[n + 1] - h->adj_ptr[n]);
    SparseAll sa{};
    {
        const int sw[NSPARSE] = {W_XPAIR, W_XLOCC, W_XLOC, W_CELL};
        const int sf[NSPARSE] = {ESM_FAM_XPAIR, ESM_FAM_XLOCC, ESM_FAM_XLOC, ESM_FAM_CELL};
        for (int q = 0; q < NSPARSE; q++) {
            sa.stats[q] = h->d_stats + (size_t)sf[q] * ESM_NSTAT * STAT_SLOTS;
            if (h->sweep_w[sw[q]] > 0) sa.mask |= 1u << q;
        }
    }
#define RLAUNCH(KU, KD, ...)                                                                           \
    do {                                                                                               \
        if (uni) hipLaunchKernelGGL(KU, dim3(nb), dim3(rb), 0, h->stream, __VA_ARGS__);                 \
        else hipLaunchKernelGGL(KD, dim3(nb), dim3(rb), 0, h->stream, __VA_ARGS__);                     \
    } while (0)
    // the steps of a class heaviest first (when a wave holds one step's replicas: several steps per
    // wave would put the heaviest together in the same waves, measured slower at 16 replicas) for
    // the launches that need more than one round of waves: spair (194 VGPRs) and the merged
    // table-driven families (an xpair evaluation reads every line and node, cells and compensated
    // pairs about half as much)
    int *perm = h->d_counts + (size_t)NW_ALL * T;   // T ints after the counts
    auto order_steps = [&](const RandArgs &a, StepWeights sw) -> const int * {
        if (!uni) return nullptr;
        hipLaunchKernelGGL(k_step_order, dim3(1), dim3(1024), 0, h->stream, h->d_counts, T, a.tau0, a.tstride, a.ntau,
                           sw, perm);
        return perm;
    };
    StepWeights w_spair{}, w_sparse{};
    w_spair.wk[W_SPAIR] = 1;
    w_sparse.wk[W_XPAIR] = 4;
    w_sparse.wk[W_XLOCC] = 2;
    w_sparse.wk[W_XLOC] = 1;
    w_sparse.wk[W_CELL] = 2;
    for (size_t c = 0; c < h->tclass.size(); c++) {
        RandArgs a{k0, k1, ctr, h->tclass[c].first, 2, h->tclass[c].second, 0, h->d_counts, nullptr, nullptr, rp};
        const unsigned nb = nblk((size_t)a.ntau * rp, (unsigned)rb);
        for (int w : order) {
            if (!(h->sweep_w[w] > 0)) continue;
            if (h->sweep_merged && (w == W_XPAIR || w == W_XLOCC || w == W_XLOC || w == W_CELL)) continue;
            a.w = w;
            auto stats = [&](int fam) { return h->d_stats + (size_t)fam * ESM_NSTAT * STAT_SLOTS; };
            switch (w) {
                case W_UNIT: {
                    a.stats = stats(ESM_FAM_UNIT);
                    ProfScope p(h, ESM_KERN_UNIT);
                    RLAUNCH((k_rand_unit<S, true>), (k_rand_unit<S, false>), m, st, a, beta, mxn);
                    break;
                }
                case W_POT: {
                    a.stats = stats(ESM_FAM_POT);
                    ProfScope p(h, ESM_KERN_POT);
                    RLAUNCH((k_rand_pot<S, 0, 4, true>), (k_rand_pot<S, 0, 4, false>), m, st, a, beta, mxn);
                    if (dmax > 4)
                        RLAUNCH((k_rand_pot<S, 4, 8, true>), (k_rand_pot<S, 4, 8, false>), m, st, a, beta, mxn);
                    if (dmax > 8)
                        RLAUNCH((k_rand_pot<S, 8, 16, true>), (k_rand_pot<S, 8, 16, false>), m, st, a, beta, mxn);
                    break;
                }
                case W_SPAIR: {
                    a.stats = stats(ESM_FAM_SPAIR);
                    ProfScope p(h, ESM_KERN_SPAIR);
                    RandArgs b = a;
                    b.perm = order_steps(a, w_spair);
                    RLAUNCH((k_rand_spair<S, true>), (k_rand_spair<S, false>), m, st, b, beta, mxn, kmin);
                    break;
                }
                case W_CELL: {
                    a.stats = stats(ESM_FAM_CELL);
                    ProfScope p(h, ESM_KERN_CELL);
                    RLAUNCH((k_sparse<S, SP_CELL, true>), (k_sparse<S, SP_CELL, false>), m, st, a, tb, beta, mxn, kmin);
                    break;
                }
                case W_XLOC: {
                    a.stats = stats(ESM_FAM_XLOC);
                    ProfScope p(h, ESM_KERN_XLOC);
                    RLAUNCH((k_sparse<S, SP_XLOC, true>), (k_sparse<S, SP_XLOC, false>), m, st, a, tb, beta, mxn, kmin);
                    break;
                }
                case W_XLOCC: {
                    a.stats = stats(ESM_FAM_XLOCC);
                    ProfScope p(h, ESM_KERN_XLOCC);
                    RLAUNCH((k_sparse<S, SP_XLOCC, true>), (k_sparse<S, SP_XLOCC, false>), m, st, a, tb, beta, mxn, kmin);
                    break;
                }
                case W_XPAIR: {
                    a.stats = stats(ESM_FAM_XPAIR);
                    ProfScope p(h, ESM_KERN_XPAIR);
                    RLAUNCH((k_sparse<S, SP_XPAIR, true>), (k_sparse<S, SP_XPAIR, false>), m, st, a, tb, beta, mxn, kmin);
                    break;
                }
                default: break;
            }
            ESM_HIP_CHECK(hipGetLastError());
        }
        if (h->sweep_merged && sa.mask) {
            ProfScope p(h, ESM_KERN_RANDOM);
            RandArgs b = a;
            b.perm = order_steps(a, w_sparse);
            const int nw = R % 64 == 0 ? sparse_waves(h) : 1;
            if (nw == 1) {
                RLAUNCH((k_sparse_all<S, true, false>), (k_sparse_all<S, false, false>), m, st, b, sa, tb, beta, mxn, kmin,
                        1);
            } else {
                const size_t lds = (size_t)nw * FSparse<S>::RED * 64 * sizeof(S);
                hipLaunchKernelGGL((k_sparse_all<S, true, true>), dim3((unsigned)((size_t)a.ntau * R / 64)), dim3(64 * nw),
                                   lds, h->stream, m, st, b, sa, tb, beta, mxn, kmin, nw);
            }
            ESM_HIP_CHECK(hipGetLastError());
        }
        if (usx) {   // the storage pairs last
            RandArgs b = a;
            b.w = W_SXPAIR;
            b.stats = h->d_stats + (size_t)ESM_FAM_SXPAIR * ESM_NSTAT * STAT_SLOTS;
            const SxArgs<S> sx = sx_args<S>(h);
            ProfScope p(h, ESM_KERN_SXPAIR);
            int ks = 1;
            const int wx = sx_lane_form(h, 32, ks);
            if (wx == 0) {
                set_error("storage pairs: a pair template is larger than the lane form holds");
                return ESM_EINVAL;
            }
            // four replicas per wave; the blocks of one step on one XCD: the grid padded to whole
            // rounds of 8 XCDs × the step's blocks (the extra blocks find no wave of theirs and exit)
            const size_t ngrp = ((size_t)R + 3) / 4, nwaves = (size_t)a.ntau * ngrp;
            size_t nblocks = (nwaves + SXP_WAVES - 1) / SXP_WAVES;
            if (ngrp % SXP_WAVES == 0) {
                const size_t round = 8 * (ngrp / SXP_WAVES);
                nblocks = (nblocks + round - 1) / round * round;
            }
            const dim3 grid((unsigned)nblocks), blk(64 * SXP_WAVES);
            if (wx == 32) hipLaunchKernelGGL((k_rand_sxpair_mix<S, 0>), grid, blk, 0, h->stream, m, st, b, sx, beta, mxn, kmin);
            else if (ks == 1) hipLaunchKernelGGL((k_rand_sxpair_mix<S, 1>), grid, blk, 0, h->stream, m, st, b, sx, beta, mxn, kmin);
            else hipLaunchKernelGGL((k_rand_sxpair_mix<S, 2>), grid, blk, 0, h->stream, m, st, b, sx, beta, mxn, kmin);
            ESM_HIP_CHECK(hipGetLastError());
        }
    }
    return 0;
}

#undef RLAUNCH

template int launch_random<float>(esm_handle *, uint32_t, uint32_t, uint32_t, double);
template int launch_random<double>(esm_handle *, uint32_t, uint32_t, uint32_t, double);

}  // namespace esm
