// k_storage.hip — the storage multiscale sweep (`storage_ms_sweep!`, src/flowmoves.jl:255-324;
// moves `prep_storage_ms!`, :187-244), which `operation_sweep!` runs after the local sweep with
// probability 1/2 (src/local.jl:2414-2418).
//
// For every dyadic block length L < T/2 (2, 4, …, 128; random order, random offset o, blocks
// o + iL without wrap, the two parity classes of blocks in random order), 8 moves per block: a
// plateau on the block for one storage unit or, with probability 1/2, for all units of one
// carrier in one country (the same change t of the state of charge at the instants b0+1 … b0+L);
// the gas unit of each node follows pro rata at the block's ends (a requirement change of +κs at
// step b0 and −κs at step b0+L, κs = k h/Δt, shifted by `soc_qadj`), with the Jacobian test. The
// flow basis' shapes do not exist in this basis, so every move is a plateau.
//
// MI355X layout: one wave per (block, group of replicas). A segment of W lanes (W the smallest
// power of two ≥ L + 1, at most 64) owns one replica: lane `sub` of the segment owns the block
// steps q ≡ sub (mod W) and evaluates their terms (kinked gas at the block ends, the storage terms
// of the moved units, the node balances) from values staged in LDS. ΔC(t) is the butterfly sum of
// the lanes' partial sums over the segment (every lane of a segment gets the same bits, so the
// segment walks as one); breakpoints and the sign changes of the secondaries are segment minima.
// The walk follows `walk!` (src/local.jl:1294-1443) with storage secondaries: the roots of the
// storage secondaries on (tp, t) split the segment, the node balances' roots are found on the
// pieces; equal roots are visited once.
#include "storage_walk.h"

namespace esm {

constexpr int SMS_SP = 3;          // block steps per lane (L + 1 ≤ 3·64)

struct SmsArgs {
    uint32_t k0, k1, ctr;
    int L, base, nblk, nrg;        // block b0 = base + 2L·blk, blk < nblk; nrg replica groups
    int nmoves, nus;               // nus: units of the largest group (LDS plane stride)
    double p_agg;
    const int *units;              // [2N] unit u: storage carrier u / N, node u % N
    int nunits;
    const int *grp_off, *grp_node, *grp_si;
    int ngroups;
    unsigned long long *stats;     // [ESM_NSTAT][STAT_SLOTS]
    // graph-stable form (sched != null): the launch is slot `slot`, class `ci` (0: the class drawn
    // first) of the schedule that k_sms_schedule wrote on the device; L, base and nblk follow from
    // it, and a launch whose segment width is not the slot's does nothing
    const int *sched;
    int slot, ci;
};


template <typename S>
struct SmsKink {
    S Q, g, ak, y0, P, e, valid;
    S gc0, gp0;                    // gas cost and output at t = 0
};

// per-(step, unit) LDS planes of a segment, each [L + 1][nus]: the staged values and the terms
// at t = 0, then the walk's cached secondaries (d at the interval ends tp and t, balances at the
// piece ends a0 and x); every plane entry is written and read by the lane that owns the step only
enum SmsPlane { P_YE, P_YE1, P_S0, P_C0, P_J0, P_PH0, P_DA, P_DB, P_BA, P_BX, SMS_NPLANE };

// the moved unit u of a segment: its storage parameters, clamped capacity, κs, node, and the
// node's gas capacity and cost (the kinked gas terms at the block's ends)
template <typename S>
struct SmsUnit {
    StoPar<S> par;
    S kst, ks;
    int node;
    S gk, gmc;
};

// one replica's move, seen by one lane of its segment; the values live in the segment's LDS
template <typename S, int W>
struct FStor {
    const DevModel<S> *m;
    int L, sub, nu, nus;                // nus: the plane stride (the largest group)
    S *pl;                             // SMS_NPLANE planes [q][u] of stride (L + 1)·nus
    const SmsKink<S> *kk_l;            // [2][u]: the kinks at steps 0 and L
    const SmsUnit<S> *un_l;            // [u]
    S kmin;
    mutable int nev = 0;
    enum { REC_NONE, REC_B, REC_X, REC_A, REC_T };
    static constexpr bool CACHED = true;               // root searches read the recorded secondaries
    static constexpr int SP = W < 64 ? 1 : SMS_SP;     // block steps per lane (L + 1 ≤ W below 64)

    __device__ S *plane(int k) const { return pl + (size_t)k * (L + 1) * nus; }
    __device__ S gen_c(int n, S y, S &p) const {
        S c;
        const S k = m->kc[ESM_KGAS * m->N + n];
        gen_cost(m->mc[ESM_KGAS * m->N + n], k, k + m->kappa0, m->dt, m->cv, y, c, p);
        return c;
    }
    __device__ S gen_cu(const SmsUnit<S> &U, S y, S &p) const {   // gen_c from the staged unit
        S c;
        gen_cost(U.gmc, U.gk, U.gk + m->kappa0, m->dt, m->cv, y, c, p);
        return c;
    }
    __device__ static S kphi(S Q, S g, S t) { return fmax(Q + g * t, S(0)) - fmax(Q, S(0)); }
    // the storage secondaries of a unit from its d: d, d − η_in (1 without charging), −d η_out − 1
    __device__ static S sec(int k, S d, const StoPar<S> &p) {
        return k == 0 ? d : (k == 1 ? (p.ein > 0 ? d - p.ein : S(1)) : -d * p.eout - S(1));
    }

    // the terms of unit u at block step q at t, in this order: its node's kinked gas (block
    // ends), its storage term, its node's balance; its d (storage secondaries) and balance
    __device__ S unit_step(int q, int u, S t, S &d1, S &S1) const {
        const SmsUnit<S> U = un_l[u];
        const int i = q * nus + u;
        S D = 0, dr = 0;
        const int kidx = q == 0 ? 0 : (q == L ? 1 : -1);
        if (kidx >= 0) {
            const SmsKink<S> kk = kk_l[kidx * nus + u];
            if (kk.valid != S(0) && kk.ak != S(0)) {
                S p1;
                const S c1 = gen_cu(U, kk.y0 + kk.ak * kphi(kk.Q, kk.g, t), p1);
                D += c1 - kk.gc0;
                dr -= p1 - kk.gp0;
            }
        }
        const S dye = q >= 1 ? S(1) : S(0), dye1 = q <= L - 1 ? S(1) : S(0);
        const S w = U.kst + m->kappa0;
        S j1;
        const S c1 = soc_cost_p(U.par, m->dt, m->cv, U.kst, w, plane(P_YE)[i] + t * dye, plane(P_YE1)[i] + t * dye1,
                                d1, j1);
        D += c1 - plane(P_C0)[i];
        dr += j1 - plane(P_J0)[i];
        S1 = plane(P_S0)[i] + dr;
        D += m->dt * (phibal(S1, m->c_shed, m->c_dump) - plane(P_PH0)[i]);
        return D;
    }

    // ΔC(t): the lanes' partial sums (own steps in increasing order), summed over the segment.
    // Recording: REC_B d and the balance at t (an interval's end), REC_X the balance (a storage
    // root); uncounted (no evaluation of the reference's walk): REC_A d and the balance at the
    // walk's start, REC_T the balance at an interval's end after a split
    __device__ S eval(S t, int rec) const {
        if (rec <= REC_X) nev++;
        S D = 0;
#pragma nounroll
        for (int i = 0; i < SP; i++) {
            const int q = sub + W * i;
            if (q > L) continue;
            S Dq = 0;
            for (int u = 0; u < nu; u++) {
                S d1, S1;
                Dq += unit_step(q, u, t, d1, S1);
                const int k = q * nus + u;
                if (rec == REC_B) plane(P_DB)[k] = d1;
                if (rec == REC_A) plane(P_DA)[k] = d1;
                if (rec != REC_NONE) plane(rec == REC_A ? P_BA : P_BX)[k] = S1;
            }
            D += Dq;
        }
        return seg_sum<S, W>(D);
    }

    // the piece's end becomes the next piece's start (balances; with d: the interval's end)
    __device__ void shift(bool d) const {
#pragma nounroll
        for (int i = 0; i < SP; i++) {
            const int q = sub + W * i;
            if (q > L) continue;
            for (int u = 0; u < nu; u++) {
                const int k = q * nus + u;
                if (d) plane(P_DA)[k] = plane(P_DB)[k];
                plane(P_BA)[k] = plane(P_BX)[k];
            }
        }
    }

    // the nearest primary breakpoint beyond tp (segment minimum)
    __device__ bool next_bp(S tp, S sgn, S &t) const {
        bool have = false;
        S best = 0;
        auto cand = [&](S b) {
            if (isfinite(b) && sgn * (b - tp) > 0 && (!have || sgn * (best - b) > 0)) {
                best = b;
                have = true;
            }
        };
#pragma nounroll
        for (int i = 0; i < SP; i++) {
            const int q = sub + W * i;
            if (q > L) continue;
            const int kidx = q == 0 ? 0 : (q == L ? 1 : -1);
            for (int u = 0; u < nu; u++) {
                if (kidx >= 0) {
                    const SmsKink<S> kk = kk_l[kidx * nus + u];
                    if (kk.valid != S(0)) {
                        cand(-kk.Q / kk.g);
                        if (kk.ak != S(0)) {
                            const S sl = kk.ak * kk.g, off = kk.ak * kk.Q;
                            cand((S(0) - kk.y0 - S(0)) / sl);
                            cand((S(1) - kk.y0 - S(0)) / sl);
                            cand((S(0) - kk.y0 - off) / sl);
                            cand((S(1) - kk.y0 - off) / sl);
                        }
                    }
                }
                if (q >= 1) {
                    const S ye = plane(P_YE)[q * nus + u];
                    cand((S(0) - ye) / S(1));
                    cand((S(1) - ye) / S(1));
                }
            }
        }
        return seg_best<S, W>(have, best, sgn, t);
    }

    // the nearest sign change beyond x, strictly inside (a, b), of the storage secondaries
    // (stor: from d at the interval's ends, P_DA and P_DB) or of the node balances (from the
    // piece's ends, P_BA and P_BX), linear in between
    __device__ bool next_root(bool stor, S a, S b, S x, S sgn, S &xr) const {
        bool have = false;
        S best = 0;
#pragma nounroll
        for (int i = 0; i < SP; i++) {
            const int q = sub + W * i;
            if (q > L) continue;
            for (int u = 0; u < nu; u++) {
                const int k0 = q * nus + u;
                const StoPar<S> par = un_l[u].par;
                const S da = plane(P_DA)[k0], db = plane(P_DB)[k0];
                const S ba = plane(P_BA)[k0], bb = plane(P_BX)[k0];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    if (!stor && k > 0) break;
                    const S A = stor ? sec(k, da, par) : ba, B = stor ? sec(k, db, par) : bb;
                    if ((A < 0) != (B < 0) && A != B) {
                        const S r = a + (b - a) * A / (A - B);
                        if (sgn * (r - a) > 0 && sgn * (b - r) > 0 && sgn * (r - x) > 0 &&
                            (!have || sgn * (best - r) > 0)) {
                            best = r;
                            have = true;
                        }
                    }
                }
            }
        }
        return seg_best<S, W>(have, best, sgn, xr);
    }

    // log Jacobian Σ (|S| − 1) log|P(t)/P| of the compensated kinks (FP64), segment sum
    __device__ double logjac(S t, bool &any) const {
        double lj = 0.0;
        int cnt = 0;
#pragma nounroll
        for (int i = 0; i < SP; i++) {
            const int q = sub + W * i;
            const int kidx = q == 0 ? 0 : (q == L ? 1 : -1);
            if (q > L || kidx < 0) continue;
            for (int u = 0; u < nu; u++) {
                const SmsKink<S> kk = kk_l[kidx * nus + u];
                if (kk.valid == S(0)) continue;
                cnt++;
                const double nm = kink_nm(*m, un_l[u].node, kmin);
                lj += nm * log(fabs((fmax((double)kk.Q + (double)kk.g * (double)t, 0.0) + (double)kk.e) / (double)kk.P));
            }
        }
        any = seg_sum<double, W>((double)cnt) > 0;
        return seg_sum<double, W>(lj);
    }
};

// words (of S) of one segment's LDS: the planes and the kinks
template <typename S>
__host__ __device__ constexpr size_t sms_seg_words(int nstep, int nus) {
    return SMS_NPLANE * (size_t)nstep * nus + sizeof(SmsKink<S>) / sizeof(S) * 2 * nus;
}
// words (of S) of the wave's moved units (nus: the largest group), shared by its segments (the
// move is the wave's)
template <typename S>
__host__ __device__ constexpr size_t sms_unit_words(int nus) {
    return (sizeof(SmsUnit<S>) + sizeof(S) - 1) / sizeof(S) * (size_t)nus;
}



// segment width of block length L: the power of two ≥ L + 1, from 4 to 64
__host__ __device__ constexpr int sms_width(int L) {
    return L + 1 <= 4 ? 4 : (L + 1 <= 8 ? 8 : (L + 1 <= 16 ? 16 : (L + 1 <= 32 ? 32 : 64)));
}

// The storage sweep's schedule on the device (the draws of sms_schedule, one thread): sched[0]
// whether it runs, sched[1] the number of block lengths, then per slot s (in sweep order) the
// length, its offset and the class drawn first at sched[2 + 3s …].
__global__ void k_sms_schedule(int T, double p, uint32_t k0, uint32_t k1, uint32_t ctr, int *sched) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t w[4];
    philox4x32(k0, k1, 0u, 0u, 9300u, ctr, w);
    sched[0] = u01(w[0], w[1]) < p ? 1 : 0;
    int len[8], n = 0;
    for (int L = 2; L <= 128; L *= 2)
        if (L < T / 2) len[n++] = L;
    for (int i = n - 1; i >= 1; i--) {   // Fisher–Yates
        philox4x32(k0, k1, (uint32_t)i, 0u, 9301u, ctr, w);
        const int jx = (int)mulhi32(w[0], (uint32_t)(i + 1));
        const int tmp = len[i];
        len[i] = len[jx];
        len[jx] = tmp;
    }
    sched[1] = n;
    for (int q = 0; q < n; q++) {
        philox4x32(k0, k1, (uint32_t)len[q], 0u, 9302u, ctr, w);
        sched[2 + 3 * q] = len[q];
        sched[3 + 3 * q] = (int)mulhi32(w[0], (uint32_t)len[q]);
        sched[4 + 3 * q] = u01(w[2], w[3]) < 0.5 ? 0 : 1;
    }
}

template <typename S, int W>
__global__ void __launch_bounds__(64, min_waves<S>(4, 2)) k_storage_ms(DevModel<S> m, SweepState<S> st, SmsArgs a, S beta, S maxnats,
                                                  S kmin) {
    extern __shared__ unsigned char sms_lds[];
    constexpr int SEGS = 64 / W;
    constexpr int SP = FStor<S, W>::SP;
    if (a.sched) {   // graph-stable form: this launch's length, offset and class from the schedule
        const int *sc = a.sched;
        if (!sc[0] || a.slot >= sc[1]) return;
        const int L = sc[2 + 3 * a.slot], o = sc[3 + 3 * a.slot], first = sc[4 + 3 * a.slot];
        if (sms_width(L) != W || o > m.T - L - 1) return;
        const int nb = (m.T - L - 1 - o) / L + 1;
        if (nb < 2) return;
        const int cls = a.ci == 0 ? first : 1 - first;
        const int nbc = (nb - cls + 1) / 2;
        if (nbc <= 0 || (int)blockIdx.x >= nbc * a.nrg) return;
        a.L = L;
        a.base = o + cls * L;
        a.nblk = nbc;
    }
    const int lane = threadIdx.x;
    const int seg = lane / W, sub = lane % W;
    const int blk = blockIdx.x / a.nrg, rg = blockIdx.x % a.nrg;
    const int L = a.L, b0 = a.base + 2 * L * blk;
    const int j = rg * SEGS + seg;
    const bool valid = j < m.R;
    const int nstep = L + 1, nus = a.nus;
    const size_t TR = (size_t)m.T * m.R;
    // LDS: the wave's units [nus], then per segment the planes [SMS_NPLANE][nstep][nus] and
    // the kinks [2][nus]
    SmsUnit<S> *un_l = (SmsUnit<S> *)sms_lds;
    S *base = (S *)sms_lds + sms_unit_words<S>(nus) + (size_t)seg * sms_seg_words<S>(nstep, nus);
    S *ye_l = base + P_YE * nstep * nus, *ye1_l = base + P_YE1 * nstep * nus;
    S *S0_l = base + P_S0 * nstep * nus, *c0_l = base + P_C0 * nstep * nus;
    S *j0_l = base + P_J0 * nstep * nus, *ph0_l = base + P_PH0 * nstep * nus;
    SmsKink<S> *kk_l = (SmsKink<S> *)(base + SMS_NPLANE * nstep * nus);
    MoveOut out;
    for (int mv = 0; mv < a.nmoves; mv++) {
        uint32_t s[4];
        philox4x32(a.k0, a.k1, (uint32_t)mv, (uint32_t)b0, 9200u + (uint32_t)L, a.ctr, s);
        // the units: a country–carrier group (probability p_agg) or one unit; lane u < nu of the
        // wave stages unit u
        int si, nu, node;
        if (u01(s[0], s[1]) < a.p_agg) {
            const int g = (int)mulhi32(s[2], (uint32_t)a.ngroups);
            si = a.grp_si[g];
            nu = a.grp_off[g + 1] - a.grp_off[g];
            node = lane < nu ? a.grp_node[a.grp_off[g] + lane] : 0;
        } else {
            const int un = a.units[(int)mulhi32(s[3], (uint32_t)a.nunits)];
            si = un / m.N;
            nu = 1;
            node = un % m.N;
        }
        if (lane < nu) {
            const int us = si * m.N + node;
            SmsUnit<S> U;
            U.par = {m.hours[us], m.eta_in[us], m.eta_out[us], m.qt[us], m.mc[(ESM_KSTOR + si) * m.N + node]};
            U.kst = m.kc[(ESM_KSTOR + si) * m.N + node];
            U.ks = U.kst * m.hours[us] / m.dt;
            U.node = node;
            U.gk = m.kc[ESM_KGAS * m.N + node];
            U.gmc = m.mc[ESM_KGAS * m.N + node];
            un_l[lane] = U;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (!valid) continue;
        FStor<S, W> f;
        f.m = &m;
        f.L = L;
        f.sub = sub;
        f.nu = nu;
        f.nus = nus;
        f.pl = base;
        f.kk_l = kk_l;
        f.un_l = un_l;
        f.kmin = kmin;
        const S *ye_g = st.v + plane((2 + si) * m.N, m.T, m.R);
        const S *yg_g = st.v + plane(m.N, m.T, m.R);
        // stage the block's values of the segment's replica (each lane its own steps)
#pragma nounroll
        for (int i = 0; i < SP; i++) {
            const int q = sub + W * i;
            if (q > L) continue;
            const int tau = b0 + q, tau1 = tau + 1 == m.T ? 0 : tau + 1;
            for (int u = 0; u < nu; u++) {
                const SmsUnit<S> U = un_l[u];
                const size_t on = (size_t)U.node * TR;
                const int i0 = q * nus + u;
                const S ye = ye_g[on + (size_t)tau * m.R + j], ye1 = ye_g[on + (size_t)tau1 * m.R + j];
                const S S0 = st.r[on + (size_t)tau * m.R + j];
                ye_l[i0] = ye;
                ye1_l[i0] = ye1;
                S0_l[i0] = S0;
                // the unit's storage term and its node's balance at t = 0
                S d0, j0;
                c0_l[i0] = soc_cost_p(U.par, m.dt, m.cv, U.kst, U.kst + m.kappa0, ye, ye1, d0, j0);
                j0_l[i0] = j0;
                ph0_l[i0] = phibal(S0, m.c_shed, m.c_dump);
            }
            // the pro-rata kinks at the block's ends (`prorata!` with `soc_qadj`)
            const int kidx = q == 0 ? 0 : (q == L ? 1 : -1);
            if (kidx >= 0)
                for (int u = 0; u < nu; u++) {
                    const SmsUnit<S> U = un_l[u];
                    SmsKink<S> kk{0, 0, 0, 0, 0, 0, 0, 0, 0};
                    const int n = U.node;
                    const S k = m.kc[ESM_KGAS * m.N + n];
                    const S g = kidx == 0 ? U.ks * (S(1) - S(0)) : U.ks * (S(0) - S(1));
                    if (fabs(k) >= kmin && g != S(0)) {
                        const S w = yg_g[(size_t)n * TR + (size_t)tau * m.R + j];
                        const S P = S(0) + k * w;
                        const S Qs = S(0) + k * clampv(w, S(0), S(1));
                        if (!(fabs(P) < S(1e-6))) {
                            // soc_qadj: nominal minus actual injection of the moved unit at (n, τ)
                            const S Y = ye_l[q * nus + u], Y1 = ye1_l[q * nus + u];
                            const S ec = clampv(Y, S(0), S(1)), ec1 = clampv(Y1, S(0), S(1));
                            const S d = U.par.h * (ec1 - ec) / m.dt - U.par.qt;
                            S pc, mcp;
                            if (d >= 0) {
                                pc = U.par.ein > 0 ? d / U.par.ein : S(0);
                                mcp = 0;
                            } else {
                                pc = 0;
                                mcp = fmin(-d * U.par.eout, S(1));
                            }
                            const S qadj = S(0) + (U.kst * U.par.h * (Y1 - Y) / m.dt - U.kst * (pc - mcp));
                            kk.Q = st.r[(size_t)n * TR + (size_t)tau * m.R + j] + Qs + qadj;
                            kk.g = g;
                            kk.e = P - fmax(kk.Q, S(0));
                            kk.P = P;
                            const S sig = k * w / P;
                            kk.ak = sig == S(0) ? S(0) : sig / k;
                            kk.y0 = w;
                            kk.valid = 1;
                            if (kk.ak != S(0)) kk.gc0 = f.gen_c(n, w, kk.gp0);
                        }
                    }
                    kk_l[kidx * nus + u] = kk;
                }
        }
        double u1, u2;
        uniforms2(a.k0, a.k1, j, (uint32_t)b0, 0x50000000u + (uint32_t)(L * 16 + mv), a.ctr, u1, u2);
        int status = 0;
        S t = sample_stor(f, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
        if (t != 0) {
            bool any;
            const double lj = f.logjac(t, any);
            if (any) {
                uint32_t w4[4];
                philox4x32(a.k0, a.k1, (uint32_t)j, (uint32_t)b0, 0x70000000u + (uint32_t)(L * 16 + mv), a.ctr, w4);
                if (!(u01(w4[0], w4[1]) < exp(fmin(0.0, lj)))) {
                    t = 0;
                    status = 2;
                }
            }
        }
        if (sub == 0) out.add(status, t != 0, f.nev);
        if (t != 0) {
            S *ye_w = st.v + plane((2 + si) * m.N, m.T, m.R);
            S *yg_w = st.v + plane(m.N, m.T, m.R);
#pragma nounroll
            for (int i = 0; i < SP; i++) {
                const int q = sub + W * i;
                if (q > L) continue;
                const int tau = b0 + q;
                const int kidx = q == 0 ? 0 : (q == L ? 1 : -1);
                for (int u = 0; u < nu; u++) {
                    S d1, S1;
                    f.unit_step(q, u, t, d1, S1);
                    const size_t o = (size_t)un_l[u].node * TR + (size_t)tau * m.R + j;
                    st.r[o] = S1;
                    if (q >= 1) ye_w[o] = ye_l[q * nus + u] + t;
                    if (kidx >= 0) {
                        const SmsKink<S> kk = kk_l[kidx * nus + u];
                        if (kk.valid != S(0) && kk.ak != S(0)) yg_w[o] = kk.y0 + kk.ak * FStor<S, W>::kphi(kk.Q, kk.g, t);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    add_counts(a.stats, out);
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// the schedule of one storage sweep (drawn on the host from the sweep's Philox counter)
struct SmsSchedule {
    bool run = false;
    std::vector<int> lengths;      // in sweep order
    std::vector<int> offset;       // per length in that order
    std::vector<int> first;        // class drawn first (0: even blocks, 1: odd)
};

static SmsSchedule sms_schedule(int T, double p, uint32_t k0, uint32_t k1, uint32_t ctr) {
    SmsSchedule sc;
    uint32_t w[4];
    philox4x32(k0, k1, 0u, 0u, 9300u, ctr, w);
    sc.run = u01(w[0], w[1]) < p;
    if (!sc.run) return sc;
    for (int L = 2; L <= 128; L *= 2)
        if (L < T / 2) sc.lengths.push_back(L);
    for (int i = (int)sc.lengths.size() - 1; i >= 1; i--) {   // Fisher–Yates
        philox4x32(k0, k1, (uint32_t)i, 0u, 9301u, ctr, w);
        const int jx = (int)mulhi32(w[0], (uint32_t)(i + 1));
        std::swap(sc.lengths[i], sc.lengths[jx]);
    }
    for (int L : sc.lengths) {
        philox4x32(k0, k1, (uint32_t)L, 0u, 9302u, ctr, w);
        sc.offset.push_back((int)mulhi32(w[0], (uint32_t)L));
        sc.first.push_back(u01(w[2], w[3]) < 0.5 ? 0 : 1);
    }
    return sc;
}

// The graph-stable form of one storage sweep (a captured batch): the schedule drawn on the device,
// then for every slot of the length order and both classes one launch per segment width, each
// sized for the largest block count and LDS of its width's lengths; the launch whose width is the
// slot's length's does that length's class, the others return at once. The same moves in the same
// order as the host-scheduled form, with a launch sequence that does not depend on the draws.
template <typename S>
static int launch_storage_ms_stable(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats) {
    const int T = h->T, R = h->R;
    hipLaunchKernelGGL(k_sms_schedule, dim3(1), dim3(64), 0, h->stream, T, h->sweep_pms, k0, k1, ctr, h->d_sched);
    ESM_HIP_CHECK(hipGetLastError());
    std::vector<int> lengths;
    for (int L = 2; L <= 128; L *= 2)
        if (L < T / 2) lengths.push_back(L);
    if (lengths.empty()) return 0;
    const DevModel<S> m = dev_model<S>(h);
    SweepState<S> st{(S *)h->d_v, (S *)h->d_r, (S *)h->d_F, (const S *)h->d_x};
    const S beta = (S)h->m.beta_v, mxn = (S)maxnats, kmin = S(1);
    const MoveTablesDev &d = h->dmt;
    int lds_max = 0, nus = 1;
    for (size_t g = 0; g + 1 < h->mt.grp_off.size(); g++) nus = std::max(nus, h->mt.grp_off[g + 1] - h->mt.grp_off[g]);
    ESM_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    // per segment width: the most blocks of a class (offset 0, the first class) and the largest
    // LDS over the lengths of that width
    const int widths[5] = {4, 8, 16, 32, 64};
    int wblocks[5] = {0, 0, 0, 0, 0};
    size_t wlds[5] = {0, 0, 0, 0, 0};
    for (int q = 0; q < 5; q++)
        for (int L : lengths) {
            if (sms_width(L) != widths[q]) continue;
            const int nb = (T - L - 1) / L + 1;
            wblocks[q] = std::max(wblocks[q], (nb + 1) / 2);
            wlds[q] = std::max(wlds[q], (sms_unit_words<S>(nus) + (size_t)(64 / widths[q]) * sms_seg_words<S>(L + 1, nus)) *
                                            sizeof(S));
        }
    for (int q = 0; q < 5; q++)
        if (wlds[q] > (size_t)lds_max) {
            set_error("storage multiscale sweep: a block length needs more LDS per workgroup than the device allows");
            return ESM_EINVAL;
        }
    ProfScope p(h, ESM_KERN_SMS);
    for (int slot = 0; slot < (int)lengths.size(); slot++)
        for (int ci = 0; ci < 2; ci++)
            for (int q = 0; q < 5; q++) {
                const int W = widths[q], blocks = wblocks[q];
                const size_t lds = wlds[q];
                if (blocks == 0) continue;
                const int nrg = (R + 64 / W - 1) / (64 / W);
                SmsArgs a{k0, k1, ctr, 0, 0, 0, nrg, h->sweep_ms_moves, nus, h->sweep_ms_pagg,
                          d.sms_units, (int)h->mt.sms_units.size(), d.grp_off, d.grp_node, d.grp_si,
                          (int)h->mt.grp_si.size(), h->d_stats + (size_t)ESM_FAM_SMS * ESM_NSTAT * STAT_SLOTS,
                          h->d_sched, slot, ci};
                const dim3 grid((unsigned)blocks * (unsigned)nrg), blk(64);
                switch (W) {
                    case 4: hipLaunchKernelGGL((k_storage_ms<S, 4>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                    case 8: hipLaunchKernelGGL((k_storage_ms<S, 8>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                    case 16: hipLaunchKernelGGL((k_storage_ms<S, 16>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                    case 32: hipLaunchKernelGGL((k_storage_ms<S, 32>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                    default: hipLaunchKernelGGL((k_storage_ms<S, 64>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                }
                ESM_HIP_CHECK(hipGetLastError());
            }
    return 0;
}

template <typename S>
int launch_storage_ms(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats) {
    const int T = h->T, R = h->R;
    if (!(h->sweep_pms > 0) || h->mt.sms_units.empty()) return 0;
    if (h->graph_stable) return launch_storage_ms_stable<S>(h, k0, k1, ctr, maxnats);
    const SmsSchedule sc = sms_schedule(T, h->sweep_pms, k0, k1, ctr);
    if (!sc.run) return 0;
    const DevModel<S> m = dev_model<S>(h);
    SweepState<S> st{(S *)h->d_v, (S *)h->d_r, (S *)h->d_F, (const S *)h->d_x};
    const S beta = (S)h->m.beta_v, mxn = (S)maxnats, kmin = S(1);
    const MoveTablesDev &d = h->dmt;
    int lds_max = 0, nus = 1;
    for (size_t g = 0; g + 1 < h->mt.grp_off.size(); g++) nus = std::max(nus, h->mt.grp_off[g + 1] - h->mt.grp_off[g]);
    ESM_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    for (size_t li = 0; li < sc.lengths.size(); li++) {
        const int L = sc.lengths[li], o = sc.offset[li];
        if (o > T - L - 1) continue;
        const int nb = (T - L - 1 - o) / L + 1;
        if (nb < 2) continue;
        int W = 4;
        while (W < L + 1 && W < 64) W *= 2;
        const int segs = 64 / W;
        const int nrg = (R + segs - 1) / segs;
        for (int ci = 0; ci < 2; ci++) {
            const int cls = ci == 0 ? sc.first[li] : 1 - sc.first[li];
            const int nbc = (nb - cls + 1) / 2;
            if (nbc <= 0) continue;
            SmsArgs a{k0, k1, ctr, L, o + cls * L, nbc, nrg, h->sweep_ms_moves, nus, h->sweep_ms_pagg,
                      d.sms_units, (int)h->mt.sms_units.size(), d.grp_off, d.grp_node, d.grp_si,
                      (int)h->mt.grp_si.size(), h->d_stats + (size_t)ESM_FAM_SMS * ESM_NSTAT * STAT_SLOTS};
            const size_t lds = (sms_unit_words<S>(nus) + (size_t)segs * sms_seg_words<S>(L + 1, nus)) * sizeof(S);
            if (lds > (size_t)lds_max) {
                set_error("storage multiscale sweep: block length " + std::to_string(L) + " needs " +
                          std::to_string(lds) + " B of LDS per workgroup, the device allows " + std::to_string(lds_max));
                return ESM_EINVAL;
            }
            const dim3 grid((unsigned)nbc * (unsigned)nrg), blk(64);
            ProfScope p(h, ESM_KERN_SMS);
            switch (W) {
                case 4: hipLaunchKernelGGL((k_storage_ms<S, 4>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                case 8: hipLaunchKernelGGL((k_storage_ms<S, 8>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                case 16: hipLaunchKernelGGL((k_storage_ms<S, 16>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                case 32: hipLaunchKernelGGL((k_storage_ms<S, 32>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
                default: hipLaunchKernelGGL((k_storage_ms<S, 64>), grid, blk, lds, h->stream, m, st, a, beta, mxn, kmin); break;
            }
            ESM_HIP_CHECK(hipGetLastError());
        }
    }
    return 0;
}

template int launch_storage_ms<float>(esm_handle *, uint32_t, uint32_t, uint32_t, double);
template int launch_storage_ms<double>(esm_handle *, uint32_t, uint32_t, uint32_t, double);


// ---------------------------------------------------------------------------------------------
// The long-range storage sweep (`storage_long_sweep!`, src/storage_long.jl:245-298), which
// `operation_sweep!` runs after the multiscale sweep with probability `storage_long` (1 by
// default, src/local.jl:2419-2423). Per cell (node at the complete level), for every storage unit
// with at least `hmin` hours and κs ≥ 1: a level move (every instant of the year rises by t) and
// `nmoves` long-range plateaus (`prep_slong!`, :77-146: the state of charge rises at the cyclic
// instants a+1 … a+L, L uniform on 1 … T−1; the node's gas unit follows pro rata at both end
// steps), shuffled. Interior steps whose two instants stay in [0, 1] change nothing (their cost
// change is exactly zero, their secondaries keep their values at t = 0: `mv_eval!`, :148-188).
//
// MI355X layout: one wave per (node, replica), running the node's move list in order (the
// cells are independent). The unit's state-of-charge trajectory is staged in LDS (T values; in
// global memory when T does not fit); lane l owns the move's steps q ≡ l (mod 64) and keeps the
// window of t over which its interior steps stay in the box, so an evaluation inside it costs the
// two end steps only. The walk is the multiscale sweep's, with W = 64: the end steps'
// secondaries are recorded by the evaluations at the interval and piece ends, the out-of-box
// interior steps' ones are recomputed there (the others do not vary). The move
// list is drawn from the sweep's counter by every wave of a node alike (Philox; see esm.h).
// ---------------------------------------------------------------------------------------------

constexpr int SLG_MAXM = 2 * 257;      // moves per node: per unit one level move and ≤ 256 plateaus
constexpr int SLG_MAXCH = 64;          // window chunks per lane

// interior steps of a lane per window chunk (at least 8), and the chunks, over a horizon of T
__host__ __device__ constexpr int slg_ch(int T) {
    return ((T + 63) / 64 + SLG_MAXCH - 1) / SLG_MAXCH > 8 ? ((T + 63) / 64 + SLG_MAXCH - 1) / SLG_MAXCH : 8;
}
__host__ __device__ constexpr int slg_nch(int T) { return ((T + 63) / 64 + slg_ch(T) - 1) / slg_ch(T); }

// float bounds of a window rounded inward (a conservative filter of the exact test)
__device__ __forceinline__ float inward_lo(double x) {
    float f = (float)x;
    return (double)f < x ? nextafterf(f, INFINITY) : f;
}
__device__ __forceinline__ float inward_hi(double x) {
    float f = (float)x;
    return (double)f > x ? nextafterf(f, -INFINITY) : f;
}

struct SlgArgs {
    uint32_t k0, k1, ctr;
    int nmoves;                        // plateaus per unit
    uint32_t elig[ESM_NSTOR][8];       // bit n of unit (s, n): takes part (hours ≥ hmin, κs ≥ 1)
    unsigned long long *st_long, *st_level;
    double pcoin;                      // ≥ 0 (graph-stable form): the sweep runs with this probability,
                                       // decided on the device from its draw (0, 0, 9410, ctr)
};

template <typename S, bool YL>
struct FLong {
    enum { REC_NONE, REC_B, REC_X, REC_A, REC_T };
    static constexpr bool CACHED = true;    // the end steps' secondaries are recorded (interior ones recomputed)
    const DevModel<S> *m;
    const S *yv;          // the unit's state of charge: LDS [T] (YL) or the global plane at (n, ·, j)
    const S *rv;          // the node's residual plane at (n, ·, j), stride R
    const S *gv;          // the node's gas plane at (n, ·, j), stride R
    int T, R, lane, L, level, a0, n;
    StoPar<S> par;
    S kst, w, kmin;
    const SmsKink<S> *kk; // [2] in LDS: the kinks at steps a and b
    S llo, lhi;           // t in [llo, lhi]: every interior step of the lane stays in the box
    // the same per chunk of `ch` consecutive interior steps of the lane (LDS [chunk][lane], bounds
    // rounded inward to float: a conservative filter, every step is still tested exactly)
    const float *cwl, *cwh;
    int ch;
    mutable int nev = 0;
    // the lane's nearest clamp point of its rising instants beyond the walk's last point `bp_tp`
    // in direction `bp_sgn` (the points are fixed during a move and a walk only moves forward, so
    // the lane scans its instants again only when the walk passes it or a new walk starts)
    mutable S bp_tp = 0, bp_sgn = 0, bp_next = 0;
    mutable bool bp_have = false;
    // the end steps' d at the interval ends (A: tp, B: t) and balances at the piece ends (A: a0,
    // X: x or t), recorded by the evaluations there; slot 0 step a, slot 1 step b
    // (scalars written by branches: an array indexed by a run-time slot, or a store through a
    // selected address, would keep the whole walker in scratch memory)
    mutable S edA0 = 0, edA1 = 0, edB0 = 0, edB1 = 0, ebA0 = 0, ebA1 = 0, ebX0 = 0, ebX1 = 0;
    __device__ void record(int e, int rec, S d1, S S1) const {
        if (rec == REC_B) {
            if (e) edB1 = d1;
            else edB0 = d1;
        } else if (rec == REC_A) {
            if (e) edA1 = d1;
            else edA0 = d1;
        }
        if (rec == REC_A) {
            if (e) ebA1 = S1;
            else ebA0 = S1;
        } else if (rec != REC_NONE) {
            if (e) ebX1 = S1;
            else ebX0 = S1;
        }
    }

    __device__ S y(int tau) const { return YL ? yv[tau] : yv[(size_t)tau * R]; }
    __device__ S r0(int tau) const { return rv[(size_t)tau * R]; }
    __device__ int tau(int q) const {
        const int x = a0 + q;
        return x >= T ? x - T : x;
    }
    __device__ int next(int tau) const { return tau + 1 == T ? 0 : tau + 1; }
    // the lane's first step at or after q0
    __device__ int first(int q0) const { return lane >= q0 ? lane : lane + 64 * ((q0 - lane + 63) / 64); }
    __device__ bool owns_a() const { return !level && lane == 0; }
    __device__ bool owns_b() const { return !level && lane == L % 64; }
    __device__ bool inside(int q, S t) const {
        const int ta = tau(q);
        const S ye = y(ta), ye1 = y(next(ta));
        return S(0) <= ye && ye <= S(1) && S(0) <= ye1 && ye1 <= S(1) && fmax(-ye, -ye1) <= t &&
               t <= fmin(S(1) - ye, S(1) - ye1);
    }
    __device__ S gen_c(S yy, S &p) const {
        S c;
        const S k = m->kc[ESM_KGAS * m->N + n];
        gen_cost(m->mc[ESM_KGAS * m->N + n], k, k + m->kappa0, m->dt, m->cv, yy, c, p);
        return c;
    }
    // an end step (q = 0: step a, its end instant rises; q = L: step b, its start instant rises)
    __device__ S end_term(int q, S t, S &d1, S &S1) const {
        const int ta = tau(q);
        const S ye = y(ta), ye1 = y(next(ta)), S0 = r0(ta);
        S D = 0, dr = 0;
        const SmsKink<S> k = kk[q == 0 ? 0 : 1];
        if (k.valid != S(0) && k.ak != S(0)) {
            S p1;
            const S c1 = gen_c(k.y0 + k.ak * FStor<S, 64>::kphi(k.Q, k.g, t), p1);
            D += c1 - k.gc0;
            dr -= p1 - k.gp0;
        }
        const S dye = q >= 1 ? S(1) : S(0), dye1 = q <= L - 1 ? S(1) : S(0);
        S d0, j0, j1;
        const S c0 = soc_cost_p(par, m->dt, m->cv, kst, w, ye, ye1, d0, j0);
        const S c1 = soc_cost_p(par, m->dt, m->cv, kst, w, ye + t * dye, ye1 + t * dye1, d1, j1);
        D += c1 - c0;
        dr += j1 - j0;
        S1 = S0 + dr;
        D += m->dt * (phibal(S1, m->c_shed, m->c_dump) - phibal(S0, m->c_shed, m->c_dump));
        return D;
    }
    // an interior step (both instants rise) at t: false when both instants stay in the box (no
    // change; with `sec`, d and the balance keep their values at t = 0: d1, S1), else its term x
    __device__ bool int_term(int q, S t, bool sec, S &x, S &d1, S &S1) const {
        const int ta = tau(q);
        const S ye = y(ta), ye1 = y(next(ta));
        const bool in = S(0) <= ye && ye <= S(1) && S(0) <= ye1 && ye1 <= S(1) && fmax(-ye, -ye1) <= t &&
                        t <= fmin(S(1) - ye, S(1) - ye1);
        if (in && !sec) return false;
        const S S0 = r0(ta);
        S d0, j0, j1;
        const S c0 = soc_cost_p(par, m->dt, m->cv, kst, w, ye, ye1, d0, j0);
        if (in) {
            d1 = d0;
            S1 = S0;
            return false;
        }
        const S c1 = soc_cost_p(par, m->dt, m->cv, kst, w, ye + t, ye1 + t, d1, j1);
        S1 = S0 + (j1 - j0);
        x = (c1 - c0) + m->dt * (phibal(S1, m->c_shed, m->c_dump) - phibal(S0, m->c_shed, m->c_dump));
        return true;
    }
    __device__ int qi0() const { return 1; }
    __device__ int qi1() const { return level ? L : L - 1; }   // interior steps qi0 … qi1

    // ΔC(t): the lane's steps in increasing q (end a, the interior steps out of the box, end b),
    // summed over the wave
    __device__ S eval(S t, int rec) const {
        if (rec <= REC_X) nev++;
        S d1, S1, xa = 0, xb = 0;
        const bool oa = owns_a(), ob = owns_b();
        // the end terms in one pass of the two owning lanes (lane 0: step a, lane L mod 64: step
        // b), a second pass for step b when lane 0 owns both (L ≡ 0 mod 64)
        if (oa || ob) {
            const int e = oa ? 0 : 1;
            const S x = end_term(e == 0 ? 0 : L, t, d1, S1);
            if (e == 0) xa = x;
            else xb = x;
            record(e, rec, d1, S1);
        }
        if (oa && ob) {
            xb = end_term(L, t, d1, S1);
            record(1, rec, d1, S1);
        }
        S D = 0;
        if (oa) D += xa;
        if (!(llo <= t && t <= lhi))
            for (int q0 = first(qi0()), c = 0; q0 <= qi1(); q0 += 64 * ch, c++) {
                if (cwl[c * 64 + lane] <= t && t <= cwh[c * 64 + lane]) continue;
                for (int q = q0; q <= qi1() && q < q0 + 64 * ch; q += 64) {
                    S x;
                    if (int_term(q, t, false, x, d1, S1)) D += x;
                }
            }
        if (ob) D += xb;
        return seg_sum<S, 64>(D);
    }
    __device__ void shift(bool d) const {
        if (d) {
            edA0 = edB0;
            edA1 = edB1;
        }
        ebA0 = ebX0;
        ebA1 = ebX1;
    }

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
            const SmsKink<S> k = kk[e];
            if (k.valid == S(0)) continue;
            cand(-k.Q / k.g);
            if (k.ak != S(0)) {
                const S sl = k.ak * k.g, off = k.ak * k.Q;
                cand((S(0) - k.y0 - S(0)) / sl);
                cand((S(1) - k.y0 - S(0)) / sl);
                cand((S(0) - k.y0 - off) / sl);
                cand((S(1) - k.y0 - off) / sl);
            }
        }
        // the clamp points of the rising instants (cached while the walk moves forward)
        if (!(bp_sgn == sgn && sgn * (tp - bp_tp) >= 0 && (!bp_have || sgn * (bp_next - tp) > 0))) {
            bool h2 = false;
            S b2 = 0;
            auto cand2 = [&](S b) {
                if (isfinite(b) && sgn * (b - tp) > 0 && (!h2 || sgn * (b2 - b) > 0)) {
                    b2 = b;
                    h2 = true;
                }
            };
            for (int q = first(1); q <= L; q += 64) {
                const S ye = y(tau(q));
                cand2((S(0) - ye) / S(1));
                cand2((S(1) - ye) / S(1));
            }
            bp_have = h2;
            bp_next = b2;
            bp_sgn = sgn;
        }
        bp_tp = tp;
        if (bp_have) cand(bp_next);
        return seg_best<S, 64>(have, best, sgn, t);
    }

    // the nearest sign change beyond x, strictly inside (a, b), of the storage secondaries (stor)
    // or of the balances, from their values at a and b (linear in between)
    __device__ bool next_root(bool stor, S a, S b, S x, S sgn, S &xr) const {
        bool have = false;
        S best = 0;
        auto test = [&](S da, S db, S Ba, S Bb) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (!stor && k > 0) break;
                const S A = stor ? FStor<S, 64>::sec(k, da, par) : Ba, B = stor ? FStor<S, 64>::sec(k, db, par) : Bb;
                if ((A < 0) != (B < 0) && A != B) {
                    const S r = a + (b - a) * A / (A - B);
                    if (sgn * (r - a) > 0 && sgn * (b - r) > 0 && sgn * (r - x) > 0 && (!have || sgn * (best - r) > 0)) {
                        best = r;
                        have = true;
                    }
                }
            }
        };
        S da, db, Ba, Bb;
        if (owns_a()) test(edA0, edB0, ebA0, ebX0);
        if (!(llo <= a && a <= lhi && llo <= b && b <= lhi))
            for (int q0 = first(qi0()), c = 0; q0 <= qi1(); q0 += 64 * ch, c++) {
                const float lo = cwl[c * 64 + lane], hi = cwh[c * 64 + lane];
                if (lo <= a && a <= hi && lo <= b && b <= hi) continue;
                for (int q = q0; q <= qi1() && q < q0 + 64 * ch; q += 64) {
                    if (inside(q, a) && inside(q, b)) continue;   // constant secondaries
                    S x;
                    int_term(q, a, true, x, da, Ba);
                    int_term(q, b, true, x, db, Bb);
                    test(da, db, Ba, Bb);
                }
            }
        if (owns_b()) test(edA1, edB1, ebA1, ebX1);
        return seg_best<S, 64>(have, best, sgn, xr);
    }

    __device__ double logjac(S t, bool &any) const {
        double lj = 0.0;
        int cnt = 0;
        for (int e = 0; e < 2; e++) {
            if (!(e == 0 ? owns_a() : owns_b())) continue;
            const SmsKink<S> k = kk[e];
            if (k.valid == S(0)) continue;
            cnt++;
            const double nm = kink_nm(*m, n, kmin);
            lj += nm * log(fabs((fmax((double)k.Q + (double)k.g * (double)t, 0.0) + (double)k.e) / (double)k.P));
        }
        any = seg_sum<double, 64>((double)cnt) > 0;
        return seg_sum<double, 64>(lj);
    }
};

// bytes of the LDS of one wave: the kinks, the chunk windows, then the trajectory (YL)
template <typename S>
__host__ __device__ constexpr size_t slg_bytes(int T, bool yl) {
    return sizeof(SmsKink<S>) * 2 + sizeof(float) * 2 * 64 * (size_t)slg_nch(T) + (yl ? sizeof(S) * (size_t)T : 0);
}

template <typename S, bool YL>
__global__ void __launch_bounds__(64) k_storage_long(DevModel<S> m, SweepState<S> st, SlgArgs a, S beta, S maxnats,
                                                    S kmin) {
    extern __shared__ unsigned char slg_lds[];
    __shared__ int msi[SLG_MAXM], ma[SLG_MAXM], mL[SLG_MAXM];
    if (a.pcoin >= 0) {
        uint32_t w[4];
        philox4x32(a.k0, a.k1, 0u, 0u, 9410u, a.ctr, w);
        if (!(u01(w[0], w[1]) < a.pcoin)) return;
    }
    const int lane = threadIdx.x;
    const int n = blockIdx.x / m.R, j = blockIdx.x % m.R;
    const int N = m.N, T = m.T, R = m.R;
    // the node's move list (the same in every wave of the node): per unit taking part a level
    // move and `nmoves` plateaus, then a Fisher–Yates shuffle
    int cnt = 0, u = 0;
    for (int si = 0; si < ESM_NSTOR; si++) {
        if (!((a.elig[si][n >> 5] >> (n & 31)) & 1u)) continue;
        uint32_t w[4];
        philox4x32(a.k0, a.k1, (uint32_t)n, (uint32_t)u, 9411u, a.ctr, w);
        if (u01(w[0], w[1]) < 1.0) {
            if (lane == 0) {
                msi[cnt] = si;
                ma[cnt] = 0;
                mL[cnt] = T;
            }
            cnt++;
        }
        for (int mm = 0; mm < a.nmoves; mm++) {
            philox4x32(a.k0, a.k1, (uint32_t)n, 65536u * (uint32_t)u + (uint32_t)mm, 9412u, a.ctr, w);
            if (lane == 0) {
                msi[cnt] = si;
                mL[cnt] = 1 + (int)mulhi32(w[0], (uint32_t)(T - 1));
                ma[cnt] = (int)mulhi32(w[1], (uint32_t)T);
            }
            cnt++;
        }
        u++;
    }
    if (cnt == 0) return;
    if (lane == 0)
        for (int i = cnt - 1; i >= 1; i--) {
            uint32_t w[4];
            philox4x32(a.k0, a.k1, (uint32_t)n, (uint32_t)i, 9413u, a.ctr, w);
            const int jx = (int)mulhi32(w[0], (uint32_t)(i + 1));
            int tmp = msi[i]; msi[i] = msi[jx]; msi[jx] = tmp;
            tmp = ma[i]; ma[i] = ma[jx]; ma[jx] = tmp;
            tmp = mL[i]; mL[i] = mL[jx]; mL[jx] = tmp;
        }
    __syncthreads();
    SmsKink<S> *kk_l = (SmsKink<S> *)slg_lds;
    const int nch = slg_nch(T);
    float *cwl = (float *)(slg_lds + sizeof(SmsKink<S>) * 2), *cwh = cwl + 64 * nch;
    S *ys = (S *)(cwh + 64 * nch);
    const size_t TR = (size_t)T * R;
    const S *rv = st.r + (size_t)n * TR + j;
    S *gv = st.v + plane(N, T, R) + (size_t)n * TR + j;
    MoveOut out_long, out_level;
    // With the trajectory in LDS (YL) the moves of a unit read and write it there, and it goes back
    // to memory once, when the wave turns to the node's other unit or has made its last move: every
    // read of the unit's instants in between is this wave's (the node's residuals and gas stay in
    // memory, written at once), and a write-back per unit instead of per move takes the stores,
    // one cache line per instant in the replica-fastest layout, from Σ L over the moves to T.
    int cur = -1;
    auto write_back = [&]() {
        if (!YL || cur < 0) return;
        S *yc = st.v + plane((2 + cur) * N, T, R) + (size_t)n * TR + j;
        for (int tau = lane; tau < T; tau += 64) yc[(size_t)tau * R] = ys[tau];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    for (int i = 0; i < cnt; i++) {
        const int si = msi[i], L = mL[i], level = L >= T;
        S *yg = st.v + plane((2 + si) * N, T, R) + (size_t)n * TR + j;
        if (YL && si != cur) {          // stage the unit's trajectory
            write_back();
            for (int tau = lane; tau < T; tau += 64) ys[tau] = yg[(size_t)tau * R];
            cur = si;
        }
        FLong<S, YL> f;
        f.m = &m;
        f.yv = YL ? ys : yg;
        f.rv = rv;
        f.gv = gv;
        f.T = T;
        f.R = R;
        f.lane = lane;
        f.L = level ? T : L;
        f.level = level;
        f.a0 = level ? T - 1 : ma[i];
        f.n = n;
        const int us = si * N + n;
        f.par = {m.hours[us], m.eta_in[us], m.eta_out[us], m.qt[us], m.mc[(ESM_KSTOR + si) * N + n]};
        f.kst = m.kc[(ESM_KSTOR + si) * N + n];
        f.w = f.kst + m.kappa0;
        f.kmin = kmin;
        f.kk = kk_l;
        f.cwl = cwl;
        f.cwh = cwh;
        f.ch = slg_ch(T);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // the kinks at the end steps (`prorata!` with `soc_qadj`), by the lanes owning them
        if (!level)
            for (int e = 0; e < 2; e++) {
                if (!(e == 0 ? f.owns_a() : f.owns_b())) continue;
                const int tau = f.tau(e == 0 ? 0 : f.L), tau1 = f.next(tau);
                SmsKink<S> k{0, 0, 0, 0, 0, 0, 0, 0, 0};
                const S ks = f.kst * m.hours[us] / m.dt;
                const S kg = m.kc[ESM_KGAS * N + n];
                const S g = e == 0 ? ks * (S(1) - S(0)) : ks * (S(0) - S(1));
                if (fabs(kg) >= kmin && g != S(0)) {
                    const S w = gv[(size_t)tau * R];
                    const S P = S(0) + kg * w;
                    const S Qs = S(0) + kg * clampv(w, S(0), S(1));
                    if (!(fabs(P) < S(1e-6))) {
                        const S Y = f.y(tau), Y1 = f.y(tau1);
                        const S ec = clampv(Y, S(0), S(1)), ec1 = clampv(Y1, S(0), S(1));
                        const S d = f.par.h * (ec1 - ec) / m.dt - f.par.qt;
                        S pc, mcp;
                        if (d >= 0) {
                            pc = f.par.ein > 0 ? d / f.par.ein : S(0);
                            mcp = 0;
                        } else {
                            pc = 0;
                            mcp = fmin(-d * f.par.eout, S(1));
                        }
                        const S qadj = S(0) + (f.kst * f.par.h * (Y1 - Y) / m.dt - f.kst * (pc - mcp));
                        k.Q = rv[(size_t)tau * R] + Qs + qadj;
                        k.g = g;
                        k.e = P - fmax(k.Q, S(0));
                        k.P = P;
                        const S sig = kg * w / P;
                        k.ak = sig == S(0) ? S(0) : sig / kg;
                        k.y0 = w;
                        k.valid = 1;
                        if (k.ak != S(0)) k.gc0 = f.gen_c(w, k.gp0);
                    }
                }
                kk_l[e] = k;
            }
        // the lane's window: t over which every one of its interior steps stays in the box
        S llo = -INFINITY, lhi = INFINITY;
        for (int q0 = f.first(f.qi0()), c = 0; q0 <= f.qi1(); q0 += 64 * f.ch, c++) {
            S clo = -INFINITY, chi = INFINITY;
            for (int q = q0; q <= f.qi1() && q < q0 + 64 * f.ch; q += 64) {
                const int tau = f.tau(q);
                const S ye = f.y(tau), ye1 = f.y(f.next(tau));
                if (S(0) <= ye && ye <= S(1) && S(0) <= ye1 && ye1 <= S(1)) {
                    clo = fmax(clo, fmax(-ye, -ye1));
                    chi = fmin(chi, fmin(S(1) - ye, S(1) - ye1));
                } else {
                    clo = INFINITY;
                    chi = -INFINITY;
                }
            }
            cwl[c * 64 + lane] = inward_lo((double)clo);
            cwh[c * 64 + lane] = inward_hi((double)chi);
            llo = fmax(llo, clo);
            lhi = fmin(lhi, chi);
        }
        f.llo = llo;
        f.lhi = lhi;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        double u1, u2;
        uniforms2(a.k0, a.k1, j, (uint32_t)n, 0x54000000u + (uint32_t)i, a.ctr, u1, u2);
        int status = 0;
        S t = sample_stor(f, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status);
        if (t != 0) {
            bool any;
            const double lj = f.logjac(t, any);
            if (any) {
                uint32_t w4[4];
                philox4x32(a.k0, a.k1, (uint32_t)j, (uint32_t)n, 0x74000000u + (uint32_t)i, a.ctr, w4);
                if (!(u01(w4[0], w4[1]) < exp(fmin(0.0, lj)))) {
                    t = 0;
                    status = 2;
                }
            }
        }
        if (lane == 0) {
            if (level) out_level.add(status, t != 0, f.nev);
            else out_long.add(status, t != 0, f.nev);
        }
        if (t != 0) {
            // from the state before the move: the residuals (and the gas at the end steps) ...
            S *rw = st.r + (size_t)n * TR + j;
            S d1, S1;
            for (int e = 0; e < 2; e++) {
                if (!(e == 0 ? f.owns_a() : f.owns_b())) continue;
                const int q = e == 0 ? 0 : f.L, tau = f.tau(q);
                f.end_term(q, t, d1, S1);
                rw[(size_t)tau * R] = S1;
                const SmsKink<S> k = kk_l[e];
                if (k.valid != S(0) && k.ak != S(0)) gv[(size_t)tau * R] = k.y0 + k.ak * FStor<S, 64>::kphi(k.Q, k.g, t);
            }
            for (int q = f.first(f.qi0()); q <= f.qi1(); q += 64) {
                S x;
                if (f.int_term(q, t, false, x, d1, S1)) rw[(size_t)f.tau(q) * R] = S1;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            // ... then the rising instants a+1 … a+L
            for (int q = f.first(1); q <= f.L; q += 64) {
                const int tau = f.tau(q);
                const S nv = f.y(tau) + t;
                if (YL) ys[tau] = nv;
                else yg[(size_t)tau * R] = nv;
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    write_back();
    add_counts(a.st_long, out_long);
    add_counts(a.st_level, out_level);
}

template <typename S>
int launch_storage_long(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats) {
    const int T = h->T, R = h->R, N = h->N;
    if (!(h->sweep_plong > 0) || T < 2) return 0;
    uint32_t w[4];
    philox4x32(k0, k1, 0u, 0u, 9410u, ctr, w);
    // graph-stable form: always launched, the kernel draws whether the sweep runs
    if (!h->graph_stable && !(u01(w[0], w[1]) < h->sweep_plong)) return 0;
    if (N > 256 || h->sweep_long_moves > 256) {
        set_error("long-range storage sweep: at most 256 nodes and 256 plateaus per unit");
        return ESM_EINVAL;
    }
    SlgArgs a{};
    a.k0 = k0;
    a.k1 = k1;
    a.ctr = ctr;
    a.nmoves = h->sweep_long_moves;
    a.pcoin = h->graph_stable ? h->sweep_plong : -1.0;
    bool any = false;
    for (int si = 0; si < ESM_NSTOR; si++)
        for (int n = 0; n < N; n++) {
            const int us = si * N + n, ku = (ESM_KSTOR + si) * N + n;
            const double kst = std::min(std::max(h->k[ku], h->klo[ku]), h->khi[ku]);
            if (h->hours[us] >= h->sweep_long_hours && kst * h->hours[us] / h->dt >= 1.0) {
                a.elig[si][n >> 5] |= 1u << (n & 31);
                any = true;
            }
        }
    if (!any) return 0;
    a.st_long = h->d_stats + (size_t)ESM_FAM_SLONG * ESM_NSTAT * STAT_SLOTS;
    a.st_level = h->d_stats + (size_t)ESM_FAM_SLEVEL * ESM_NSTAT * STAT_SLOTS;
    int lds_max = 0;
    ESM_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    const size_t stat_lds = 3 * sizeof(int) * SLG_MAXM;
    const bool yl = slg_bytes<S>(T, true) + stat_lds <= (size_t)lds_max;
    const size_t lds = slg_bytes<S>(T, yl);
    if (lds + stat_lds > (size_t)lds_max) {
        set_error("long-range storage sweep: the horizon needs more LDS than the device has");
        return ESM_EINVAL;
    }
    const DevModel<S> m = dev_model<S>(h);
    SweepState<S> st{(S *)h->d_v, (S *)h->d_r, (S *)h->d_F, (const S *)h->d_x};
    const dim3 grid((unsigned)(N * R)), blk(64);
    {
        ProfScope p(h, ESM_KERN_SLONG);
        if (yl)
            hipLaunchKernelGGL((k_storage_long<S, true>), grid, blk, lds, h->stream, m, st, a, (S)h->m.beta_v,
                               (S)maxnats, S(1));
        else
            hipLaunchKernelGGL((k_storage_long<S, false>), grid, blk, lds, h->stream, m, st, a, (S)h->m.beta_v,
                               (S)maxnats, S(1));
        ESM_HIP_CHECK(hipGetLastError());
    }
    // then the interval storage pairs (`storage_cross_phase!`, src/storage_long.jl:293)
    return launch_sxlong<S>(h, k0, k1, ctr, maxnats, a.pcoin);
}

template int launch_storage_long<float>(esm_handle *, uint32_t, uint32_t, uint32_t, double);
template int launch_storage_long<double>(esm_handle *, uint32_t, uint32_t, uint32_t, double);

}  // namespace esm
