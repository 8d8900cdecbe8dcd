// k_follow.hip — dispatch following of a capacity change (follow_capacities!,
// src/replicas.jl:448-472) on every replica.
//
// The dispatchable unit (gas) keeps its output: its utilization at every step is scaled by
// c(k0)/c(k1) (set_dispatchable_capacity!, :412-435). Then every renewable and storage unit of
// the node, in unit order (solar, onwind, offwind, battery, H2), moves along its kop curve
// (kop_term / kop_eval / kop_commit!, src/kop.jl:85-178, :254-271): its utilization (renewable
// output share, state-of-charge fractions) stays fixed, so its injection changes by
// j_τ (kc1 − kc0), and the node's gas unit follows share-preservingly,
//     P1 = (Q − Δinj)₊ + e,  Q = r + κ clamp(ω),  e = κω − Q₊,  ω' = P1/κ,
// leaving the node residual unchanged inside the box. A unit below κmin moves its capacity
// alone. Per (node, step, replica) the map touches only that node and step, so one thread runs
// it for all units of the node in order.
#include "esm_internal.h"

namespace esm {

template <typename S>
struct FollowArgs {
    const S *kc0, *kc1;     // [ESM_NK][N] clamped capacities before and after
    const int *moves;       // [ESM_NK][N] 1: kop curve with compensation, 0: capacity alone
    S kmin;
};

template <typename S>
__global__ void __launch_bounds__(256) k_follow(DevModel<S> m, FollowArgs<S> fa, S *__restrict__ v,
                                                const S *__restrict__ x, S *__restrict__ r) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t TR = (size_t)m.T * m.R;
    if (g >= (size_t)m.N * TR) return;
    const int n = (int)(g / TR);
    const size_t o = g - (size_t)n * TR;            // t·R + j
    const int t = (int)(o / m.R);
    const int t1 = t + 1 == m.T ? 0 : t + 1;
    const size_t o1 = (size_t)t1 * m.R + (o - (size_t)t * m.R);
    S *yg = v + plane(m.N, m.T, m.R) + (size_t)n * TR;
    const size_t pn = (size_t)n * TR + o;
    S y = yg[o], rr = r[pn];
    // gas at fixed output: ω ← ω + ω (sc − 1), sc = c(k0)/c(k1)
    const S kg0 = fa.kc0[ESM_KGAS * m.N + n], kg = fa.kc1[ESM_KGAS * m.N + n];
    if (kg0 != kg) {
        const S sc = kg0 / kg;
        const S y1 = y + y * (sc - 1);
        rr -= kg * clampv(y1, S(0), S(1)) - kg0 * clampv(y, S(0), S(1));
        y = y1;
    }
    // renewables and storage along their kop curves
    for (int u = 0; u < ESM_NK; u++) {
        if (u == ESM_KGAS) continue;
        const int q = u * m.N + n;
        const S dkc = fa.kc1[q] - fa.kc0[q];
        S jinj;
        if (u < ESM_NREN) {
            jinj = clampv(x[(size_t)q * TR + o], S(0), S(1));
        } else {
            const int s = u - ESM_KSTOR, us = s * m.N + n;
            const S *ye = v + plane((2 + s) * m.N, m.T, m.R) + (size_t)n * TR;
            const S ec = clampv(ye[o], S(0), S(1)), ec1 = clampv(ye[o1], S(0), S(1));
            const S d = m.hours[us] * (ec1 - ec) / m.dt - m.qt[us];
            S pc, mcp;
            if (d >= 0) {
                pc = m.eta_in[us] > 0 ? d / m.eta_in[us] : S(0);
                mcp = 0;
            } else {
                pc = 0;
                mcp = fmin(-d * m.eta_out[us], S(1));
            }
            jinj = mcp - pc;
        }
        if (dkc == 0) continue;
        const S dinj = jinj * dkc;
        S dS = -dinj;
        const S kap = kg;
        if (fa.moves[q] && fabs(kap) >= fa.kmin) {
            const S Pt = kap * y;
            if (fabs(Pt) >= S(1e-6)) {
                const S Q = rr + kap * clampv(y, S(0), S(1));
                const S e = Pt - fmax(Q, S(0));
                const S P1 = fmax(Q - dinj, S(0)) + e;
                const S sig = kap * y / Pt;
                const S y1 = sig * P1 / kap;
                dS -= kap * (clampv(y1, S(0), S(1)) - clampv(y, S(0), S(1)));
                y = y1;
            }
        }
        rr += dS;
    }
    yg[o] = y;
    r[pn] = rr;
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename S>
int launch_follow(esm_handle *h, const void *d_kc0, const void *d_kc1, const int *d_moves, double kmin) {
    const DevModel<S> m = dev_model<S>(h);
    FollowArgs<S> fa{(const S *)d_kc0, (const S *)d_kc1, d_moves, (S)kmin};
    const size_t n = (size_t)h->N * h->T * h->R;
    ProfScope ps(h, ESM_KERN_FOLLOW);
    hipLaunchKernelGGL((k_follow<S>), dim3(nblk(n, 256)), dim3(256), 0, h->stream, m, fa, (S *)h->d_v,
                       (const S *)h->d_x, (S *)h->d_r);
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

template int launch_follow<float>(esm_handle *, const void *, const void *, const int *, double);
template int launch_follow<double>(esm_handle *, const void *, const void *, const int *, double);

}  // namespace esm
