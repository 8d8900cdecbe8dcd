// k_cross.hip — the interval storage pairs `sxlong` of the long-range storage sweep
// (`storage_cross_phase!`, src/storage_cross.jl:501-544), which runs after the long-range plateaus
// whenever the long-range sweep runs.
//
// For every unit of the pairs (hours ≥ hmin) with κ = k h/Δt ≥ 1, `nmoves` interval pairs with a
// partner by `sx_partner`, start a uniform on 0 … T−1 and length L uniform on 1 … T−1; the list
// is shuffled, then grouped by first fit into rounds of disjoint footprints (`sx_rounds`,
// :470-492: the nodes of the pair's template at any step), whose moves commute. The selection
// depends on the capacities and the sweep's counter only, so it is the same in every replica.
//
// MI355X layout: k_sx_schedule (one wave) draws the list and its rounds on the device (no host
// draws, so a captured batch graph does not change with them): lane 0 draws and shuffles, the 64
// lanes test the rounds of a move's first fit in parallel. k_sxlong runs one workgroup of
// SXL_WAVES waves per replica: round by round, wave w takes moves w, w + SXL_WAVES, … of the
// round, one move per wave (cross_move.h); a workgroup barrier
// separates the rounds. The move's lanes hold the items of its end steps' template and its
// interior steps (FXLane). Draws (Philox under the sweep's key and counter c): move m of the unit
// with index u among the units of the pairs (u, m, 9420, c) — partner u01(w0, w1), a = ⌊w2 T⌋,
// L = 1 + ⌊w3 (T − 1)⌋; the shuffle (i, 0, 9421, c); the i-th move of the round-ordered list:
// line draw (replica, i, 0x5C000000, c), Jacobian test (replica, i, 0x7C000000, c).
#include "cross_move.h"

namespace esm {

constexpr int SXL_WAVES = 4;   // waves per replica's workgroup

// layout of the schedule buffer (ints): [0] moves, [1] rounds, then round offsets [cap + 1], the
// round-ordered moves [cap][8], the drawn moves [cap][8], the round of each move [cap], the
// rounds' node marks [cap][8] (bit n of word n / 32)
__host__ __device__ inline size_t sxm_roff(size_t) { return 2; }
__host__ __device__ inline size_t sxm_mv(size_t cap) { return 2 + cap + 1; }
__host__ __device__ inline size_t sxm_tmp(size_t cap) { return sxm_mv(cap) + 8 * cap; }
__host__ __device__ inline size_t sxm_rid(size_t cap) { return sxm_tmp(cap) + 8 * cap; }
__host__ __device__ inline size_t sxm_mark(size_t cap) { return sxm_rid(cap) + cap; }
size_t sxm_ints(size_t cap) { return sxm_mark(cap) + 8 * cap; }

template <typename S>
__global__ void __launch_bounds__(64) k_sx_schedule(DevModel<S> m, SxArgs<S> sx, int nmoves, int cap, uint32_t k0,
                                                    uint32_t k1, uint32_t ctr, double pcoin, int *buf) {
    const int lane = threadIdx.x, N = m.N, T = m.T;
    __shared__ int cnt_s, nr_s;
    int *roff = buf + sxm_roff(cap), *mv = buf + sxm_mv(cap), *tmp = buf + sxm_tmp(cap), *rid = buf + sxm_rid(cap);
    unsigned *mark = (unsigned *)(buf + sxm_mark(cap));
    if (lane == 0) {
        int cnt = 0;
        bool run = true;
        if (pcoin >= 0) {   // graph-stable form: the long-range sweep's own draw
            uint32_t w[4];
            philox4x32(k0, k1, 0u, 0u, 9410u, ctr, w);
            run = u01(w[0], w[1]) < pcoin;
        }
        for (int ui = 0; run && ui < sx.nsx; ui++) {
            const int ua = sx.units[ui], sa = ua / N, na = ua % N;
            if (m.kc[(ESM_KSTOR + sa) * N + na] * m.hours[ua] / m.dt < S(1)) continue;
            for (int mm = 0; mm < nmoves && cnt < cap; mm++) {
                uint32_t w[4];
                philox4x32(k0, k1, (uint32_t)ui, (uint32_t)mm, 9420u, ctr, w);
                int slot, sb, nb;
                if (!sx_partner(m, sx, sa, na, u01(w[0], w[1]), slot, sb, nb)) continue;
                int *o = tmp + 8 * cnt;
                o[0] = sa;
                o[1] = na;
                o[2] = sb;
                o[3] = nb;
                o[4] = slot;
                o[5] = (int)mulhi32(w[2], (uint32_t)T);
                o[6] = 1 + (int)mulhi32(w[3], (uint32_t)(T - 1));
                o[7] = 0;
                cnt++;
            }
        }
        for (int i = cnt - 1; i >= 1; i--) {   // Fisher–Yates
            uint32_t w[4];
            philox4x32(k0, k1, (uint32_t)i, 0u, 9421u, ctr, w);
            const int jx = (int)mulhi32(w[0], (uint32_t)(i + 1));
            for (int q = 0; q < 8; q++) {
                const int x = tmp[8 * i + q];
                tmp[8 * i + q] = tmp[8 * jx + q];
                tmp[8 * jx + q] = x;
            }
        }
        cnt_s = cnt;
        nr_s = 0;
    }
    __syncthreads();
    const int cnt = cnt_s;
    // first fit: the lanes test the existing rounds in parallel, the lowest free one takes the move
    for (int i = 0; i < cnt; i++) {
        const int sl = tmp[8 * i + 4], f0 = sx.tb.ft_off[sl], f1 = sx.tb.ft_off[sl + 1], nr = nr_s;
        unsigned fp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = f0; q < f1; q++) {
            const int n = sx.tb.ft_node[q];
#pragma unroll
            for (int w = 0; w < 8; w++)
                if (n >> 5 == w) fp[w] |= 1u << (n & 31);
        }
        int best = nr;
        for (int r = lane; r < nr; r += 64) {
            bool clash = false;
#pragma unroll
            for (int w = 0; w < 8; w++) clash = clash || (mark[8 * r + w] & fp[w]) != 0u;
            if (!clash) {
                best = r;
                break;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) best = min(best, __shfl_xor(best, off));
        if (lane == 0) {
            if (best == nr) {
#pragma unroll
                for (int w = 0; w < 8; w++) mark[8 * nr + w] = 0u;
                nr_s = nr + 1;
            }
#pragma unroll
            for (int w = 0; w < 8; w++) mark[8 * best + w] |= fp[w];
            rid[i] = best;
        }
        __syncthreads();
    }
    // the moves round by round, each round in the drawn order: a counting sort on the round (the
    // marks are done, their words hold the rounds' cursors)
    if (lane == 0) {
        const int nr = nr_s;
        int *cur = (int *)mark;
        for (int r = 0; r <= nr; r++) cur[r] = 0;
        for (int i = 0; i < cnt; i++) cur[rid[i] + 1]++;
        for (int r = 0; r < nr; r++) cur[r + 1] += cur[r];
        for (int r = 0; r <= nr; r++) roff[r] = cur[r];
        for (int i = 0; i < cnt; i++) {
            const int k = cur[rid[i]]++;
            for (int q = 0; q < 8; q++) mv[8 * k + q] = tmp[8 * i + q];
        }
        buf[0] = cnt;
        buf[1] = nr;
    }
}

template <typename S, int KS>
__global__ void __launch_bounds__(64 * SXL_WAVES) k_sxlong(DevModel<S> m, SweepState<S> st, SxArgs<S> sx,
                                                           const int *buf, int cap, uint32_t k0, uint32_t k1,
                                                           uint32_t ctr, unsigned long long *stats, S beta, S maxnats,
                                                           S kmin) {
    __shared__ XlLds<S, 64, KS> shm[SXL_WAVES];
    __shared__ float cwin[SXL_WAVES][2 * XL_MAXCH * 64];
    // replica j: a contiguous eighth of the replicas per XCD (neighbouring replicas share cache lines)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = (int)xcd_block(blockIdx.x, gridDim.x, gridDim.x % 8 == 0 ? gridDim.x / 8 : 0);
    const int N = m.N, T = m.T, R = m.R;
    const size_t TR = (size_t)T * R;
    const int cnt = buf[0], nr = buf[1];
    const int *roff = buf + sxm_roff(cap), *mv = buf + sxm_mv(cap);
    MoveOut out;
    for (int r = 0; r < nr && cnt > 0; r++) {
        for (int i = roff[r] + wave; i < roff[r + 1]; i += SXL_WAVES) {
            const int *o = mv + 8 * i;
            const int sa = o[0], na = o[1], sb = o[2], nb = o[3], slot = o[4], a = o[5], L = o[6];
            if (!xl_fits<S, 64, KS>(sx.tb, slot)) {   // larger than the lane form: a guard (t = 0)
                if (lane == 0) out.add(2, false, 0);
                continue;
            }
            FXLane<S, true, 64, KS> f;
            if (!xl_prepare(f, m, st, sx.tb, &shm[wave], sa, na, sb, nb, slot, a, L, j, lane, kmin, cwin[wave])) continue;
            double u1, u2;
            uniforms2(k0, k1, j, (uint32_t)i, 0x5C000000u, ctr, u1, u2);
            int status = 0;
            S t = sample_stor_pts<XL_NP>(f, beta, maxnats, unit_open<S>(u1), unit_open<S>(u2), status, shm[wave].pts, lane);
            if (t != 0) {
                bool any;
                const double lj = f.logjac(t, any);
                if (any) {
                    uint32_t w[4];
                    philox4x32(k0, k1, (uint32_t)j, (uint32_t)i, 0x7C000000u, ctr, w);
                    if (!(u01(w[0], w[1]) < exp(fmin(0.0, lj)))) {
                        t = 0;
                        status = 2;
                    }
                }
            }
            if (lane == 0) out.add(status, t != 0, f.nev);
            if (t != 0)
                f.commit(t, st.r + j, st.F + j, st.v + plane(N, T, R) + j, st.v + j,
                         st.v + plane((2 + sa) * N, T, R) + (size_t)na * TR + j,
                         st.v + plane((2 + sb) * N, T, R) + (size_t)nb * TR + j);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        __syncthreads();   // the next round reads this round's states
    }
    add_counts(stats, out);
}

template <typename S>
int launch_sxlong(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats, double pcoin) {
    if (h->sweep_nxlong <= 0 || h->nsx == 0 || h->T < 2 || !h->d_sxm) return 0;
    const int cap = (int)h->sxm_cap;
    const DevModel<S> m = dev_model<S>(h);
    const SxArgs<S> sx = sx_args<S>(h);
    SweepState<S> st{(S *)h->d_v, (S *)h->d_r, (S *)h->d_F, (const S *)h->d_x};
    ProfScope p(h, ESM_KERN_SXLONG);
    hipLaunchKernelGGL((k_sx_schedule<S>), dim3(1), dim3(64), 0, h->stream, m, sx, h->sweep_nxlong, cap, k0, k1, ctr,
                       pcoin, h->d_sxm);
    ESM_HIP_CHECK(hipGetLastError());
    // one kink slot per lane unless some template has more than 32 leak nodes per end; a template
    // beyond two slots (none in the networks built here) makes its moves guards
    int ks = 1;
    if (sx_lane_form(h, 64, ks) == 0) ks = 2;
    unsigned long long *stats = h->d_stats + (size_t)ESM_FAM_SXLONG * ESM_NSTAT * STAT_SLOTS;
    if (ks == 1)
        hipLaunchKernelGGL((k_sxlong<S, 1>), dim3((unsigned)h->R), dim3(64 * SXL_WAVES), 0, h->stream, m, st, sx,
                           (const int *)h->d_sxm, cap, k0, k1, ctr, stats, (S)h->m.beta_v, (S)maxnats, S(1));
    else
        hipLaunchKernelGGL((k_sxlong<S, 2>), dim3((unsigned)h->R), dim3(64 * SXL_WAVES), 0, h->stream, m, st, sx,
                           (const int *)h->d_sxm, cap, k0, k1, ctr, stats, (S)h->m.beta_v, (S)maxnats, S(1));
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

template int launch_sxlong<float>(esm_handle *, uint32_t, uint32_t, uint32_t, double, double);
template int launch_sxlong<double>(esm_handle *, uint32_t, uint32_t, uint32_t, double, double);

}  // namespace esm
