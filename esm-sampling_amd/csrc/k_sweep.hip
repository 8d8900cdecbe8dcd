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
// Atom families (the complete level of the synthetic system, one dispatchable unit per node):
//   GEN    gas utilization atom at (n, τ)
//   THETA  voltage angle of node n at τ
//   SOC    state of charge of storage s at instant τ
// A round is a set of sites whose terms are disjoint, so every site of a round reads the state
// before the round and the order of sites inside a round does not matter: GEN all (n, τ);
// THETA the nodes of one colour of a distance-2 colouring, every τ; SOC one storage carrier and
// one time-parity class (τ and τ ± 1 share a storage term). After the atoms the sweep runs the
// random plain families (k_random.hip), with probability 1/2 the storage multiscale sweep and
// with probability `storage_long` the long-range storage sweep (k_storage.hip).
//
// Random numbers: per (site, replica) two uniforms from Philox4x32-10 with counter
// (replica, n·T + τ, round id, sweep counter).
#include "sweep_core.h"

namespace esm {

template <typename S>
__global__ void __launch_bounds__(256) k_gen(DevModel<S> m, SweepState<S> st, SweepArgs a, S beta, S maxnats) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nsite = (size_t)m.N * a.ntau;
    MoveOut out;
    if (g < nsite * m.R) {
        const int j = (int)(g % m.R);
        const size_t site = g / m.R;
        const int n = (int)(site / a.ntau);
        const int tau = a.tau0 + (int)(site % a.ntau) * a.tstride;
        double u1, u2;
        uniforms2(a.k0, a.k1, j, (uint32_t)(n * m.T + tau), a.round, a.ctr, u1, u2);
        site_gen(m, st, n, tau, j, u1, u2, beta, maxnats, out);
    }
    add_counts(a.stats, out);
}

template <typename S, int D>
__global__ void __launch_bounds__(256, min_waves<S>(D <= 8 ? 3 : 1)) k_theta(DevModel<S> m, SweepState<S> st, SweepArgs a, S beta,
                                                                S maxnats) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nsite = (size_t)a.nnodes * a.ntau;
    MoveOut out;
    if (g < nsite * m.R) {
        const int j = (int)(g % m.R);
        const size_t site = g / m.R;
        const int n = a.nodes[site / a.ntau];
        const int tau = a.tau0 + (int)(site % a.ntau) * a.tstride;
        double u1, u2;
        uniforms2(a.k0, a.k1, j, (uint32_t)(n * m.T + tau), a.round, a.ctr, u1, u2);
        site_theta<S, D>(m, st, n, tau, j, u1, u2, beta, maxnats, out);
    }
    add_counts(a.stats, out);
}

template <typename S>
__global__ void __launch_bounds__(256, min_waves<S>(4, 2)) k_soc(DevModel<S> m, SweepState<S> st, SweepArgs a, S beta, S maxnats, S kmin) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nsite = (size_t)m.N * a.ntau;
    MoveOut out;
    if (g < nsite * m.R) {
        const int j = (int)(g % m.R);
        const size_t site = g / m.R;
        const int n = (int)(site / a.ntau);
        const int tau = a.tau0 + (int)(site % a.ntau) * a.tstride;
        double u1, u2;
        uniforms2(a.k0, a.k1, j, (uint32_t)(n * m.T + tau), a.round, a.ctr, u1, u2);
        site_soc<S, false>(m, st, a.s, n, tau, j, u1, u2, beta, maxnats, kmin, out);
    }
    add_counts(a.stats, out);
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
        if (h->sweep_atoms) {
            // GEN: every (n, τ)
            a.round = 0;
            a.tau0 = 0;
            a.tstride = 1;
            a.ntau = T;
            a.stats = h->d_stats + ESM_FAM_GEN * ESM_NSTAT * STAT_SLOTS;
            {
                ProfScope pg(h, ESM_KERN_GEN);
                hipLaunchKernelGGL((k_gen<S>), dim3(nblk((size_t)N * T * R, B)), dim3(B), 0, h->stream, m, st, a, beta,
                                   mxn);
            }
            // THETA: colour by colour, degree buckets
            for (size_t rnd = 0; rnd < h->theta_rounds.size(); rnd++) {
                const auto &tr = h->theta_rounds[rnd];
                a.round = 1000u + (uint32_t)tr.colour;
                a.nodes = h->d_theta_nodes + tr.off;
                a.nnodes = tr.count;
                a.stats = h->d_stats + ESM_FAM_THETA * ESM_NSTAT * STAT_SLOTS;
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
            // SOC: per storage carrier and time-parity class
            for (int s = 0; s < ESM_NSTOR; s++) {
                for (size_t c = 0; c < h->tclass.size(); c++) {
                    a.s = s;
                    a.tau0 = h->tclass[c].first;
                    a.tstride = 2;
                    a.ntau = h->tclass[c].second;
                    a.round = 2000u + (uint32_t)(s * 4 + (int)c);
                    a.stats = h->d_stats + ESM_FAM_SOC * ESM_NSTAT * STAT_SLOTS;
                    const unsigned nb = nblk((size_t)N * a.ntau * R, B);
                    ProfScope pf(h, ESM_KERN_SOC);
                    hipLaunchKernelGGL((k_soc<S>), dim3(nb), dim3(B), 0, h->stream, m, st, a, beta, mxn, kmin);
                }
            }
        }
        // the random families, per time-parity class, then (each with its probability) the storage
        // multiscale sweep and the long-range storage sweep (`operation_sweep!`, src/local.jl:2413-2423)
        int rc = launch_random<S>(h, k0, k1, a.ctr, maxnats);
        if (rc) return rc;
        if ((rc = launch_storage_ms<S>(h, k0, k1, a.ctr, maxnats))) return rc;
        if ((rc = launch_storage_long<S>(h, k0, k1, a.ctr, maxnats))) return rc;
    }
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

template int launch_sweeps<float>(esm_handle *, int, uint32_t, uint32_t, uint32_t, double);
template int launch_sweeps<double>(esm_handle *, int, uint32_t, uint32_t, uint32_t, double);

}  // namespace esm
