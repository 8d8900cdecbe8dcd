// k_energy.hip — network energy, flows and residual caches; capacity score; Langevin step.
//
// Layout: every state plane [T][R] is contiguous with the replica fastest, so thread g of a
// launch over T*R elements reads element g of every plane: all loads and stores are coalesced
// for any replica count. Terms are evaluated in the state precision S, per-step and whole-
// horizon totals in FP64 (fixed-order reductions: the results do not depend on scheduling).
#include "esm_internal.h"

namespace esm {

// One thread per (step t, replica j): flows of every line, the residual of every node and the
// per-step cost parts. Restates energy_cost (src/energy.jl:551-677) per step and rebuild_step!
// (src/local.jl:332-374). Per-step parts: 0 gen, 1 storage running cost, 2 shed MWh, 3 dump MWh,
// 4 penalty, 5 co2. drift: part 0 receives max_n |r_cache − r| instead (rebuild_flat).
// STAGE: the thread's N angles are read from memory once into its own LDS column ([n][thread]:
// consecutive threads, consecutive banks), and the lines' ends and the nodes' neighbours read
// them there (each angle is used by 1 + 2·degree terms); no thread reads another's column.
template <typename S, bool COST, bool DRIFT, bool STAGE>
__global__ void __launch_bounds__(256) k_energy(DevModel<S> m, const S *__restrict__ v, const S *__restrict__ x,
                                                S *__restrict__ r, S *__restrict__ F, double *__restrict__ hp) {
    extern __shared__ unsigned char en_lds[];
    const size_t TR = (size_t)m.T * m.R;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= TR) return;
    const int t = (int)(g / m.R);
    const int t1 = (t + 1 == m.T) ? 0 : t + 1;
    const size_t g1 = (size_t)t1 * m.R + (g - (size_t)t * m.R);
    const S *th = v;
    const S *yg = v + plane(m.N, m.T, m.R);
    const S *ye0 = v + plane(2 * m.N, m.T, m.R);
    const S *ye1 = v + plane(3 * m.N, m.T, m.R);
    double gen = 0, stor = 0, shed = 0, dump = 0, pen = 0, co2 = 0, drift = 0;
    const S cvdt = m.cv * m.dt;
    S *col = (S *)en_lds + threadIdx.x;   // the thread's column, stride blockDim.x
    if (STAGE)
        for (int n = 0; n < m.N; n++) col[(size_t)n * blockDim.x] = th[plane(n, m.T, m.R) + g];
    auto theta = [&](int n) -> S { return STAGE ? col[(size_t)n * blockDim.x] : th[plane(n, m.T, m.R) + g]; };
    for (int e = 0; e < m.L; e++) {
        const int a = m.bus0[e], b = m.bus1[e];
        const S f = m.lb[e] * (theta(a) - theta(b));
        F[plane(e, m.T, m.R) + g] = f;
        if (COST) pen += (double)(cvdt * fmax(fabs(f) - m.ls[e], S(0)));
    }
    for (int n = 0; n < m.N; n++) {
        const size_t pn = plane(n, m.T, m.R) + g;
        CSum<S> rn(fmax(m.load[(size_t)n * m.T + t], S(0)));
        const S thn = theta(n);
        for (int q = m.adj_ptr[n]; q < m.adj_ptr[n + 1]; q++) rn.add(m.adj_b[q] * (thn - theta(m.adj_other[q])));
#pragma unroll
        for (int c = 0; c < ESM_NREN; c++) {
            const S A = clampv(x[plane(c * m.N + n, m.T, m.R) + g], S(0), S(1));
            const S p = A * m.kc[c * m.N + n];
            if (COST) gen += (double)(m.dt * (m.mc[c * m.N + n] * p));
            rn.add(-p);
        }
        {
            const S Y = yg[pn], Yc = clampv(Y, S(0), S(1));
            const S kg = m.kc[ESM_KGAS * m.N + n];
            const S p = kg * Yc;
            if (COST) {
                gen += (double)(m.dt * (m.mc[ESM_KGAS * m.N + n] * p));
                co2 += (double)(m.dt * m.ef[n] * p);
                const S ex = fabs(Y - Yc);
                if (ex > 0) pen += (double)(cvdt * (kg + m.kappa0) * ex);
            }
            rn.add(-p);
        }
#pragma unroll
        for (int s = 0; s < ESM_NSTOR; s++) {
            const S *ye = s == 0 ? ye0 : ye1;
            const S kv = m.kc[(ESM_KSTOR + s) * m.N + n];
            S d, inj, run;
            const S c = soc_cost(m, s, n, kv, kv + m.kappa0, ye[pn], ye[plane(n, m.T, m.R) + g1], d, inj, run);
            if (COST) {
                stor += (double)run;
                pen += (double)c - (double)run;
            }
            rn.add(inj);
        }
        const S rv = rn.get();
        if (DRIFT) drift = fmax(drift, (double)fabs(r[pn] - rv));
        r[pn] = rv;
        if (COST) {
            if (rv > 0) shed += (double)(m.dt * rv);
            else dump -= (double)(m.dt * rv);
        }
    }
    if (COST) {
        hp[0 * TR + g] = gen;
        hp[1 * TR + g] = stor;
        hp[2 * TR + g] = shed;
        hp[3 * TR + g] = dump;
        hp[4 * TR + g] = pen;
        hp[5 * TR + g] = co2;
    }
    if (DRIFT) hp[g] = drift;
}

// Fixed-order compensated reduction over steps: partial[c][p][j] = Σ_{t in chunk c} hp[p][t][j].
__global__ void k_reduce_chunks(const double *__restrict__ hp, double *__restrict__ part, int T, int R, int np,
                                int chunk, int nchunk, int op_max) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nchunk * np * R;
    if (g >= total) return;
    const int j = (int)(g % R);
    const int p = (int)((g / R) % np);
    const int c = (int)(g / ((size_t)R * np));
    const size_t TR = (size_t)T * R;
    const int t0 = c * chunk, t1 = min(T, t0 + chunk);
    double s = 0, comp = 0;
    for (int t = t0; t < t1; t++) {
        const double x = hp[p * TR + (size_t)t * R + j];
        if (op_max) {
            s = fmax(s, x);
        } else {
            const double u = s + x;
            comp += fabs(s) >= fabs(x) ? (s - u) + x : (x - u) + s;
            s = u;
        }
    }
    part[g] = op_max ? s : s + comp;
}

// out[j][p] = Σ_c partial[c][p][j] (or max).
__global__ void k_reduce_final(const double *__restrict__ part, double *__restrict__ out, int R, int np, int nchunk,
                               int op_max) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= R * np) return;
    const int j = g / np, p = g % np;
    double s = 0, comp = 0;
    for (int c = 0; c < nchunk; c++) {
        const double x = part[((size_t)c * np + p) * R + j];
        if (op_max) {
            s = fmax(s, x);
        } else {
            const double u = s + x;
            comp += fabs(s) >= fabs(x) ? (s - u) + x : (x - u) + s;
            s = u;
        }
    }
    out[(size_t)j * np + p] = op_max ? s : s + comp;
}

// Capacity score partials: thread (chunk c, node n, replica j) sums ∂C/∂k of the six carriers of
// node n over the steps of chunk c (gen_grad src/energy.jl:835-854, soc_grad :466-502, the
// renewables with Yc = 1). part[c][f][n][j].
template <typename S>
__global__ void __launch_bounds__(256) k_score_part(DevModel<S> m, const S *__restrict__ v, const S *__restrict__ x,
                                                    const S *__restrict__ r, double *__restrict__ part, int chunk,
                                                    int nchunk) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t NR = (size_t)m.N * m.R;
    if (g >= (size_t)nchunk * NR) return;
    const int j = (int)(g % m.R);
    const int n = (int)((g / m.R) % m.N);
    const int c = (int)(g / NR);
    const int t0 = c * chunk, t1 = min(m.T, t0 + chunk);
    const S *yg = v + plane(m.N, m.T, m.R);
    double acc[ESM_NK];
#pragma unroll
    for (int f = 0; f < ESM_NK; f++) acc[f] = 0;
    const S cvdt = m.cv * m.dt;
    S mcr[ESM_NREN];
#pragma unroll
    for (int q = 0; q < ESM_NREN; q++) mcr[q] = m.mc[q * m.N + n];
    const S mcg = m.mc[ESM_KGAS * m.N + n];
    for (int t = t0; t < t1; t++) {
        const size_t o = ((size_t)n * m.T + t) * m.R + j;
        const int tn = (t + 1 == m.T) ? 0 : t + 1;
        const size_t o1 = ((size_t)n * m.T + tn) * m.R + j;
        const S rn = r[o];
        const S Gr = m.dt * (rn > 0 ? m.c_shed : (rn < 0 ? -m.c_dump : S(0)));
#pragma unroll
        for (int q = 0; q < ESM_NREN; q++) {
            const S A = clampv(x[plane(q * m.N, m.T, m.R) + o], S(0), S(1));
            acc[q] += (double)(m.dt * (mcr[q] * A)) - (double)(A * Gr);
        }
        {
            const S Y = yg[o], Yc = clampv(Y, S(0), S(1));
            if (!(S(0) < Y && Y < S(1))) acc[ESM_KGAS] += (double)(cvdt * fabs(Y - Yc));
            acc[ESM_KGAS] += (double)(m.dt * (mcg * Yc)) - (double)(Yc * Gr);
        }
#pragma unroll
        for (int s = 0; s < ESM_NSTOR; s++) {
            const S *ye = v + plane((2 + s) * m.N, m.T, m.R);
            const int u = s * m.N + n;
            const S h = m.hours[u];
            const S y = ye[o];
            const S ec = clampv(y, S(0), S(1)), ec1 = clampv(ye[o1], S(0), S(1));
            const S d = h * (ec1 - ec) / m.dt - m.qt[u];
            const S mcs = m.mc[(ESM_KSTOR + s) * m.N + n];
            double a;
            if (d >= 0 && m.eta_in[u] <= 0) {
                a = (double)(cvdt * d);
            } else if (d >= 0) {
                const S pc = d / m.eta_in[u];
                a = (double)(m.dt * mcs * pc) + (double)(Gr * pc) + (pc > 1 ? (double)(cvdt * (pc - 1)) : 0.0);
            } else {
                const S mq = -d * m.eta_out[u];
                const S mcc = fmin(mq, S(1));
                a = (double)(m.dt * mcs * mcc) - (double)(Gr * mcc) +
                    ((mq > 1 && m.qt[u] <= 0) ? (double)(cvdt * (mq - 1)) : 0.0);
            }
            if (!(S(0) < y && y < S(1))) a += (double)(m.cv * h * fabs(y - ec));
            acc[ESM_KSTOR + s] += a;
        }
    }
#pragma unroll
    for (int f = 0; f < ESM_NK; f++) part[(((size_t)c * ESM_NK + f) * m.N + n) * m.R + j] = acc[f];
}

// sk[j][f][n] = −β_v mask Σ_c part[c][f][n][j] + ∂_k log J; fk[f][n] = β_k (capex + γ_k capex sign(k − kc)).
// With the physical reference measure (nmeas = T, src/energy.jl:811-829) the dispatchable unit
// (gas) and the two storage units of a node add n_u/(k + κ0), n_u = T utilization coordinates
// each, where the capacity is inside its box (mask); renewables keep the utilization measure.
__global__ void k_score_final(const double *__restrict__ part, const double *__restrict__ k, const double *__restrict__ klo,
                              const double *__restrict__ khi, const double *__restrict__ capex, double beta_v,
                              double beta_k, double gamma_k, double kappa0, double nmeas, int N, int R, int nchunk,
                              double *__restrict__ sk, double *__restrict__ fk) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= R * ESM_NK * N) return;
    const int j = g / (ESM_NK * N);
    const int u = g % (ESM_NK * N);
    double s = 0, comp = 0;
    for (int c = 0; c < nchunk; c++) {
        const double x = part[((size_t)c * ESM_NK * N + u) * R + j];
        const double w = s + x;
        comp += fabs(s) >= fabs(x) ? (s - w) + x : (x - w) + s;
        s = w;
    }
    s += comp;
    const double mask = (klo[u] < k[u] && k[u] < khi[u]) ? 1.0 : 0.0;
    const double gj = (u >= ESM_KGAS * N && nmeas > 0) ? mask * (nmeas / (k[u] + kappa0)) : 0.0;
    sk[(size_t)j * ESM_NK * N + u] = -beta_v * (mask * s) + gj;
    if (j == 0) {
        const double kc = fmin(fmax(k[u], klo[u]), khi[u]);
        const double ex = k[u] - kc;
        const double sg = ex > 0 ? 1.0 : (ex < 0 ? -1.0 : 0.0);
        fk[u] = beta_k * (capex[u] + gamma_k * capex[u] * sg);
    }
}

// langevin_update (src/replicas.jl:368-381) of every (field, node) in FP64.
__global__ void k_langevin(const double *__restrict__ k, const double *__restrict__ sk, const double *__restrict__ fk,
                           const double *__restrict__ khi, const double *__restrict__ w, const double *__restrict__ xi,
                           const double *__restrict__ eta, const double *__restrict__ ks, double rr, double kmin,
                           double clip, int box, int N, int R, double *__restrict__ knew, int *__restrict__ nclip) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= ESM_NK * N) return;
    const int f = u / N;
    if (eta[f] == 0) {
        knew[u] = k[u];
        return;
    }
    double sbar = 0;
    for (int j = 0; j < R; j++) sbar += w[j] * sk[(size_t)j * ESM_NK * N + u];
    const double g = -fk[u] + rr * sbar;
    const double kss = ks[f];
    const double w0 = log(k[u] + kss);
    const double dw0 = eta[f] * ((k[u] + kss) * g + 1);
    if (fabs(dw0) > clip) atomicAdd(nclip, 1);
    double w1 = w0 + fmin(fmax(dw0, -clip), clip) + sqrt(2 * eta[f]) * xi[u];
    const double wmin = log(kmin + kss);
    if (w1 < wmin) w1 = 2 * wmin - w1;
    if (box) {
        const double wm = log(fmax(khi[u], kmin + 1.0) + kss);
        if (w1 > wm && w0 <= wm) w1 = fmax(2 * wm - w1, wmin);
    }
    knew[u] = exp(w1) - kss;
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename S>
int launch_energy(esm_handle *h, bool cost, bool drift) {
    const DevModel<S> m = dev_model<S>(h);
    const size_t TR = (size_t)h->T * h->R;
    const S *v = (const S *)h->d_v;
    const S *x = (const S *)h->d_x;
    S *r = (S *)h->d_r;
    S *F = (S *)h->d_F;
    // the angles staged in LDS when the columns of a block of at least 128 threads fit 40 KB (four
    // blocks per CU): N = 40 with 256 threads in FP32, 128 in FP64; not at N = 256
    unsigned B = 256;
    while (B > 128 && (size_t)h->N * B * sizeof(S) > 40960) B /= 2;
    const bool stage = (size_t)h->N * B * sizeof(S) <= 40960;
    if (!stage) B = 256;
    const size_t lds = stage ? (size_t)h->N * B * sizeof(S) : 0;
    ProfScope ps(h, ESM_KERN_ENERGY);
#define ENERGY_LAUNCH(C, D)                                                                                        \
    do {                                                                                                           \
        if (stage)                                                                                                 \
            hipLaunchKernelGGL((k_energy<S, C, D, true>), dim3(nblk(TR, B)), dim3(B), lds, h->stream, m, v, x, r, F, \
                               h->d_hparts);                                                                       \
        else                                                                                                       \
            hipLaunchKernelGGL((k_energy<S, C, D, false>), dim3(nblk(TR, B)), dim3(B), 0, h->stream, m, v, x, r, F, \
                               h->d_hparts);                                                                       \
    } while (0)
    if (cost) ENERGY_LAUNCH(true, false);
    else if (drift) ENERGY_LAUNCH(false, true);
    else ENERGY_LAUNCH(false, false);
#undef ENERGY_LAUNCH
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

static int reduce(esm_handle *h, int np, int op_max, double *d_out) {
    const int chunk = 64;
    const int nchunk = (h->T + chunk - 1) / chunk;
    const size_t n1 = (size_t)nchunk * np * h->R;
    if (n1 > h->part2_len) {
        set_error("reduction scratch too small");
        return ESM_EINVAL;
    }
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(n1, 256)), dim3(256), 0, h->stream, h->d_hparts, h->d_part2, h->T,
                       h->R, np, chunk, nchunk, op_max);
    hipLaunchKernelGGL(k_reduce_final, dim3(nblk((size_t)h->R * np, 256)), dim3(256), 0, h->stream, h->d_part2,
                       d_out, h->R, np, nchunk, op_max);
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_reduce_parts(esm_handle *h, double *d_parts) { return reduce(h, 6, 0, d_parts); }
int launch_reduce_max(esm_handle *h, double *d_drift) { return reduce(h, 1, 1, d_drift); }

template <typename S>
int launch_scores(esm_handle *h) {
    const DevModel<S> m = dev_model<S>(h);
    const int chunk = 64;
    const int nchunk = (h->T + chunk - 1) / chunk;
    const size_t n1 = (size_t)nchunk * h->N * h->R;
    if (n1 * ESM_NK > h->part2_len) {
        set_error("score scratch too small");
        return ESM_EINVAL;
    }
    {
    ProfScope ps(h, ESM_KERN_SCORE);
    hipLaunchKernelGGL((k_score_part<S>), dim3(nblk(n1, 256)), dim3(256), 0, h->stream, m, (const S *)h->d_v,
                       (const S *)h->d_x, (const S *)h->d_r, h->d_part2, chunk, nchunk);
    }
    hipLaunchKernelGGL(k_score_final, dim3(nblk((size_t)h->R * ESM_NK * h->N, 256)), dim3(256), 0, h->stream,
                       h->d_part2, h->d_k, h->d_klo, h->d_khi,
                       h->d_capex, h->m.beta_v, h->m.beta_k, h->m.gamma_k, h->m.kappa0,
                       h->measure == ESM_MEASURE_PHYSICAL ? (double)h->T : 0.0, h->N, h->R, nchunk, h->d_sk,
                       h->d_fk);
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_langevin(esm_handle *h, const double *d_w, const double *d_xi, const double *d_eta, const double *d_ks,
                    double kmin, double clip, int box, double *d_knew, int *d_nclip) {
    const int n = ESM_NK * h->N;
    hipLaunchKernelGGL(k_langevin, dim3(nblk(n, 256)), dim3(256), 0, h->stream, h->d_k, h->d_sk, h->d_fk, h->d_khi,
                       d_w, d_xi, d_eta, d_ks, h->m.beta_k / h->m.beta_v, kmin, clip, box, h->N, h->R, d_knew,
                       d_nclip);
    ESM_HIP_CHECK(hipGetLastError());
    return 0;
}

template int launch_energy<float>(esm_handle *, bool, bool);
template int launch_energy<double>(esm_handle *, bool, bool);
template int launch_scores<float>(esm_handle *);
template int launch_scores<double>(esm_handle *);

}  // namespace esm
