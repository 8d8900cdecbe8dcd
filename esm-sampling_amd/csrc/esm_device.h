// esm_device.h — device-side model view and term functions of the per-sweep path (gfx950).
//
// The term functions restate the reference's cost terms for the synthetic system (hard line
// limits, no losses, no links, no vintages, no unit commitment, complete level): gen_cost
// (src/local.jl:760-764), soc_cost / soc_power (src/energy.jl:437-463), the line overload
// c_viol Δt (|f| − s)₊ (src/energy.jl:132-176 with ψ = (u − 1)₊) and the balance
// φ(x) = c_shed x₊ + c_dump x₋ (src/local.jl:773).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/esm.h"

namespace esm {

// Offsets of the state planes: plane p of size T*R starts at p*T*R. v holds ESM_NV fields of N
// planes each, x holds ESM_NREN fields of N planes each, r holds N planes, F holds L planes.
__host__ __device__ __forceinline__ size_t plane(int p, int T, int R) { return (size_t)p * T * R; }

// Device view of the static model in the state precision S. Per-unit tables are [carrier][node].
template <typename S>
struct DevModel {
    int N, L, T, R;
    S dt, cv, c_shed, c_dump, kappa0;
    const int *bus0, *bus1;     // [L]
    const S *lb, *ls;           // [L]
    const int *adj_ptr;         // [N+1] CSR of the lines incident to every node
    const int *adj_line;        // [2L]
    const int *adj_other;       // [2L] the node at the other end
    const S *adj_b;             // [2L] susceptance of the line
    const S *load;              // [N][T]
    const S *mc;                // [ESM_NK][N]
    const S *ef;                // [N]
    const S *kc;                // [ESM_NK][N] clamped capacities
    const S *hours, *eta_in, *eta_out, *qt;   // [ESM_NSTOR][N]
};

// XCD-aware block order. The workgroups of a 1-D launch go to the GPU's 8 XCDs (each with its own
// L2) in turn: block b runs on XCD b mod 8. For a launch whose consecutive groups of `grp` blocks
// read the same cache lines (the replicas of one step, neighbouring replicas), the logical block of
// physical block b puts each group on one XCD and deals the groups to the XCDs in order:
// XCD x runs groups x, x + 8, x + 16, … (a bijection when nb is a multiple of 8·grp; else b).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb, unsigned grp) {
    if (grp == 0 || nb % (8u * grp) != 0) return b;
    const unsigned x = b % 8u, k = b / 8u;
    return ((k / grp) * 8u + x) * grp + k % grp;
}

template <typename S>
__device__ __forceinline__ S clampv(S x, S lo, S hi) { return x < lo ? lo : (x > hi ? hi : x); }

// |S| − 1 of a pro-rata kink at node n (`prorata!`, src/local.jl:1784-1802), S the node's
// dispatchable units with |κ| ≥ κmin. The model's one dispatchable unit per node is its gas unit
// (renewables enter as load), so this is 0 at every node that carries a kink.
template <typename S>
__device__ __forceinline__ double kink_nm(const DevModel<S> &m, int n, S kmin) {
    const int nu = fabs(m.kc[ESM_KGAS * m.N + n]) >= kmin ? 1 : 0;
    return (double)(nu - 1);
}

// Neumaier compensated sum.
template <typename S>
struct CSum {
    S s, c;
    __device__ __forceinline__ CSum(S a) : s(a), c(0) {}
    __device__ __forceinline__ void add(S x) {
        S t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ S get() const { return s + c; }
};

// soc_cost of storage carrier s at node n: returns the cost; writes d and the injection.
template <typename S>
__device__ __forceinline__ S soc_cost(const DevModel<S> &m, int s, int n, S k, S w, S ye, S ye1, S &d,
                                      S &inj, S &run) {
    const int u = s * m.N + n;
    const S ec = clampv(ye, S(0), S(1)), ec1 = clampv(ye1, S(0), S(1));
    const S h = m.hours[u];
    d = h * (ec1 - ec) / m.dt - m.qt[u];
    const S eta = m.eta_in[u];
    S pc, mcp;
    if (d >= 0) {
        pc = eta > 0 ? d / eta : S(0);
        mcp = 0;
    } else {
        pc = 0;
        mcp = fmin(-d * m.eta_out[u], S(1));
    }
    run = m.dt * m.mc[(ESM_KSTOR + s) * m.N + n] * k * (pc + mcp);
    S c = run + m.cv * w * h * fabs(ye - ec);
    if (pc > 1) c += m.cv * m.dt * w * (pc - 1);
    if (d < 0 && m.qt[u] <= 0) {
        const S ex = -d * m.eta_out[u] - S(1);
        if (ex > 0) c += m.cv * m.dt * w * ex;
    }
    if (d >= 0 && eta <= 0) c += m.cv * m.dt * w * d;
    inj = k * (pc - mcp);
    return c;
}

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
    return (uint32_t)(((uint64_t)a * b) >> 32);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11); also used by the host for the per-sweep
// schedule draws.
__host__ __device__ __forceinline__ void philox4x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                                                    uint32_t c3, uint32_t out[4]) {
#pragma unroll
    for (int i = 0; i < 10; i++) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = mulhi32(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = mulhi32(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Uniform in [0, 1) from two 32-bit words (53 bits).
__host__ __device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) {
    const uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 11) * 0x1.0p-53;
}

}  // namespace esm
