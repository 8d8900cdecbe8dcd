// storage_walk.h — the exact line sampler of the storage moves whose ΔC(t) is a sum over the
// lanes of a segment (one replica per segment of W lanes, W = 1 … 64): segment sums and minima,
// and `walk!` (src/local.jl:1294-1443) with storage secondaries driven through a move functor
// (eval, next_bp, next_root, shift, the recording protocol when CACHED). Included by
// k_storage.hip (multiscale and long-range storage moves) and k_cross.hip (storage pairs).
#pragma once
#include "sweep_core.h"

namespace esm {

constexpr int SMS_PMAX = 4096;     // walk points per side

template <typename S, int W>
__device__ __forceinline__ S seg_sum(S v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// segment minimum in walking order of a candidate (have = false: none)
template <typename S, int W>
__device__ __forceinline__ bool seg_best(bool have, S x, S sgn, S &out) {
    S key = have ? sgn * x : S(INFINITY);
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) key = fmin(key, __shfl_xor(key, off));
    if (!(key < S(INFINITY))) return false;
    out = sgn * key;
    return true;
}

template <class V, typename S>
__device__ __forceinline__ int sms_visit(V &vis, S x, S cx, S &cmin, int &np) {
    if (++np > SMS_PMAX) return -1;
    if (vis(x, cx)) return 2;
    cmin = fmin(cmin, cx);
    return 0;
}

// walk! with storage secondaries: 1 proper, 0 improper, -1 point cap, 2 stopped by the visitor.
// The secondaries at the ends of each interval and piece are the ones recorded by the evaluations
// there (the walk's start records t = 0), so the root searches evaluate nothing. Each pass of the
// outer loop evaluates ΔC at exactly one point per walking segment: the inner loop runs the cheap
// part (breakpoint and root searches, visits of points whose cost is known) until the next point
// to evaluate, so the segments of a wave stay in step at the one evaluation site.
//   BP   next primary breakpoint t (else a ray probe, then T); evaluate t
//   T    evaluate the interval's end t
//   STOR next storage root x on (tp, t): evaluate x, then BAL on (a0, x); none: BAL on (a0, t)
//   BAL  next balance root y on the piece: evaluate and visit y; none: visit the piece's end
template <class F, typename S, class V>
__device__ __forceinline__ int walk_stor(const F &f, S beta, S maxnats, S sgn, V &vis) {
    enum { ST_START, ST_BP, ST_RAY, ST_T, ST_STOR, ST_BAL };
    S cmin = 0, tp = 0, delta = 0, clast = 0, t = 0, c = 0, a0 = 0, xs = 0, x = 0, cx = 0, xb = 0, hp = 0;
    int nx = 0, np = 0, w;
    bool split = false, inx = false;
    int st = ST_START;
    while (true) {
        S qt = 0, r = 0;
        int qm = F::REC_NONE;
        while (true) {
            if (st == ST_START) {
                if (!F::CACHED) {               // nothing to record at t = 0
                    st = ST_BP;
                    continue;
                }
                qm = F::REC_A;
                break;
            }
            if (st == ST_BP) {
                if (f.next_bp(tp, sgn, t)) {
                    st = ST_T;
                } else {
                    hp = S(1e-6) * fmax(S(1), fabs(tp));
                    qt = tp + sgn * hp;
                    st = ST_RAY;
                    break;
                }
            }
            if (st == ST_T) {
                qt = t;
                qm = F::REC_B;
                break;
            }
            const bool stor = st == ST_STOR;
            if (f.next_root(stor, stor ? tp : a0, stor ? t : (inx ? x : t), stor ? xs : xb, sgn, r)) {
                qt = r;
                qm = stor ? F::REC_X : F::REC_NONE;
                break;
            }
            if (stor) {
                inx = false;
                xb = a0;
                st = ST_BAL;
                if (F::CACHED && split) {   // the balance at t, overwritten by the storage roots
                    qt = t;
                    qm = F::REC_T;
                    break;
                }
                continue;
            }
            // the piece is done: visit its end (a storage root or the interval's end)
            if ((w = sms_visit(vis, inx ? x : t, inx ? cx : c, cmin, np))) return w;
            if (inx) {
                f.shift(false);
                a0 = x;
                xs = x;
                st = ST_STOR;
            } else {
                clast = c;
                if (c - cmin > maxnats) return 1;
                tp = t;
                f.shift(true);
                st = ST_BP;
            }
        }
        const S val = f.eval(qt, qm);
        if (st == ST_START) {
            st = ST_BP;
        } else if (st == ST_RAY) {
            const S sl = (beta * val - clast) / hp;
            if (sl > 0) delta = fmin(2 * maxnats / sl, S(1e3) * fmax(fabs(tp), S(1e-3)));
            else delta = delta == 0 ? fmax(fabs(tp), S(1e-6)) : 2 * delta;
            t = tp + sgn * delta;
            if (++nx > 80) return 0;
            if (!isfinite(t)) return 0;
            st = ST_T;
        } else if (st == ST_T) {
            c = beta * val;
            a0 = tp;
            xs = tp;
            split = false;
            st = ST_STOR;
        } else if (st == ST_STOR) {             // a storage root
            x = r;
            cx = beta * val;
            split = true;
            inx = true;
            xb = a0;
            st = ST_BAL;
        } else if (qm == F::REC_NONE) {         // ST_BAL: a balance root
            if ((w = sms_visit(vis, r, beta * val, cmin, np))) return w;
            xb = r;
        }                                       // ST_BAL after REC_T: search the last piece
    }
}

template <typename S>
struct VisS {
    int mode = 0;
    MassAcc<S> ma;
    SegSel<S> ss;
    __device__ bool operator()(S t, S c) { return mode == 0 ? ma(t, c) : ss(t, c); }
};

template <class F, typename S>
__device__ __forceinline__ S sample_stor(const F &f, S beta, S maxnats, S useg, S uoff, int &status) {
    VisS<S> vis;
    S Mx0 = S(-INFINITY), Mx1 = S(-INFINITY), Ms0 = 0, Ms1 = 0;   // the right (0) and left (1) walks
#pragma nounroll
    for (int k = 0; k < 3; k++) {
        S sgn;
        if (k == 2) {
            const S MX = fmax(Mx0, Mx1);
            const S Mr = Ms0 * exp(Mx0 - MX), Ml = Ms1 * exp(Mx1 - MX);
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
        const int w = walk_stor(f, beta, maxnats, sgn, vis);
        if (k < 2) {
            if (w != 1) {
                status = w == 0 ? 1 : 2;
                return 0;
            }
            if (k == 0) {
                Mx0 = vis.ma.mx;
                Ms0 = vis.ma.s;
            } else {
                Mx1 = vis.ma.mx;
                Ms1 = vis.ma.s;
            }
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

// The same with the points of the two mass walks kept (t, c in walk order, up to NP per side, in
// the segment's LDS `pts` [2][NP][2]): the selection then reads the chosen side's points instead of
// walking it again, and walks again only when that side had more than NP points. Same draws as
// sample_stor: the selection sees the same points in the same order.
template <typename S>
struct VisSP {
    VisS<S> v;
    S *pts = nullptr;   // this side's [NP][2]
    int np = 0, cap = 0, sub = 0;
    bool rec = false;
    __device__ __forceinline__ bool operator()(S t, S c) {
        if (rec) {
            if (np < cap && sub == 0) {
                pts[2 * np] = t;
                pts[2 * np + 1] = c;
            }
            np++;
        }
        return v(t, c);
    }
};

template <int NP, class F, typename S>
__device__ __forceinline__ S sample_stor_pts(const F &f, S beta, S maxnats, S useg, S uoff, int &status, S *pts,
                                             int sub) {
    VisSP<S> vis;
    vis.cap = NP;
    vis.sub = sub;
    S Mx0 = S(-INFINITY), Mx1 = S(-INFINITY), Ms0 = 0, Ms1 = 0;
    int n0 = 0, n1 = 0;
#pragma nounroll
    for (int k = 0; k < 2; k++) {
        vis.v.mode = 0;
        vis.v.ma = MassAcc<S>();
        vis.rec = true;
        vis.np = 0;
        vis.pts = pts + 2 * NP * k;
        const int w = walk_stor(f, beta, maxnats, k == 0 ? S(1) : S(-1), vis);
        if (w != 1) {
            status = w == 0 ? 1 : 2;
            return 0;
        }
        if (k == 0) {
            Mx0 = vis.v.ma.mx;
            Ms0 = vis.v.ma.s;
            n0 = vis.np;
        } else {
            Mx1 = vis.v.ma.mx;
            Ms1 = vis.v.ma.s;
            n1 = vis.np;
        }
    }
    const S MX = fmax(Mx0, Mx1);
    const S Mr = Ms0 * exp(Mx0 - MX), Ml = Ms1 * exp(Mx1 - MX);
    const S u = useg * (Ml + Mr);
    SegSel<S> ss;
    ss.MX = MX;
    ss.left = u <= Ml;
    ss.target = ss.left ? Ml - u : u - Ml;
    const int cnt = ss.left ? n1 : n0;
    if (cnt <= NP) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const S *ps = pts + 2 * NP * (ss.left ? 1 : 0);
        for (int i = 0; i < cnt; i++)
            if (ss(ps[2 * i], ps[2 * i + 1])) break;
        __builtin_amdgcn_wave_barrier();   // the next move's walks write the points again
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        vis.rec = false;
        vis.v.mode = 1;
        vis.v.ss = ss;
        walk_stor(f, beta, maxnats, ss.left ? S(-1) : S(1), vis);
        ss = vis.v.ss;
    }
    status = 0;
    const S L = ss.tb - ss.ta;
    const S g = (ss.cb - ss.ca) / L;
    const S t = (g >= 0 || fabs(g * L) < S(1e-10)) ? ss.ta + trexp_offset(g, L, uoff) : ss.tb - trexp_offset(-g, L, uoff);
    if (!isfinite(t)) {
        status = 2;
        return 0;
    }
    return t;
}

}  // namespace esm
