// esm_api.hip — the C ABI of libesmhip.so (include/esm.h): handle, state transfer, energy,
// scores and the Langevin step. Host code; the kernels live in the k_*.hip files.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "esm_internal.h"

namespace esm {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// The staging of the allocation-free Langevin step (esm_langevin_noalloc): weights [R], normals
// [6N], steps [6] and offsets [6], the new capacities [6N] and the clip counter, in this fixed
// device array instead of a per-call allocation. Every device holds its own copy at its own
// address, looked up once per device (esm_create, with the device current) and kept in
// g_scratch; host code uses that address, never the array's name. Calls on one device share the
// array and keep the last call's contents: each call clears the counter, and calls that overlap
// on one device race.
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[ESM_NOALLOC_DOUBLES];
static double *g_scratch[SCRATCH_DEVICES] = {};
static std::mutex g_scratch_mu;

static int scratch_init(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) {
        set_error("device index beyond the allocation-free scratch table");
        return ESM_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return ESM_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress of the allocation-free scratch: ") + hipGetErrorString(e));
        return ESM_EHIP;
    }
    g_scratch[device] = (double *)p;
    return ESM_OK;
}

__global__ void k_zero_i32(int *p) { *p = 0; }

template <typename S>
static void *upload_conv(const std::vector<double> &src, hipStream_t st, int *rc) {
    void *p = nullptr;
    if (src.empty()) return nullptr;
    if (hipMalloc(&p, src.size() * sizeof(S)) != hipSuccess) {
        *rc = ESM_ENOMEM;
        return nullptr;
    }
    std::vector<S> tmp(src.begin(), src.end());
    if (hipMemcpyAsync(p, tmp.data(), tmp.size() * sizeof(S), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        *rc = ESM_EHIP;
    return p;
}

template <typename S>
DevModel<S> dev_model(const esm_handle *h) {
    DevModel<S> m;
    m.N = h->N;
    m.L = h->L;
    m.T = h->T;
    m.R = h->R;
    m.dt = (S)h->dt;
    m.cv = (S)h->m.c_viol;
    m.c_shed = (S)h->m.c_shed;
    m.c_dump = (S)h->m.c_dump;
    m.kappa0 = (S)h->m.kappa0;
    m.bus0 = h->d_bus0;
    m.bus1 = h->d_bus1;
    m.lb = (const S *)h->mb.lb;
    m.ls = (const S *)h->mb.ls;
    m.adj_ptr = h->d_adj_ptr;
    m.adj_line = h->d_adj_line;
    m.adj_other = h->d_adj_other;
    m.adj_b = (const S *)h->mb.adj_b;
    m.load = (const S *)h->mb.load;
    m.mc = (const S *)h->mb.mc;
    m.ef = (const S *)h->mb.ef;
    m.kc = (const S *)h->mb.kc;
    m.hours = (const S *)h->mb.hours;
    m.eta_in = (const S *)h->mb.eta_in;
    m.eta_out = (const S *)h->mb.eta_out;
    m.qt = (const S *)h->mb.qt;
    return m;
}
template DevModel<float> dev_model<float>(const esm_handle *);
template DevModel<double> dev_model<double>(const esm_handle *);

// The clamped capacities in the state precision (re-uploaded whenever k changes).
static int upload_k(esm_handle *h) {
    const int n = ESM_NK * h->N;
    std::vector<double> kc(n);
    for (int u = 0; u < n; u++) kc[u] = std::min(std::max(h->k[u], h->klo[u]), h->khi[u]);
    ESM_HIP_CHECK(hipMemcpyAsync(h->d_k, h->k.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (h->prec == ESM_FP64) {
        ESM_HIP_CHECK(hipMemcpyAsync(h->mb.kc, kc.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    } else {
        std::vector<float> kf(kc.begin(), kc.end());
        ESM_HIP_CHECK(hipMemcpyAsync(h->mb.kc, kf.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    h->scores_valid = false;
    return 0;
}

// Convert a host FP64 array to the state precision and copy it to the device.
static int put_state(esm_handle *h, void *dst, const double *src, size_t n) {
    if (h->prec == ESM_FP64) {
        ESM_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        return 0;
    }
    const size_t chunk = (size_t)1 << 24;
    std::vector<float> tmp(std::min(n, chunk));
    for (size_t o = 0; o < n; o += chunk) {
        const size_t c = std::min(chunk, n - o);
        for (size_t i = 0; i < c; i++) tmp[i] = (float)src[o + i];
        ESM_HIP_CHECK(hipMemcpyAsync((float *)dst + o, tmp.data(), c * sizeof(float), hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

static int get_state(esm_handle *h, double *dst, const void *src, size_t n) {
    if (h->prec == ESM_FP64) {
        ESM_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        return 0;
    }
    const size_t chunk = (size_t)1 << 24;
    std::vector<float> tmp(std::min(n, chunk));
    for (size_t o = 0; o < n; o += chunk) {
        const size_t c = std::min(chunk, n - o);
        ESM_HIP_CHECK(hipMemcpyAsync(tmp.data(), (const float *)src + o, c * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        for (size_t i = 0; i < c; i++) dst[o + i] = tmp[i];
    }
    return 0;
}

static int rebuild(esm_handle *h) {
    return h->prec == ESM_FP64 ? launch_energy<double>(h, false, false) : launch_energy<float>(h, false, false);
}

// forget the recorded intervals (their events return to the pool); with `destroy`, free the pool
static void clear_prof(esm_handle *h, bool destroy = false) {
    for (int i = 0; i < ESM_NKERN; i++) h->prof_ev[i].clear();
    h->prof_next = 0;
    if (destroy) {
        for (hipEvent_t e : h->prof_pool) (void)hipEventDestroy(e);
        h->prof_pool.clear();
    }
}

// Copy the move templates to the device: the integer arrays in one buffer, the values in the
// state precision in another.
template <typename S>
static int upload_move_tables(esm_handle *h) {
    const MoveTablesHost &mt = h->mt;
    MoveTablesDev &d = h->dmt;
    const std::vector<const std::vector<int> *> iv = {&mt.th_off, &mt.th_node, &mt.en_off, &mt.en_node, &mt.en_div,
                                                      &mt.kk_off, &mt.kk_node, &mt.ln_off, &mt.ln_line, &mt.nd_off,
                                                      &mt.nd_node, &mt.nd_ent, &mt.nd_kk, &mt.nd_l, &mt.nl_pos,
                                                      &mt.nl_sgn, &mt.xp_off, &mt.lev_off, &mt.sms_units,
                                                      &mt.grp_off, &mt.grp_node, &mt.grp_si, &mt.sx_poff, &mt.sx_part,
                                                      &mt.sx_th_off, &mt.sx_th_node, &mt.sx_kk_off, &mt.sx_kk_node,
                                                      &mt.sx_ln_off, &mt.sx_ln_line, &mt.sx_nd_off, &mt.sx_nd_node,
                                                      &mt.sx_nd_kk, &mt.sx_nd_l, &mt.sx_nl_pos, &mt.sx_nl_sgn,
                                                      &mt.sx_ft_off, &mt.sx_ft_node};
    std::vector<int> ib;
    std::vector<size_t> io;
    for (auto *v : iv) {
        io.push_back(ib.size());
        ib.insert(ib.end(), v->begin(), v->end());
        ib.push_back(0);   // keeps every array non-empty
    }
    const std::vector<const std::vector<double> *> vv = {&mt.th_val,    &mt.en_c,    &mt.kk_g,     &mt.ln_df,
                                                         &mt.sx_th_val, &mt.sx_kk_g, &mt.sx_ln_ta, &mt.sx_ln_tb};
    std::vector<S> vb;
    std::vector<size_t> vo;
    for (auto *v : vv) {
        vo.push_back(vb.size());
        for (double x : *v) vb.push_back((S)x);
        vb.push_back(S(0));
    }
    ESM_HIP_CHECK(hipMalloc((void **)&d.ibuf, ib.size() * sizeof(int)));
    ESM_HIP_CHECK(hipMemcpy(d.ibuf, ib.data(), ib.size() * sizeof(int), hipMemcpyHostToDevice));
    ESM_HIP_CHECK(hipMalloc(&d.vbuf, vb.size() * sizeof(S)));
    ESM_HIP_CHECK(hipMemcpy(d.vbuf, vb.data(), vb.size() * sizeof(S), hipMemcpyHostToDevice));
    ESM_HIP_CHECK(hipMalloc((void **)&d.qcum, mt.qcum.size() * sizeof(double)));
    ESM_HIP_CHECK(hipMemcpy(d.qcum, mt.qcum.data(), mt.qcum.size() * sizeof(double), hipMemcpyHostToDevice));
    int **dst[] = {&d.th_off, &d.th_node, &d.en_off, &d.en_node, &d.en_div, &d.kk_off, &d.kk_node,
                               &d.ln_off, &d.ln_line, &d.nd_off, &d.nd_node, &d.nd_ent, &d.nd_kk, &d.nd_l,
                               &d.nl_pos, &d.nl_sgn, &d.xp_off, &d.lev_off, &d.sms_units, &d.grp_off,
                               &d.grp_node, &d.grp_si, &d.sx_poff, &d.sx_part, &d.sx_th_off, &d.sx_th_node,
                               &d.sx_kk_off, &d.sx_kk_node, &d.sx_ln_off, &d.sx_ln_line, &d.sx_nd_off, &d.sx_nd_node,
                               &d.sx_nd_kk, &d.sx_nd_l, &d.sx_nl_pos, &d.sx_nl_sgn, &d.sx_ft_off, &d.sx_ft_node};
    for (size_t q = 0; q < iv.size(); q++) *dst[q] = d.ibuf + io[q];
    const S *vbase = (const S *)d.vbuf;
    d.th_val = vbase + vo[0];
    d.en_c = vbase + vo[1];
    d.kk_g = vbase + vo[2];
    d.ln_df = vbase + vo[3];
    d.sx_th_val = vbase + vo[4];
    d.sx_kk_g = vbase + vo[5];
    d.sx_ln_ta = vbase + vo[6];
    d.sx_ln_tb = vbase + vo[7];
    d.xp0 = mt.xp0;
    d.cl0 = mt.cl0;
    d.nlev = mt.nlev;
    return 0;
}

static void free_handle(esm_handle *h) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    clear_prof(h, true);
    mlp_free(h);
    void *ps[] = {h->d_bus0, h->d_bus1, h->d_adj_ptr, h->d_adj_line, h->d_adj_other, h->mb.lb, h->mb.ls, h->mb.adj_b,
                  h->mb.load, h->mb.mc, h->mb.ef, h->mb.kc, h->mb.hours, h->mb.eta_in, h->mb.eta_out, h->mb.qt,
                  h->d_k, h->d_klo, h->d_khi, h->d_capex, h->d_v, h->d_x, h->d_r, h->d_F, h->d_hparts, h->d_part2,
                  h->d_out, h->d_sk, h->d_fk, h->d_z, h->d_ytmp, h->d_carry, h->d_zin, h->d_mu, h->d_xbar, h->d_m,
                  h->d_mnew, h->d_hod, h->d_moved, h->d_harm, h->d_wpar, h->d_Ptr, h->d_pi0, h->d_theta_nodes, h->d_stats,
                  h->d_counts, h->dmt.ibuf, h->dmt.vbuf, h->dmt.qcum, h->d_mu64, h->d_ffbs, h->d_ffbs_mask, h->d_sched,
                  h->d_sxu, h->d_sxm, h->d_feat};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // namespace esm

using namespace esm;

extern "C" {

int32_t esm_version(void) { return ESM_ABI_VERSION; }
const char *esm_last_error(void) { return g_err.c_str(); }

int32_t esm_create(const esm_model *mdl, int32_t nrep, int32_t precision, esm_handle **out) {
    if (!mdl || !out) {
        set_error("esm_create: null model or output pointer");
        return ESM_EINVAL;
    }
    *out = nullptr;
    if (precision != ESM_FP64 && precision != ESM_FP32) {
        set_error("esm_create: precision must be ESM_FP64 (64) or ESM_FP32 (32)");
        return ESM_EINVAL;
    }
    if (mdl->N < 2 || mdl->L < 1 || mdl->T < 2 || nrep < 1 || !(mdl->dt > 0)) {
        set_error("esm_create: need N >= 2, L >= 1, T >= 2, nrep >= 1 and dt > 0");
        return ESM_EINVAL;
    }
    const double *req[] = {mdl->line_b, mdl->line_s, mdl->load, mdl->mc, mdl->ef, mdl->capex, mdl->klo, mdl->khi,
                           mdl->stor_hours, mdl->stor_eta_in, mdl->stor_eta_out, mdl->stor_qt};
    for (const double *p : req)
        if (!p || !mdl->line_bus0 || !mdl->line_bus1) {
            set_error("esm_create: a model table is NULL");
            return ESM_EINVAL;
        }
    const int N = mdl->N, L = mdl->L, T = mdl->T;
    for (int e = 0; e < L; e++) {
        const int a = mdl->line_bus0[e], b = mdl->line_bus1[e];
        if (a < 0 || a >= N || b < 0 || b >= N || a == b) {
            set_error("esm_create: line " + std::to_string(e) + " has an invalid end");
            return ESM_EINVAL;
        }
    }
    if (mdl->ncountry < 0 || (mdl->ncountry > 0 && !mdl->country)) {
        set_error("esm_create: ncountry > 0 needs the country labels");
        return ESM_EINVAL;
    }
    for (int n = 0; mdl->ncountry > 0 && n < N; n++)
        if (mdl->country[n] < 0 || mdl->country[n] >= mdl->ncountry) {
            set_error("esm_create: country label of node " + std::to_string(n) + " out of range");
            return ESM_EINVAL;
        }
    esm_handle *h = new esm_handle();
    h->N = N;
    h->L = L;
    h->T = T;
    h->R = nrep;
    h->prec = precision;
    h->es = precision == ESM_FP64 ? 8 : 4;
    h->dt = mdl->dt;
    h->m = *mdl;
    auto cp = [](const double *p, size_t n) { return std::vector<double>(p, p + n); };
    h->bus0.assign(mdl->line_bus0, mdl->line_bus0 + L);
    h->bus1.assign(mdl->line_bus1, mdl->line_bus1 + L);
    h->line_b = cp(mdl->line_b, L);
    h->line_s = cp(mdl->line_s, L);
    h->load = cp(mdl->load, (size_t)N * T);
    h->mc = cp(mdl->mc, ESM_NK * N);
    h->ef = cp(mdl->ef, N);
    h->capex = cp(mdl->capex, ESM_NK * N);
    h->klo = cp(mdl->klo, ESM_NK * N);
    h->khi = cp(mdl->khi, ESM_NK * N);
    h->hours = cp(mdl->stor_hours, ESM_NSTOR * N);
    h->eta_in = cp(mdl->stor_eta_in, ESM_NSTOR * N);
    h->eta_out = cp(mdl->stor_eta_out, ESM_NSTOR * N);
    h->qt = cp(mdl->stor_qt, ESM_NSTOR * N);
    h->k.assign(ESM_NK * N, 0.0);
    h->ncountry = mdl->ncountry;
    if (mdl->ncountry > 0) h->country.assign(mdl->country, mdl->country + N);
    // CSR adjacency: lines incident to every node, in line order
    h->adj_ptr.assign(N + 1, 0);
    for (int e = 0; e < L; e++) {
        h->adj_ptr[h->bus0[e] + 1]++;
        h->adj_ptr[h->bus1[e] + 1]++;
    }
    for (int n = 0; n < N; n++) h->adj_ptr[n + 1] += h->adj_ptr[n];
    h->adj_line.assign(2 * L, 0);
    h->adj_other.assign(2 * L, 0);
    h->adj_bv.assign(2 * L, 0.0);
    {
        std::vector<int> fill(h->adj_ptr.begin(), h->adj_ptr.end() - 1);
        for (int e = 0; e < L; e++) {
            int a = h->bus0[e], b = h->bus1[e];
            int qa = fill[a]++, qb = fill[b]++;
            h->adj_line[qa] = e;
            h->adj_other[qa] = b;
            h->adj_bv[qa] = h->line_b[e];
            h->adj_line[qb] = e;
            h->adj_other[qb] = a;
            h->adj_bv[qb] = h->line_b[e];
        }
    }
    // sweep schedule: distance-2 colouring of the nodes (greedy in node order: two nodes of one
    // colour share no line and no neighbour, so their angle moves touch disjoint terms), degree
    // buckets of the angle kernels, and the time-parity classes of the storage moves
    // (src/local.jl:1948-1950: odd and even steps, three classes when T is odd)
    {
        std::vector<std::vector<int>> nbr(N);
        for (int e = 0; e < L; e++) {
            nbr[h->bus0[e]].push_back(h->bus1[e]);
            nbr[h->bus1[e]].push_back(h->bus0[e]);
        }
        h->colour.assign(N, -1);
        std::vector<char> used;
        for (int n = 0; n < N; n++) {
            used.assign(N + 1, 0);
            for (int a : nbr[n]) {
                if (h->colour[a] >= 0) used[h->colour[a]] = 1;
                for (int b : nbr[a])
                    if (b != n && h->colour[b] >= 0) used[h->colour[b]] = 1;
            }
            int c = 0;
            while (used[c]) c++;
            h->colour[n] = c;
            h->ncolour = std::max(h->ncolour, c + 1);
        }
        std::vector<int> flat;
        const int buckets[3] = {4, 8, 16};
        for (int n = 0; n < N; n++)
            if ((int)nbr[n].size() > 16) {
                set_error("esm_create: node " + std::to_string(n) + " has more than 16 lines");
                delete h;
                return ESM_EINVAL;
            }
        for (int c = 0; c < h->ncolour; c++)
            for (int bi = 0; bi < 3; bi++) {
                const int lo = bi == 0 ? 0 : buckets[bi - 1] + 1, hi = buckets[bi];
                esm_handle::ThetaRound tr{c, hi, (int)flat.size(), 0};
                for (int n = 0; n < N; n++) {
                    const int d = (int)nbr[n].size();
                    if (h->colour[n] == c && d >= lo && d <= hi && d > 0) {
                        flat.push_back(n);
                        tr.count++;
                    }
                }
                if (tr.count) h->theta_rounds.push_back(tr);
            }
        h->tclass.clear();
        if (T % 2 == 0) {
            h->tclass.push_back({0, T / 2});
            h->tclass.push_back({1, T / 2});
        } else {
            h->tclass.push_back({0, (T - 1) / 2});
            h->tclass.push_back({1, (T - 1) / 2});
            h->tclass.push_back({T - 1, 1});
        }
        h->theta_nodes_h = flat;
    }
    {
        std::string err;
        if (!build_move_tables(h, h->mt, err)) {
            set_error("esm_create: " + err);
            delete h;
            return ESM_EINVAL;
        }
    }
    int rc = 0;
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
        set_error("esm_create: no HIP device or stream");
        free_handle(h);
        return ESM_EHIP;
    }
    if ((rc = scratch_init(h->device))) {   // the device is current: its copy of the scratch array
        free_handle(h);
        return rc;
    }
    auto imalloc = [&](int **p, const std::vector<int> &src) {
        if (rc) return;
        if (hipMalloc((void **)p, src.size() * sizeof(int)) != hipSuccess) { rc = ESM_ENOMEM; return; }
        if (hipMemcpy(*p, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) rc = ESM_EHIP;
    };
    auto dmalloc = [&](double **p, const std::vector<double> &src) {
        if (rc) return;
        if (hipMalloc((void **)p, src.size() * sizeof(double)) != hipSuccess) { rc = ESM_ENOMEM; return; }
        if (hipMemcpy(*p, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = ESM_EHIP;
    };
    imalloc(&h->d_bus0, std::vector<int>(h->bus0.begin(), h->bus0.end()));
    imalloc(&h->d_bus1, std::vector<int>(h->bus1.begin(), h->bus1.end()));
    imalloc(&h->d_adj_ptr, h->adj_ptr);
    imalloc(&h->d_adj_line, h->adj_line);
    imalloc(&h->d_adj_other, h->adj_other);
    if (!h->theta_nodes_h.empty()) imalloc(&h->d_theta_nodes, h->theta_nodes_h);
    dmalloc(&h->d_k, h->k);
    dmalloc(&h->d_klo, h->klo);
    dmalloc(&h->d_khi, h->khi);
    dmalloc(&h->d_capex, h->capex);
    std::vector<double> kc0(ESM_NK * N, 0.0);
    for (int u = 0; u < ESM_NK * N; u++) kc0[u] = std::min(std::max(0.0, h->klo[u]), h->khi[u]);
    if (!rc) {
        if (precision == ESM_FP64) {
            h->mb.lb = upload_conv<double>(h->line_b, h->stream, &rc);
            h->mb.ls = upload_conv<double>(h->line_s, h->stream, &rc);
            h->mb.adj_b = upload_conv<double>(h->adj_bv, h->stream, &rc);
            h->mb.load = upload_conv<double>(h->load, h->stream, &rc);
            h->mb.mc = upload_conv<double>(h->mc, h->stream, &rc);
            h->mb.ef = upload_conv<double>(h->ef, h->stream, &rc);
            h->mb.kc = upload_conv<double>(kc0, h->stream, &rc);
            h->mb.hours = upload_conv<double>(h->hours, h->stream, &rc);
            h->mb.eta_in = upload_conv<double>(h->eta_in, h->stream, &rc);
            h->mb.eta_out = upload_conv<double>(h->eta_out, h->stream, &rc);
            h->mb.qt = upload_conv<double>(h->qt, h->stream, &rc);
        } else {
            h->mb.lb = upload_conv<float>(h->line_b, h->stream, &rc);
            h->mb.ls = upload_conv<float>(h->line_s, h->stream, &rc);
            h->mb.adj_b = upload_conv<float>(h->adj_bv, h->stream, &rc);
            h->mb.load = upload_conv<float>(h->load, h->stream, &rc);
            h->mb.mc = upload_conv<float>(h->mc, h->stream, &rc);
            h->mb.ef = upload_conv<float>(h->ef, h->stream, &rc);
            h->mb.kc = upload_conv<float>(kc0, h->stream, &rc);
            h->mb.hours = upload_conv<float>(h->hours, h->stream, &rc);
            h->mb.eta_in = upload_conv<float>(h->eta_in, h->stream, &rc);
            h->mb.eta_out = upload_conv<float>(h->eta_out, h->stream, &rc);
            h->mb.qt = upload_conv<float>(h->qt, h->stream, &rc);
        }
    }
    const size_t TR = (size_t)T * nrep;
    auto zmalloc = [&](void **p, size_t bytes) {
        if (rc) return;
        if (hipMalloc(p, bytes) != hipSuccess) { rc = ESM_ENOMEM; return; }
        if (hipMemset(*p, 0, bytes) != hipSuccess) rc = ESM_EHIP;
    };
    zmalloc(&h->d_v, (size_t)ESM_NV * N * TR * h->es);
    zmalloc(&h->d_x, (size_t)ESM_NREN * N * TR * h->es);
    zmalloc(&h->d_r, (size_t)N * TR * h->es);
    zmalloc(&h->d_F, (size_t)L * TR * h->es);
    zmalloc((void **)&h->d_hparts, 6 * TR * sizeof(double));
    const int nchunk = (T + 63) / 64;
    h->part2_len = (size_t)nchunk * std::max(6, ESM_NK * N) * nrep;
    zmalloc((void **)&h->d_part2, h->part2_len * sizeof(double));
    zmalloc((void **)&h->d_out, (size_t)std::max(nrep * 8, ESM_NK * N * 2 + 64) * sizeof(double));
    zmalloc((void **)&h->d_sk, (size_t)nrep * ESM_NK * N * sizeof(double));
    zmalloc((void **)&h->d_fk, (size_t)ESM_NK * N * sizeof(double));
    zmalloc((void **)&h->d_stats, (size_t)ESM_NFAM * ESM_NSTAT * STAT_SLOTS * sizeof(unsigned long long));
    zmalloc((void **)&h->d_counts, (size_t)(ESM_NWEIGHT + 2) * T * sizeof(int));   // sxpair's counts, the step order
    zmalloc((void **)&h->d_sched, 32 * sizeof(int));
    if (!rc) rc = precision == ESM_FP64 ? upload_move_tables<double>(h) : upload_move_tables<float>(h);
    // weather process (optional: a model without weather tables has none)
    h->has_weather = mdl->Ptr && mdl->pi0 && mdl->mu && mdl->xbar && mdl->K >= 1 && mdl->K <= ESM_MAXK &&
                     mdl->sigma > 0 && mdl->sigma0 > 0;
    if (h->has_weather && !rc) {
        const int K = mdl->K, nser = ESM_NREN * N;
        const int nch = (T + h->wx_chunk - 1) / h->wx_chunk;
        zmalloc(&h->d_z, (size_t)nser * TR * h->es);
        zmalloc(&h->d_ytmp, (size_t)nser * TR * h->es);
        zmalloc(&h->d_carry, (size_t)nch * nser * nrep * h->es);
        zmalloc(&h->d_zin, (size_t)nch * nser * nrep * h->es);
        zmalloc((void **)&h->d_m, TR * sizeof(int));
        zmalloc((void **)&h->d_mnew, TR * sizeof(int));
        zmalloc((void **)&h->d_moved, (size_t)nrep * sizeof(int));
        zmalloc((void **)&h->d_wpar, 3 * (size_t)nrep * sizeof(double));
        std::vector<double> mu(mdl->mu, mdl->mu + (size_t)K * nser * 3);
        h->xbar_h.assign(mdl->xbar, mdl->xbar + (size_t)nser * 24);
        h->hod_h.resize(T);
        std::vector<double> harm(2 * (size_t)T);
        for (int t = 0; t < T; t++) {
            h->hod_h[t] = (int)fmod(t * mdl->dt, 24.0);
            harm[2 * t] = cos(2 * M_PI * (t + 1) / T);
            harm[2 * t + 1] = sin(2 * M_PI * (t + 1) / T);
        }
        dmalloc(&h->d_harm, harm);
        h->Ptr_h.assign(mdl->Ptr, mdl->Ptr + K * K);
        h->pi0_h.assign(mdl->pi0, mdl->pi0 + K);
        dmalloc(&h->d_Ptr, h->Ptr_h);
        dmalloc(&h->d_pi0, h->pi0_h);
        dmalloc(&h->d_mu64, mu);
        imalloc(&h->d_hod, h->hod_h);
        if (!rc) {
            if (precision == ESM_FP64) {
                h->d_mu = upload_conv<double>(mu, h->stream, &rc);
                h->d_xbar = upload_conv<double>(h->xbar_h, h->stream, &rc);
            } else {
                h->d_mu = upload_conv<float>(mu, h->stream, &rc);
                h->d_xbar = upload_conv<float>(h->xbar_h, h->stream, &rc);
            }
        }
    }
    if (rc) {
        set_error(rc == ESM_ENOMEM ? "esm_create: device allocation failed" : "esm_create: HIP copy failed");
        free_handle(h);
        return rc;
    }
    *out = h;
    return ESM_OK;
}

int32_t esm_destroy(esm_handle *h) {
    if (!h) return ESM_OK;
    (void)hipStreamSynchronize(h->stream);
    free_handle(h);
    return ESM_OK;
}

int32_t esm_upload_state(esm_handle *h, const double *v, const double *k, const double *x) {
    if (!h) { set_error("esm_upload_state: null handle"); return ESM_EINVAL; }
    const size_t TR = (size_t)h->T * h->R;
    int rc;
    if (k) {
        for (int u = 0; u < ESM_NK * h->N; u++)
            if (!std::isfinite(k[u])) { set_error("esm_upload_state: non-finite capacity"); return ESM_EINVAL; }
        h->k.assign(k, k + ESM_NK * h->N);
        if ((rc = upload_k(h))) return rc;
    }
    if (v && (rc = put_state(h, h->d_v, v, (size_t)ESM_NV * h->N * TR))) return rc;
    if (x && (rc = put_state(h, h->d_x, x, (size_t)ESM_NREN * h->N * TR))) return rc;
    h->scores_valid = false;
    if ((v || k || x) && (rc = rebuild(h))) return rc;
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    return ESM_OK;
}

int32_t esm_heuristic_state(esm_handle *h) {
    if (!h) { set_error("esm_heuristic_state: null handle"); return ESM_EINVAL; }
    int rc;
    if ((rc = launch_heuristic(h))) return rc;
    h->scores_valid = false;
    if ((rc = rebuild(h))) return rc;
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    return ESM_OK;
}

int32_t esm_download_state(esm_handle *h, double *v, double *k, double *x) {
    if (!h) { set_error("esm_download_state: null handle"); return ESM_EINVAL; }
    const size_t TR = (size_t)h->T * h->R;
    int rc;
    if (v && (rc = get_state(h, v, h->d_v, (size_t)ESM_NV * h->N * TR))) return rc;
    if (x && (rc = get_state(h, x, h->d_x, (size_t)ESM_NREN * h->N * TR))) return rc;
    if (k) std::copy(h->k.begin(), h->k.end(), k);
    return ESM_OK;
}

int32_t esm_energy(esm_handle *h, double *parts, double *per_hour) {
    if (!h || !parts) { set_error("esm_energy: null handle or output"); return ESM_EINVAL; }
    int rc = h->prec == ESM_FP64 ? launch_energy<double>(h, true, false) : launch_energy<float>(h, true, false);
    if (rc) return rc;
    if ((rc = launch_reduce_parts(h, h->d_out))) return rc;
    const int R = h->R;
    std::vector<double> p6((size_t)R * 6);
    ESM_HIP_CHECK(hipMemcpyAsync(p6.data(), h->d_out, p6.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    std::vector<double> ph;
    if (per_hour) {
        // per-step op = gen + storage + c_shed shed + c_dump dump + penalty
        std::vector<double> hp(6 * (size_t)h->T * R);
        ESM_HIP_CHECK(hipMemcpyAsync(hp.data(), h->d_hparts, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        const size_t TR = (size_t)h->T * R;
        for (size_t g = 0; g < TR; g++)
            per_hour[g] = hp[g] + hp[TR + g] + h->m.c_shed * hp[2 * TR + g] + h->m.c_dump * hp[3 * TR + g] + hp[4 * TR + g];
    }
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    // capex from the shared capacities (src/energy.jl:557-568), FP64 on the host
    double capex = 0;
    for (int f = 0; f < ESM_NK; f++) {
        double dot = 0, pen = 0;
        for (int n = 0; n < h->N; n++) {
            const int u = f * h->N + n;
            const double kc = std::min(std::max(h->k[u], h->klo[u]), h->khi[u]);
            dot += h->capex[u] * h->k[u];
            pen += h->capex[u] * fabs(h->k[u] - kc);
        }
        capex += dot + h->m.gamma_k * pen;
    }
    for (int j = 0; j < R; j++) {
        const double *q = &p6[(size_t)j * 6];
        double *o = parts + (size_t)j * ESM_NPARTS;
        const double cbal = h->m.c_shed * q[2] + h->m.c_dump * q[3];
        o[0] = capex;
        o[1] = q[0] + q[1] + cbal + q[4];
        o[2] = q[0];
        o[3] = q[1];
        o[4] = q[2];
        o[5] = q[3];
        o[6] = q[4];
        o[7] = q[5];
    }
    return ESM_OK;
}

int32_t esm_rebuild_residual(esm_handle *h, double *drift) {
    if (!h) { set_error("esm_rebuild_residual: null handle"); return ESM_EINVAL; }
    int rc = h->prec == ESM_FP64 ? launch_energy<double>(h, false, true) : launch_energy<float>(h, false, true);
    if (rc) return rc;
    if ((rc = launch_reduce_max(h, h->d_out))) return rc;
    if (drift) ESM_HIP_CHECK(hipMemcpyAsync(drift, h->d_out, h->R * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    return ESM_OK;
}

int32_t esm_download_caches(esm_handle *h, double *r, double *F) {
    if (!h) { set_error("esm_download_caches: null handle"); return ESM_EINVAL; }
    const size_t TR = (size_t)h->T * h->R;
    int rc;
    if (r && (rc = get_state(h, r, h->d_r, (size_t)h->N * TR))) return rc;
    if (F && (rc = get_state(h, F, h->d_F, (size_t)h->L * TR))) return rc;
    return ESM_OK;
}

int32_t esm_scores(esm_handle *h, double *sk, double *fk) {
    if (!h) { set_error("esm_scores: null handle"); return ESM_EINVAL; }
    int rc = h->prec == ESM_FP64 ? launch_scores<double>(h) : launch_scores<float>(h);
    if (rc) return rc;
    if (sk) ESM_HIP_CHECK(hipMemcpyAsync(sk, h->d_sk, (size_t)h->R * ESM_NK * h->N * sizeof(double),
                                         hipMemcpyDeviceToHost, h->stream));
    if (fk) ESM_HIP_CHECK(hipMemcpyAsync(fk, h->d_fk, (size_t)ESM_NK * h->N * sizeof(double), hipMemcpyDeviceToHost,
                                         h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->scores_valid = true;
    return ESM_OK;
}

int32_t esm_score_measure(esm_handle *h, int32_t measure) {
    if (!h) { set_error("esm_score_measure: null handle"); return ESM_EINVAL; }
    if (measure != ESM_MEASURE_UTILIZATION && measure != ESM_MEASURE_PHYSICAL) {
        set_error("esm_score_measure: measure must be ESM_MEASURE_UTILIZATION or ESM_MEASURE_PHYSICAL");
        return ESM_EINVAL;
    }
    h->measure = measure;
    h->scores_valid = false;
    return ESM_OK;
}

int32_t esm_langevin(esm_handle *h, const double *weights, const double *xi, const double *eta,
                     const double *kappa_s, double kmin, double clip, int32_t box, double *knew, int32_t *nclip) {
    if (!h || !weights || !xi || !eta || !kappa_s || !knew) {
        set_error("esm_langevin: null argument");
        return ESM_EINVAL;
    }
    if (!h->scores_valid) {
        set_error("esm_langevin: no scores of the current state (call esm_scores first)");
        return ESM_ESTATE;
    }
    const int n = ESM_NK * h->N, R = h->R;
    // staging in d_out: weights [R], xi [n], eta [6], kappa_s [6], knew [n], nclip (int)
    std::vector<double> stage((size_t)R + n + 2 * ESM_NK);
    std::copy(weights, weights + R, stage.begin());
    std::copy(xi, xi + n, stage.begin() + R);
    std::copy(eta, eta + ESM_NK, stage.begin() + R + n);
    std::copy(kappa_s, kappa_s + ESM_NK, stage.begin() + R + n + ESM_NK);
    for (int f = 0; f < ESM_NK; f++)
        if (!(eta[f] >= 0) || !(kappa_s[f] > -kmin)) {
            set_error("esm_langevin: eta must be >= 0 and kappa_s + kmin > 0");
            return ESM_EINVAL;
        }
    double *d = nullptr;
    const size_t bytes = (stage.size() + n + 1) * sizeof(double);
    ESM_HIP_CHECK(hipMallocAsync((void **)&d, bytes, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(d, stage.data(), stage.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int *d_nc = (int *)(d + stage.size() + n);
    ESM_HIP_CHECK(hipMemsetAsync(d_nc, 0, sizeof(int), h->stream));
    int rc = launch_langevin(h, d, d + R, d + R + n, d + R + n + ESM_NK, kmin, clip, box, d + stage.size(), d_nc);
    if (rc) return rc;
    int nc = 0;
    ESM_HIP_CHECK(hipMemcpyAsync(knew, d + stage.size(), n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(&nc, d_nc, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipFreeAsync(d, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (nclip) *nclip = nc;
    return ESM_OK;
}

int32_t esm_langevin_noalloc(esm_handle *h, const double *weights, const double *xi, const double *eta,
                             const double *kappa_s, double kmin, double clip, int32_t box, double *knew,
                             int32_t *nclip) {
    if (!h || !weights || !xi || !eta || !kappa_s || !knew) {
        set_error("esm_langevin_noalloc: null argument");
        return ESM_EINVAL;
    }
    const int n = ESM_NK * h->N, R = h->R;
    // layout in the scratch: weights [R] | xi [n] | eta [6] | kappa_s [6] | knew [n] | counter
    const size_t need = (size_t)R + 2 * (size_t)n + 2 * ESM_NK + 1;
    if (need > (size_t)ESM_NOALLOC_DOUBLES) {
        set_error("esm_langevin_noalloc: R + 12 N + 13 = " + std::to_string(need) + " doubles exceed the fixed scratch of " +
                  std::to_string(ESM_NOALLOC_DOUBLES) + " (use esm_langevin)");
        return ESM_EINVAL;
    }
    if (!h->scores_valid) {
        set_error("esm_langevin_noalloc: no scores of the current state (call esm_scores first)");
        return ESM_ESTATE;
    }
    for (int f = 0; f < ESM_NK; f++)
        if (!(eta[f] >= 0) || !(kappa_s[f] > -kmin)) {
            set_error("esm_langevin_noalloc: eta must be >= 0 and kappa_s + kmin > 0");
            return ESM_EINVAL;
        }
    double *d = h->device >= 0 && h->device < SCRATCH_DEVICES ? g_scratch[h->device] : nullptr;
    if (!d) {
        set_error("esm_langevin_noalloc: the scratch of this device was not set up");
        return ESM_ESTATE;
    }
    double *d_w = d, *d_xi = d + R, *d_eta = d + R + n, *d_ks = d_eta + ESM_NK, *d_knew = d_ks + ESM_NK;
    int *d_nc = (int *)(d_knew + n);
    ESM_HIP_CHECK(hipMemcpyAsync(d_w, weights, (size_t)R * sizeof(double), hipMemcpyHostToDevice, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(d_xi, xi, (size_t)n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(d_eta, eta, ESM_NK * sizeof(double), hipMemcpyHostToDevice, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(d_ks, kappa_s, ESM_NK * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, h->stream, d_nc);
    ESM_HIP_CHECK(hipGetLastError());
    int rc = launch_langevin(h, d_w, d_xi, d_eta, d_ks, kmin, clip, box, d_knew, d_nc);
    if (rc) return rc;
    int nc = 0;
    ESM_HIP_CHECK(hipMemcpyAsync(knew, d_knew, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipMemcpyAsync(&nc, d_nc, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (nclip) *nclip = nc;
    return ESM_OK;
}

int32_t esm_sweeps(esm_handle *h, int32_t nsweeps, const uint32_t key[2], uint32_t counter, double maxnats,
                   int64_t *stats) {
    if (!h || !key) { set_error("esm_sweeps: null argument"); return ESM_EINVAL; }
    if (nsweeps < 0 || !(maxnats > 0)) { set_error("esm_sweeps: nsweeps must be >= 0 and maxnats > 0"); return ESM_EINVAL; }
    constexpr int NST = ESM_NFAM * ESM_NSTAT;
    ESM_HIP_CHECK(hipMemsetAsync(h->d_stats, 0, (size_t)NST * STAT_SLOTS * sizeof(unsigned long long), h->stream));
    int rc = h->prec == ESM_FP64 ? launch_sweeps<double>(h, nsweeps, key[0], key[1], counter, maxnats)
                                 : launch_sweeps<float>(h, nsweeps, key[0], key[1], counter, maxnats);
    if (rc) return rc;
    if (stats) {
        std::vector<unsigned long long> st((size_t)NST * STAT_SLOTS);
        ESM_HIP_CHECK(hipMemcpyAsync(st.data(), h->d_stats, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                     h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
        for (int i = 0; i < NST; i++) {
            unsigned long long acc = 0;
            for (int q = 0; q < STAT_SLOTS; q++) acc += st[(size_t)i * STAT_SLOTS + q];
            stats[i] = (int64_t)acc;
        }
    }
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->scores_valid = false;
    return ESM_OK;
}

int32_t esm_sweep_options(esm_handle *h, double extra, const double *weights, int32_t atoms, double p_storage_ms) {
    if (!h) { set_error("esm_sweep_options: null handle"); return ESM_EINVAL; }
    if (!(extra >= 0) || extra > 1e4) { set_error("esm_sweep_options: extra must be in [0, 1e4]"); return ESM_EINVAL; }
    static const double wdef[ESM_NWEIGHT] = {0.3, 0.25, 0.03, 0.08, 0.04, 0.1, 0.1, 0.1, 0.3};
    const double *w = weights ? weights : wdef;
    double tot = 0.0;
    for (int q = 0; q < ESM_NWEIGHT; q++) {
        if (!(w[q] >= 0) || !std::isfinite(w[q])) { set_error("esm_sweep_options: weights must be finite and >= 0"); return ESM_EINVAL; }
        tot += w[q];
    }
    (void)tot;   // all zero: the plain families make no draws (the storage pairs may still)
    if (!(p_storage_ms >= 0 && p_storage_ms <= 1)) { set_error("esm_sweep_options: p_storage_ms must be in [0, 1]"); return ESM_EINVAL; }
    h->sweep_extra = extra;
    for (int q = 0; q < ESM_NWEIGHT; q++) h->sweep_w[q] = w[q];
    h->sweep_atoms = atoms != 0;
    h->sweep_pms = p_storage_ms;
    return ESM_OK;
}

int32_t esm_sweep_launch(esm_handle *h, int32_t merged) {
    if (!h) { set_error("esm_sweep_launch: null handle"); return ESM_EINVAL; }
    if (merged != 0 && merged != 1 && merged != 2 && merged != 4 && merged != 8 && merged != 16) {
        set_error("esm_sweep_launch: 0 (one launch per family), 1 (merged), or 2, 4, 8, 16 (merged, waves per step)");
        return ESM_EINVAL;
    }
    if (merged > 1 && h->R % 64) {
        set_error("esm_sweep_launch: waves per step need a replica count that is a multiple of 64");
        return ESM_EINVAL;
    }
    h->sweep_merged = merged != 0;
    h->sweep_waves = merged > 1 ? merged : 0;
    return ESM_OK;
}

// The units of the storage pairs (hours ≥ hmin) on the device and the buffer of the sxlong
// schedule, for the options in force; refuses units that only discharge (the loss share of
// `sx_loss` is not built: every unit of the pairs must be able to charge).
static int sx_refresh(esm_handle *h, double pxpair, int nxlong, double hmin, const char *who) {
    const int N = h->N;
    std::vector<int> us;
    for (int n = 0; n < N; n++)
        for (int si = 0; si < ESM_NSTOR; si++)
            if (h->hours[si * N + n] >= hmin) us.push_back(si * N + n);
    if (pxpair > 0 || nxlong > 0) {
        if (N > 256) {
            set_error(std::string(who) + ": the storage pairs are built for at most 256 nodes");
            return ESM_EINVAL;
        }
        for (int u : us)
            if (!(h->eta_in[u] > 0)) {
                set_error(std::string(who) + ": a storage unit of the pairs cannot charge (eta_in = 0); the loss share "
                                             "of such units is not built");
                return ESM_EINVAL;
            }
    }
    if (h->d_sxu) (void)hipFree(h->d_sxu);
    if (h->d_sxm) (void)hipFree(h->d_sxm);
    h->d_sxu = nullptr;
    h->d_sxm = nullptr;
    h->nsx = 0;
    h->sxm_cap = 0;
    if (!us.empty()) {
        ESM_HIP_CHECK(hipMalloc((void **)&h->d_sxu, us.size() * sizeof(int)));
        ESM_HIP_CHECK(hipMemcpy(h->d_sxu, us.data(), us.size() * sizeof(int), hipMemcpyHostToDevice));
        h->nsx = (int)us.size();
    }
    if (nxlong > 0 && !us.empty()) {
        h->sxm_cap = us.size() * (size_t)nxlong;
        ESM_HIP_CHECK(hipMalloc((void **)&h->d_sxm, esm::sxm_ints(h->sxm_cap) * sizeof(int)));
        ESM_HIP_CHECK(hipMemset(h->d_sxm, 0, esm::sxm_ints(h->sxm_cap) * sizeof(int)));
    }
    h->sweep_pxpair = pxpair;
    h->sweep_nxlong = nxlong;
    return ESM_OK;
}

int32_t esm_storage_long_options(esm_handle *h, double p, int32_t nmoves, double hmin) {
    if (!h) { set_error("esm_storage_long_options: null handle"); return ESM_EINVAL; }
    if (!(p >= 0 && p <= 1)) { set_error("esm_storage_long_options: p must be in [0, 1]"); return ESM_EINVAL; }
    if (nmoves < 0 || nmoves > 256) { set_error("esm_storage_long_options: nmoves must be in [0, 256]"); return ESM_EINVAL; }
    if (!(hmin >= 0) || !std::isfinite(hmin)) { set_error("esm_storage_long_options: hmin must be finite and >= 0"); return ESM_EINVAL; }
    const int rc = sx_refresh(h, h->sweep_pxpair, h->sweep_nxlong, hmin, "esm_storage_long_options");
    if (rc) return rc;
    h->sweep_plong = p;
    h->sweep_long_moves = nmoves;
    h->sweep_long_hours = hmin;
    return ESM_OK;
}

int32_t esm_storage_cross_options(esm_handle *h, double p_xpair, int32_t nxlong) {
    if (!h) { set_error("esm_storage_cross_options: null handle"); return ESM_EINVAL; }
    if (!(p_xpair >= 0) || !std::isfinite(p_xpair)) {
        set_error("esm_storage_cross_options: p_xpair must be finite and >= 0");
        return ESM_EINVAL;
    }
    if (nxlong < 0 || nxlong > 256) { set_error("esm_storage_cross_options: nxlong must be in [0, 256]"); return ESM_EINVAL; }
    return sx_refresh(h, p_xpair, nxlong, h->sweep_long_hours, "esm_storage_cross_options");
}

int32_t esm_follow_capacities(esm_handle *h, const double *knew, double kmin, double *kout) {
    if (!h || !knew) { set_error("esm_follow_capacities: null argument"); return ESM_EINVAL; }
    const int N = h->N, n = ESM_NK * N;
    std::vector<double> k0 = h->k, k1(n), knext(n), kc0(n), kc1(n);
    std::vector<int> moves(n, 0);
    for (int u = 0; u < n; u++) {
        if (!(knew[u] > 0) || !std::isfinite(knew[u])) {
            set_error("esm_follow_capacities: capacities must be positive and finite");
            return ESM_EINVAL;
        }
        const int f = u / N;
        if (f != ESM_KGAS && k0[u] >= kmin) {
            k1[u] = k0[u] * exp(log(knew[u] / k0[u]));   // the kop curve at u = log(k'/k)
            moves[u] = 1;
        } else {
            k1[u] = knew[u];
        }
        knext[u] = k0[u] + (k1[u] - k0[u]);
        kc0[u] = std::min(std::max(k0[u], h->klo[u]), h->khi[u]);
        kc1[u] = std::min(std::max(k1[u], h->klo[u]), h->khi[u]);
    }
    // device staging: kc0, kc1 in the state precision, then the move flags
    void *d = nullptr;
    const size_t bytes = 2 * n * h->es + n * sizeof(int);
    ESM_HIP_CHECK(hipMallocAsync(&d, bytes, h->stream));
    if (h->prec == ESM_FP64) {
        ESM_HIP_CHECK(hipMemcpyAsync(d, kc0.data(), n * 8, hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipMemcpyAsync((char *)d + n * 8, kc1.data(), n * 8, hipMemcpyHostToDevice, h->stream));
    } else {
        std::vector<float> a(kc0.begin(), kc0.end()), b(kc1.begin(), kc1.end());
        ESM_HIP_CHECK(hipMemcpyAsync(d, a.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipMemcpyAsync((char *)d + n * 4, b.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    int *d_moves = (int *)((char *)d + 2 * n * h->es);
    ESM_HIP_CHECK(hipMemcpyAsync(d_moves, moves.data(), n * sizeof(int), hipMemcpyHostToDevice, h->stream));
    int rc = h->prec == ESM_FP64 ? launch_follow<double>(h, d, (char *)d + n * 8, d_moves, kmin)
                                 : launch_follow<float>(h, d, (char *)d + n * 4, d_moves, kmin);
    if (rc) return rc;
    ESM_HIP_CHECK(hipFreeAsync(d, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->k = knext;
    if ((rc = upload_k(h))) return rc;
    if ((rc = rebuild(h))) return rc;
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (kout) std::copy(knext.begin(), knext.end(), kout);
    h->scores_valid = false;
    return ESM_OK;
}

int32_t esm_schedule(esm_handle *h, int32_t *colour, int32_t *ncolours) {
    if (!h) { set_error("esm_schedule: null handle"); return ESM_EINVAL; }
    if (colour) std::copy(h->colour.begin(), h->colour.end(), colour);
    if (ncolours) *ncolours = h->ncolour;
    return ESM_OK;
}

static int weather_call(esm_handle *h, bool init, int B, int compensate, const uint32_t key[2], uint32_t ctr) {
    return h->prec == ESM_FP64 ? launch_weather<double>(h, init, B, compensate, key[0], key[1], ctr)
                               : launch_weather<float>(h, init, B, compensate, key[0], key[1], ctr);
}

int32_t esm_weather_init(esm_handle *h, const uint32_t key[2], uint32_t counter) {
    if (!h || !key) { set_error("esm_weather_init: null argument"); return ESM_EINVAL; }
    if (!h->has_weather) { set_error("esm_weather_init: the model has no weather tables"); return ESM_ESTATE; }
    const int R = h->R;
    std::vector<double> wpar(3 * (size_t)R, 0.0);
    for (int j = 0; j < R; j++) wpar[R + j] = 1.0;    // c = 0, s = 1: z from fresh innovations
    ESM_HIP_CHECK(hipMemcpyAsync(h->d_wpar, wpar.data(), wpar.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int rc = weather_call(h, true, 0, 0, key, counter);
    if (rc) return rc;
    if ((rc = rebuild(h))) return rc;
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->weather_ready = true;
    h->scores_valid = false;
    return ESM_OK;
}

int32_t esm_weather_step(esm_handle *h, const double *theta, const double *win_prob, int32_t B, int32_t compensate,
                         const uint32_t key[2], uint32_t counter, int32_t *moved) {
    if (!h || !theta || !win_prob || !key) { set_error("esm_weather_step: null argument"); return ESM_EINVAL; }
    if (!h->weather_ready) {
        set_error("esm_weather_step: no weather drawn or uploaded (esm_weather_init / esm_upload_weather)");
        return ESM_ESTATE;
    }
    if (B < 0 || B > ESM_MAXWIN) { set_error("esm_weather_step: window length out of range"); return ESM_EINVAL; }
    const int R = h->R;
    std::vector<double> wpar(3 * (size_t)R);
    for (int j = 0; j < R; j++) {
        if (!(theta[j] >= 0) || !(win_prob[j] >= 0)) { set_error("esm_weather_step: negative angle or probability"); return ESM_EINVAL; }
        wpar[j] = theta[j] > 0 ? cos(theta[j]) : 1.0;
        wpar[R + j] = theta[j] > 0 ? sin(theta[j]) : 0.0;
        wpar[2 * R + j] = win_prob[j];
    }
    ESM_HIP_CHECK(hipMemcpyAsync(h->d_wpar, wpar.data(), wpar.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int rc = weather_call(h, false, B, compensate, key, counter);
    if (rc) return rc;
    if (moved) ESM_HIP_CHECK(hipMemcpyAsync(moved, h->d_moved, R * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->scores_valid = false;
    return ESM_OK;
}

int32_t esm_weather_ffbs(esm_handle *h, const int32_t *mask, const uint32_t key[2], uint32_t counter) {
    if (!h || !key) { set_error("esm_weather_ffbs: null argument"); return ESM_EINVAL; }
    if (!h->weather_ready) {
        set_error("esm_weather_ffbs: no weather drawn or uploaded (esm_weather_init / esm_upload_weather)");
        return ESM_ESTATE;
    }
    std::vector<int> mk((size_t)h->R, 1);
    if (mask)
        for (int j = 0; j < h->R; j++) mk[j] = mask[j] != 0;
    int rc = h->prec == ESM_FP64 ? launch_ffbs<double>(h, mk.data(), key[0], key[1], counter)
                                 : launch_ffbs<float>(h, mk.data(), key[0], key[1], counter);
    if (rc) return rc;
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    return ESM_OK;
}

int32_t esm_replica_features(esm_handle *h, const double *kref, double *out) {
    if (!h || !kref || !out) { set_error("esm_replica_features: null argument"); return ESM_EINVAL; }
    if (!h->weather_ready) {
        set_error("esm_replica_features: no weather drawn or uploaded (esm_weather_init / esm_upload_weather)");
        return ESM_ESTATE;
    }
    for (int q = 0; q < ESM_NREN * h->N; q++)
        if (!std::isfinite(kref[q])) { set_error("esm_replica_features: non-finite reference capacity"); return ESM_EINVAL; }
    return h->prec == ESM_FP64 ? launch_features<double>(h, kref, out) : launch_features<float>(h, kref, out);
}

int32_t esm_download_weather(esm_handle *h, int32_t *m, double *z) {
    if (!h) { set_error("esm_download_weather: null handle"); return ESM_EINVAL; }
    if (!h->has_weather) { set_error("esm_download_weather: the model has no weather tables"); return ESM_ESTATE; }
    const size_t TR = (size_t)h->T * h->R;
    if (m) ESM_HIP_CHECK(hipMemcpyAsync(m, h->d_m, TR * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    int rc;
    if (z && (rc = get_state(h, z, h->d_z, (size_t)ESM_NREN * h->N * TR))) return rc;
    return ESM_OK;
}

int32_t esm_upload_weather(esm_handle *h, const int32_t *m, const double *z) {
    if (!h || !m || !z) { set_error("esm_upload_weather: null argument"); return ESM_EINVAL; }
    if (!h->has_weather) { set_error("esm_upload_weather: the model has no weather tables"); return ESM_ESTATE; }
    const int T = h->T, R = h->R, nser = ESM_NREN * h->N;
    const size_t TR = (size_t)T * R;
    for (size_t g = 0; g < TR; g++)
        if (m[g] < 0 || m[g] >= h->m.K) { set_error("esm_upload_weather: regime out of range"); return ESM_EINVAL; }
    ESM_HIP_CHECK(hipMemcpyAsync(h->d_m, m, TR * sizeof(int), hipMemcpyHostToDevice, h->stream));
    int rc;
    if ((rc = put_state(h, h->d_z, z, (size_t)nser * TR))) return rc;
    std::vector<double> x((size_t)nser * TR);
    for (int i = 0; i < nser; i++)
        for (int t = 0; t < T; t++) {
            const double xb = h->xbar_h[(size_t)i * 24 + h->hod_h[t]];
            for (int j = 0; j < R; j++) {
                const size_t o = ((size_t)i * T + t) * R + j;
                x[o] = 1.0 / (1.0 + exp(-(xb + z[o])));
            }
        }
    if ((rc = put_state(h, h->d_x, x.data(), x.size()))) return rc;
    if ((rc = rebuild(h))) return rc;
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->weather_ready = true;
    h->scores_valid = false;
    return ESM_OK;
}

int32_t esm_copy_replica(esm_handle *src, int32_t js, esm_handle *dst, int32_t jd) {
    if (!src || !dst) { set_error("esm_copy_replica: null handle"); return ESM_EINVAL; }
    if (src->N != dst->N || src->L != dst->L || src->T != dst->T || src->prec != dst->prec) {
        set_error("esm_copy_replica: the handles differ in shape or precision");
        return ESM_EINVAL;
    }
    if (js < 0 || js >= src->R || jd < 0 || jd >= dst->R) { set_error("esm_copy_replica: replica index out of range"); return ESM_EINVAL; }
    if (src == dst && js == jd) return ESM_OK;
    ESM_HIP_CHECK(hipStreamSynchronize(src->stream));
    const bool weather = src->has_weather && dst->has_weather && src->weather_ready;
    int rc;
    if ((rc = launch_copy_replica(src, js, dst, jd, weather))) return rc;
    if (dst != src) {
        dst->k = src->k;
        if ((rc = upload_k(dst))) return rc;
    }
    if (weather && dst->R == 1) dst->weather_ready = true;
    if ((rc = rebuild(dst))) return rc;
    ESM_HIP_CHECK(hipStreamSynchronize(dst->stream));
    dst->scores_valid = false;
    return ESM_OK;
}

int32_t esm_profile(esm_handle *h, int32_t enable) {
    if (!h) { set_error("esm_profile: null handle"); return ESM_EINVAL; }
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (enable) {
        // intervals of earlier profiles are dropped; the pool holds enough events for a replica
        // step's launches, so the timed work creates none (creating them on the way would keep the
        // device waiting for the host between launches)
        clear_prof(h);
        while (h->prof_pool.size() < 8192) {
            hipEvent_t e = nullptr;
            ESM_HIP_CHECK(hipEventCreate(&e));
            h->prof_pool.push_back(e);
        }
    }
    h->prof = enable != 0;
    return ESM_OK;
}

int32_t esm_profile_read(esm_handle *h, int32_t kernel, double *total_ms, int64_t *launches) {
    if (!h || kernel < 0 || kernel >= ESM_NKERN) { set_error("esm_profile_read: bad handle or kernel id"); return ESM_EINVAL; }
    ESM_HIP_CHECK(hipStreamSynchronize(h->stream));
    double tot = 0;
    for (auto &e : h->prof_ev[kernel]) {
        float ms = 0;
        ESM_HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (int64_t)h->prof_ev[kernel].size();
    return ESM_OK;
}

}  // extern "C"
