// esm_internal.h — the handle and the kernel launchers of libesmhip.so (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/esm.h"
#include "esm_device.h"

namespace esm {

constexpr int ESM_MAXK = 4;       // largest regime count of the weather model
constexpr int ESM_MAXWIN = 256;   // longest regime window
constexpr int SMS_NU = 64;        // units of one storage multiscale move (the largest country–carrier
                                  // group): one lane of the wave stages each unit
constexpr int STAT_SLOTS = 256;   // sweep counters are spread over this many addresses per statistic

// Minimum waves per SIMD asked of a latency-bound sweep kernel (second __launch_bounds__
// argument), per precision, each the occupancy measured fastest: the caps limit the registers
// and accept a few spilled bytes per lane. The FP64 instances need about twice the registers;
// most keep the compiler's own allocation (1), which the FP32 caps would turn into hundreds of
// spilled bytes per lane.
template <typename S>
constexpr int min_waves(int fp32_waves, int fp64_waves = 1) { return sizeof(S) == 4 ? fp32_waves : fp64_waves; }

void set_error(const std::string &msg);

#define ESM_HIP_CHECK(expr)                                                                   \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            ::esm::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));              \
            return ESM_EHIP;                                                                  \
        }                                                                                     \
    } while (0)

// Device copies of the model tables in one precision.
struct ModelBuf {
    void *lb = nullptr, *ls = nullptr, *adj_b = nullptr, *load = nullptr, *mc = nullptr, *ef = nullptr,
         *kc = nullptr, *hours = nullptr, *eta_in = nullptr, *eta_out = nullptr, *qt = nullptr;
};

}  // namespace esm

namespace esm {

// Lane-uniform templates of the table-driven random moves (cell, xloc / xlocc, xpair), built
// from the network at esm_create. Template q lists, each in a canonical order:
//   θ entries      th_node/th_val [th_off[q], th_off[q+1])      angle change per unit t
//   gas entries    en_node/en_c/en_div [en_off..)               slope en_c, divided by the node's
//                                                               gas capacity when en_div
//   kinks          kk_node/kk_g [kk_off..)                      leak nodes and their requirement
//                                                               change g per unit t (xlocc)
//   lines          ln_line/ln_df [ln_off..)                     flow change per unit t (sorted)
//   touched nodes  nd_node/nd_ent/nd_kk [nd_off..)              sorted; the node's gas entry and
//                                                               kink (positions, −1 none); its
//                                                               lines nl_pos/nl_sgn [nd_l[slot],
//                                                               nd_l[slot+1]) (+1: the node is bus0)
// Template ids: xloc (n, p) at xp_off[n] + p (p-th partner of n); xpair (n, m) at xp0 + n·N + m
// (empty when n = m or in different components); the cells of partition level l at
// cl0 + lev_off[l] + c.
struct MoveTablesHost {
    std::vector<int> th_off, th_node, en_off, en_node, en_div, kk_off, kk_node, ln_off, ln_line, nd_off, nd_node,
        nd_ent, nd_kk, nd_l, nl_pos, nl_sgn, xp_off, lev_off;
    std::vector<double> th_val, en_c, kk_g, ln_df, qcum;
    int xp0 = 0, cl0 = 0, ntpl = 0, nlev = 0;
    // storage multiscale moves: every storage unit (carrier·N + node, carrier-major) and the
    // country–carrier groups (sorted by country, then carrier; nodes increasing)
    std::vector<int> sms_units, grp_off, grp_node, grp_si;
    // storage pairs (sx_templates): slots sx_poff[n] … sx_poff[n+1]−1 of node n (the node itself,
    // then its neighbours increasing), per slot the partner, the angle entries (unscaled), the leak
    // nodes and their requirement change per unit, the lines at the angle nodes with the angle
    // change at each end, the touched nodes (with their lines and leak) and the footprint
    std::vector<int> sx_poff, sx_part, sx_th_off, sx_th_node, sx_kk_off, sx_kk_node, sx_ln_off, sx_ln_line, sx_nd_off,
        sx_nd_node, sx_nd_kk, sx_nd_l, sx_nl_pos, sx_nl_sgn, sx_ft_off, sx_ft_node;
    std::vector<double> sx_th_val, sx_kk_g, sx_ln_ta, sx_ln_tb;
};

struct MoveTablesDev {
    int *ibuf = nullptr;     // all integer arrays, one allocation
    void *vbuf = nullptr;    // all value arrays in the state precision
    double *qcum = nullptr;
    int *th_off = nullptr, *th_node = nullptr, *en_off = nullptr, *en_node = nullptr, *en_div = nullptr,
              *kk_off = nullptr, *kk_node = nullptr, *ln_off = nullptr, *ln_line = nullptr, *nd_off = nullptr,
              *nd_node = nullptr, *nd_ent = nullptr, *nd_kk = nullptr, *nd_l = nullptr, *nl_pos = nullptr,
              *nl_sgn = nullptr, *xp_off = nullptr, *lev_off = nullptr;
    const void *th_val = nullptr, *en_c = nullptr, *kk_g = nullptr, *ln_df = nullptr;
    int xp0 = 0, cl0 = 0, nlev = 0;
    int *sms_units = nullptr, *grp_off = nullptr, *grp_node = nullptr, *grp_si = nullptr;
    int *sx_poff = nullptr, *sx_part = nullptr, *sx_th_off = nullptr, *sx_th_node = nullptr, *sx_kk_off = nullptr,
        *sx_kk_node = nullptr, *sx_ln_off = nullptr, *sx_ln_line = nullptr, *sx_nd_off = nullptr, *sx_nd_node = nullptr,
        *sx_nd_kk = nullptr, *sx_nd_l = nullptr, *sx_nl_pos = nullptr, *sx_nl_sgn = nullptr, *sx_ft_off = nullptr,
        *sx_ft_node = nullptr;
    const void *sx_th_val = nullptr, *sx_kk_g = nullptr, *sx_ln_ta = nullptr, *sx_ln_tb = nullptr;
};

// Build the templates from the model (host, FP64); false with a message on a bad network.
bool build_move_tables(const esm_handle *h, MoveTablesHost &mt, std::string &err);

}  // namespace esm

namespace esm {
struct MlpState;   // k_mlp.hip
}

struct esm_handle {
    int N = 0, L = 0, T = 0, R = 0, prec = 64;
    double dt = 1.0;
    size_t es = 8;  // element size of the state precision
    hipStream_t stream = nullptr;
    int device = 0;
    // host copy of the model scalars and tables (FP64)
    esm_model m{};
    std::vector<int32_t> bus0, bus1;
    std::vector<double> line_b, line_s, load, mc, ef, capex, klo, khi, hours, eta_in, eta_out, qt;
    std::vector<int> adj_ptr, adj_line, adj_other;
    std::vector<double> adj_bv;
    std::vector<double> k;   // [ESM_NK][N] raw capacities (shared)
    int ncountry = 0;
    std::vector<int> country;   // [N] middle partition level of the cell moves (empty: none)
    // device model
    int *d_bus0 = nullptr, *d_bus1 = nullptr, *d_adj_ptr = nullptr, *d_adj_line = nullptr, *d_adj_other = nullptr;
    esm::ModelBuf mb;
    double *d_k = nullptr;       // [ESM_NK][N] FP64 raw capacities
    double *d_klo = nullptr, *d_khi = nullptr, *d_capex = nullptr;
    // device state (precision S)
    void *d_v = nullptr;   // [ESM_NV][N][T][R]
    void *d_x = nullptr;   // [ESM_NREN][N][T][R]
    void *d_r = nullptr;   // [N][T][R]
    void *d_F = nullptr;   // [L][T][R]
    // FP64 scratch
    double *d_hparts = nullptr;   // [6][T][R] per-step parts (gen, stor, shed, dump, pen, co2) or drift
    double *d_part2 = nullptr;    // partial sums (reductions)
    size_t part2_len = 0;
    double *d_out = nullptr;      // small outputs
    double *d_sk = nullptr;       // [R][ESM_NK][N]
    double *d_fk = nullptr;       // [ESM_NK][N]
    bool scores_valid = false;
    int measure = ESM_MEASURE_UTILIZATION;   // reference measure of the capacity score (esm_score_measure)
    // weather process
    bool has_weather = false, weather_ready = false;
    int wx_chunk = 128;
    void *d_z = nullptr;        // [3N][T][R] anomalies (S)
    void *d_ytmp = nullptr;     // [3N][T][R] chunk-local recurrence (S)
    void *d_carry = nullptr;    // [nchunk][3N][R] (S)
    void *d_zin = nullptr;      // [nchunk][3N][R] (S)
    void *d_mu = nullptr;       // [K][3N][3] (S)
    void *d_xbar = nullptr;     // [3N][24] (S)
    int *d_m = nullptr, *d_mnew = nullptr;   // [T][R]
    int *d_hod = nullptr;       // [T]
    int *d_moved = nullptr;     // [R]
    double *d_harm = nullptr;   // [T][2]
    double *d_wpar = nullptr;   // [3][R] cos θ, sin θ, window probability
    double *d_Ptr = nullptr, *d_pi0 = nullptr;
    std::vector<double> xbar_h;  // [3N][24]
    std::vector<int> hod_h;      // [T]
    std::vector<double> Ptr_h, pi0_h;   // [K][K], [K] (host copies for the FFBS's log tables)
    double *d_mu64 = nullptr;    // [K][3N][3] regime means in FP64 (FFBS emissions)
    double *d_ffbs = nullptr;    // [K][T][R] FFBS emissions, then the forward filter (allocated at first use)
    double *d_feat = nullptr;    // replica features: kref [3][N], D [T], rows [R][nf] (allocated at first use)
    size_t feat_len = 0;
    int *d_ffbs_mask = nullptr;  // [R]
    // sweep schedule
    struct ThetaRound {
        int colour, dmax, off, count;
    };
    std::vector<int> colour;                      // [N] distance-2 colouring
    int ncolour = 0;
    std::vector<ThetaRound> theta_rounds;         // per (colour, degree bucket)
    std::vector<int> theta_nodes_h;
    int *d_theta_nodes = nullptr;
    std::vector<std::pair<int, int>> tclass;      // time-parity classes: (first step, count), stride 2
    unsigned long long *d_stats = nullptr;        // [ESM_NFAM][ESM_NSTAT][STAT_SLOTS]
    // random families of the sweep (esm_sweep_options) and their move templates
    bool sweep_atoms = true;
    double sweep_extra = 4.0;
    double sweep_w[ESM_NWEIGHT] = {0.3, 0.25, 0.03, 0.08, 0.04, 0.1, 0.1, 0.1, 0.3};
    double sweep_pms = 0.5;      // probability of the storage multiscale sweep per operation sweep
    int sweep_ms_moves = 8;      // moves per block
    double sweep_ms_pagg = 0.5;  // probability of a country–carrier group
    double sweep_plong = 1.0;       // probability of the long-range storage sweep per operation sweep
    int sweep_long_moves = 32;      // long-range plateaus per storage unit and sweep
    double sweep_long_hours = 24.0; // ... for the units with at least this many hours of storage
    bool sweep_merged = true;       // the random families of a parity class in one launch
    int sweep_waves = 0;            // waves per step of that launch (0: by network size)
    int rand_block = 256;           // workgroup size of the one-wave random launches (64, 128, 256)
    bool rand_pad = true;           // ... each step on whole waves (R rounded up to a multiple of 64)
    // cross-node storage pairs (esm_storage_cross_options): the weight of sxpair in the local
    // sweep, the sxlong moves per unit in the long-range sweep, the units taking part (hours ≥
    // sweep_long_hours) and the buffer of the sxlong schedule
    double sweep_pxpair = 0.0;
    int sweep_nxlong = 0;
    int *d_sxu = nullptr, nsx = 0;
    int *d_sxm = nullptr;
    size_t sxm_cap = 0;             // moves the schedule buffer holds
    bool graph_stable = false;      // while a batch captures: a launch sequence independent of the draws
    int *d_sched = nullptr;         // [32] the storage sweep's schedule in the graph-stable form
    int *d_counts = nullptr;                      // [ESM_NWEIGHT][T] moves per family and step, then T ints
                                                  // (a parity class's steps in launch order)
    esm::MoveTablesHost mt;                       // templates (host, FP64)
    esm::MoveTablesDev dmt;                       // the same on the device (values in the state precision)
    // the learned dispatch-block proposal (esm_mlp_*; null until esm_mlp_init)
    struct esm::MlpState *mlp = nullptr;
    // kernel timing (esm_profile)
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev[ESM_NKERN];
    std::vector<hipEvent_t> prof_pool;   // events created ahead of the timed work (handed out in order)
    size_t prof_next = 0;
    hipEvent_t prof_event() {            // an event from the pool (a new one when it is used up)
        if (prof_next < prof_pool.size()) return prof_pool[prof_next++];
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        prof_pool.push_back(e);
        prof_next = prof_pool.size();
        return e;
    }
};

namespace esm {

// RAII bracket of one launch with HIP events on the handle's stream (when profiling).
struct ProfScope {
    esm_handle *h;
    int id;
    hipEvent_t b = nullptr;
    ProfScope(esm_handle *h_, int id_) : h(h_), id(id_) {   // id < 0: not timed
        if (h->prof && id >= 0) {
            hipEvent_t a = h->prof_event();
            b = a ? h->prof_event() : nullptr;
            if (b) {
                (void)hipEventRecord(a, h->stream);
                h->prof_ev[id].push_back({a, b});
            }
        }
    }
    ~ProfScope() {
        if (b) (void)hipEventRecord(b, h->stream);
    }
};

template <typename S>
DevModel<S> dev_model(const esm_handle *h);

// k_energy.hip
template <typename S>
int launch_energy(esm_handle *h, bool cost, bool drift);
int launch_reduce_parts(esm_handle *h, double *d_parts /*[R][6]*/);
int launch_reduce_max(esm_handle *h, double *d_drift /*[R]*/);
template <typename S>
int launch_scores(esm_handle *h);
int launch_langevin(esm_handle *h, const double *d_w, const double *d_xi, const double *d_eta, const double *d_ks,
                    double kmin, double clip, int box, double *d_knew, int *d_nclip);

// k_sweep.hip
template <typename S>
int launch_sweeps(esm_handle *h, int nsweeps, uint32_t k0, uint32_t k1, uint32_t ctr0, double maxnats);

// k_random.hip: the random families of one sweep (counter ctr), class by class
template <typename S>
int launch_random(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats);

// k_storage.hip: the storage multiscale sweep (with its probability) of one operation sweep
template <typename S>
int launch_storage_ms(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats);

// k_mlp.hip: release the learned proposal's network and buffers
void mlp_free(esm_handle *h);
template <typename S>
int launch_storage_long(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats);
template <typename S>
int launch_sxlong(esm_handle *h, uint32_t k0, uint32_t k1, uint32_t ctr, double maxnats, double pcoin);
// k_random.hip: the storage pairs' lane form (lanes 32 or 64, kink slots ks) that every template fits
int sx_lane_form(const esm_handle *h, int wmin, int &ks);
size_t sxm_ints(size_t cap);   // ints of the sxlong schedule buffer for `cap` moves (k_cross.hip)

// k_follow.hip
template <typename S>
int launch_follow(esm_handle *h, const void *d_kc0, const void *d_kc1, const int *d_moves, double kmin);

// k_weather.hip
template <typename S>
int launch_weather(esm_handle *h, bool init, int B, int compensate, uint32_t k0, uint32_t k1, uint32_t ctr);
// the centred FFBS of the regimes of the replicas whose mask [R] (host) is set
template <typename S>
int launch_ffbs(esm_handle *h, const int *mask, uint32_t k0, uint32_t k1, uint32_t ctr);

// k_features.hip: the weather features of every replica into out [R][3 + 5 + K] (host)
template <typename S>
int launch_features(esm_handle *h, const double *kref, double *out);

// Copy replica js of hs into replica jd of hd (operation, raw weather and, with `weather`, the
// anomalies and regimes) on hd's stream. Both handles have the same shape and precision.
int launch_copy_replica(const esm_handle *hs, int js, esm_handle *hd, int jd, bool weather);
// The heuristic start state of every replica from the capacities and the raw weather (k_copy.hip).
int launch_heuristic(esm_handle *h);

}  // namespace esm
