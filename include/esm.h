/*
 * esm.h — C ABI of libesmhip.so, the MI355X per-sweep path of the replica sampler.
 *
 * The reference (FHell/ESM-Sampling, Julia) has no FFI for this path: its boundary is the
 * Julia function API that `run_replicas` calls. Every entry point below replaces one of those
 * functions; the citation next to each declaration names it (file:line in the reference).
 *
 * Conventions
 *   - every function returns 0 on success and a negative ESM_E* code on failure; the message
 *     of the last failure of the calling thread is returned by esm_last_error();
 *   - host arrays are caller-owned, FP64, and laid out as documented per argument. Arrays of
 *     the device state use the layout [field][node][hour][replica] (replica fastest), i.e. a
 *     Julia Array{Float64,4} of size (R, T, N, F), the reference device port's [chain, step,
 *     item] order (device/src/core.jl);
 *   - device buffers are owned by the handle; a handle is not thread-safe (one handle per host
 *     thread); all work is ordered on the handle's HIP stream and every call returns after it
 *     has finished (host results are valid on return);
 *   - node values are stored, not Haar coefficients: at the spatially complete level every
 *     basis is orthonormal and square, so the two carry the same information
 *     (docs/system/device.md:29-38). The angle field keeps all N node angles; the gauge mode
 *     it carries does not enter any flow.
 *
 * Tech set of a node (the synthetic system): three renewable carriers as load (solar, onwind,
 * offwind: output A k, curtailment is nodal dump, src/energy.jl:331-337), one dispatchable
 * carrier (gas, utilization field), two storage carriers in the state-of-charge form
 * (battery, H2; src/energy.jl:382-398).
 */
#ifndef ESM_H
#define ESM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ESM_ABI_VERSION 9

/* carriers */
#define ESM_NREN 3   /* renewables as load: 0 solar, 1 onwind, 2 offwind              */
#define ESM_NSTOR 2  /* storage (SoC form): 0 battery, 1 H2                            */
#define ESM_NK 6     /* capacity fields: solar, onwind, offwind, gas, battery, H2      */
#define ESM_NV 4     /* operation fields: 0 theta, 1 gas utilization, 2 battery SoC, 3 H2 SoC */
#define ESM_KGAS 3   /* index of the gas capacity field */
#define ESM_KSTOR 4  /* index of the first storage capacity field */
#define ESM_NPARTS 8 /* CostParts: capex, op, gen, storage, shed, dump, penalty, co2 (src/energy.jl:527-536) */

/* error codes */
#define ESM_OK 0
#define ESM_EINVAL -1    /* invalid argument (shape, pointer, range) */
#define ESM_EHIP -2      /* HIP runtime error */
#define ESM_ENOMEM -3    /* device allocation failed */
#define ESM_ESTATE -4    /* call out of order (e.g. weather step before the weather is drawn) */

/* precision of the device state */
#define ESM_FP64 64
#define ESM_FP32 32

/*
 * The static model of one complete level. All pointers are host FP64 arrays (or int32 for the
 * line ends), read once by esm_create. Layouts: [carrier][node] for per-unit tables; [node][hour]
 * for the load.
 */
typedef struct esm_model {
    int32_t N;                 /* nodes                                                      */
    int32_t L;                 /* AC lines                                                   */
    int32_t T;                 /* steps of the horizon (cyclic: step T is followed by step 1) */
    double dt;                 /* hours per step (Δt)                                        */
    const int32_t *line_bus0;  /* [L] 0-based                                                */
    const int32_t *line_bus1;  /* [L] 0-based                                                */
    const double *line_b;      /* [L] susceptance: F = b (θ_a − θ_b) in MW                   */
    const double *line_s;      /* [L] thermal limit (MW), hard: cost c_viol Δt (|F| − s)₊      */
    const double *load;        /* [N][T] raw load (MW); the data map is max(x, 0)            */
    const double *mc;          /* [ESM_NK][N] marginal cost (€/MWh) of every carrier          */
    const double *ef;          /* [N] gas emission factor (t CO2 per MWh)                    */
    const double *capex;       /* [ESM_NK][N] annualised capex (€/MW)                        */
    const double *klo;         /* [ESM_NK][N] capacity box, lower (MW)                       */
    const double *khi;         /* [ESM_NK][N] capacity box, upper (MW)                       */
    const double *stor_hours;  /* [ESM_NSTOR][N] energy-to-power ratio (h)                   */
    const double *stor_eta_in; /* [ESM_NSTOR][N] charging efficiency                         */
    const double *stor_eta_out;/* [ESM_NSTOR][N] discharging efficiency                      */
    const double *stor_qt;     /* [ESM_NSTOR][N] inflow per unit power                       */
    double c_shed;             /* €/MWh of positive residual                                 */
    double c_dump;             /* €/MWh of negative residual                                 */
    double c_viol;             /* price of constraint violations                             */
    double gamma_k;            /* capacity box penalty factor                                */
    double kappa0;             /* capacity floor of the penalty weights (MW)                 */
    double beta_v;             /* operational temperature (1/€)                              */
    double beta_k;             /* investment temperature (1/€)                               */
    /* weather: ESM_NREN*N capacity-factor series i = c*N + n, logit CF = xbar + z with z a
     * K-regime VAR(1), z_τ = φ z_{τ-1} + μ_{m_τ}(τ) + σ e_τ, z_1 = σ0 e_1 (src/weather.jl:461-481
     * with Φ = φ I, Σ_k = σ² I, Σ0 = σ0² I and harmonics (1, cos 2πτ/T, sin 2πτ/T)).      */
    int32_t K;                 /* regimes                                                    */
    const double *Ptr;         /* [K][K] regime transition matrix (row: from)               */
    const double *pi0;         /* [K] initial regime law                                     */
    double phi;                /* AR coefficient                                             */
    double sigma;              /* innovation standard deviation                             */
    double sigma0;             /* standard deviation of z at step 1                         */
    const double *mu;          /* [K][ESM_NREN*N][3] regime intercept, cos and sin coefficients */
    const double *xbar;        /* [ESM_NREN*N][24] climatological logit by hour of day      */
    /* partitions of the graph-Haar hierarchy between the whole system and the single nodes
     * (E.parts, src/energy.jl:292: [one cell, country, node]): the cell moves act on the cells
     * of every level. country [N] in 0..ncountry-1; NULL (or ncountry = 0): no middle level.  */
    int32_t ncountry;
    const int32_t *country;
} esm_model;

typedef struct esm_handle esm_handle;

/* Library version (ESM_ABI_VERSION) and the message of the last failure on this thread. */
int32_t esm_version(void);
const char *esm_last_error(void);

/*
 * Create a handle for `nrep` replicas (chains × replicas) of one model, with the device state in
 * `precision` (ESM_FP64 or ESM_FP32). Replaces the set-up of `run_replicas`
 * (src/replicas.jl:841-900: `setup`, `LocalModel`, `init_replicas`) as far as the device state
 * goes. The state starts at zero; upload one with esm_upload_state.
 */
int32_t esm_create(const esm_model *model, int32_t nrep, int32_t precision, esm_handle **out);
int32_t esm_destroy(esm_handle *h);

/*
 * Device state transfer. v: [ESM_NV][N][T][R] node values of the operation fields; k: [ESM_NK][N]
 * raw node capacities shared by every replica (src/replicas.jl:702 keeps them bitwise equal);
 * x: [ESM_NREN][N][T][R] raw node weather (capacity factors before the data map clamp).
 * Any pointer may be NULL to leave that part unchanged / not downloaded. Uploading v, k or x
 * rebuilds the residual and flow caches.
 */
int32_t esm_upload_state(esm_handle *h, const double *v, const double *k, const double *x);
int32_t esm_download_state(esm_handle *h, double *v, double *k, double *x);

/*
 * The heuristic start of every replica on the device (the role of `heuristic_start`,
 * src/run.jl:469, for this system): zero angles, the gas utilization covering the node's net
 * load, clamp((load − Σ_c k_c x_c)/k_gas, 0, 1) from the raw capacities and the raw weather of
 * the replica, both states of charge 0.5; the caches are rebuilt. No host arrays: for replica
 * sets whose state is too large to stage on the host (e.g. 256 nodes × 192,720 steps × 32
 * replicas). Usually after esm_weather_init.
 */
int32_t esm_heuristic_state(esm_handle *h);

/*
 * `energy_cost` of every replica (src/energy.jl:551-677) and `LocalCost` rebuild
 * (src/local.jl:313-374): rebuilds the residual r [N][T][R] and flow F [L][T][R] caches from
 * the node arrays, and writes per replica the CostParts [R][ESM_NPARTS] (capex, op, gen,
 * storage, shed MWh, dump MWh, penalty, co2). `per_hour` ([T][R], may be NULL) receives the
 * operational cost of every step (`cache_cost` terms, src/local.jl:2065-2098). Terms are
 * evaluated in the state precision; totals are accumulated in FP64.
 */
int32_t esm_energy(esm_handle *h, double *parts, double *per_hour);

/*
 * Rebuild the residual and flow caches from the node arrays (`rebuild_flat`,
 * device/src/residual.jl:16) and report per replica the largest drift max|r_cache − r| [R]
 * (may be NULL).
 */
int32_t esm_rebuild_residual(esm_handle *h, double *drift);

/* Download the caches: r [N][T][R] and F [L][T][R] (either may be NULL). */
int32_t esm_download_caches(esm_handle *h, double *r, double *F);

/*
 * `capacity_scores` (src/replicas.jl:329-338 via `energy_cost(grad = true)`,
 * src/energy.jl:680-757): per replica the sampled score ŝ_k = −β_v ∂C/∂k (node capacity
 * coordinates; 0 where the capacity is clamped), sk [R][ESM_NK][N], and the capex force
 * f_k = β_k ∂K/∂k, fk [ESM_NK][N] (replica independent). Uses the residual cache, so call
 * esm_energy or esm_rebuild_residual first when the caches may have drifted.
 */
int32_t esm_scores(esm_handle *h, double *sk, double *fk);

/*
 * The reference measure of the capacity score (`measure`, src/run.jl:36; `log_measure_grad`,
 * src/energy.jl:811-829, added to the score by `capacity_scores`, src/replicas.jl:334):
 * ESM_MEASURE_UTILIZATION (the default) leaves ŝ_k = −β_v ∂C/∂k; ESM_MEASURE_PHYSICAL measures
 * the dispatch of the dispatchable unit and of the storage units in MW (dp = k dω), which adds
 * ∂_k log J = T/(k + κ0) to their fields (gas, battery, H2) where the capacity is strictly inside
 * its box (T utilization coordinates per unit; renewables keep the utilization measure). Applies
 * to the following esm_scores calls.
 */
#define ESM_MEASURE_UTILIZATION 0
#define ESM_MEASURE_PHYSICAL 1
int32_t esm_score_measure(esm_handle *h, int32_t measure);

/*
 * `langevin_capacities` / `langevin_update` (src/replicas.jl:368-403) of the shared
 * capacities, in FP64 on the device: per k field f with eta[f] > 0,
 * w = log(k + κ_s), w' = w + clamp(η((k + κ_s) g + 1), ±clip) + √(2η) ξ, with
 * g = −fk + (β_k/β_v) Σ_j ω_j sk_j, reflected at log(kmin + κ_s) and (box != 0) at
 * log(max(khi, kmin + 1) + κ_s) when crossed from below. weights [R] (ω_j; all replicas),
 * xi [ESM_NK][N] standard normals (fields with eta = 0 draw none and are ignored), eta and
 * kappa_s [ESM_NK]. Reads the scores of the last esm_scores call. Writes knew [ESM_NK][N] and
 * *nclip (may be NULL); does not change the state's k (see esm_follow_capacities).
 */
int32_t esm_langevin(esm_handle *h, const double *weights, const double *xi, const double *eta,
                     const double *kappa_s, double kmin, double clip, int32_t box,
                     double *knew, int32_t *nclip);

/*
 * esm_langevin without a device allocation. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The same step, kernel, arguments and results as esm_langevin; its
 * staging (weights, normals, steps, offsets, the new capacities, the clip counter) lives in a
 * fixed device array of ESM_NOALLOC_DOUBLES doubles (64 KiB, held by every device the library
 * runs on until the process ends, its address looked up by esm_create) instead of a per-call
 * allocation. Limit: R + 12 N + 13 <= ESM_NOALLOC_DOUBLES, else ESM_EINVAL. Calls on one device
 * share the array: calls that overlap on one device (several threads or streams) race; every
 * call clears the clip counter.
 */
#define ESM_NOALLOC_DOUBLES 8192
int32_t esm_langevin_noalloc(esm_handle *h, const double *weights, const double *xi, const double *eta,
                             const double *kappa_s, double kmin, double clip, int32_t box,
                             double *knew, int32_t *nclip);

/* move families of the sweep statistics */
#define ESM_FAM_GEN 0     /* gas utilization atoms                                   */
#define ESM_FAM_THETA 1   /* voltage-angle atoms                                     */
#define ESM_FAM_SOC 2     /* state-of-charge atoms                                   */
#define ESM_FAM_SPAIR 3   /* random: storage–gas pair (prep_spair!)                  */
#define ESM_FAM_UNIT 4    /* random: one unit at one node (prep_random! :unit)       */
#define ESM_FAM_POT 5     /* random: the angle of one node (:potential)              */
#define ESM_FAM_CELL 6    /* random: one carrier over a cell of a partition (:cell)  */
#define ESM_FAM_XLOC 7    /* random: localized cross-node pair, leak to the residuals */
#define ESM_FAM_XLOCC 8   /* random: localized pair, leak compensated pro rata        */
#define ESM_FAM_XPAIR 9   /* random: cross-node pair, global re-routing               */
#define ESM_FAM_SMS 10    /* storage multiscale: block plateaus (storage_ms_sweep!)   */
#define ESM_FAM_SLONG 11  /* long-range storage plateau (slong, storage_long_sweep!)  */
#define ESM_FAM_SLEVEL 12 /* level move of a storage trajectory (slevel)              */
#define ESM_FAM_SXPAIR 13 /* random: cross-node storage pair at one instant (sxpair)  */
#define ESM_FAM_SXLONG 14 /* interval storage pair of the long-range sweep (sxlong)   */
#define ESM_NFAM 15
#define ESM_NSTAT 5       /* attempted, accepted, improper, guard, ΔC evaluations     */
/* the plain random families in the reference's order and default weights
 * (plain_families, src/local.jl:1912-1915): unit 0.3, pair 0.25, xpair 0.03, xlocc 0.08,
 * xloc 0.04, potential 0.1, cellpair 0.1, cell 0.1, spair 0.3                            */
#define ESM_NWEIGHT 9

/*
 * `operation_sweep!` -> `local_sweep!` (src/local.jl:1927-2062, :2401-2425) of every replica,
 * `nsweeps` times, at fixed capacities and weather: exact line-conditional moves drawn from
 * exp(−β_v ΔC(t)) (`sample_line!`, :1501-1560; `maxnats` is the walk's stop, 40 in the
 * reference). One sweep runs, in this order:
 *   atoms (unless switched off by esm_sweep_options), in conflict-free rounds: gas
 *     utilization at every (node, step); voltage angles, one round per colour of a distance-2
 *     colouring; states of charge per storage carrier and time-parity class;
 *   random moves, per time-parity class: at every step round(extra·N) draws over the plain
 *     families with their weights (pair and cellpair are void in a system with one
 *     dispatchable unit per node and one injecting carrier per cell), run family by family,
 *     the draws of a step in sequence; the selection of a move (family, node, partner, cell)
 *     depends on the step and the sweep only, the line draw on the replica.
 * Uniforms are Philox4x32-10 under `key`: atoms with counter (replica, node·T + step, round
 * id, counter + sweep index); the family draws with (2^31 + draw, step, 9000, ·), the
 * selection of the index-th move of family w (its position in the weight list) with
 * (2^31 + w·2^20 + index, step, 9001, ·), its line draw with (replica, step,
 * 2^30 + w·2^20 + index, ·) and the Jacobian test of the compensated pairs with (replica,
 * step, 3·2^29 + w·2^20 + index, ·). The residual and flow caches
 * are updated incrementally. stats ([ESM_NFAM][ESM_NSTAT], may be NULL): per family the moves
 * attempted, accepted (t ≠ 0), improper lines (t = 0), walks cut at the point cap, drawn
 * non-finite or rejected by the Jacobian test (t = 0), and the ΔC(t) evaluations made.
 */
int32_t esm_sweeps(esm_handle *h, int32_t nsweeps, const uint32_t key[2], uint32_t counter, double maxnats,
                   int64_t *stats);

/*
 * Options of the operation sweep (`local_sweep!` keywords, src/local.jl:1927-1929): `extra`
 * random moves per node and step (`local_extra`, 4 by default, src/run.jl:21), the weights of
 * the plain families (ESM_NWEIGHT, in the order above; NULL: the reference's), whether the
 * atom rounds run (`atoms`), and the probability of the storage multiscale sweep after them
 * (`operation_sweep!`, src/local.jl:2414-2418: 1/2). Defaults hold from esm_create on.
 *
 * The storage multiscale sweep (`storage_ms_sweep!`, src/flowmoves.jl:255-324): for every block
 * length L in {2, 4, …, 128} below T/2, in an order drawn per sweep, with an offset o drawn in
 * [0, L), the blocks b0 = o + iL (steps b0 … b0 + L, no wrap) in two parity classes drawn in
 * either order; 8 plateau moves per block on one storage unit or (probability 1/2) one carrier's
 * units in one country, compensated pro rata at the block ends, with the Jacobian test. Draws
 * (Philox under the sweep's key and counter c): whether it runs (0, 0, 9300, c); the length
 * order (i, 0, 9301, c), Fisher–Yates; offset and class order (L, 0, 9302, c); the units of
 * move m of the block at b0 (m, b0, 9200 + L, c); its line draw (replica, b0,
 * 0x50000000 + 16L + m, c) and Jacobian test (replica, b0, 0x70000000 + 16L + m, c).
 */
int32_t esm_sweep_options(esm_handle *h, double extra, const double *weights, int32_t atoms, double p_storage_ms);

/*
 * Options of the long-range storage sweep (`storage_long_sweep!`, src/storage_long.jl:245-298),
 * which the operation sweep runs after the multiscale sweep with probability `p`
 * (`storage_long`, 1 by default, src/run.jl:27-29): for every storage unit with at least `hmin`
 * hours (`storage_long_hours`, 24) and κs = k h/Δt ≥ 1, a level move (every instant of the year
 * rises by t) and `nmoves` (`storage_long_moves`, 32; at most 256) long-range plateaus (the state
 * of charge rises at the cyclic instants a+1 … a+L, L uniform on 1 … T−1, a uniform; the node's
 * gas unit follows pro rata at both end steps, with the Jacobian test), per node in a shuffled
 * order. Draws (Philox under the sweep's key and counter c): whether it runs (0, 0, 9410, c);
 * for node n and its u-th unit taking part the level move (n, u, 9411, c) and plateau m's L and a
 * (n, 65536u + m, 9412, c); the node's list shuffled by Fisher–Yates with (n, i, 9413, c); move
 * i's line draw (replica, n, 0x54000000 + i, c) and Jacobian test (replica, n,
 * 0x74000000 + i, c). Statistics: families ESM_FAM_SLONG (plateaus) and ESM_FAM_SLEVEL.
 */
int32_t esm_storage_long_options(esm_handle *h, double p, int32_t nmoves, double hmin);

/*
 * Cross-node storage pairs (src/storage_cross.jl; `operation_sweep!` keywords `sxpair` and
 * `nxmoves`, src/local.jl:2414-2422; configured by `storage_xpair` 0.1 and `storage_xlong` 8,
 * src/run.jl:30-31; both 0, off, from esm_create on). A pair moves energy between unit A and a
 * unit B at A's node or a neighbour (partner hops 1, ball radius 1: `storage_xpair_hops`,
 * `storage_xpair_ball`): A's state of charge rises by t and B's falls by ρ t (ρ = E_A/E_B) at the
 * instants a+1 … a+L; at the end steps the flow between them is re-routed on the ball of radius
 * one around the pair (the whole component when the ball covers it) and the leak into the first
 * layer outside is compensated pro rata by the gas unit of each leak node, with the Jacobian test.
 * The units taking part have at least the long-range sweep's `hmin` hours; B is drawn with
 * probability ∝ min(k_A, k_B) among those with κ = k h/Δt ≥ 1 (`sx_partner`, :62-90).
 *   p_xpair  weight of `sxpair` (L = 1, instant τ: steps τ−1 and τ) among the plain families of
 *            the local sweep; the draws per step are scaled by Σ weights / Σ plain weights so that
 *            every plain family keeps its rate (local_sweep!, :1934-1941). Family weight position
 *            9 in the draw counters: selection (2^31 + 9·2^20 + index, step, 9001, ·) with A from
 *            word 0 and the partner from words 2–3; line draw and Jacobian test as the other
 *            families. Statistics ESM_FAM_SXPAIR.
 *   nxlong   `sxlong` moves per unit of the pairs with κ ≥ 1 in the long-range sweep, after its
 *            plateaus whenever it runs (0 … 256): partner, a uniform on 0 … T−1 and L uniform on
 *            1 … T−1 from (u, m, 9420, c) for move m of the unit with index u among the units of
 *            the pairs; the list shuffled with (i, 0, 9421, c) and run in rounds of disjoint
 *            footprints (first fit, `sx_rounds` :470-492); the i-th move of the round-ordered list
 *            draws its line with (replica, i, 0x5C000000, c) and its Jacobian test with (replica,
 *            i, 0x7C000000, c). Statistics ESM_FAM_SXLONG.
 * Units that only discharge (eta_in = 0) cannot take part (the loss share of `sx_loss` is not
 * built): ESM_EINVAL. At most 256 nodes.
 */
int32_t esm_storage_cross_options(esm_handle *h, double p_xpair, int32_t nxlong);

/*
 * How the random families of an operation sweep are launched (no counterpart in the reference,
 * which runs them on CPU threads): merged = 1 (the default) runs the table-driven families of a
 * step (xpair, xlocc, xloc, cell) in one launch per parity class; merged = 0 launches one kernel
 * per family and parity class (per-family kernel times for esm_profile_read); merged = 2, 4, 8 or
 * 16 runs the merged launch with that many waves per (step, 64 replicas), which split every
 * move's template entries (replica counts that are multiples of 64; merged = 1 uses one wave).
 * States and statistics of 0 and 1 are identical; more waves sum the terms of ΔC in another
 * order.
 */
int32_t esm_sweep_launch(esm_handle *h, int32_t merged);

/*
 * The learned dispatch-block proposal (no counterpart in the reference; the extension that the
 * plain line-conditional moves judge): a move replaces the gas utilization of node n over the B
 * hours of block d (steps dB … dB+B−1, T/B blocks per node) of one replica by a draw from
 * q(·|c), a product over the hours of mixtures of K Gaussians given by an MLP of the block's
 * context c (the
 * requirement r + κ cl(y) the unit meets, per hour over κ; the node's three capacity factors per
 * hour; its six capacities over the system means; a constant; per hour the node's net line
 * inflow over κ and the largest loading |F|/s of its lines — nothing that depends on the
 * block's y), accepted with probability min(1, exp(−β_v ΔC(y') + log q(y|c) − log q(y'|c)))
 * (ΔC: the block's gas terms and node balances, exact; FP64 ratio), so the target is unchanged.
 * Network: 160 inputs (6B + 7 used), three hidden layers of 256 (ReLU), dout = 3KB outputs rounded up to a
 * multiple of 64 ([logits K·B][means K·B around cl(Q/κ)][log sd K·B, clamped to [−7, 2]],
 * index k·B + b; the weights of hour b are the softmax over k), BF16 weights and activations
 * with FP32 accumulation (MFMA), FP32 master weights, Adam.
 *
 * esm_mlp_init: a new network for blocks of B hours (1 … min(24, T)) with K components (1 … 8),
 * weights from Philox under `key`. esm_mlp_dims: [B, K, din, dh, dout, blocks per node, number
 * of parameters]. esm_mlp_params: get and/or set the FP32 master parameters (per layer
 * l = 1 … 4: Wt [out][in] then the bias [out]); either pointer may be NULL.
 * esm_mlp_forward: the outputs [M][dout] of M ≤ 65536 explicit input rows x [M][160] (rounded
 * to BF16). esm_mlp_gradient: for explicit rows x, the blocks' y [M][B] and the centres mu0
 * [M][B], the loss Σ −log q(y|c) and the gradient of its mean over the rows (parameter layout).
 * esm_mlp_train: `steps` Adam steps (β = 0.9, 0.999) with learning rate lr on minibatches of
 * `batch` ≤ 65536 instances drawn from the current state of every replica (instance i of step s:
 * Philox (i, s, 0x5c000000, counter)); loss [steps] (may be NULL): each step's mean loss.
 * esm_mlp_propose: one learned move of every (node, block, replica), uniforms Philox (replica,
 * node·nb + block, ·, counter): 0x5a000000 + b/2 for the normals of hours b, b+1, 0x5e000000 +
 * b/2 for their components, 0x5b000000 (words 2–3) for the acceptance; stats [ESM_NSTAT]
 * (attempted, accepted, 0, 0, network evaluations); out [n_inst][dout], yprop [n_inst][B],
 * logratio [n_inst] (may be NULL) receive the network outputs, proposals and MH log-ratios,
 * instance i = (node·nb + block)·R + replica.
 */
int32_t esm_mlp_init(esm_handle *h, int32_t block_hours, int32_t ncomp, const uint32_t key[2]);
int32_t esm_mlp_dims(esm_handle *h, int32_t *dims);
int32_t esm_mlp_params(esm_handle *h, double *get, const double *set);
/* esm_mlp_optimizer: get and/or set Adam's state [m (np)][v (np)][t] (np from esm_mlp_dims; t the
 * steps taken), so that a snapshot resumes training exactly; either pointer may be NULL. The
 * learned proposal has no counterpart in the reference; this serves its driver's snapshot
 * (`write_replica_snapshot`, src/replicas.jl:1082-1124, extended by the network's state). */
int32_t esm_mlp_optimizer(esm_handle *h, double *get, const double *set);
/* esm_mlp_options: the forward pass of esm_mlp_propose and esm_mlp_forward as one fused kernel
 * per 128-row tile (fused = 1, the default) or as one kernel per layer (0); the outputs are the
 * same bit for bit. No counterpart in the reference (the learned proposal is new). */
int32_t esm_mlp_options(esm_handle *h, int32_t fused);
int32_t esm_mlp_forward(esm_handle *h, const double *x, int64_t M, double *out);
int32_t esm_mlp_gradient(esm_handle *h, const double *x, const double *y, const double *mu0, int64_t M, double *grad,
                         double *loss);
int32_t esm_mlp_train(esm_handle *h, int32_t steps, int32_t batch, double lr, const uint32_t key[2], uint32_t counter,
                      double *loss);
int32_t esm_mlp_propose(esm_handle *h, const uint32_t key[2], uint32_t counter, int64_t *stats, double *out,
                        double *yprop, double *logratio);

/*
 * `follow_capacities!` (src/replicas.jl:448-472) of every replica: move the shared capacities
 * from the current ones to knew [ESM_NK][N] with the dispatch following: gas at fixed output
 * (utilizations × c(k0)/c(k1), :412-435), then every renewable and storage unit with
 * k0 >= kmin along its kop curve at u = log(knew/k0) (src/kop.jl:200-271: fixed utilizations,
 * the node's gas unit following share-preservingly); units below kmin move their capacity
 * alone. The caches are rebuilt. kout [ESM_NK][N] (may be NULL) receives the capacities as
 * committed: k0 + (k0 e^u − k0), the same bits in every replica and every call with the same
 * inputs (the reference asserts this, :702).
 */
int32_t esm_follow_capacities(esm_handle *h, const double *knew, double kmin, double *kout);

/* The sweep schedule: the colour of every node [N] and the number of colours. */
int32_t esm_schedule(esm_handle *h, int32_t *colour, int32_t *ncolours);

/*
 * Weather process of every replica (src/replicas.jl:29-236, src/weather.jl:461-481), with
 * counter-based Philox4x32-10 random numbers under `key`; `counter` must differ between calls
 * that should draw independent numbers (the streams of the reference are per replica,
 * src/replicas.jl:872-874; here a replica's numbers are addressed by its index).
 *
 * esm_weather_init: an exact prior draw per replica (`sample_weather`): the regime chain from
 * π0 and Ptr, then the VAR(1) recursion; sets the raw weather x = sigmoid(xbar + z) and
 * rebuilds the caches. The operation state v is left as it is.
 */
int32_t esm_weather_init(esm_handle *h, const uint32_t key[2], uint32_t counter);

/*
 * esm_weather_step: one step of every replica's weather process (`replica_weather!` ->
 * `weather_process_step`, src/replicas.jl:179-197 and :586-607): a pCN step of angle theta[j]
 * in the innovations at fixed regimes (theta[j] = 0: none), then, with probability
 * win_prob[j] (one uniform of the replica), a regime window move of length B (exact Gibbs
 * redraw of the regimes in a uniformly placed window by backward messages, z re-built from its
 * innovations at the new regimes). With `compensate`, `compensate_weather!` follows: at every
 * node and step the gas unit absorbs the change of the residual (up to its headroom, or its
 * output). moved (may be NULL) receives [R] flags: 1 pCN only, 2 window (and pCN if theta > 0),
 * 0 untouched.
 */
int32_t esm_weather_step(esm_handle *h, const double *theta, const double *win_prob, int32_t B,
                         int32_t compensate, const uint32_t key[2], uint32_t counter, int32_t *moved);

/*
 * esm_weather_ffbs: the centred FFBS of the regimes (`replica_ffbs = true`: `fine_regime_step!`,
 * src/weather_fine.jl:290-300, by `ffbs`, src/weather.jl:160-174) of every replica whose
 * mask[j] != 0 (mask NULL: all): the regime path is redrawn from p(m | z) given the anomalies —
 * the emission log-likelihoods of every step under every regime (`emission_loglik`, :117-133),
 * the forward filter, then the backward draws (`sample_log`, :176-185) with the uniforms Philox
 * (replica, τ, 0x46464253, counter) words 0–1, in FP64. The anomalies, the weather and the
 * dispatch are unchanged (at the complete level there are no detail levels to re-derive).
 */
int32_t esm_weather_ffbs(esm_handle *h, const int32_t *mask, const uint32_t key[2], uint32_t counter);

/* Regimes m [T][R] (0-based) and the VAR anomalies z [ESM_NREN*N][T][R] (either may be NULL). */
int32_t esm_download_weather(esm_handle *h, int32_t *m, double *z);

/*
 * esm_replica_features: the weather features of every replica (`replica_features`,
 * src/replicas.jl:270-273, over `weather_features`, :250-267), the rows of the driver's
 * `replicas/<c>/features` record. kref [ESM_NK][N]: the fixed reference capacities (the
 * renewable fields 0 … ESM_NREN−1 are read). out [R][ESM_NREN + 5 + K] (FP64), per replica in
 * the order of `replica_feature_names`: load_GW (mean over t of Σ_n max(load, 0), /1e3), cf of
 * solar, onwind, offwind (mean of clamp(x, 0, 1) over nodes and steps), residual load
 * res_t = (D_t − Σ kref·A_t)/1e3: its mean, peak and worst cyclic 7-step mean (GW), the
 * fraction of steps in each regime 1 … K, the number of regime switches.
 */
int32_t esm_replica_features(esm_handle *h, const double *kref, double *out);
/* Set the regimes and anomalies; x = sigmoid(xbar + z) follows and the caches are rebuilt. */
int32_t esm_upload_weather(esm_handle *h, const int32_t *m, const double *z);

/*
 * Copy replica js of `src` into replica jd of `dst` (same N, L, T and precision; the handles may
 * be the same): the operation fields, the raw weather and, when both hold a weather process,
 * the anomalies and regimes; `dst` takes the capacities of `src` and its caches are rebuilt.
 * Replaces `copy(s)` of one replica's State for the lag probe, which sweeps a copy of a replica
 * at frozen (k, x) (src/replicas.jl:733-750).
 */
int32_t esm_copy_replica(esm_handle *src, int32_t js, esm_handle *dst, int32_t jd);

/*
 * Kernel timing with HIP events recorded on the handle's stream around every launch of the
 * named kernel (the stage timing of the reference, src/stages.jl:26-53, at kernel grain).
 * esm_profile(h, 1) clears the records and starts recording, esm_profile(h, 0) stops;
 * esm_profile_read returns the summed event time (ms) and the launch count of one kernel.
 */
#define ESM_KERN_ENERGY 0  /* energy, flows and residual caches          */
#define ESM_KERN_SCORE 1   /* capacity score partial sums                */
#define ESM_KERN_SWEEP 2   /* exact line moves (all rounds of a sweep)   */
#define ESM_KERN_WEATHER 3 /* weather process                            */
#define ESM_KERN_GEN 4     /* one round of gas atoms                     */
#define ESM_KERN_THETA 5   /* one round of potential atoms (one bucket)  */
#define ESM_KERN_SOC 6     /* one round of SoC atoms                     */
#define ESM_KERN_SPAIR 7   /* random storage-gas pairs of one parity class */
#define ESM_KERN_FOLLOW 8  /* dispatch following                         */
#define ESM_KERN_UNIT 9    /* random unit moves of one parity class      */
#define ESM_KERN_POT 10    /* random potential moves                     */
#define ESM_KERN_CELL 11   /* random cell moves                          */
#define ESM_KERN_XLOC 12   /* random localized pairs                     */
#define ESM_KERN_XLOCC 13  /* random compensated localized pairs         */
#define ESM_KERN_XPAIR 14  /* random re-routed pairs                     */
#define ESM_KERN_SMS 15    /* storage multiscale moves of one block length and class */
#define ESM_KERN_SLONG 16  /* long-range storage moves (one launch per sweep)         */
#define ESM_KERN_MLPF 17   /* one MFMA GEMM of the learned proposal's forward pass (4 per chunk) */
#define ESM_KERN_MLPP 18   /* the learned moves of one chunk (draw, MH ratio, commit)         */
#define ESM_KERN_MLPT 19   /* one training step (features, forward, head, backward, Adam)     */
#define ESM_KERN_RANDOM 20 /* the table-driven random families of one parity class, merged    */
#define ESM_KERN_SXPAIR 21 /* the storage pairs sxpair of one parity class                     */
#define ESM_KERN_SXLONG 22 /* the interval storage pairs sxlong of a long-range sweep          */
#define ESM_NKERN 23
int32_t esm_profile(esm_handle *h, int32_t enable);
int32_t esm_profile_read(esm_handle *h, int32_t kernel, double *total_ms, int64_t *launches);

/*
 * Batched sweeps and the launch autotuner (no counterpart in the reference, whose chains run on
 * CPU threads; the batching extension of this library).
 *
 * A batch holds up to ESM_BATCH_MAX handles on one device, of any network size, horizon,
 * replica count and precision. esm_batch_sweeps runs esm_sweeps on every handle, with its own
 * key (keys[2i], keys[2i+1]) and counter (counters[i]), as ONE HIP graph: the launches of each
 * handle are captured on its own stream, forked from and joined into the first handle's stream,
 * so the replica sets run as concurrent branches of one graph launch. The states, caches and
 * statistics are exactly those of esm_sweeps on each handle (the same kernels, arguments and
 * order). stats ([n][ESM_NFAM][ESM_NSTAT], may be NULL) as esm_sweeps. The executable graph is
 * kept and updated in place on later calls (instantiated again when the launch sequence
 * changes); esm_batch_counts gives (graph launches, instantiations, in-place updates). The
 * handles must not be used by other threads during the call, and kernel profiling
 * (esm_profile) must be off on them. Destroying a batch leaves its handles alone; destroy it
 * before any of them.
 *
 * esm_sweep_block sets the workgroup size of the one-wave random launches (64, 128 or 256, the
 * default); esm_sweep_pad(h, 1) gives every step of those launches whole waves (R rounded up to a
 * multiple of 64, the extra lanes idle), so that a wave holds the replicas of one step (the
 * wave-uniform selection and the heaviest-first order of steps) when R is not a multiple of 64
 * (the default; 0: R threads per step). States and statistics depend on neither.
 *
 * esm_autotune picks the launch of the random families for this handle's shape by timing 12
 * candidates: with R a multiple of 64, c = 3·f + b over the forms f of esm_sweep_launch (0, 1, 2,
 * 4) and the blocks b (64, 128, 256); otherwise c = 6·p + 3·f + b over the padding p (0, 1), the
 * forms 0 and 1 and the blocks. Each runs `nsweeps` sweeps (1 … 64) under `key` and `counter` from
 * the handle's current state (after one untimed sweep), timed with HIP events. The state is
 * restored afterwards: the call leaves v, the caches and the chain unchanged (the statistics of
 * esm_sweeps are overwritten). Sets the fastest candidate on the handle and returns it in best[0]
 * (form), best[1] (block) and best[2] (padding); ms[c] (may be NULL; 12 entries) the mean ms per
 * sweep of each candidate, *ncand (may be NULL) their number.
 */
#define ESM_BATCH_MAX 64
typedef struct esm_batch esm_batch;
int32_t esm_batch_create(esm_handle *const *handles, int32_t n, esm_batch **out);
int32_t esm_batch_destroy(esm_batch *b);
int32_t esm_batch_sweeps(esm_batch *b, int32_t nsweeps, const uint32_t *keys, const uint32_t *counters, double maxnats,
                         int64_t *stats);
int32_t esm_batch_counts(const esm_batch *b, int64_t *counts);
int32_t esm_sweep_block(esm_handle *h, int32_t block);
int32_t esm_sweep_pad(esm_handle *h, int32_t pad);
int32_t esm_autotune(esm_handle *h, int32_t nsweeps, const uint32_t key[2], uint32_t counter, double maxnats,
                     int32_t *best, double *ms, int32_t *ncand);

#ifdef __cplusplus
}
#endif

#endif /* ESM_H */
