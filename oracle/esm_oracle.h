/*
 * esm_oracle.h — FP64 CPU restatement of the reference's per-sweep path (TEST INFRASTRUCTURE).
 *
 * This library is the checker of the HIP path: only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load it. It is never part of the product path.
 *
 * Inputs use the data format of include/esm.h (esm_model) and the state layout
 * [field][node][hour][replica] with replica fastest.
 */
#ifndef ESM_ORACLE_H
#define ESM_ORACLE_H

#include <stdint.h>
#include "../include/esm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* energy_cost (value) of every replica plus the LocalCost caches r [N][T][R] and F [L][T][R]
 * (either may be NULL) and the per-step operational cost per_hour [T][R] (may be NULL). */
int orc_energy(const esm_model *m, int R, const double *v, const double *k, const double *x,
               double *parts, double *per_hour, double *r, double *F);

/* capacity_scores: sk [R][ESM_NK][N], fk [ESM_NK][N]; sk_abs (may be NULL) receives the sums of
 * the absolute values of the summed terms (the scale of the round-off of each entry). */
int orc_scores(const esm_model *m, int R, const double *v, const double *k, const double *x,
               double *sk, double *fk, double *sk_abs);

/* langevin_update of every k field with eta[f] > 0 (see esm_langevin). */
/* log_measure / log_measure_grad (src/energy.jl:790-829); physical = 0: the utilization measure */
double orc_log_measure(const esm_model *m, const double *k, int physical);
void orc_log_measure_grad(const esm_model *m, const double *k, int physical, double *g);
int orc_langevin(const esm_model *m, int R, const double *k, const double *sk, const double *fk,
                 const double *weights, const double *xi, const double *eta, const double *kappa_s,
                 double kmin, double clip, int box, double *knew, int *nclip);

/* Weather process (oracle_weather.c): prior draw of regimes mreg [T][R], anomalies z [3N][T][R]
 * and raw weather x [3][N][T][R]; one weather step of every replica with the compensated
 * re-dispatch of the gas utilization in v (see esm_weather_step). */
int orc_weather_init(const esm_model *m, int R, const uint32_t key[2], uint32_t ctr, int *mreg, double *z, double *x);
int orc_weather_step(const esm_model *m, int R, const double *theta, const double *win_prob, int B, int compensate,
                     const uint32_t key[2], uint32_t ctr, int *mreg, double *z, double *x, double *v, const double *k,
                     int *moved);
/* Explicit-uniform building blocks (the reference's own explicit forms): categorical inversion
 * (0-based result) and the regime window redraw of one regime path m [T] over [a, b] (1-based). */
/* Centred FFBS of the regimes of the masked replicas (mask NULL: all) given z: m [T][R] redrawn
 * from p(m | z); loglik [R] (may be NULL): log p(z) of the forward filter (src/weather.jl:160-174). */
int orc_weather_ffbs(const esm_model *m, int R, const double *z, int *mreg, const int *mask, const uint32_t key[2],
                     uint32_t ctr, double *loglik);
int orc_categorical(const double *p, int K, double u);
void orc_regime_window_path(const esm_model *m, int *path, int a, int b, const double *us);

/* Operation sweeps (oracle_sweep.c), in place on v and the caches r, F (built by orc_energy),
 * with the options of esm_sweep_options (weights NULL: the reference's); stats [ESM_NFAM][5] as
 * esm_sweeps (the evaluation count is the oracle's own); nthreads OpenMP threads over the replicas. */
/* The reference's own sweep schedule (local_sweep! with the plain families, per replica, shuffled
 * per step; no storage sweeps): the reference chain of the chain-marginal exactness check. */
int orc_ref_sweeps(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int nsweeps,
                   const uint32_t key[2], uint32_t ctr0, double maxnats, double extra, const double *weights,
                   int64_t *stats, int nthreads);
int orc_sweeps(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int nsweeps,
               const uint32_t key[2], uint32_t ctr0, double maxnats, double extra, const double *weights, int atoms,
               double p_storage_ms, double p_storage_long, int long_moves, double long_hours, double p_xpair,
               int xlong_moves, int64_t *stats, int nthreads);
/* Cross-node storage pairs (oracle_sweep.c): slot k of node n's pair templates (partner, angle
 * entries, leak nodes and their requirement change per unit, footprint); the partner draw of unit
 * (sa, na) with the uniform u (1, *sb, *nb; 0 none); one pair move of replica j (units A, B at
 * the slot of B's node among A's; instants a+1 … a+L; W = 1 the sxpair summation, 64 sxlong's):
 * op 0 ΔC(t), op 1 draws with given uniforms, op 2 commit t[0]. */
int orc_sx_slot(const esm_model *m, int n, int k, int *part, int *nth, int *thn, double *thv, int *nlk, int *lkn,
                double *lkg, int *nft, int *ft);
int orc_sx_partner(const esm_model *m, const double *k, double hmin, int sa, int na, double u, int *sb, int *nb);
int orc_sx_move(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int sa, int na,
                int sb, int nb, int a, int L, int W, int op, const double *t, int nt, double maxnats, const double *useg,
                const double *uoff, double *out, int *status);
/* One storage multiscale plateau (oracle_sweep.c) on the block b0 … b0 + L of replica j for the
 * units of carrier si at nodes[nu]: ΔC(t), draws with given uniforms (no commit), the commit. */
int orc_sms_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int L, int b0,
                 int si, const int *nodes, int nu, const double *t, int nt, double *dc);
int orc_sms_draw_many(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int L,
                      int b0, int si, const int *nodes, int nu, double maxnats, const double *useg, const double *uoff,
                      int nd, double *t_out, int *status);
int orc_sms_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int L, int b0,
                  int si, const int *nodes, int nu, double t);
/* long-range storage moves of unit (si, n) of replica j: a plateau on instants a+1 … a+L (cyclic,
 * a 0-based), L = T the level move (storage_long.jl) */
int orc_slong_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int si, int n,
                   int a, int L, const double *t, int nt, double *dc);
int orc_slong_draw_many(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int si,
                        int n, int a, int L, double maxnats, const double *useg, const double *uoff, int nd,
                        double *t_out, int *status);
int orc_slong_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int si, int n,
                    int a, int L, double t);
/* Templates of the table-driven random moves: info [8] (templates, first xpair, first cell,
 * levels, cells per level), xp_off [N + 1]; the touched and leak nodes of one template. */
int orc_tables_info(const esm_model *m, int *info, int *xp_off);
int orc_template_nodes(const esm_model *m, int tid, int *nodes, int *nn, int *kk, double *kg, int *nk);
/* One table-driven move (kind 0 cell, 1 xloc, 2 xlocc, 3 xpair) of template tid at (τ, j):
 * the draw with given uniforms (status 3: not valid), ΔC(t), and the commit at t. */
int orc_sparse_draw(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                    int tid, int tau, double maxnats, double useg, double uoff, double *t_out, int *status);
int orc_sparse_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                    int tid, int tau, const double *t, int nt, double *dc);
int orc_sparse_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                     int tid, int tau, double t);
int orc_sparse_draw_many(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int kind,
                         int tid, int tau, double maxnats, const double *useg, const double *uoff, int nd,
                         double *t_out, int *status);
int orc_colouring(const esm_model *m, int *colour);
/* One move of family fam (0 gas, 1 potential, 2 SoC, 3 storage–gas pair) at (n, τ) of replica j
 * with given uniforms (no commit); its ΔC(t) at given t; its commit at t. */
int orc_move_draw(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int fam, int s,
                  int n, int tau, double maxnats, double useg, double uoff, double *t_out, int *status);
int orc_move_cost(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int fam, int s,
                  int n, int tau, const double *t, int nt, double *dc);
int orc_move_apply(const esm_model *m, int R, double *v, const double *k, double *r, double *F, int j, int fam, int s,
                   int n, int tau, double t);
double orc_logseg(double L, double x);
double orc_trexp_offset(double g, double L, double u);

/* follow_capacities! (oracle_follow.c): in place on v (gas utilization) and k. */
int orc_follow(const esm_model *m, int R, double *v, double *k, const double *x, const double *knew, double kmin);

/* Philox4x32-10 (Salmon et al., SC'11): four 32-bit outputs of counter ctr under key key. */
void orc_philox(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif
