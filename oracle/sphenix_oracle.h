/*
 * sphenix_oracle.h -- CPU oracle for the SPHENIX neighbour pass.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (swift_amd/, include/)
 * includes, links or calls this code; only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py use it, as the checker.
 *
 * It restates, in plain C, SWIFT's SPHENIX density -> ghost -> gradient ->
 * extra ghost -> force -> end-force pass (no cosmology, Wendland C2, 3D,
 * gamma = 5/3) in two forms:
 *
 *  - orc_pass_fp64: the "FP64 rerun of the same pass".  Given the final
 *    smoothing lengths h, every accumulation and every per-particle finisher
 *    is evaluated in double precision.  Only the neighbour test uses the
 *    FP32 pair contract, so that the neighbour sets are the ones an FP32
 *    implementation forms.
 *  - orc_pass_fp32: an FP32 restatement of the same pass including the
 *    smoothing-length Newton-Raphson / bisection ghost with subset re-runs
 *    (runner_ghost.c:1113-1628).
 *
 * The FP32 pair contract (shared, by specification, with the HIP kernels)
 * is the expression of the reference's own loops and brute-force checkers
 * for top-level cells (cells of 1/cdim of the box; a particle's cell is
 * (int)(x * iwidth), space_cell_index.c:114-128):
 *   same cell   dx_k = (float)(x_i,k - x_j,k)             (DOSELF1 :2459,
 *                                                  self_all_density tools.c:617)
 *   other cells dx_k = (float)(x_i,k - oi_k) - (float)(x_j,k - oj_k), where
 *               oi, oj are the corner of the "upper" cell of the pair (the
 *               second cell after space_getsid_and_swap_cells, space_getsid.h:
 *               47-80) with the periodic shift applied on the side of the
 *               other cell                       (DOPAIR1 :1327-1346, DOPAIR2
 *                                          :1672-1781, pairs_all_* tools.c:218)
 *   r2     = fmaf(dx2, dx2, fmaf(dx1, dx1, dx0 * dx0))  (the sum of squares as
 *            gcc contracts it under -ffp-contract=fast)
 *   H2_i   = (h_i * h_i) * kernel_gamma2        (two FP32 roundings)
 *   density / gradient neighbour : r2 < H2_i
 *   force neighbour              : r2 < H2_i || r2 < H2_j
 * The reference's ghost re-runs (DOPAIR_SUBSET :891-914, DOSELF_SUBSET
 * :1135-1163) use the two expressions the other way round; the brute-force
 * checkers, which define the neighbour sets here, do not.
 */
#ifndef SWSPH_ORACLE_H
#define SWSPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run-time parameters read on the path (hydro_properties.c:348-390 defaults,
 * hydro_parameters.h:59-101 SPHENIX defaults). */
typedef struct orc_config {
  double dim[3];          /* periodic box */
  int ncell[3];           /* top-level grid; cell width must be >= max H */
  float eta;              /* SPH:resolution_eta */
  float h_tolerance;      /* SPH:h_tolerance */
  float h_max, h_min;     /* SPH:h_max, h_min */
  int max_iter;           /* SPH:max_ghost_iterations */
  float alpha_max, alpha_min, visc_length; /* viscosity_global_data */
  float diff_beta, diff_alpha_max, diff_alpha_min; /* diffusion_global_data */
  double time_base;       /* engine time_base; dt = 2^(bin+1) * time_base */
  int max_active_bin;     /* part_is_active: time_bin <= max_active_bin */
  int nthreads;           /* OpenMP threads (0 = default) */
} orc_config;

/* Particle state, structure of arrays, n particles.  Vectors are [3n]
 * interleaved (x0,y0,z0,x1,...).  Fields marked (io) are read for inactive
 * neighbours and overwritten for active particles. */
typedef struct orc_parts {
  int n;
  const double *x;
  const float *v;
  const float *m;
  const float *u;
  const int8_t *time_bin;
  float *h;                 /* (io) */
  float *rho;               /* (io) */
  float *pressure;          /* (io) force.pressure */
  float *soundspeed;        /* (io) force.soundspeed */
  float *f;                 /* (io) force.f (grad-h term) */
  float *balsara;           /* (io) force.balsara */
  float *alpha_visc;        /* (io) viscosity.alpha */
  float *alpha_diff;        /* (io) diffusion.alpha */
  float *div_v_prev;        /* (io) viscosity.div_v_previous_step */
} orc_parts;

/* Outputs of orc_pass_fp64 (all [n] or [3n], double). */
typedef struct orc_out64 {
  double *rho, *rho_dh, *wcount, *wcount_dh, *div_v, *rot_v;
  double *pressure, *soundspeed, *f, *balsara;
  double *v_sig, *laplace_u, *alpha_visc_max_ngb;
  double *alpha_visc, *alpha_diff, *div_v_dt;
  double *a_hydro, *u_dt, *h_dt;
  int32_t *n_density, *n_gradient, *n_force;
  int8_t *min_ngb_time_bin;
  /* Magnitude sums sum_j |term_j| of the cancelling accumulations, scaled by
   * the same finisher factors as the value: the error scale of a sum. */
  double *s_rho_dh, *s_wcount_dh, *s_div_v, *s_rot_v, *s_laplace_u;
  double *s_a_hydro, *s_u_dt, *s_h_dt;
} orc_out64;

/* Outputs of orc_pass_fp32 (float), plus the converged h written to
 * parts->h and the (io) fields of parts. */
typedef struct orc_out32 {
  float *rho_dh, *wcount, *wcount_dh, *div_v, *rot_v;
  float *v_sig, *laplace_u, *alpha_visc_max_ngb, *div_v_dt;
  float *a_hydro, *u_dt, *h_dt;
  int32_t *n_density, *n_gradient, *n_force;
  int8_t *min_ngb_time_bin;
  int32_t *ghost_iters;     /* [1]: number of ghost rounds used */
} orc_out32;

/* Pair contract and kernel helpers (exported for the pinning tests). */
float orc_kernel_gamma(void);
float orc_kernel_gamma2(void);
float orc_kernel_root(void);
float orc_kernel_norm(void);
void orc_kernel_deval32(float u, float *w, float *dw_dx);
void orc_kernel_deval64(double u, double *w, double *dw_dx);
float orc_cbrt32(float x);
/* Frame of a pair of top-level cells (corners loc_i of i's cell, loc_j of
 * j's cell): self != 0 for one cell; else x_i is taken relative to oi and
 * x_j relative to oj, the upper cell's corner (space_getsid.h:47-80). */
typedef struct orc_frame {
  int self;
  double oi[3], oj[3];
} orc_frame;
void orc_pair_frame(const double loc_i[3], const double loc_j[3],
                    const double dim[3], orc_frame *f);
void orc_pair_dx32(const double xi[3], const double xj[3], const orc_frame *f,
                   float dx[3], float *r2);
int orc_contract_flips(const orc_config *cfg, const orc_parts *p,
                       int64_t out[4]);

/* Single-pair interactions in FP32, restating hydro_iact.h.  The particle
 * record is a flat float array, see ORC_P_* in sphenix_oracle.c. */
int orc_part_record_size(void);
void orc_iact_density32(float r2, const float dx[3], float hi, float hj,
                        float *pi, float *pj, int symmetric);
void orc_iact_gradient32(float r2, const float dx[3], float hi, float hj,
                         float *pi, float *pj, int symmetric);
void orc_iact_force32(float r2, const float dx[3], float hi, float hj,
                      float *pi, float *pj, int symmetric);

/* Whole passes. Return 0 on success, non-zero on error (message printed). */
int orc_pass_fp64(const orc_config *cfg, const orc_parts *p, orc_out64 *out);
int orc_pass_fp32(const orc_config *cfg, orc_parts *p, orc_out32 *out);

/* O(N^2) brute-force neighbour counts over every pair of particles, in the
 * reference checkers' expression (pins the 27-cell search on small inputs).
 * mode 0 = density (r2 < H2_i), 1 = force. */
int orc_count_bruteforce(const orc_config *cfg, const orc_parts *p, int mode,
                         int32_t *counts);

/* ---- time integration (gas only, no cosmology) -------------------------
 * Restates drift_part + hydro_predict_extra (src/drift.h:138,
 * src/hydro/SPHENIX/hydro.h:1029), kick_part + hydro_kick_extra
 * (src/kick.h, hydro.h:1120) with the half-kick intervals of
 * runner_do_kick1 / runner_do_kick2 (src/runner_time_integration.c:87, 360)
 * and hydro_reset_predicted_values (hydro.h:983) after kick2, and
 * runner_do_timestep's get_part_timestep / make_integer_timestep
 * (src/timestep.h) on the time line of src/timeline.h. */
typedef struct orc_gas {
  int n;
  double *x;                       /* [3n] */
  float *v, *v_full, *a_hydro;     /* [3n] */
  float *u, *u_full, *u_dt, *h, *h_dt, *rho, *pressure, *soundspeed, *v_sig;
  int8_t *time_bin;
  const int8_t *min_ngb_time_bin;
} orc_gas;

typedef struct orc_integ {
  double time_base;
  float cfl, log_max_h_change, min_u;
  double dt_min, dt_max;
  int fixed_bins;
} orc_integ;

int orc_max_active_bin(long long ti);
long long orc_make_integer_timestep(float new_dt, int old_bin, int min_ngb_bin,
                                    long long ti_current, double time_base_inv);
void orc_drift(orc_gas *g, const orc_integ *c, long long ti_old,
               long long ti_current);
/* mode 1: kick1 of the particles starting a step; 2: kick2 of the active
 * particles, then hydro_reset_predicted_values. */
void orc_kick(orc_gas *g, const orc_integ *c, long long ti_current, int mode);
/* New bins of the active particles; returns the number of particles whose
 * step falls below dt_min or to zero (their bin is kept). */
int orc_timestep(orc_gas *g, const orc_integ *c, long long ti_current,
                 long long *ti_end_min, long long *n_active);

#ifdef __cplusplus
}
#endif
#endif
