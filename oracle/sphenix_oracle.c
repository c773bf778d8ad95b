/*
 * sphenix_oracle.c -- CPU oracle for the SPHENIX neighbour pass.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  The
 * product never includes, links or calls this file.
 *
 * Every function cites the reference lines it restates (paths relative to
 * the SWIFT source tree).  Build: oracle/Makefile (-O2 -ffp-contract=off, so
 * that every FP32 expression rounds exactly as written; the FMAs of the pair
 * contract are explicit fmaf calls).
 *
 * The same file, built with -DORC_SORTED and SWIFT's production flags
 * (-O3 -ffast-math -funroll-loops), is the threaded CPU baseline
 * (libcpuport.so): its candidate loops walk per-cell lists sorted along the
 * 13 cell-pair axes and stop at the projected distance, as the sorted
 * DOPAIR loops do (runner_doiact_functions_hydro.h:1297-1414, with the sort
 * of runner_sort.c:203-473).  It is timed as the CPU baseline and checked
 * against the oracle proper; it never serves as the checker.
 */
#include "sphenix_oracle.h"

/* Planted mistakes, for the tests that prove each pinning check can fail:
 * tests/test_oracle_pins.py builds this file a second time per mistake with
 * -DORC_MUTATION=k (into oracle/_mut/); the checker itself is always built
 * with ORC_MUTATION 0, where every MUT(k, good, bad) is `good`. */
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif
#define MUT(k, good, bad) (ORC_MUTATION == (k) ? (bad) : (good))

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Constants: Wendland C2, 3D (kernel_hydro.h:121-146, 203-241),
 * dimension.h:40-41, adiabatic_index.h:41-49 (gamma = 5/3),
 * hydro_parameters.h:53 (beta = 3), timeline.h:36 (num_time_bins).        */
/* ------------------------------------------------------------------------ */
#define KG ((float)(1.936492))
#define KCONST ((float)(21. * M_1_PI / 2.))
#define KG_INV ((float)(1. / KG))
#define KG2 ((float)(KG * KG))
#define KG_INV_DIM ((float)(1. / (KG * KG * KG)))
#define KG_INV_DIM1 ((float)(1. / (KG * KG * KG * KG)))
#define KROOT ((float)(1.f) * KCONST * KG_INV_DIM)
#define KNORM ((float)(((float)(4. * M_PI / 3.)) * ((float)(KG * KG * KG))))
#define HDIM 3.f
#define HDIM_INV 0.3333333333f
#define GAS_GAMMA 1.66666666666666667f
#define GAS_GM1 0.66666666666666667f
#define VISC_BETA 3.0f
#define NUM_TIME_BINS 56
#define TIME_BIN_INHIBITED (NUM_TIME_BINS + 2)

static const float kcoeffs[12] = {4.f, -15.f, 20.f, -10.f, 0.f, 1.f,
                                  0.f, 0.f,   0.f,  0.f,   0.f, 0.f};

float orc_kernel_gamma(void) { return KG; }
float orc_kernel_gamma2(void) { return KG2; }
float orc_kernel_root(void) { return KROOT; }
float orc_kernel_norm(void) { return KNORM; }

/* kernel_deval, FP32, same operation order (kernel_hydro.h:257-285). */
void orc_kernel_deval32(float u, float *W, float *dW_dx) {
  const float x = u * KG_INV;
  const int temp = (int)(x * 1.0f);
  const int ind = temp > 1 ? 1 : temp;
  const float *c = &kcoeffs[ind * 6];
  float w = c[0] * x + c[1];
  float dw = c[0];
  for (int k = 2; k <= 5; k++) {
    dw = dw * x + w;
    w = x * w + c[k];
  }
  w = fmaxf(w, 0.f);
  dw = fminf(dw, 0.f);
  *W = w * KCONST * KG_INV_DIM;
  *dW_dx = dw * KCONST * KG_INV_DIM1;
}

/* kernel_deval in FP64: closed form of the same polynomial,
 * w(q) = (1-q)^4 (1+4q), dw/dq = -20 q (1-q)^3 for q < 1, else 0. */
void orc_kernel_deval64(double u, double *W, double *dW_dx) {
  const double x = u * (double)KG_INV;
  if (x >= 1.0) {
    *W = 0.0;
    *dW_dx = 0.0;
    return;
  }
  const double om = 1.0 - x;
  const double w = om * om * om * om * (1.0 + 4.0 * x);
  const double dw = -20.0 * x * om * om * om;
  *W = w * (double)KCONST * (double)KG_INV_DIM;
  *dW_dx = dw * (double)KCONST * (double)KG_INV_DIM1;
}

/* Frame of the pair of top-level cells with corners loc_i (cell of i) and
 * loc_j (cell of j): space_getsid_and_swap_cells (space_getsid.h:47-80)
 * finds the shift that brings cell j next to cell i (loc differences beyond
 * half the box wrap) and the sort id of the offset; runner_flip swaps the
 * cells so that the second cell of the pair ("cj", the upper cell) lies in
 * the positive half-space of the offset (sort_part.h:59-61).  Both
 * directions of the pair then use cj's corner: particles of cj relative to
 * cj->loc, those of ci relative to cj->loc + shift
 * (runner_doiact_functions_hydro.h:1327-1341, 1446-1465, 1672-1674;
 * tools.c:218-246).  The same cell (sort id 13) is a self pair. */
void orc_pair_frame(const double loc_i[3], const double loc_j[3],
                    const double dim[3], orc_frame *f) {
  double shift[3], d[3];
  for (int k = 0; k < 3; k++) {
    d[k] = loc_j[k] - loc_i[k];
    if (d[k] < -dim[k] / 2)
      shift[k] = dim[k];
    else if (d[k] > dim[k] / 2)
      shift[k] = -dim[k];
    else
      shift[k] = 0.0;
    d[k] += shift[k];
  }
  int sid = 0;
  for (int k = 0; k < 3; k++)
    sid = 3 * sid + ((d[k] < 0.0) ? 0 : ((d[k] > 0.0) ? 2 : 1));
  f->self = sid == 13;
  if (f->self) {
    for (int k = 0; k < 3; k++) f->oi[k] = f->oj[k] = 0.0;
  } else if (sid > 13) {
    /* no flip: ci = cell of i, cj = cell of j */
    for (int k = 0; k < 3; k++) {
      f->oi[k] = loc_j[k] + shift[k];
      f->oj[k] = loc_j[k];
    }
  } else {
    /* flip: ci = cell of j, cj = cell of i, shift negated */
    for (int k = 0; k < 3; k++) {
      f->oi[k] = loc_i[k];
      f->oj[k] = loc_i[k] + (-shift[k]);
    }
  }
}

/* The FP32 pair contract: the expression of the reference's loops and of
 * its brute-force checker for the pair's frame.  Self pairs round the FP64
 * difference once (DOSELF1 :2459, DOSELF2 :2784, self_all_density
 * tools.c:617-619); cell pairs round both cell-frame offsets, then subtract
 * in FP32 (DOPAIR1 :1327-1346, DOPAIR2 :1748-1781, pairs_all_density
 * tools.c:232-246).  r2 sums the squares as gcc contracts
 * dx0*dx0 + dx1*dx1 + dx2*dx2 under -ffp-contract=fast. */
void orc_pair_dx32(const double xi[3], const double xj[3], const orc_frame *f,
                   float dx[3], float *r2) {
  for (int k = 0; k < 3; k++) {
    if (f->self) {
      dx[k] = (float)(xi[k] - xj[k]);
    } else {
      const float pix = (float)(xi[k] - f->oi[k]);
      const float pjx = (float)(xj[k] - f->oj[k]);
      dx[k] = pix - pjx;
    }
  }
  *r2 = fmaf(dx[2], dx[2], fmaf(dx[1], dx[1], dx[0] * dx[0]));
}

/* The contract of the first build of this oracle: one rounding of the FP64
 * difference for every pair, dx = (float)((x_i - t) - x_j).  Kept only to
 * count the mask decisions on which the two contracts differ. */
static void pair_dx32_fp64(const double xi[3], const double xj[3],
                           const double t[3], float dx[3], float *r2) {
  for (int k = 0; k < 3; k++) dx[k] = (float)((xi[k] - t[k]) - xj[k]);
  *r2 = fmaf(dx[2], dx[2], fmaf(dx[1], dx[1], dx[0] * dx[0]));
}

static inline float hig2_of(float h) { return (h * h) * KG2; }

/* pow_inv_dimension for 3D (dimension.h:98-102): the C library's cbrtf. */
float orc_cbrt32(float x) {
  return MUT(17, cbrtf(x), expf(logf(x) / 3.f));
}

/* ------------------------------------------------------------------------ */
/* Single-pair FP32 interactions on a flat particle record                  */
/* ------------------------------------------------------------------------ */
enum {
  P_V0, P_V1, P_V2, P_M, P_H, P_U, P_RHO, P_RHO_DH, P_WC, P_WC_DH, P_DIVV,
  P_ROT0, P_ROT1, P_ROT2, P_F, P_P, P_C, P_BAL, P_VSIG, P_LAPU, P_AVMAX,
  P_ALPHA, P_ADIFF, P_A0, P_A1, P_A2, P_UDT, P_HDT, P_NDEN, P_NGRAD,
  P_NFORCE, P_TB, P_MINTB, P_COUNT
};
int orc_part_record_size(void) { return P_COUNT; }

/* runner_iact_nonsym_density (hydro_iact.h:141-193), updates pi only;
 * runner_iact_density (hydro_iact.h:48-127) when symmetric. */
void orc_iact_density32(float r2, const float dx[3], float hi, float hj,
                        float *pi, float *pj, int symmetric) {
  const float r = sqrtf(r2);
  const float mi = pi[P_M], mj = pj[P_M];
  const float hi_inv = 1.f / hi;
  const float ui = r * hi_inv;
  float wi, wi_dx;
  orc_kernel_deval32(ui, &wi, &wi_dx);
  pi[P_RHO] += mj * wi;
  pi[P_RHO_DH] -= mj * (HDIM * wi + ui * wi_dx);
  pi[P_WC] += wi;
  pi[P_WC_DH] -= (HDIM * wi + ui * wi_dx);
  float wj = 0.f, wj_dx = 0.f, uj = 0.f;
  if (symmetric) {
    const float hj_inv = 1.f / hj;
    uj = r * hj_inv;
    orc_kernel_deval32(uj, &wj, &wj_dx);
    pj[P_RHO] += mi * wj;
    pj[P_RHO_DH] -= mi * (HDIM * wj + uj * wj_dx);
    pj[P_WC] += wj;
    pj[P_WC_DH] -= (HDIM * wj + uj * wj_dx);
  }
  const float r_inv = r ? 1.0f / r : 0.0f;
  const float faci = mj * wi_dx * r_inv;
  const float dv[3] = {pi[P_V0] - pj[P_V0], pi[P_V1] - pj[P_V1],
                       pi[P_V2] - pj[P_V2]};
  const float dvdr = dv[0] * dx[0] + dv[1] * dx[1] + dv[2] * dx[2];
  pi[P_DIVV] -= faci * dvdr;
  const float curl[3] = {dv[1] * dx[2] - dv[2] * dx[1],
                         dv[2] * dx[0] - dv[0] * dx[2],
                         dv[0] * dx[1] - dv[1] * dx[0]};
  pi[P_ROT0] += faci * curl[0];
  pi[P_ROT1] += faci * curl[1];
  pi[P_ROT2] += faci * curl[2];
  pi[P_NDEN] += 1.f;
  if (symmetric) {
    const float facj = mi * wj_dx * r_inv;
    pj[P_DIVV] -= facj * dvdr;
    pj[P_ROT0] += facj * curl[0];
    pj[P_ROT1] += facj * curl[1];
    pj[P_ROT2] += facj * curl[2];
    pj[P_NDEN] += 1.f;
  }
}

/* runner_iact_nonsym_gradient (hydro_iact.h:291-344) / runner_iact_gradient
 * (hydro_iact.h:211-272), a = 1, H = 0. */
void orc_iact_gradient32(float r2, const float dx[3], float hi, float hj,
                         float *pi, float *pj, int symmetric) {
  const float r = sqrtf(r2);
  const float r_inv = r ? 1.0f / r : 0.0f;
  const float fac_mu = 1.f, a2_Hubble = 0.f;
  const float dvdr = (pi[P_V0] - pj[P_V0]) * dx[0] +
                     (pi[P_V1] - pj[P_V1]) * dx[1] +
                     (pi[P_V2] - pj[P_V2]) * dx[2];
  const float dvdr_Hubble = dvdr + a2_Hubble * r2;
  const float omega_ij = fminf(dvdr_Hubble, 0.f);
  const float mu_ij = fac_mu * r_inv * omega_ij;
  const float new_v_sig =
      pi[P_C] + pj[P_C] - MUT(13, VISC_BETA, 2.f) * mu_ij;
  pi[P_VSIG] = fmaxf(pi[P_VSIG], new_v_sig);
  float wi, wi_dx;
  const float ui = r / hi;
  orc_kernel_deval32(ui, &wi, &wi_dx);
  const float delta_u_factor = (pi[P_U] - pj[P_U]) * r_inv;
  pi[P_LAPU] += pj[P_M] * delta_u_factor * wi_dx / pj[P_RHO];
  pi[P_AVMAX] = fmaxf(pi[P_AVMAX], pj[P_ALPHA]);
  pi[P_NGRAD] += 1.f;
  if (symmetric) {
    pj[P_VSIG] = fmaxf(pj[P_VSIG], new_v_sig);
    float wj, wj_dx;
    const float uj = r / hj;
    orc_kernel_deval32(uj, &wj, &wj_dx);
    pj[P_LAPU] -= pi[P_M] * delta_u_factor * wj_dx / pi[P_RHO];
    pj[P_AVMAX] = fmaxf(pj[P_AVMAX], pi[P_ALPHA]);
    pj[P_NGRAD] += 1.f;
  }
}

/* runner_iact_nonsym_force (hydro_iact.h:507-632) / runner_iact_force
 * (hydro_iact.h:358-493) + runner_iact_{,nonsym_}timebin
 * (timestep_limiter_iact.h:34-70), a = 1, H = 0. */
void orc_iact_force32(float r2, const float dx[3], float hi, float hj,
                      float *pi, float *pj, int symmetric) {
  const float fac_mu = 1.f, a2_Hubble = 0.f;
  const float r = sqrtf(r2);
  const float r_inv = r ? 1.0f / r : 0.0f;
  const float mj = pj[P_M], mi = pi[P_M];
  const float rhoi = pi[P_RHO], rhoj = pj[P_RHO];
  const float pressurei = pi[P_P], pressurej = pj[P_P];
  const float hi_inv = 1.0f / hi;
  const float hi_inv2 = hi_inv * hi_inv;
  const float hid_inv = hi_inv2 * hi_inv2;
  const float xi = r * hi_inv;
  float wi, wi_dx;
  orc_kernel_deval32(xi, &wi, &wi_dx);
  const float wi_dr = hid_inv * wi_dx;
  const float hj_inv = 1.0f / hj;
  const float hj_inv2 = hj_inv * hj_inv;
  const float hjd_inv = hj_inv2 * hj_inv2;
  const float xj = r * hj_inv;
  float wj, wj_dx;
  orc_kernel_deval32(xj, &wj, &wj_dx);
  const float wj_dr = hjd_inv * wj_dx;
  const float dvdr = (pi[P_V0] - pj[P_V0]) * dx[0] +
                     (pi[P_V1] - pj[P_V1]) * dx[1] +
                     (pi[P_V2] - pj[P_V2]) * dx[2];
  const float dvdr_Hubble = dvdr + a2_Hubble * r2;
  const float omega_ij = fminf(dvdr_Hubble, 0.f);
  const float mu_ij = fac_mu * r_inv * omega_ij;
  const float v_sig = pi[P_C] + pj[P_C] - VISC_BETA * mu_ij;
  const float f_ij = 1.f - pi[P_F] / mj;
  const float f_ji = 1.f - pj[P_F] / mi;
  const float rho_ij = rhoi + rhoj;
  const float alpha = pi[P_ALPHA] + pj[P_ALPHA];
  const float visc = MUT(12, -0.25f, -0.5f) * alpha * v_sig * mu_ij *
                     (pi[P_BAL] + pj[P_BAL]) / rho_ij;
  const float visc_acc_term =
      0.5f * visc * (wi_dr * f_ij + wj_dr * f_ji) * r_inv;
  const float P_over_rho2_i = pressurei / (rhoi * rhoi) * f_ij;
  const float P_over_rho2_j = pressurej / (rhoj * rhoj) * f_ji;
  const float sph_acc_term =
      (P_over_rho2_i * wi_dr + P_over_rho2_j * wj_dr) * r_inv;
  const float acc = sph_acc_term + visc_acc_term;
  pi[P_A0] -= mj * acc * dx[0];
  pi[P_A1] -= mj * acc * dx[1];
  pi[P_A2] -= mj * acc * dx[2];
  const float sph_du_term_i = P_over_rho2_i * dvdr * r_inv * wi_dr;
  const float visc_du_term = 0.5f * visc_acc_term * dvdr_Hubble;
  const float alpha_diff =
      (pressurei * pi[P_ADIFF] + pressurej * pj[P_ADIFF]) /
      (pressurei + pressurej);
  const float v_diff =
      alpha_diff * MUT(14, 0.5f, 1.f) *
      (sqrtf(2.f * fabsf(pressurei - pressurej) / rho_ij) +
       fabsf(fac_mu * r_inv * dvdr_Hubble));
  const float diff_du_term = v_diff * (pi[P_U] - pj[P_U]) *
                             (f_ij * wi_dr / rhoi + f_ji * wj_dr / rhoj);
  const float du_dt_i = sph_du_term_i + visc_du_term + diff_du_term;
  pi[P_UDT] += du_dt_i * mj;
  pi[P_HDT] -= mj * dvdr * r_inv / rhoj * wi_dr;
  if (pj[P_TB] > 0.f) pi[P_MINTB] = fminf(pi[P_MINTB], pj[P_TB]);
  pi[P_NFORCE] += 1.f;
  if (symmetric) {
    pj[P_A0] += mi * acc * dx[0];
    pj[P_A1] += mi * acc * dx[1];
    pj[P_A2] += mi * acc * dx[2];
    const float sph_du_term_j = P_over_rho2_j * dvdr * r_inv * wj_dr;
    const float du_dt_j = sph_du_term_j + visc_du_term - diff_du_term;
    pj[P_UDT] += du_dt_j * mi;
    pj[P_HDT] -= mi * dvdr * r_inv / rhoi * wj_dr;
    if (pi[P_TB] > 0.f) pj[P_MINTB] = fminf(pj[P_MINTB], pi[P_TB]);
    pj[P_NFORCE] += 1.f;
  }
}

/* ------------------------------------------------------------------------ */
/* Top-level grid (mirrors the binning rule of the product's cell build).   */
/* ------------------------------------------------------------------------ */
typedef struct grid {
  int nc[3];
  int ncells;
  double w[3];
  int *cell_of;  /* [n] */
  int *start;    /* [ncells+1] */
  int *index;    /* [n] particle ids sorted by cell */
#ifdef ORC_SORTED
  /* per axis s (13 cell-pair directions): each cell's particle ids sorted by
   * the projection x . n_s, and the projections in that order */
  int *sidx[13];
  double *skey[13];
  float *hgmax; /* [ncells] largest search radius gamma h of each cell */
#endif
} grid;

/* space_parts_get_cell_index (space_cell_index.c:114-128): the cell of a
 * position in [0, dim) is (int)(x * iwidth), iwidth = 1 / (dim / cdim)
 * (space_regrid.c:221-222); the clamp only guards positions the caller
 * failed to wrap. */
static int cell_coord(double x, double inv_w, int n) {
  int c = (int)(x * inv_w);
  if (x < 0.0) c = 0;
  if (c < 0) c = 0;
  if (c >= n) c = n - 1;
  return c;
}

static int grid_build(const orc_config *cfg, const orc_parts *p, grid *g) {
  for (int k = 0; k < 3; k++) {
    g->nc[k] = cfg->ncell[k];
    if (g->nc[k] < 3) {
      fprintf(stderr, "oracle: need >= 3 top-level cells per axis\n");
      return 1;
    }
    g->w[k] = cfg->dim[k] / g->nc[k];
  }
  g->ncells = g->nc[0] * g->nc[1] * g->nc[2];
#ifdef ORC_SORTED
  for (int s = 0; s < 13; s++) {
    g->sidx[s] = NULL;
    g->skey[s] = NULL;
  }
  g->hgmax = NULL;
#endif
  g->cell_of = (int *)malloc(sizeof(int) * (size_t)(p->n > 0 ? p->n : 1));
  g->start = (int *)calloc((size_t)g->ncells + 1, sizeof(int));
  g->index = (int *)malloc(sizeof(int) * (size_t)(p->n > 0 ? p->n : 1));
  if (!g->cell_of || !g->start || !g->index) return 1;
  double inv_w[3];
  for (int k = 0; k < 3; k++) inv_w[k] = 1.0 / g->w[k];
  for (int i = 0; i < p->n; i++) {
    int c[3];
    for (int k = 0; k < 3; k++)
      c[k] = cell_coord(p->x[3 * i + k], inv_w[k], g->nc[k]);
    const int cid = (c[0] * g->nc[1] + c[1]) * g->nc[2] + c[2];
    g->cell_of[i] = cid;
    g->start[cid + 1]++;
  }
  for (int c = 0; c < g->ncells; c++) g->start[c + 1] += g->start[c];
  int *fill = (int *)malloc(sizeof(int) * (size_t)g->ncells);
  memcpy(fill, g->start, sizeof(int) * (size_t)g->ncells);
  for (int i = 0; i < p->n; i++) g->index[fill[g->cell_of[i]]++] = i;
  free(fill);
  return 0;
}

static void grid_free(grid *g) {
  free(g->cell_of);
  free(g->start);
  free(g->index);
#ifdef ORC_SORTED
  for (int s = 0; s < 13; s++) {
    free(g->sidx[s]);
    free(g->skey[s]);
  }
  free(g->hgmax);
#endif
}

#ifdef ORC_SORTED
/* Unit vector of stencil position q (0..26, ox-major; 13 is the cell itself):
 * offset (q / 9 - 1, q / 3 % 3 - 1, q % 3 - 1) normalised.  The 13 positions
 * q > 13 are the sort axes (SWIFT's sid with flip folds the other 13 onto
 * them, sort_part.h:32-93). */
static void axis_of(int q, double n[3]) {
  const int o[3] = {q / 9 - 1, q / 3 % 3 - 1, q % 3 - 1};
  const double len = sqrt((double)(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
  for (int k = 0; k < 3; k++) n[k] = o[k] / len;
}

typedef struct skey_entry {
  double d;
  int i;
} skey_entry;

static int skey_cmp(const void *a, const void *b) {
  const double da = ((const skey_entry *)a)->d, db = ((const skey_entry *)b)->d;
  return (da > db) - (da < db);
}

/* runner_do_hydro_sort (runner_sort.c:203-473): every cell sorted along
 * every axis. */
static int grid_sort(const orc_parts *p, grid *g) {
  const size_t n = (size_t)(p->n > 0 ? p->n : 1);
  g->hgmax = (float *)calloc((size_t)g->ncells, sizeof(float));
  if (!g->hgmax) return 1;
  for (int s = 0; s < 13; s++) {
    g->sidx[s] = (int *)malloc(sizeof(int) * n);
    g->skey[s] = (double *)malloc(sizeof(double) * n);
    if (!g->sidx[s] || !g->skey[s]) return 1;
  }
#pragma omp parallel
  {
    int cap = 0;
    skey_entry *buf = NULL;
#pragma omp for schedule(dynamic, 16)
    for (int c = 0; c < g->ncells; c++) {
      const int a = g->start[c], e = g->start[c + 1];
      if (e - a > cap) {
        cap = e - a;
        buf = (skey_entry *)realloc(buf, sizeof(skey_entry) * (size_t)cap);
      }
      for (int s = 0; s < 13; s++) {
        double nv[3];
        axis_of(14 + s, nv);
        for (int k = a; k < e; k++) {
          const int i = g->index[k];
          buf[k - a].i = i;
          buf[k - a].d = p->x[3 * i] * nv[0] + p->x[3 * i + 1] * nv[1] +
                         p->x[3 * i + 2] * nv[2];
        }
        qsort(buf, (size_t)(e - a), sizeof(skey_entry), skey_cmp);
        for (int k = a; k < e; k++) {
          g->sidx[s][k] = buf[k - a].i;
          g->skey[s][k] = buf[k - a].d;
        }
      }
    }
    free(buf);
  }
  return 0;
}

/* Largest search radius of every cell (the hj_max of DOPAIR2). */
static void grid_hmax(const orc_parts *p, grid *g) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < g->ncells; c++) {
    float m = 0.f;
    for (int k = g->start[c]; k < g->start[c + 1]; k++)
      m = fmaxf(m, KG * p->h[g->index[k]]);
    g->hgmax[c] = m;
  }
}
#endif


/* The 27 neighbour cells of cell cid with the image translation t applied
 * to positions of j (x_j + t is the image nearest to cell cid), and the
 * frame of the pair contract for (cell cid, neighbour). */
typedef struct nbr {
  int cell;
  double t[3];
  orc_frame fr;
} nbr;

static void neighbours27(const grid *g, const double dim[3], int cid,
                         nbr out[27]) {
  const int cz = cid % g->nc[2];
  const int cy = (cid / g->nc[2]) % g->nc[1];
  const int cx = cid / (g->nc[2] * g->nc[1]);
  const int c[3] = {cx, cy, cz};
  int m = 0;
  for (int ox = -1; ox <= 1; ox++)
    for (int oy = -1; oy <= 1; oy++)
      for (int oz = -1; oz <= 1; oz++) {
        const int o[3] = {ox, oy, oz};
        int cc[3];
        for (int k = 0; k < 3; k++) {
          int v = c[k] + o[k];
          out[m].t[k] = 0.0;
          if (v < 0) {
            v += g->nc[k];
            out[m].t[k] = -dim[k];
          } else if (v >= g->nc[k]) {
            v -= g->nc[k];
            out[m].t[k] = dim[k];
          }
          cc[k] = v;
        }
        out[m].cell = (cc[0] * g->nc[1] + cc[1]) * g->nc[2] + cc[2];
        /* cell corners as space_regrid sets them: loc = index * width
         * (space_regrid.c:307) */
        double li[3], lj[3];
        for (int k = 0; k < 3; k++) {
          li[k] = c[k] * g->w[k];
          lj[k] = cc[k] * g->w[k];
        }
        orc_pair_frame(li, lj, dim, &out[m].fr);
        m++;
      }
}

/* Candidates j of particle i (position xi) in stencil neighbour q with image
 * translation b->t, given the search radius R: ids idx[a..e).  The oracle
 * proper takes the whole cell.  The sorted build keeps the prefix (or
 * suffix, for the flipped axes) of the cell's sorted list whose projected
 * distance (x_j + t - x_i) . n is below R (DOPAIR1: sort_j[pjd].d < di,
 * runner_doiact_functions_hydro.h:1322-1332); the projection never exceeds
 * the distance, and the 1e-4 widening covers the FP32 rounding of r2. */
static inline void cand_range(const grid *g, const nbr *b, int q,
                              const double *xi, float R, const int **idx,
                              int *a, int *e) {
  const int c = b->cell;
  *a = g->start[c];
  *e = g->start[c + 1];
  *idx = g->index;
#ifdef ORC_SORTED
  if (q == 13) return;
  const int qa = q > 13 ? q : 26 - q;
  const int s = qa - 14;
  double nv[3];
  axis_of(qa, nv);
  const double pi = xi[0] * nv[0] + xi[1] * nv[1] + xi[2] * nv[2];
  const double pt = b->t[0] * nv[0] + b->t[1] * nv[1] + b->t[2] * nv[2];
  const double Rw = (double)R * 1.0001 + 1e-12;
  const double *key = g->skey[s];
  *idx = g->sidx[s];
  int lo = *a, hi = *e;
  if (q > 13) {
    /* keep key + pt - pi < Rw: the first index with key >= cut ends it */
    const double cut = pi - pt + Rw;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (key[m] < cut) lo = m + 1; else hi = m;
    }
    *e = lo;
  } else {
    /* keep pi - (key + pt) < Rw: the first index with key > cut starts it */
    const double cut = pi - pt - Rw;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (key[m] <= cut) lo = m + 1; else hi = m;
    }
    *a = lo;
  }
#else
  (void)q;
  (void)xi;
  (void)R;
#endif
}

static int is_active(const orc_config *cfg, int8_t tb) {
  return tb <= cfg->max_active_bin; /* active.h:349-366 */
}
static int is_inhibited(int8_t tb) {
  return tb == TIME_BIN_INHIBITED; /* active.h:518-521, timeline.h:42 */
}

/* dt of a particle's time-step (timeline.h:59-95): 2^(bin+1) * time_base. */
static double bin_dt(int8_t bin, double time_base) {
  if (bin <= 0) return 0.0;
  return (double)(1LL << (bin + 1)) * time_base;
}

static int check_h_fits(const orc_config *cfg, const grid *g,
                        const orc_parts *p) {
  float hmax = 0.f;
  for (int i = 0; i < p->n; i++)
    if (!is_inhibited(p->time_bin[i])) hmax = fmaxf(hmax, p->h[i]);
  const double H = (double)hmax * (double)KG;
  for (int k = 0; k < 3; k++)
    if (H > g->w[k]) {
      fprintf(stderr,
              "oracle: kernel support %g exceeds top-level cell width %g\n", H,
              g->w[k]);
      return 1;
    }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 rerun of the pass at fixed h                                       */
/* ------------------------------------------------------------------------ */
int orc_pass_fp64(const orc_config *cfg, const orc_parts *p, orc_out64 *o) {
  grid g;
  if (grid_build(cfg, p, &g)) return 1;
  if (check_h_fits(cfg, &g, p)) {
    grid_free(&g);
    return 1;
  }
#ifdef _OPENMP
  if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
  const int n = p->n;
  const double kr = (double)KROOT;

  /* Density loop (DOSELF1/DOPAIR1 with runner_iact_nonsym_density),
   * then hydro_end_density (hydro.h:613-662). */
#pragma omp parallel for schedule(dynamic, 1)
  for (int cid = 0; cid < g.ncells; cid++) {
    nbr nb[27];
    neighbours27(&g, cfg->dim, cid, nb);
    for (int ii = g.start[cid]; ii < g.start[cid + 1]; ii++) {
      const int i = g.index[ii];
      if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
        continue;
      const float hi = p->h[i];
      const float hig2 = hig2_of(hi);
      const double h = (double)hi;
      double rho = 0, rho_dh = 0, wc = 0, wc_dh = 0, divv = 0;
      double rot[3] = {0, 0, 0};
      double s_rdh = 0, s_wdh = 0, s_dv = 0, s_rot[3] = {0, 0, 0};
      int cnt = 0;
      for (int q = 0; q < 27; q++) {
        const nbr *b = &nb[q];
        for (int jj = g.start[b->cell]; jj < g.start[b->cell + 1]; jj++) {
          const int j = g.index[jj];
          if (j == i || is_inhibited(p->time_bin[j])) continue;
          float dxf[3], r2f;
          orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &b->fr, dxf, &r2f);
          if (!(r2f < hig2)) continue;
          double dx[3];
          for (int k = 0; k < 3; k++)
            dx[k] = (p->x[3 * i + k] - b->t[k]) - p->x[3 * j + k];
          const double r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
          const double r = sqrt(r2);
          const double ui = r / h;
          double wi, wi_dx;
          orc_kernel_deval64(ui, &wi, &wi_dx);
          const double mj = p->m[j];
          rho += mj * wi;
          rho_dh -= mj * (3.0 * wi + ui * wi_dx);
          wc += wi;
          wc_dh -= (3.0 * wi + ui * wi_dx);
          const double r_inv = r > 0 ? 1.0 / r : 0.0;
          const double faci = mj * wi_dx * r_inv;
          const double dv[3] = {(double)p->v[3 * i] - p->v[3 * j],
                                (double)p->v[3 * i + 1] - p->v[3 * j + 1],
                                (double)p->v[3 * i + 2] - p->v[3 * j + 2]};
          const double dvdr = dv[0] * dx[0] + dv[1] * dx[1] + dv[2] * dx[2];
          divv -= faci * dvdr;
          rot[0] += faci * (dv[1] * dx[2] - dv[2] * dx[1]);
          rot[1] += faci * (dv[2] * dx[0] - dv[0] * dx[2]);
          rot[2] += faci * (dv[0] * dx[1] - dv[1] * dx[0]);
          s_rdh += fabs(mj * 3.0 * wi) + fabs(mj * ui * wi_dx);
          s_wdh += fabs(3.0 * wi) + fabs(ui * wi_dx);
          s_dv += fabs(faci * dvdr);
          s_rot[0] += fabs(faci * dv[1] * dx[2]) + fabs(faci * dv[2] * dx[1]);
          s_rot[1] += fabs(faci * dv[2] * dx[0]) + fabs(faci * dv[0] * dx[2]);
          s_rot[2] += fabs(faci * dv[0] * dx[1]) + fabs(faci * dv[1] * dx[0]);
          cnt++;
        }
      }
      /* hydro_end_density */
      const double mi = p->m[i];
      const double h_inv = 1.0 / h;
      const double hid = h_inv * h_inv * h_inv;
      const double hid1 = hid * h_inv;
      rho = (rho + mi * kr) * hid;
      rho_dh = (rho_dh - 3.0 * mi * kr) * hid1;
      wc = (wc + kr) * hid;
      wc_dh = (wc_dh - 3.0 * kr) * hid1;
      const double rho_inv = 1.0 / rho;
      for (int k = 0; k < 3; k++) o->rot_v[3 * i + k] = rot[k] * hid1 * rho_inv;
      o->div_v[i] = divv * hid1 * rho_inv;
      o->s_rho_dh[i] = (s_rdh + 3.0 * mi * kr) * hid1;
      o->s_wcount_dh[i] = (s_wdh + 3.0 * kr) * hid1;
      o->s_div_v[i] = s_dv * hid1 * rho_inv;
      for (int k = 0; k < 3; k++)
        o->s_rot_v[3 * i + k] = s_rot[k] * hid1 * rho_inv;
      if (cnt == 0 && hi >= cfg->h_max) {
        /* The ghost ends a particle that finds no neighbour at h_max in
         * hydro_part_has_no_neighbours (runner_ghost.c:1217-1227, 1432-1442;
         * hydro.h:791-825): rho, wcount as above (self term only), the
         * h-derivatives zeroed. */
        rho_dh = 0.0;
        wc_dh = 0.0;
        o->s_rho_dh[i] = 0.0;
        o->s_wcount_dh[i] = 0.0;
      }
      o->rho[i] = rho;
      o->rho_dh[i] = rho_dh;
      o->wcount[i] = wc;
      o->wcount_dh[i] = wc_dh;
      o->n_density[i] = cnt;
    }
  }

  /* Per-particle state after the ghost: hydro_prepare_gradient +
   * hydro_reset_gradient (hydro.h:671-750).  Inactive particles keep the
   * (io) values they came with. */
  double *rho_all = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  double *c_all = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  double *P_all = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  double *f_all = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  double *bal_all = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  double *alpha_all = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  double *adiff_all = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  for (int i = 0; i < n; i++) {
    if (is_active(cfg, p->time_bin[i]) && !is_inhibited(p->time_bin[i])) {
      const double h = p->h[i];
      const double rho = o->rho[i];
      const double curl_v =
          sqrt(o->rot_v[3 * i] * o->rot_v[3 * i] +
               o->rot_v[3 * i + 1] * o->rot_v[3 * i + 1] +
               o->rot_v[3 * i + 2] * o->rot_v[3 * i + 2]);
      const double abs_div_v = fabs(o->div_v[i]);
      const double P = (double)GAS_GM1 * (double)p->u[i] * rho;
      const double c = sqrt((double)GAS_GAMMA * P / rho);
      const double bal =
          abs_div_v / (abs_div_v + curl_v + 0.0001 * c / h);
      const double common = h * (1.0 / 3.0) / o->wcount[i];
      double fgh;
      if (h > 0.9999 * (double)cfg->h_max)
        fgh = 0.0;
      else {
        const double gW = common * o->wcount_dh[i];
        fgh = (gW < -0.9999) ? 0.0 : common * o->rho_dh[i] / (1.0 + gW);
      }
      rho_all[i] = rho;
      c_all[i] = c;
      P_all[i] = P;
      f_all[i] = fgh;
      bal_all[i] = bal;
      o->pressure[i] = P;
      o->soundspeed[i] = c;
      o->f[i] = fgh;
      o->balsara[i] = bal;
    } else {
      rho_all[i] = p->rho[i];
      c_all[i] = p->soundspeed[i];
      P_all[i] = p->pressure[i];
      f_all[i] = p->f[i];
      bal_all[i] = p->balsara[i];
    }
    alpha_all[i] = p->alpha_visc[i];
    adiff_all[i] = p->alpha_diff[i];
  }

  /* Gradient loop (runner_iact_nonsym_gradient), TYPE1 mask r < H_i,
   * then hydro_end_gradient (hydro.h:762-780) and hydro_prepare_force
   * (hydro.h:840-954) in FP64 with the same branch structure. */
#pragma omp parallel for schedule(dynamic, 1)
  for (int cid = 0; cid < g.ncells; cid++) {
    nbr nb[27];
    neighbours27(&g, cfg->dim, cid, nb);
    for (int ii = g.start[cid]; ii < g.start[cid + 1]; ii++) {
      const int i = g.index[ii];
      if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
        continue;
      const float hi = p->h[i];
      const float hig2 = hig2_of(hi);
      const double h = hi;
      double vsig = 2.0 * c_all[i];
      double avmax = alpha_all[i];
      double lapu = 0.0, s_lapu = 0.0;
      int cnt = 0;
      for (int q = 0; q < 27; q++) {
        const nbr *b = &nb[q];
        for (int jj = g.start[b->cell]; jj < g.start[b->cell + 1]; jj++) {
          const int j = g.index[jj];
          if (j == i || is_inhibited(p->time_bin[j])) continue;
          float dxf[3], r2f;
          orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &b->fr, dxf, &r2f);
          if (!(r2f < hig2)) continue;
          double dx[3];
          for (int k = 0; k < 3; k++)
            dx[k] = (p->x[3 * i + k] - b->t[k]) - p->x[3 * j + k];
          const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
          const double r_inv = r > 0 ? 1.0 / r : 0.0;
          const double dvdr =
              ((double)p->v[3 * i] - p->v[3 * j]) * dx[0] +
              ((double)p->v[3 * i + 1] - p->v[3 * j + 1]) * dx[1] +
              ((double)p->v[3 * i + 2] - p->v[3 * j + 2]) * dx[2];
          const double mu =
              r_inv * MUT(8, (dvdr < 0 ? dvdr : 0.0), (dvdr > 0 ? dvdr : 0.0));
          const double nv = c_all[i] + c_all[j] - MUT(7, 3.0, 2.0) * mu;
          if (nv > vsig) vsig = nv;
          double wi, wi_dx;
          orc_kernel_deval64(r / h, &wi, &wi_dx);
          const double lt = (double)p->m[j] * ((double)p->u[i] - p->u[j]) *
                            MUT(6, r_inv, 1.0) * wi_dx / rho_all[j];
          lapu += lt;
          s_lapu += fabs(lt);
          if (alpha_all[j] > avmax) avmax = alpha_all[j];
          cnt++;
        }
      }
      const double h_inv = 1.0 / h;
      const double hid1 = h_inv * h_inv * h_inv * h_inv;
      lapu *= MUT(5, 2.0, 1.0) * hid1;
      o->s_laplace_u[i] = s_lapu * MUT(5, 2.0, 1.0) * hid1;
      o->v_sig[i] = vsig;
      o->laplace_u[i] = lapu;
      o->alpha_visc_max_ngb[i] = avmax;
      o->n_gradient[i] = cnt;

      /* hydro_prepare_force, a = 1 */
      const double dt_alpha = (double)(float)bin_dt(p->time_bin[i], cfg->time_base);
      const double ksp = h * (double)KG;
      const double ksp_inv = 1.0 / ksp;
      const double c = c_all[i];
      const double sct_inv = c * ksp_inv;
      const double div_v = o->div_v[i];
      const double div_v_dt =
          dt_alpha == 0.0 ? 0.0 : (div_v - (double)p->div_v_prev[i]) / dt_alpha;
      const double S = div_v < 0.0 ? MUT(9, ksp * ksp, h * h) * fmax(0.0, -div_v_dt)
                                   : 0.0;
      const double alpha_loc = (double)cfg->alpha_max * S / (c * c + S);
      double alpha = alpha_all[i];
      if (alpha_loc > alpha)
        alpha = alpha_loc;
      else {
        const double tr =
            dt_alpha * sct_inv * MUT(10, (double)cfg->visc_length, 1.0);
        alpha += alpha_loc * tr;
        alpha /= (1.0 + tr);
      }
      alpha = fmax(alpha, (double)cfg->alpha_min);
      o->alpha_visc[i] = alpha;
      o->div_v_dt[i] = div_v_dt;
      const double dtpi = vsig * ksp_inv;
      const double sqrt_u_inv =
          MUT(11, 1.0 / sqrt((double)p->u[i]), 1.0 / (double)p->u[i]);
      double adt = (double)cfg->diff_beta * ksp * lapu * sqrt_u_inv;
      adt -= (adiff_all[i] - (double)cfg->diff_alpha_min) * dtpi;
      double na = adiff_all[i] + adt * dt_alpha;
      na = fmax(na, (double)cfg->diff_alpha_min);
      const double lim = (double)cfg->diff_alpha_max *
                         (1.0 - avmax / (double)cfg->alpha_max);
      na = fmin(na, lim);
      o->alpha_diff[i] = na;
    }
  }
  /* New alphas are visible to the force loop. */
  for (int i = 0; i < n; i++)
    if (is_active(cfg, p->time_bin[i]) && !is_inhibited(p->time_bin[i])) {
      alpha_all[i] = o->alpha_visc[i];
      adiff_all[i] = o->alpha_diff[i];
    }

  /* Force loop (runner_iact_nonsym_force + timebin), symmetric mask, then
   * hydro_end_force (hydro.h:1097-1101). */
#pragma omp parallel for schedule(dynamic, 1)
  for (int cid = 0; cid < g.ncells; cid++) {
    nbr nb[27];
    neighbours27(&g, cfg->dim, cid, nb);
    for (int ii = g.start[cid]; ii < g.start[cid + 1]; ii++) {
      const int i = g.index[ii];
      if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
        continue;
      const float hi = p->h[i];
      const float hig2 = hig2_of(hi);
      const double h_i = hi;
      double a[3] = {0, 0, 0}, udt = 0, hdt = 0;
      double s_a[3] = {0, 0, 0}, s_udt = 0, s_hdt = 0;
      int mintb = NUM_TIME_BINS + 1;
      int cnt = 0;
      const double mi = p->m[i], rhoi = rho_all[i], Pi = P_all[i];
      const double fi = f_all[i];
      for (int q = 0; q < 27; q++) {
        const nbr *b = &nb[q];
        for (int jj = g.start[b->cell]; jj < g.start[b->cell + 1]; jj++) {
          const int j = g.index[jj];
          if (j == i || is_inhibited(p->time_bin[j])) continue;
          float dxf[3], r2f;
          orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &b->fr, dxf, &r2f);
          const float hj = p->h[j];
          if (!(r2f < hig2 || r2f < hig2_of(hj))) continue;
          double dx[3];
          for (int k = 0; k < 3; k++)
            dx[k] = (p->x[3 * i + k] - b->t[k]) - p->x[3 * j + k];
          const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
          const double r_inv = r > 0 ? 1.0 / r : 0.0;
          const double mj = p->m[j], rhoj = rho_all[j], Pj = P_all[j];
          const double hi_inv = 1.0 / h_i, hj_inv = 1.0 / (double)hj;
          double wi, wi_dx, wj, wj_dx;
          orc_kernel_deval64(r * hi_inv, &wi, &wi_dx);
          orc_kernel_deval64(r * hj_inv, &wj, &wj_dx);
          const double wi_dr = hi_inv * hi_inv * hi_inv * hi_inv * wi_dx;
          const double wj_dr = hj_inv * hj_inv * hj_inv * hj_inv * wj_dx;
          const double dvdr =
              ((double)p->v[3 * i] - p->v[3 * j]) * dx[0] +
              ((double)p->v[3 * i + 1] - p->v[3 * j + 1]) * dx[1] +
              ((double)p->v[3 * i + 2] - p->v[3 * j + 2]) * dx[2];
          const double omega = dvdr < 0 ? dvdr : 0.0;
          const double mu = r_inv * omega;
          const double vsig = c_all[i] + c_all[j] - MUT(16, 3.0, 2.0) * mu;
          const double f_ij = 1.0 - fi / mj;
          const double f_ji = 1.0 - f_all[j] / mi;
          const double rho_ij = rhoi + rhoj;
          const double alpha = alpha_all[i] + alpha_all[j];
          const double visc = -0.25 * alpha * vsig * mu *
                              (bal_all[i] + bal_all[j]) /
                              MUT(2, rho_ij, 2.0 * rhoi);
          const double visc_acc = 0.5 * visc * (wi_dr * f_ij + wj_dr * f_ji) * r_inv;
          const double Pri = Pi / (rhoi * rhoi) * f_ij;
          const double Prj = Pj / (rhoj * rhoj) * f_ji;
          const double acc = (Pri * wi_dr + Prj * wj_dr) * r_inv + visc_acc;
          for (int k = 0; k < 3; k++) {
            a[k] -= mj * acc * dx[k];
            s_a[k] += fabs(mj * dx[k] * r_inv) *
                      (fabs(Pri * wi_dr) + fabs(Prj * wj_dr) + fabs(visc_acc * r));
          }
          const double sph_du = Pri * dvdr * r_inv * MUT(15, wi_dr, wj_dr);
          const double visc_du = MUT(1, 0.5, 1.0) * visc_acc * dvdr;
          const double ad =
              (Pi * adiff_all[i] + Pj * MUT(3, adiff_all[j], adiff_all[i])) /
              (Pi + Pj);
          const double v_diff =
              ad * 0.5 * (sqrt(2.0 * fabs(Pi - Pj) / rho_ij) + fabs(r_inv * dvdr));
          const double diff_du =
              v_diff *
              MUT(4, (double)p->u[i] - p->u[j], (double)p->u[j] - p->u[i]) *
              (f_ij * wi_dr / rhoi + f_ji * wj_dr / rhoj);
          udt += (sph_du + visc_du + diff_du) * mj;
          s_udt += (fabs(sph_du) + fabs(visc_du) + fabs(diff_du)) * mj;
          hdt -= mj * dvdr * r_inv / rhoj * wi_dr;
          s_hdt += fabs(mj * dvdr * r_inv / rhoj * wi_dr);
          if (p->time_bin[j] > 0 && p->time_bin[j] < mintb)
            mintb = p->time_bin[j];
          cnt++;
        }
      }
      for (int k = 0; k < 3; k++) {
        o->a_hydro[3 * i + k] = a[k];
        o->s_a_hydro[3 * i + k] = s_a[k];
      }
      o->s_u_dt[i] = s_udt;
      o->s_h_dt[i] = s_hdt * h_i * (1.0 / 3.0);
      o->u_dt[i] = udt;
      o->h_dt[i] = hdt * h_i * (1.0 / 3.0);
      o->min_ngb_time_bin[i] = (int8_t)mintb;
      o->n_force[i] = cnt;
    }
  }
  free(rho_all);
  free(c_all);
  free(P_all);
  free(f_all);
  free(bal_all);
  free(alpha_all);
  free(adiff_all);
  grid_free(&g);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP32 restatement including the ghost                                    */
/* ------------------------------------------------------------------------ */

/* Density sums of particle i at smoothing length hi (DOSELF1/DOPAIR1 +
 * runner_iact_nonsym_density), FP32, neighbours in cell order. */
static void density32(const orc_config *cfg, const grid *g, const orc_parts *p,
                      const nbr nb[27], int i, float hi, float acc[9],
                      int *cnt) {
  const float hig2 = hig2_of(hi);
  for (int k = 0; k < 9; k++) acc[k] = 0.f;
  int c = 0;
  const float h_inv = 1.f / hi;
  for (int q = 0; q < 27; q++) {
    const nbr *b = &nb[q];
    const int *jl;
    int ja, je;
    cand_range(g, b, q, &p->x[3 * i], KG * hi, &jl, &ja, &je);
    for (int jj = ja; jj < je; jj++) {
      const int j = jl[jj];
      if (j == i || is_inhibited(p->time_bin[j])) continue;
      float dx[3], r2;
      orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &b->fr, dx, &r2);
      if (!(r2 < hig2)) continue;
      const float r = sqrtf(r2);
      const float mj = p->m[j];
      const float ui = r * h_inv;
      float wi, wi_dx;
      orc_kernel_deval32(ui, &wi, &wi_dx);
      acc[0] += mj * wi;
      acc[1] -= mj * (HDIM * wi + ui * wi_dx);
      acc[2] += wi;
      acc[3] -= (HDIM * wi + ui * wi_dx);
      const float r_inv = r ? 1.0f / r : 0.0f;
      const float faci = mj * wi_dx * r_inv;
      const float dv[3] = {p->v[3 * i] - p->v[3 * j],
                           p->v[3 * i + 1] - p->v[3 * j + 1],
                           p->v[3 * i + 2] - p->v[3 * j + 2]};
      const float dvdr = dv[0] * dx[0] + dv[1] * dx[1] + dv[2] * dx[2];
      acc[4] -= faci * dvdr;
      acc[5] += faci * (dv[1] * dx[2] - dv[2] * dx[1]);
      acc[6] += faci * (dv[2] * dx[0] - dv[0] * dx[2]);
      acc[7] += faci * (dv[0] * dx[1] - dv[1] * dx[0]);
      c++;
    }
  }
  (void)cfg;
  *cnt = c;
}

int orc_pass_fp32(const orc_config *cfg, orc_parts *p, orc_out32 *o) {
  grid g;
#ifdef _OPENMP
  if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
  if (grid_build(cfg, p, &g)) return 1;
#ifdef ORC_SORTED
  if (grid_sort(p, &g)) return 1;
#endif
  const int n = p->n;
  const float eta_dim = cfg->eta * cfg->eta * cfg->eta;
  const float eps = cfg->h_tolerance;
  int max_rounds = 0;
  int failed = 0;

  /* Density + ghost (runner_ghost.c:1113-1628).  A particle's density sums
   * depend only on its own h, so the cell-level redo rounds of the reference
   * are equivalent to iterating each particle on its own. */
  float *c_all = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1));
#pragma omp parallel for schedule(dynamic, 1) reduction(max : max_rounds) reduction(+ : failed)
  for (int cid = 0; cid < g.ncells; cid++) {
    nbr nb[27];
    neighbours27(&g, cfg->dim, cid, nb);
    for (int ii = g.start[cid]; ii < g.start[cid + 1]; ii++) {
      const int i = g.index[ii];
      if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
        continue;
      float left = 0.f, right = cfg->h_max;
      float acc[9];
      int cnt = 0;
      int converged = 0;
      int rounds = 0;
      float rho = 0, rho_dh = 0, wc = 0, wc_dh = 0, divv = 0, rot[3] = {0, 0, 0};
      for (int it = 0; it < cfg->max_iter && !converged; it++) {
        rounds = it + 1;
        const float h_old = p->h[i];
        density32(cfg, &g, p, nb, i, h_old, acc, &cnt);
        rho = acc[0];
        rho_dh = acc[1];
        wc = acc[2];
        wc_dh = acc[3];
        divv = acc[4];
        rot[0] = acc[5];
        rot[1] = acc[6];
        rot[2] = acc[7];
        const float h_old_dim = h_old * h_old * h_old;
        const float h_old_dim_m1 = h_old * h_old;
        float h_new;
        int has_no_ngb = 0;
        if (wc < 1.e-5 * KROOT) {
          has_no_ngb = 1;
          h_new = 2.f * h_old;
        } else {
          /* hydro_end_density (hydro.h:613-662) */
          const float h_inv = 1.0f / h_old;
          const float hid = h_inv * h_inv * h_inv;
          const float hid1 = hid * h_inv;
          const float mi = p->m[i];
          rho += mi * KROOT;
          rho_dh -= HDIM * mi * KROOT;
          wc += KROOT;
          wc_dh -= HDIM * KROOT;
          rho *= hid;
          rho_dh *= hid1;
          wc *= hid;
          wc_dh *= hid1;
          const float rho_inv = 1.f / rho;
          for (int k = 0; k < 3; k++) rot[k] *= hid1 * 1.f * rho_inv;
          divv *= hid1 * rho_inv * 1.f;
          divv += 0.f * HDIM;
          const float n_sum = wc * h_old_dim;
          const float f = n_sum - eta_dim;
          const float f_prime = wc_dh * h_old_dim + HDIM * wc * h_old_dim_m1;
          if (n_sum < eta_dim)
            left = fmaxf(left, h_old);
          else if (n_sum > eta_dim)
            right = fminf(right, h_old);
          if (((p->h[i] >= cfg->h_max) && (f < 0.f)) ||
              ((p->h[i] <= cfg->h_min) && (f > 0.f))) {
            converged = 1;
            break;
          }
          h_new = h_old - f / (f_prime + FLT_MIN);
          h_new = fminf(h_new, 2.f * h_old);
          h_new = fmaxf(h_new, 0.5f * h_old);
          h_new = fmaxf(h_new, left);
          h_new = fminf(h_new, right);
        }
        if (fabsf(h_new - h_old) > eps * h_old) {
          if ((h_new == left && h_old == right) ||
              (h_old == left && h_new == right))
            p->h[i] =
                orc_cbrt32(0.5f * (left * left * left + right * right * right));
          else
            p->h[i] = h_new;
          if (p->h[i] < cfg->h_max && p->h[i] > cfg->h_min) continue;
          if (p->h[i] <= cfg->h_min)
            p->h[i] = cfg->h_min;
          else {
            p->h[i] = cfg->h_max;
            if (has_no_ngb) {
              /* hydro_part_has_no_neighbours (hydro.h:791-825) */
              const float hinv = 1.0f / p->h[i];
              const float hid = hinv * hinv * hinv;
              rho = p->m[i] * KROOT * hid;
              wc = KROOT * hid;
              rho_dh = 0.f;
              wc_dh = 0.f;
              rot[0] = rot[1] = rot[2] = 0.f;
              divv = 0.f;
            }
          }
        }
        converged = 1;
      }
      if (!converged) failed++;
      if (rounds > max_rounds) max_rounds = rounds;
      /* hydro_prepare_gradient + hydro_reset_gradient (hydro.h:671-750) */
      const float h = p->h[i];
      const float curl_v =
          sqrtf(rot[0] * rot[0] + rot[1] * rot[1] + rot[2] * rot[2]);
      const float abs_div_v = fabsf(divv);
      const float P = GAS_GM1 * p->u[i] * rho;
      const float c = sqrtf(GAS_GAMMA * P / rho);
      const float bal = abs_div_v / (abs_div_v + curl_v + 0.0001f * c * 1.f / h);
      const float common = h * HDIM_INV / wc;
      float fgh;
      if (h > 0.9999f * cfg->h_max)
        fgh = 0.f;
      else {
        const float gW = common * wc_dh;
        fgh = (gW < -0.9999f) ? 0.f : common * rho_dh / (1.f + gW);
      }
      p->rho[i] = rho;
      p->pressure[i] = P;
      p->soundspeed[i] = c;
      p->f[i] = fgh;
      p->balsara[i] = bal;
      o->rho_dh[i] = rho_dh;
      o->wcount[i] = wc;
      o->wcount_dh[i] = wc_dh;
      o->div_v[i] = divv;
      for (int k = 0; k < 3; k++) o->rot_v[3 * i + k] = rot[k];
      o->n_density[i] = cnt;
      c_all[i] = c;
    }
  }
  if (failed) {
    fprintf(stderr, "oracle: smoothing length failed to converge on %d particles\n",
            failed);
    free(c_all);
    grid_free(&g);
    return 2;
  }
  if (check_h_fits(cfg, &g, p)) {
    free(c_all);
    grid_free(&g);
    return 1;
  }
  for (int i = 0; i < n; i++)
    if (!(is_active(cfg, p->time_bin[i]) && !is_inhibited(p->time_bin[i])))
      c_all[i] = p->soundspeed[i];

  /* Gradient loop + hydro_end_gradient + hydro_prepare_force (FP32). */
  float *alpha_new = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1));
  float *adiff_new = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1));
#pragma omp parallel for schedule(dynamic, 1)
  for (int cid = 0; cid < g.ncells; cid++) {
    nbr nb[27];
    neighbours27(&g, cfg->dim, cid, nb);
    for (int ii = g.start[cid]; ii < g.start[cid + 1]; ii++) {
      const int i = g.index[ii];
      alpha_new[i] = p->alpha_visc[i];
      adiff_new[i] = p->alpha_diff[i];
      if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
        continue;
      const float hi = p->h[i];
      const float hig2 = hig2_of(hi);
      float vsig = 2.f * c_all[i];
      float avmax = p->alpha_visc[i];
      float lapu = 0.f;
      int cnt = 0;
      for (int q = 0; q < 27; q++) {
        const nbr *b = &nb[q];
        const int *jl;
        int ja, je;
        cand_range(&g, b, q, &p->x[3 * i], KG * hi, &jl, &ja, &je);
        for (int jj = ja; jj < je; jj++) {
          const int j = jl[jj];
          if (j == i || is_inhibited(p->time_bin[j])) continue;
          float dx[3], r2;
          orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &b->fr, dx, &r2);
          if (!(r2 < hig2)) continue;
          const float r = sqrtf(r2);
          const float r_inv = r ? 1.0f / r : 0.0f;
          const float dvdr = (p->v[3 * i] - p->v[3 * j]) * dx[0] +
                             (p->v[3 * i + 1] - p->v[3 * j + 1]) * dx[1] +
                             (p->v[3 * i + 2] - p->v[3 * j + 2]) * dx[2];
          const float omega = fminf(dvdr, 0.f);
          const float mu = 1.f * r_inv * omega;
          const float nv = c_all[i] + c_all[j] - VISC_BETA * mu;
          vsig = fmaxf(vsig, nv);
          float wi, wi_dx;
          orc_kernel_deval32(r / hi, &wi, &wi_dx);
          const float duf = (p->u[i] - p->u[j]) * r_inv;
          lapu += p->m[j] * duf * wi_dx / p->rho[j];
          avmax = fmaxf(avmax, p->alpha_visc[j]);
          cnt++;
        }
      }
      const float h_inv = 1.0f / hi;
      const float hid = h_inv * h_inv * h_inv;
      lapu *= 2.f * (hid * h_inv);
      o->v_sig[i] = vsig;
      o->laplace_u[i] = lapu;
      o->alpha_visc_max_ngb[i] = avmax;
      o->n_gradient[i] = cnt;
      /* hydro_prepare_force (hydro.h:840-954), a = 1 */
      const float dt_alpha = (float)bin_dt(p->time_bin[i], cfg->time_base);
      const float ksp = hi * 1.f * KG;
      const float ksp_inv = 1.f / ksp;
      const float P = GAS_GM1 * p->u[i] * p->rho[i];
      const float c = sqrtf(GAS_GAMMA * P / p->rho[i]) * 1.f;
      const float sct_inv = c * ksp_inv;
      const float div_v_dt =
          dt_alpha == 0.f ? 0.f : (o->div_v[i] - p->div_v_prev[i]) / dt_alpha;
      const float S =
          o->div_v[i] < 0.f ? ksp * ksp * fmaxf(0.f, -1.f * div_v_dt) : 0.f;
      const float c2 = c * c;
      const float alpha_loc = cfg->alpha_max * S / (c2 + S);
      float alpha = p->alpha_visc[i];
      if (alpha_loc > alpha)
        alpha = alpha_loc;
      else {
        const float tr = dt_alpha * sct_inv * cfg->visc_length;
        alpha += alpha_loc * tr;
        alpha /= (1.f + tr);
      }
      alpha = fmaxf(alpha, cfg->alpha_min);
      o->div_v_dt[i] = div_v_dt;
      const float dtpi = vsig * 1.f * ksp_inv;
      const float sqrt_u_inv = 1.f / sqrtf(p->u[i]);
      float adt = cfg->diff_beta * ksp * lapu * 1.f * sqrt_u_inv * 1.f;
      adt -= (p->alpha_diff[i] - cfg->diff_alpha_min) * dtpi;
      float na = p->alpha_diff[i];
      na += adt * dt_alpha;
      na = fmaxf(na, cfg->diff_alpha_min);
      const float lim = cfg->diff_alpha_max * (1.f - avmax / cfg->alpha_max);
      na = fminf(na, lim);
      alpha_new[i] = alpha;
      adiff_new[i] = na;
    }
  }
  for (int i = 0; i < n; i++) {
    if (is_active(cfg, p->time_bin[i]) && !is_inhibited(p->time_bin[i]))
      p->div_v_prev[i] = o->div_v[i];
    p->alpha_visc[i] = alpha_new[i];
    p->alpha_diff[i] = adiff_new[i];
  }

  /* Force loop (runner_iact_nonsym_force + timebin) + hydro_end_force. */
#ifdef ORC_SORTED
  grid_hmax(p, &g);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int cid = 0; cid < g.ncells; cid++) {
    nbr nb[27];
    neighbours27(&g, cfg->dim, cid, nb);
    float pi[P_COUNT], pj[P_COUNT];
    for (int ii = g.start[cid]; ii < g.start[cid + 1]; ii++) {
      const int i = g.index[ii];
      if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
        continue;
      const float hi = p->h[i];
      const float hig2 = hig2_of(hi);
      memset(pi, 0, sizeof(pi));
      pi[P_V0] = p->v[3 * i];
      pi[P_V1] = p->v[3 * i + 1];
      pi[P_V2] = p->v[3 * i + 2];
      pi[P_M] = p->m[i];
      pi[P_U] = p->u[i];
      pi[P_RHO] = p->rho[i];
      pi[P_F] = p->f[i];
      pi[P_P] = p->pressure[i];
      pi[P_C] = p->soundspeed[i];
      pi[P_BAL] = p->balsara[i];
      pi[P_ALPHA] = p->alpha_visc[i];
      pi[P_ADIFF] = p->alpha_diff[i];
      pi[P_TB] = p->time_bin[i];
      pi[P_MINTB] = NUM_TIME_BINS + 1;
      for (int q = 0; q < 27; q++) {
        const nbr *b = &nb[q];
        const int *jl;
        int ja, je;
#ifdef ORC_SORTED
        const float Rq = fmaxf(KG * hi, g.hgmax[b->cell]);
#else
        const float Rq = KG * hi;
#endif
        cand_range(&g, b, q, &p->x[3 * i], Rq, &jl, &ja, &je);
        for (int jj = ja; jj < je; jj++) {
          const int j = jl[jj];
          if (j == i || is_inhibited(p->time_bin[j])) continue;
          float dx[3], r2;
          orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &b->fr, dx, &r2);
          const float hj = p->h[j];
          if (!(r2 < hig2 || r2 < hig2_of(hj))) continue;
          pj[P_V0] = p->v[3 * j];
          pj[P_V1] = p->v[3 * j + 1];
          pj[P_V2] = p->v[3 * j + 2];
          pj[P_M] = p->m[j];
          pj[P_U] = p->u[j];
          pj[P_RHO] = p->rho[j];
          pj[P_F] = p->f[j];
          pj[P_P] = p->pressure[j];
          pj[P_C] = p->soundspeed[j];
          pj[P_BAL] = p->balsara[j];
          pj[P_ALPHA] = p->alpha_visc[j];
          pj[P_ADIFF] = p->alpha_diff[j];
          pj[P_TB] = p->time_bin[j];
          orc_iact_force32(r2, dx, hi, hj, pi, pj, 0);
        }
      }
      o->a_hydro[3 * i] = pi[P_A0];
      o->a_hydro[3 * i + 1] = pi[P_A1];
      o->a_hydro[3 * i + 2] = pi[P_A2];
      o->u_dt[i] = pi[P_UDT];
      o->h_dt[i] = pi[P_HDT] * (hi * HDIM_INV);
      o->n_force[i] = (int32_t)pi[P_NFORCE];
      o->min_ngb_time_bin[i] = (int8_t)pi[P_MINTB];
    }
  }
  if (o->ghost_iters) o->ghost_iters[0] = max_rounds;
  free(alpha_new);
  free(adiff_new);
  free(c_all);
  grid_free(&g);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* O(N^2) brute-force counts: every pair of particles, in the reference's   */
/* brute-force checkers' own expression (self_all_density tools.c:590-650, */
/* pairs_all_density / pairs_all_force tools.c:207-470): a pair in one      */
/* top-level cell rounds the FP64 difference, any other pair works in the   */
/* frame of its two cells (orc_pair_frame) and folds dx with nearest()      */
/* (periodic.h:66-72).  No neighbour list, no stencil.                      */
/* ------------------------------------------------------------------------ */
int orc_count_bruteforce(const orc_config *cfg, const orc_parts *p, int mode,
                         int32_t *counts) {
  grid g;
  if (grid_build(cfg, p, &g)) return 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; i++) {
    counts[i] = 0;
    if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
      continue;
    const float hig2 = hig2_of(p->h[i]);
    const int ci = g.cell_of[i];
    const int cix[3] = {ci / (g.nc[2] * g.nc[1]), (ci / g.nc[2]) % g.nc[1],
                        ci % g.nc[2]};
    double loc_i[3];
    for (int k = 0; k < 3; k++) loc_i[k] = cix[k] * g.w[k];
    int c = 0;
    for (int j = 0; j < p->n; j++) {
      if (j == i || is_inhibited(p->time_bin[j])) continue;
      const int cj = g.cell_of[j];
      const int cjx[3] = {cj / (g.nc[2] * g.nc[1]), (cj / g.nc[2]) % g.nc[1],
                          cj % g.nc[2]};
      double loc_j[3];
      for (int k = 0; k < 3; k++) loc_j[k] = cjx[k] * g.w[k];
      orc_frame fr;
      orc_pair_frame(loc_i, loc_j, cfg->dim, &fr);
      float dx[3], r2;
      orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &fr, dx, &r2);
      if (!fr.self) {
        for (int k = 0; k < 3; k++) {
          const double b = cfg->dim[k];
          dx[k] = (float)(dx[k] > 0.5 * b ? dx[k] - b
                                          : (dx[k] < -0.5 * b ? dx[k] + b
                                                              : (double)dx[k]));
        }
        r2 = fmaf(dx[2], dx[2], fmaf(dx[1], dx[1], dx[0] * dx[0]));
      }
      const int hit = mode == 0 ? (r2 < hig2)
                                : (r2 < hig2 || r2 < hig2_of(p->h[j]));
      c += hit;
    }
    counts[i] = c;
  }
  grid_free(&g);
  return 0;
}

/* Mask decisions of the density loop (r2 < (gamma_K h_i)^2 over the 27-cell
 * stencil) on which the reference's contract (orc_pair_dx32) and the single
 * FP64-difference rounding (float)((x_i - t) - x_j) disagree: out[0]
 * neighbours only under the reference's, out[1] only under the other,
 * out[2] candidate pairs examined, out[3] pairs inside under both. */
int orc_contract_flips(const orc_config *cfg, const orc_parts *p,
                       int64_t out[4]) {
  grid g;
  if (grid_build(cfg, p, &g)) return 1;
  int64_t a = 0, b = 0, tot = 0, both = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : a, b, tot, both)
  for (int cid = 0; cid < g.ncells; cid++) {
    nbr nb[27];
    neighbours27(&g, cfg->dim, cid, nb);
    for (int ii = g.start[cid]; ii < g.start[cid + 1]; ii++) {
      const int i = g.index[ii];
      if (!is_active(cfg, p->time_bin[i]) || is_inhibited(p->time_bin[i]))
        continue;
      const float hig2 = hig2_of(p->h[i]);
      for (int q = 0; q < 27; q++) {
        const nbr *bq = &nb[q];
        for (int jj = g.start[bq->cell]; jj < g.start[bq->cell + 1]; jj++) {
          const int j = g.index[jj];
          if (j == i || is_inhibited(p->time_bin[j])) continue;
          float dx[3], r2, dx2[3], r22;
          orc_pair_dx32(&p->x[3 * i], &p->x[3 * j], &bq->fr, dx, &r2);
          pair_dx32_fp64(&p->x[3 * i], &p->x[3 * j], bq->t, dx2, &r22);
          const int in1 = r2 < hig2, in2 = r22 < hig2;
          a += in1 && !in2;
          b += in2 && !in1;
          both += in1 && in2;
          tot++;
        }
      }
    }
  }
  out[0] = a;
  out[1] = b;
  out[2] = tot;
  out[3] = both;
  grid_free(&g);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Time integration.                                                         */
/* ------------------------------------------------------------------------ */
#define ORC_NUM_TIME_BINS 56
#define ORC_MAX_NR_TIMESTEPS (1LL << (ORC_NUM_TIME_BINS + 1))
#define ORC_BIN_INHIBITED (ORC_NUM_TIME_BINS + 2)

/* timeline.h: get_integer_timestep, get_time_bin, get_integer_time_begin,
 * get_integer_time_end, get_max_active_bin. */
static long long tl_step(int bin) { return bin <= 0 ? 0 : (1LL << (bin + 1)); }
static int tl_bin(long long dti) {
  return (8 * (int)sizeof(long long) - 2) - __builtin_clzll((unsigned long long)dti);
}
static long long tl_begin(long long ti, int bin) {
  const long long dti = tl_step(bin);
  return dti == 0 ? 0 : dti * ((ti - 1) / dti);
}
static long long tl_end(long long ti, int bin) {
  const long long dti = tl_step(bin);
  if (dti == 0) return 0;
  const long long mod = ti % dti;
  return mod == 0 ? ti : ti - mod + dti;
}
int orc_max_active_bin(long long ti) {
  if (ti == 0) return ORC_NUM_TIME_BINS;
  int bin = 1;
  while (!((1LL << (bin + 1)) & ti)) ++bin;
  return bin;
}

/* approx_math.h:35 */
static float orc_approx_expf(float x) {
  return 1.f + x * (1.f + x * (0.5f + x * (1.f / 6.f + 1.f / 24.f * x)));
}

/* timestep.h:49-89 (0 for a step that rounds to zero time-line units) */
long long orc_make_integer_timestep(float new_dt, int old_bin, int min_ngb_bin,
                                    long long ti_current, double time_base_inv) {
  long long new_dti = (long long)(new_dt * time_base_inv);
  if (new_dti <= 0) return 0;
  int new_bin = tl_bin(new_dti);
  if (min_ngb_bin + 2 < new_bin) new_bin = min_ngb_bin + 2;
  new_dti = tl_step(new_bin);
  const long long current_dti = tl_step(old_bin);
  const long long ti_end = tl_end(ti_current, old_bin);
  if (old_bin > 0 && 2 * current_dti < new_dti) new_dti = 2 * current_dti;
  long long dti_timeline = ORC_MAX_NR_TIMESTEPS;
  while (new_dti < dti_timeline) dti_timeline /= 2;
  new_dti = dti_timeline;
  if (new_dti > current_dti) {
    if ((ORC_MAX_NR_TIMESTEPS - ti_end) % new_dti > 0) new_dti = current_dti;
  }
  return new_dti;
}

/* runner_do_drift_part -> drift_part (drift.h:138-200) with
 * hydro_predict_extra (hydro.h:1029-1078); no gravity, no cosmology:
 * dt_drift = dt_kick_hydro = dt_therm = (ti_current - ti_old) time_base. */
void orc_drift(orc_gas *g, const orc_integ *c, long long ti_old,
               long long ti_current) {
  const double dt = (double)(ti_current - ti_old) * c->time_base;
  for (int i = 0; i < g->n; i++) {
    if (g->time_bin[i] == ORC_BIN_INHIBITED) continue;
    for (int k = 0; k < 3; k++) {
      g->x[3 * i + k] += g->v_full[3 * i + k] * dt;
      g->v[3 * i + k] += g->a_hydro[3 * i + k] * dt;
    }
    const float dt_drift = (float)dt, dt_therm = (float)dt;
    g->u[i] += g->u_dt[i] * dt_therm;
    const float h_inv = 1.f / g->h[i];
    const float w1 = g->h_dt[i] * h_inv * dt_drift;
    if (fabsf(w1) < 0.2f)
      g->h[i] *= orc_approx_expf(w1);
    else
      g->h[i] *= expf(w1);
    const float w2 = -HDIM * w1;
    if (fabsf(w2) < 0.2f)
      g->rho[i] *= orc_approx_expf(w2);
    else
      g->rho[i] *= expf(w2);
    /* no entropy floor (floor u = 0); absolute minimum */
    g->u[i] = fmaxf(g->u[i], 0.f);
    g->u[i] = fmaxf(g->u[i], c->min_u);
    const float P = GAS_GM1 * g->u[i] * g->rho[i];
    const float cs = sqrtf(GAS_GAMMA * P / g->rho[i]);
    g->pressure[i] = P;
    g->soundspeed[i] = cs;
    g->v_sig[i] = fmaxf(g->v_sig[i], 2.f * cs);
  }
}

/* runner_do_kick1 (runner_time_integration.c:143-183) / runner_do_kick2
 * (:408-441) -> kick_part (kick.h) -> hydro_kick_extra (hydro.h:1120-1151);
 * kick2 ends with hydro_reset_predicted_values (hydro.h:983-1006). */
void orc_kick(orc_gas *g, const orc_integ *c, long long ti_current, int mode) {
  const int max_active_bin = orc_max_active_bin(ti_current);
  for (int i = 0; i < g->n; i++) {
    const int tb = g->time_bin[i];
    if (tb == ORC_BIN_INHIBITED) continue;
    const long long dti = tl_step(tb);
    long long ti_begin, ti_end;
    if (mode == 1) {
      ti_begin = tl_begin(ti_current + 1, tb);
      if (ti_begin != ti_current) continue;
      ti_end = ti_begin + dti / 2;
    } else {
      if (tb > max_active_bin) continue;
      ti_begin = tl_begin(ti_current, tb) + dti / 2;
      ti_end = ti_begin + dti / 2;
    }
    const double dt = (double)(ti_end - ti_begin) * c->time_base;
    for (int k = 0; k < 3; k++) g->v_full[3 * i + k] += g->a_hydro[3 * i + k] * dt;
    const float dt_therm = (float)dt;
    const float delta_u = g->u_dt[i] * dt_therm;
    g->u_full[i] = fmaxf(g->u_full[i] + delta_u, 0.5f * g->u_full[i]);
    const float energy_min = fmaxf(c->min_u, 0.f);
    if (g->u_full[i] < energy_min) {
      g->u_full[i] = energy_min;
      g->u_dt[i] = 0.f;
    }
    if (mode == 2) {
      for (int k = 0; k < 3; k++) g->v[3 * i + k] = g->v_full[3 * i + k];
      g->u[i] = g->u_full[i];
      const float P = GAS_GM1 * g->u[i] * g->rho[i];
      const float cs = sqrtf(GAS_GAMMA * P / g->rho[i]);
      g->pressure[i] = P;
      g->soundspeed[i] = cs;
      g->v_sig[i] = fmaxf(g->v_sig[i], 2.f * cs);
    }
  }
}

/* runner_do_timestep (runner_time_integration.c:639-760) with
 * get_part_timestep (timestep.h) and hydro_compute_timestep
 * (hydro.h:475-487): a = 1, a_factor_sound_speed = 1, no other criteria. */
int orc_timestep(orc_gas *g, const orc_integ *c, long long ti_current,
                 long long *ti_end_min, long long *n_active) {
  const int max_active_bin = orc_max_active_bin(ti_current);
  long long tmin = ORC_MAX_NR_TIMESTEPS, nact = 0;
  int bad = 0;
  for (int i = 0; i < g->n; i++) {
    const int tb = g->time_bin[i];
    if (tb == ORC_BIN_INHIBITED) continue;
    long long te;
    if (tb <= max_active_bin) {
      nact++;
      long long new_dti;
      if (c->fixed_bins) {
        new_dti = tl_step(tb);
        if (new_dti == 0) {
          bad++;
          new_dti = 1;
        }
      } else {
        const double a = 1.0, a_fss = 1.0;
        const float dt_cfl =
            2.f * KG * c->cfl * a * g->h[i] / (a_fss * g->v_sig[i]);
        float new_dt = dt_cfl;
        const float dt_h_change =
            g->h_dt[i] != 0.0f
                ? fabsf(c->log_max_h_change * g->h[i] / g->h_dt[i])
                : FLT_MAX;
        if (dt_h_change < new_dt) new_dt = dt_h_change;
        new_dt = (float)((double)new_dt < c->dt_max ? (double)new_dt : c->dt_max);
        if ((double)new_dt < c->dt_min) {
          bad++;
          new_dti = tl_step(tb);
        } else {
          new_dti = orc_make_integer_timestep(new_dt, tb, g->min_ngb_time_bin[i],
                                              ti_current, 1.0 / c->time_base);
          if (new_dti == 0) {
            bad++;
            new_dti = tl_step(tb);
          } else {
            g->time_bin[i] = (int8_t)tl_bin(new_dti);
          }
        }
      }
      te = ti_current + new_dti;
    } else {
      te = tl_end(ti_current, tb);
    }
    if (te < tmin) tmin = te;
  }
  if (ti_end_min) *ti_end_min = tmin;
  if (n_active) *n_active = nact;
  return bad;
}
