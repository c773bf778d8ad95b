/*
 * abi_27cells.c -- the test27cells call sequence through the C ABI of
 * libswsph.so (include/swsph.h), checked against the CPU oracle
 * (oracle/sphenix_oracle.h, the checker).
 *
 * Like the reference's tests/test27cells.c (a C program driving the runner
 * functions on 27 cells and comparing with a brute-force rerun,
 * tests/test27cells.c:634-655), it builds a periodic box of 3^3 top-level
 * cells (a jittered lattice with random velocities and internal energies),
 * drives the drop-in task entry points in the reference's dependency order
 * -- hydro_init_part, the self and pair density tasks, the ghost (split in
 * two shares on every other cell, as runner_do_ghost's offset / ntasks),
 * gradient, extra ghost, force, end force -- writes the results of every
 * cell into an array of structs shaped like struct part with
 * swsph_cell_write_back, and compares them with the oracle's FP64 rerun at
 * the converged smoothing lengths: neighbour counts and min_ngb_time_bin
 * exactly; density quantities within 1e-5 and force quantities within 1e-4
 * of max(|value|, sum of |terms|).
 *
 * Usage: abi_27cells [n_per_axis]   (exit status 0 on success)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/swsph.h"
#include "../../oracle/sphenix_oracle.h"

/* The caller's particle record: an array of structs, as SWIFT's
 * struct part (src/hydro/SPHENIX/hydro_part.h:105-324), trimmed to the
 * fields of this path. */
struct part {
  double x[3];
  float v[3];
  float m, u, h;
  float rho, pressure, soundspeed, f, balsara, alpha_visc, alpha_diff;
  float div_v_prev;
  float wcount, div_v, rot_v[3], v_sig, laplace_u;
  float a_hydro[3], u_dt, h_dt;
  int32_t n_density, n_force;
  int8_t time_bin, min_ngb_time_bin;
};

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static double uniform01(void) { /* xorshift64* */
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return (double)((rng_state * 2685821657736338717ull) >> 11) * 0x1p-53;
}

static int failures = 0;
#define CALL(x)                                                             \
  do {                                                                      \
    if ((x) != 0) {                                                         \
      fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x,             \
              swsph_last_error());                                          \
      exit(2);                                                              \
    }                                                                       \
  } while (0)

static int cid(int cx, int cy, int cz) {
  cx = (cx + 3) % 3;
  cy = (cy + 3) % 3;
  cz = (cz + 3) % 3;
  return (cx * 3 + cy) * 3 + cz;
}

/* Relative error of a value against the FP64 rerun, on the scale
 * max(|ref|, sum of |terms|). */
static void check_rel(const char *name, int i, double got, double ref,
                      double scale, double tol, double *worst) {
  const double s = fmax(fabs(ref), scale);
  const double e = s > 0 ? fabs(got - ref) / s : fabs(got - ref);
  if (e > *worst) *worst = e;
  if (e > tol && failures < 20) {
    fprintf(stderr, "%s[%d]: %.9g vs %.9g (rel %.3g > %.1g)\n", name, i, got,
            ref, e, tol);
    failures++;
  } else if (e > tol) {
    failures++;
  }
}

int main(int argc, char **argv) {
  const int n1 = argc > 1 ? atoi(argv[1]) : 15;
  const int n = n1 * n1 * n1;
  const double sp = 1.0 / n1;
  struct part *parts = calloc((size_t)n, sizeof(struct part));
  if (!parts) return 2;
  for (int a = 0, k = 0; a < n1; a++)
    for (int b = 0; b < n1; b++)
      for (int c = 0; c < n1; c++, k++) {
        const int ijk[3] = {a, b, c};
        for (int d = 0; d < 3; d++) {
          double x = (ijk[d] + 0.5 + 0.2 * (uniform01() - 0.5)) * sp;
          if (x >= 1.0) x -= 1.0;
          if (x < 0.0) x += 1.0;
          parts[k].x[d] = x;
          parts[k].v[d] = (float)(0.2 * (uniform01() - 0.5));
        }
        parts[k].m = (float)(1.0 / n);
        parts[k].u = (float)(0.5 + uniform01());
        parts[k].h = (float)(1.2348 * sp);
        parts[k].alpha_visc = 0.1f;
        parts[k].time_bin = 1;
      }

  /* The particles as the ABI reads them: strided views would need copies,
   * so the shim-side SoA arrays are gathered once. */
  double *x = malloc(sizeof(double) * 3 * n);
  float *v = malloc(sizeof(float) * 3 * n), *m = malloc(sizeof(float) * n);
  float *u = malloc(sizeof(float) * n), *h = malloc(sizeof(float) * n);
  float *av = malloc(sizeof(float) * n);
  int8_t *tb = malloc((size_t)n);
  for (int i = 0; i < n; i++) {
    memcpy(&x[3 * i], parts[i].x, sizeof(parts[i].x));
    memcpy(&v[3 * i], parts[i].v, sizeof(parts[i].v));
    m[i] = parts[i].m;
    u[i] = parts[i].u;
    h[i] = parts[i].h;
    av[i] = parts[i].alpha_visc;
    tb[i] = parts[i].time_bin;
  }
  swsph_props props = {.dim = {1.0, 1.0, 1.0}, .ncell = {3, 3, 3},
                       .eta = 1.2348f, .h_tolerance = 1e-4f,
                       .h_max = 3.402823466e38f, .h_min = 0.f,
                       .max_iter = 30, .alpha_max = 2.f, .alpha_min = 0.f,
                       .visc_length = 0.05f, .diff_beta = 1.f,
                       .diff_alpha_max = 1.f, .diff_alpha_min = 0.f,
                       .time_base = 0.0, .max_active_bin = 56};
  swsph_host_parts hp = {.n = n, .x = x, .v = v, .m = m, .u = u,
                         .time_bin = tb, .h = h, .alpha_visc = av};
  swsph_engine *e = NULL;
  CALL(swsph_engine_create(&props, &e));
  CALL(swsph_set_particles(e, &hp));

  /* The task graph of one step for the 27 cells, in dependency order
   * (engine_make_hydro_loops_dependencies, src/engine_maketasks.c:
   * 2546-2573); every unordered pair of adjacent cells once (the upper
   * half of each cell's stencil). */
  for (int c = 0; c < 27; c++) CALL(swsph_init_cell(e, c));
  for (int loop = 0; loop < 3; loop++) {
    for (int c = 0; c < 27; c++) {
      const int cx = c / 9, cy = (c / 3) % 3, cz = c % 3;
      if (loop == 0) CALL(swsph_doself1_density(e, c));
      if (loop == 1) CALL(swsph_doself1_gradient(e, c));
      if (loop == 2) CALL(swsph_doself2_force(e, c));
      for (int o = 14; o < 27; o++) {
        const int d = cid(cx + o / 9 - 1, cy + (o / 3) % 3 - 1, cz + o % 3 - 1);
        if (loop == 0) CALL(swsph_dopair1_density(e, c, d));
        if (loop == 1) CALL(swsph_dopair1_gradient(e, c, d));
        if (loop == 2) CALL(swsph_dopair2_force(e, c, d));
      }
    }
    for (int c = 0; c < 27; c++) {
      if (loop == 0) {
        if (c % 2) {
          CALL(swsph_do_ghost(e, c, 0, 2));
          CALL(swsph_do_ghost(e, c, 1, 2));
        } else {
          CALL(swsph_do_ghost(e, c, 0, 1));
        }
      }
      if (loop == 1) CALL(swsph_do_extra_ghost(e, c));
      if (loop == 2) CALL(swsph_do_end_hydro_force(e, c));
    }
  }

  /* Results into the array of structs, cell by cell. */
#define DST(fid, member) \
  {fid, &parts[0].member, (int64_t)sizeof(struct part)}
  const swsph_field_dst dst[] = {
      DST(SWSPH_H, h), DST(SWSPH_RHO, rho), DST(SWSPH_WCOUNT, wcount),
      DST(SWSPH_DIV_V, div_v), DST(SWSPH_ROT_V, rot_v),
      DST(SWSPH_PRESSURE, pressure), DST(SWSPH_SOUNDSPEED, soundspeed),
      DST(SWSPH_F, f), DST(SWSPH_BALSARA, balsara), DST(SWSPH_V_SIG, v_sig),
      DST(SWSPH_LAPLACE_U, laplace_u), DST(SWSPH_ALPHA_VISC, alpha_visc),
      DST(SWSPH_ALPHA_DIFF, alpha_diff), DST(SWSPH_A_HYDRO, a_hydro),
      DST(SWSPH_U_DT, u_dt), DST(SWSPH_H_DT, h_dt),
      DST(SWSPH_N_DENSITY, n_density), DST(SWSPH_N_FORCE, n_force),
      DST(SWSPH_MIN_NGB_TIME_BIN, min_ngb_time_bin)};
  const int ndst = (int)(sizeof(dst) / sizeof(dst[0]));
  int total = 0;
  for (int c = 0; c < 27; c++) {
    int cnt = 0;
    CALL(swsph_cell_count(e, c, &cnt));
    total += cnt;
    CALL(swsph_cell_write_back(e, c, dst, ndst));
  }
  if (total != n) {
    fprintf(stderr, "cells hold %d of %d particles\n", total, n);
    return 1;
  }
  CALL(swsph_engine_destroy(e));

  /* The oracle's FP64 rerun at the converged h, with the initial (io)
   * state of the pass. */
  float *hh = malloc(sizeof(float) * n);
  float *zero = calloc((size_t)n, sizeof(float));
  float *rho0 = calloc((size_t)n, sizeof(float));
  float *av0 = malloc(sizeof(float) * n);
  for (int i = 0; i < n; i++) {
    hh[i] = parts[i].h;
    av0[i] = 0.1f;
  }
  orc_config cfg = {.dim = {1.0, 1.0, 1.0}, .ncell = {3, 3, 3},
                    .eta = props.eta, .h_tolerance = props.h_tolerance,
                    .h_max = props.h_max, .h_min = props.h_min,
                    .max_iter = props.max_iter, .alpha_max = props.alpha_max,
                    .alpha_min = props.alpha_min,
                    .visc_length = props.visc_length,
                    .diff_beta = props.diff_beta,
                    .diff_alpha_max = props.diff_alpha_max,
                    .diff_alpha_min = props.diff_alpha_min,
                    .time_base = props.time_base,
                    .max_active_bin = props.max_active_bin, .nthreads = 0};
  orc_parts op = {.n = n, .x = x, .v = v, .m = m, .u = u, .time_bin = tb,
                  .h = hh, .rho = rho0, .pressure = zero,
                  .soundspeed = zero, .f = zero, .balsara = zero,
                  .alpha_visc = av0, .alpha_diff = zero, .div_v_prev = zero};
  orc_out64 o;
  double *buf[32];
  int nb = 0;
#define ALLOC(field, w)                                        \
  do {                                                         \
    buf[nb] = calloc((size_t)(w) * n, sizeof(double));         \
    o.field = buf[nb++];                                       \
  } while (0)
  ALLOC(rho, 1); ALLOC(rho_dh, 1); ALLOC(wcount, 1); ALLOC(wcount_dh, 1);
  ALLOC(div_v, 1); ALLOC(rot_v, 3); ALLOC(pressure, 1); ALLOC(soundspeed, 1);
  ALLOC(f, 1); ALLOC(balsara, 1); ALLOC(v_sig, 1); ALLOC(laplace_u, 1);
  ALLOC(alpha_visc_max_ngb, 1); ALLOC(alpha_visc, 1); ALLOC(alpha_diff, 1);
  ALLOC(div_v_dt, 1); ALLOC(a_hydro, 3); ALLOC(u_dt, 1); ALLOC(h_dt, 1);
  ALLOC(s_rho_dh, 1); ALLOC(s_wcount_dh, 1); ALLOC(s_div_v, 1);
  ALLOC(s_rot_v, 3); ALLOC(s_laplace_u, 1); ALLOC(s_a_hydro, 3);
  ALLOC(s_u_dt, 1); ALLOC(s_h_dt, 1);
  o.n_density = malloc(sizeof(int32_t) * n);
  o.n_gradient = malloc(sizeof(int32_t) * n);
  o.n_force = malloc(sizeof(int32_t) * n);
  o.min_ngb_time_bin = malloc((size_t)n);
  if (orc_pass_fp64(&cfg, &op, &o) != 0) {
    fprintf(stderr, "oracle failed\n");
    return 2;
  }

  double w_rho = 0, w_wc = 0, w_dv = 0, w_a = 0, w_udt = 0, w_hdt = 0,
         w_vs = 0, w_p = 0;
  long nd_sum = 0, nf_sum = 0;
  for (int i = 0; i < n; i++) {
    const struct part *p = &parts[i];
    if (p->n_density != o.n_density[i] || p->n_force != o.n_force[i] ||
        p->min_ngb_time_bin != o.min_ngb_time_bin[i]) {
      if (failures < 20)
        fprintf(stderr, "counts[%d]: density %d/%d force %d/%d bin %d/%d\n",
                i, p->n_density, o.n_density[i], p->n_force, o.n_force[i],
                p->min_ngb_time_bin, o.min_ngb_time_bin[i]);
      failures++;
    }
    nd_sum += p->n_density;
    nf_sum += p->n_force;
    check_rel("rho", i, p->rho, o.rho[i], 0.0, 1e-5, &w_rho);
    check_rel("wcount", i, p->wcount, o.wcount[i], 0.0, 1e-5, &w_wc);
    check_rel("div_v", i, p->div_v, o.div_v[i], o.s_div_v[i], 1e-5, &w_dv);
    check_rel("pressure", i, p->pressure, o.pressure[i], 0.0, 1e-4, &w_p);
    check_rel("v_sig", i, p->v_sig, o.v_sig[i], 0.0, 1e-4, &w_vs);
    for (int d = 0; d < 3; d++)
      check_rel("a_hydro", i, p->a_hydro[d], o.a_hydro[3 * i + d],
                o.s_a_hydro[3 * i + d], 1e-4, &w_a);
    check_rel("u_dt", i, p->u_dt, o.u_dt[i], o.s_u_dt[i], 1e-4, &w_udt);
    check_rel("h_dt", i, p->h_dt, o.h_dt[i], o.s_h_dt[i], 1e-4, &w_hdt);
  }
  printf("abi_27cells: %d particles, N_density %ld, N_force %ld; worst rel: "
         "rho %.2e wcount %.2e div_v %.2e P %.2e v_sig %.2e a_hydro %.2e "
         "u_dt %.2e h_dt %.2e; %d failures\n",
         n, nd_sum, nf_sum, w_rho, w_wc, w_dv, w_p, w_vs, w_a, w_udt, w_hdt,
         failures);
  return failures ? 1 : 0;
}
