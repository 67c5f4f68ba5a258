/*
 * meridional.c -- CPU restatement of the meridional advection-diffusion
 * sediment solve (the producer of the per-bin r-z grid of sediment model 3).
 *
 * TEST INFRASTRUCTURE ONLY, under the same rules as oracle.c: it may be
 * built, linked or executed only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product calls it.
 *
 * Every function follows the reference's CUDA kernels and driver loop in
 * CLARITAS/CLARITAS.py (line numbers per function) in plain sequential
 * double precision; in particular each bin's mass is summed cell by cell in
 * (i, k) order as meridional_normalize_kernel does.  The bins are
 * independent, so they are relaxed in parallel between residual checks.
 *
 * Pinned by: the reference's golden velocities and closed-form velocity
 * identities (validate_claritas.py:294-356), its mask and mass-closure
 * checks (validate_claritas.py:1265-1289, 1405-1458), the closed-form
 * discrete equilibrium of a flat cell with no circulation (product of
 * upwind face ratios), the flux antisymmetry of the assembled coefficients,
 * and a direct null-space solve of the assembled system (tests/).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Vortex-surface diagnostics, CLARITAS.py:708-792 (host). out: wall_z,
 * center_z, h_inf, h_bar, vtheta_peak.  fsm 1 parabolic, 2 Scully. */
int orc_vortex_geometry(int fsm, double zmean, double R, double D, double a, double g, double out[5]) {
  if (!(R > 0.0) || !(D >= 0.0) || !(g > 0.0)) return -1;
  double R2 = R * R;
  if (fsm == 0 || D == 0.0) {
    out[0] = zmean; out[1] = zmean; out[2] = 0.0; out[3] = 0.0; out[4] = 0.0;
    return 0;
  }
  if (fsm == 1) {
    double omega = sqrt(2.0 * g * D / R2);
    out[0] = zmean + 0.5 * D; out[1] = zmean - 0.5 * D; out[2] = D; out[3] = 0.5 * D; out[4] = omega * R;
    return 0;
  }
  if (fsm != 2 || !(a > 0.0)) return -1;
  double a2 = a * a, h_inf = D * (a2 + R2) / R2;
  double h_bar = h_inf * (1.0 - (a2 / R2) * log1p(R2 / a2));
  double omega_c = sqrt(2.0 * g * h_inf / a2);
  out[0] = zmean - h_bar + D; out[1] = zmean - h_bar; out[2] = h_inf; out[3] = h_bar; out[4] = 0.5 * omega_c * a;
  return 0;
}

/* meridional_surface_geometry_cuda, CLARITAS.py:4046-4057 */
static void surface_geometry(int fsm, double r, double zmean, double R, double D, double a, double* zs, double* hp) {
  double rr = fmin(fmax(r, 0.0), R), rr2 = rr * rr, R2 = R * R;
  *zs = zmean; *hp = 0.0;
  if (!(D > 0.0) || fsm == 0) return;
  if (fsm == 1) { *zs = zmean + D * (rr2 / R2 - 0.5); *hp = 2.0 * D * rr / R2; return; }
  double a2 = a * a, Hinf = D * (a2 + R2) / R2, hbar = Hinf * (1.0 - (a2 / R2) * log1p(R2 / a2));
  *zs = zmean - hbar + Hinf * rr2 / (a2 + rr2);
  *hp = 2.0 * Hinf * a2 * rr / ((a2 + rr2) * (a2 + rr2));
}

/* meridional_velocity_cuda, CLARITAS.py:4059-4073 */
static void velocity(double r, double z, double zmin, double zmean, double R, int fsm, double D, double a,
                     double upeak, double* ur, double* uz) {
  *ur = 0.0; *uz = 0.0;
  if (!(upeak > 0.0) || !(R > 0.0) || !(D > 0.0) || (fsm == 2 && !(a > 0.0))) return;
  double rr = fmin(fmax(r, 0.0), R), R2 = R * R, rr2 = rr * rr, zs, hp;
  surface_geometry(fsm, rr, zmean, R, D, a, &zs, &hp);
  double H = zs - zmin;
  if (!(H > 0.0)) return;
  double eta = (z - zmin) / H;
  if (eta < 0.0) eta = 0.0;
  if (eta > 1.0) eta = 1.0;
  double s = rr2 / R2, f = eta * (1.0 - eta), fp = 1.0 - 2.0 * eta, A = 2.0 * upeak, om = 1.0 - s;
  *ur = A * rr * om * om * fp / H;
  *uz = -A * (2.0 * om * (1.0 - 3.0 * s) * f - rr * om * om * fp * eta * hp / H);
}

void orc_meridional_velocity(int n, const double* r, const double* z, double zmin, double zmean, double R, int fsm,
                             double D, double a, double upeak, double* ur, double* uz) {
  for (int i = 0; i < n; ++i) velocity(r[i], z[i], zmin, zmean, R, fsm, D, a, upeak, ur + i, uz + i);
}

/* radial_slip_lookup_cuda, CLARITAS.py:4075-4081 */
static double radial_slip(const double* WR, int nrad, double r, double R) {
  if (nrad < 2 || !(R > 0.0)) return 0.0;
  double q = fmin(fmax(r / R, 0.0), 1.0) * (double)(nrad - 1);
  int j = (int)floor(q);
  double t = q - (double)j;
  if (j >= nrad - 1) { j = nrad - 2; t = 1.0; }
  return (1.0 - t) * WR[j] + t * WR[j + 1];
}

/* meridional_physical_fluid, CLARITAS.py:4083-4088 */
static int physical_fluid(double r, double z, double zmin, double zmean, double R, int fsm, double D, double a) {
  if (z < zmin) return 0;
  double zs, hp;
  surface_geometry(fsm, r, zmean, R, D, a, &zs, &hp);
  return z < zs;
}

/* model-2 analytic scale at a float point (sediment_local_scale with
 * TRANSPORT_MODEL 2, CLARITAS.py:4228-4262): the solver's initial guess */
static double analytic_scale(const double* radlog, int nrad, double B, double norm, float x, float z, float R,
                             float zmin) {
  double rr2 = (double)x * (double)x, R2 = (double)R * (double)R;
  if (rr2 > R2) rr2 = R2;
  double rr = sqrt(fmax(rr2, 0.0));
  double u = (R > 0.0f) ? rr / (double)R * (double)(nrad - 1) : 0.0;
  int j = (int)floor(u);
  if (j < 0) j = 0;
  if (j > nrad - 2) j = nrad - 2;
  double t = u - (double)j, rl = (1.0 - t) * radlog[j] + t * radlog[j + 1];
  if (rl > 0.0) rl = 0.0;
  double h = fmax((double)z - (double)zmin, 0.0);
  if (!(norm > 0.0)) return 0.0;
  return exp(rl - B * h) / norm;
}

/* sediment_grid_sample on a double grid, CLARITAS.py:4205-4226 */
static double grid_sample(const double* G, int NR, int NZ, float x, float z, float R, float zmin, float zmax) {
  if (NR < 2 || NZ < 2 || !(zmax > zmin)) return 1.0;
  double rr = sqrt(fmax((double)x * (double)x, 0.0));
  if (rr > (double)R) rr = (double)R;
  double dr = (double)R / (double)NR, dz = ((double)zmax - (double)zmin) / (double)NZ;
  double ur = rr / dr - 0.5, uz = ((double)z - (double)zmin) / dz - 0.5;
  if (ur < 0.0) ur = 0.0;
  if (uz < 0.0) uz = 0.0;
  if (ur > (double)(NR - 1)) ur = (double)(NR - 1);
  if (uz > (double)(NZ - 1)) uz = (double)(NZ - 1);
  int i = (int)floor(ur), k = (int)floor(uz);
  double tr = ur - (double)i, tz = uz - (double)k;
  if (i >= NR - 1) { i = NR - 2; tr = 1.0; }
  if (k >= NZ - 1) { k = NZ - 2; tz = 1.0; }
  double q00 = G[(size_t)i * NZ + k], q10 = G[(size_t)(i + 1) * NZ + k];
  double q01 = G[(size_t)i * NZ + k + 1], q11 = G[(size_t)(i + 1) * NZ + k + 1];
  double q0 = (1.0 - tr) * q00 + tr * q10, q1 = (1.0 - tr) * q01 + tr * q11;
  return fmax((1.0 - tz) * q0 + tz * q1, 0.0);
}

typedef struct {
  int NR, NZ;
  double R, zmin, zgrid, zmean, target;
  const unsigned char* active;
} geo_t;

/* meridional_normalize_kernel, CLARITAS.py:4166-4177 (one bin) */
static int normalize_bin(const geo_t* g, double* G) {
  double dr = g->R / (double)g->NR, dz = (g->zgrid - g->zmin) / (double)g->NZ, mass = 0.0;
  for (int i = 0; i < g->NR; ++i) {
    double rw = (double)i * dr, re = (double)(i + 1) * dr, vol = 0.5 * (re * re - rw * rw) * dz;
    for (int k = 0; k < g->NZ; ++k) {
      int c = i * g->NZ + k;
      if (g->active[c]) mass += G[c] * vol;
    }
  }
  if (!(mass > 0.0) || !(g->target > 0.0) || !isfinite(mass)) return 41;
  double scale = g->target / mass;
  for (int c = 0; c < g->NR * g->NZ; ++c)
    if (g->active[c]) G[c] *= scale;
  return 0;
}

/* neighbour sum of the five-point stencil, CLARITAS.py:4156-4162 */
static double stencil(const geo_t* g, const double* const A[5], size_t t, const double* G, int i, int k) {
  int NR = g->NR, NZ = g->NZ;
  double num = 0.0;
  if (i + 1 < NR) num += A[1][t] * G[(size_t)(i + 1) * NZ + k];
  if (i > 0) num += A[2][t] * G[(size_t)(i - 1) * NZ + k];
  if (k + 1 < NZ) num += A[3][t] * G[(size_t)i * NZ + k + 1];
  if (k > 0) num += A[4][t] * G[(size_t)i * NZ + k - 1];
  return num;
}

/* meridional_relax_kernel, CLARITAS.py:4150-4164 (one bin, one colour) */
static void relax_bin(const geo_t* g, const double* const A[5], double* G, int parity, double omega) {
  for (int i = 0; i < g->NR; ++i)
    for (int k = 0; k < g->NZ; ++k) {
      int c = i * g->NZ + k;
      if (!g->active[c] || ((i + k) & 1) != parity || !(A[0][c] > 0.0)) continue;
      double num = stencil(g, A, (size_t)c, G, i, k);
      double target = num / A[0][c], nv = (1.0 - omega) * G[c] + omega * target;
      if (!isfinite(nv) || nv < 1e-14) nv = 1e-14;
      G[c] = nv;
    }
}

/* meridional_residual_kernel, CLARITAS.py:4180-4188 (one bin) */
static double residual_bin(const geo_t* g, const double* const A[5], const double* G) {
  double mr = 0.0;
  for (int i = 0; i < g->NR; ++i)
    for (int k = 0; k < g->NZ; ++k) {
      int c = i * g->NZ + k;
      if (!g->active[c] || !(A[0][c] > 0.0)) continue;
      double num = stencil(g, A, (size_t)c, G, i, k);
      double den = fmax(fabs(A[0][c] * G[c]), 1e-30), q = fabs(A[0][c] * G[c] - num) / den;
      if (q > mr) mr = q;
    }
  return mr;
}

int orc_meridional_solve(const clr_config* cfg, const clr_particles* p, const clr_sediment* sed,
                         const clr_meridional* mp, clr_meridional_out* out) {
  if (!cfg || !p || !sed || !mp || !out || p->n <= 0 || !p->diameter) return CLR_E_ARG;
  const int fsm = cfg->free_surface_model, n = p->n, NR = mp->nr, NZ = mp->nz;
  const double R = cfg->rin, zmin = cfg->zmin, zmean = cfg->zmean, D = cfg->vortex_delta_h,
               a = cfg->vortex_core_radius;
  /* CLARITAS.py:5873-5887, 5911-5921 */
  if (!(mp->circulation_ratio >= 0.0 && mp->circulation_ratio <= 0.5)) return CLR_E_ARG;
  if (NR < 16 || NZ < 32 || mp->max_iterations < 250 || mp->check_interval < 10) return CLR_E_ARG;
  if (!(mp->tolerance > 0.0 && mp->tolerance < 1e-2) || !(mp->sor >= 1.0 && mp->sor < 1.95)) return CLR_E_ARG;
  if ((fsm != 1 && fsm != 2) || !(D > 0.0)) return CLR_E_ARG;
  double vg[5];
  if (orc_vortex_geometry(fsm, zmean, R, D, a, sed->gravity, vg)) return CLR_E_ARG;
  if (!(vg[1] > zmin)) return CLR_E_ARG;
  const double zgrid = vg[0], upeak = mp->circulation_ratio * vg[4];
  out->grid_zmax = zgrid;
  out->vtheta_peak = vg[4];
  out->peak_speed = upeak;
  out->iterations = 0;

  /* per-bin tables with the finite-Re (Schiller-Naumann) slips */
  const int nrad = sed->n_radial;
  clr_config c3 = *cfg;
  c3.sediment_model = CLR_SED_GRID_PER_BIN;
  double *sA = calloc((size_t)n, 8), *sB = calloc((size_t)n, 8), *sN = calloc((size_t)n, 8), *sW = calloc((size_t)n, 8);
  double *rl = calloc((size_t)n * nrad, 8), *wr = calloc((size_t)n * nrad, 8);
  const size_t plane = (size_t)NR * NZ, total = plane * (size_t)n;
  unsigned char* act = calloc(plane, 1);
  double* G = calloc(total, 8);
  double* A = calloc(5 * total, 8);
  double* res = calloc((size_t)n, 8);
  int err = orc_sediment_bins(&c3, sed, n, p->diameter, sA, sB, sN, sW, rl, wr);
  if (err) goto done;

  const double dr = R / (double)NR, dz = (zgrid - zmin) / (double)NZ;
  /* meridional_active_mask_kernel, CLARITAS.py:4091-4098 */
  for (int i = 0; i < NR; ++i)
    for (int k = 0; k < NZ; ++k) {
      double r = ((double)i + 0.5) * dr, z = zmin + ((double)k + 0.5) * dz;
      act[i * NZ + k] = (unsigned char)physical_fluid(r, z, zmin, zmean, R, fsm, D, a);
    }
  geo_t g = {NR, NZ, R, zmin, zgrid, zmean, 0.5 * R * R * (zmean - zmin), act};
  const double eddy = sed->eddy_diffusivity;
  for (int pb = 0; pb < n; ++pb) {
    double* Gb = G + (size_t)pb * plane;
    const double* WR = wr + (size_t)pb * nrad;
    for (int i = 0; i < NR; ++i)
      for (int k = 0; k < NZ; ++k) {
        const size_t c = (size_t)i * NZ + k, t = (size_t)pb * plane + c;
        double *AP = A + t, *AE = A + total + t, *AW = A + 2 * total + t, *AN = A + 3 * total + t,
               *AS = A + 4 * total + t;
        *AP = *AE = *AW = *AN = *AS = 0.0;
        if (!act[c]) { Gb[c] = 0.0; continue; }
        /* meridional_grid_initialize_kernel, CLARITAS.py:4100-4110 */
        float xf = (float)(((double)i + 0.5) * dr), zf32 = (float)(zmin + ((double)k + 0.5) * dz);
        Gb[c] = analytic_scale(rl + (size_t)pb * nrad, nrad, sB[pb], sN[pb], xf, zf32, (float)R, (float)zmin);
        /* meridional_build_coefficients_kernel, CLARITAS.py:4112-4148 */
        double rc = ((double)i + 0.5) * dr, zc = zmin + ((double)k + 0.5) * dz;
        double rw = (double)i * dr, re = (double)(i + 1) * dr;
        double Az = 0.5 * (re * re - rw * rw), ap = 0.0, ur = 0.0, uz = 0.0, V, F, diff;
        if (i + 1 < NR && act[(size_t)(i + 1) * NZ + k] && physical_fluid(re, zc, zmin, zmean, R, fsm, D, a)) {
          velocity(re, zc, zmin, zmean, R, fsm, D, a, upeak, &ur, &uz);
          V = ur + radial_slip(WR, nrad, re, R); F = V * re * dz; diff = eddy * re * dz / dr;
          *AE = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
        }
        if (i > 0 && act[(size_t)(i - 1) * NZ + k] && physical_fluid(rw, zc, zmin, zmean, R, fsm, D, a)) {
          velocity(rw, zc, zmin, zmean, R, fsm, D, a, upeak, &ur, &uz);
          V = ur + radial_slip(WR, nrad, rw, R); F = -V * rw * dz; diff = eddy * rw * dz / dr;
          *AW = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
        }
        double zfa = zc + 0.5 * dz;
        if (k + 1 < NZ && act[(size_t)i * NZ + k + 1] && physical_fluid(rc, zfa, zmin, zmean, R, fsm, D, a)) {
          velocity(rc, zfa, zmin, zmean, R, fsm, D, a, upeak, &ur, &uz);
          V = uz - sW[pb]; F = V * Az; diff = eddy * Az / dz;
          *AN = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
        }
        zfa = zc - 0.5 * dz;
        if (k > 0 && act[(size_t)i * NZ + k - 1] && physical_fluid(rc, zfa, zmin, zmean, R, fsm, D, a)) {
          velocity(rc, zfa, zmin, zmean, R, fsm, D, a, upeak, &ur, &uz);
          V = uz - sW[pb]; F = -V * Az; diff = eddy * Az / dz;
          *AS = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
        }
        *AP = ap;
      }
  }
  for (int pb = 0; pb < n && !err; ++pb) err = normalize_bin(&g, G + (size_t)pb * plane);
  if (err) goto done;

  /* driver loop, CLARITAS.py:6568-6700 */
  const int maxit = mp->max_iterations, check = mp->check_interval;
  const int min_it = maxit < 1000 ? maxit : 1000;
  int it = 0, last_checked = 0, iterations = 0;
  double maxres = 0.0;
  while (it < maxit) {
    int next = (it / check + 1) * check;
    if (next > maxit) next = maxit;
    int e41 = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : e41)
    for (int pb = 0; pb < n; ++pb) {
      double* Gb = G + (size_t)pb * plane;
      const double* Ab[5];
      for (int q = 0; q < 5; ++q) Ab[q] = A + (size_t)q * total + (size_t)pb * plane;
      int bad = 0;
      for (int j = it + 1; j <= next && !bad; ++j) {
        relax_bin(&g, Ab, Gb, 0, mp->sor);
        relax_bin(&g, Ab, Gb, 1, mp->sor);
        if (j % 20 == 0) bad = normalize_bin(&g, Gb);
      }
      if (!bad) bad = normalize_bin(&g, Gb);
      res[pb] = bad ? INFINITY : residual_bin(&g, Ab, Gb);
      e41 |= bad ? 1 : 0;
    }
    it = next;
    last_checked = it;
    if (e41) { err = 41; goto done; }
    maxres = 0.0;
    for (int pb = 0; pb < n; ++pb) if (res[pb] > maxres) maxres = res[pb];
    if (it >= min_it && maxres <= mp->tolerance) { iterations = it; break; }
  }
  if (iterations == 0) iterations = last_checked ? last_checked : maxit;
  out->iterations = iterations;

  for (int pb = 0; pb < n; ++pb) {
    double* Gb = G + (size_t)pb * plane;
    /* meridional_extend_grid_kernel, CLARITAS.py:4190-4196 */
    for (int i = 0; i < NR; ++i) {
      int top = -1;
      for (int k = NZ - 1; k >= 0; --k) if (act[i * NZ + k]) { top = k; break; }
      if (top < 0) { for (int k = 0; k < NZ; ++k) Gb[(size_t)i * NZ + k] = 1.0; continue; }
      double q = Gb[(size_t)i * NZ + top];
      for (int k = top + 1; k < NZ; ++k) Gb[(size_t)i * NZ + k] = q;
    }
    /* meridional_finalize_kernel, CLARITAS.py:4198-4209 */
    double mx = 0.0, mn = 1e300;
    for (size_t c = 0; c < plane; ++c)
      if (act[c]) { double q = Gb[c]; if (q > mx) mx = q; if (q < mn) mn = q; }
    if (!(mx > 0.0)) { mx = 1.0; mn = 1.0; }
    if (out->scale_max) out->scale_max[pb] = mx;
    if (out->scale_min) out->scale_min[pb] = mn;
    if (out->point_scales) {
      const float Rf = (float)R, zf = (float)zmin, zt = (float)zgrid;
      double* ps = out->point_scales + 4 * (size_t)pb;
      ps[0] = grid_sample(Gb, NR, NZ, 0.0f, 0.0f, Rf, zf, zt);
      ps[1] = grid_sample(Gb, NR, NZ, (float)(0.95 * R), 0.0f, Rf, zf, zt);
      ps[2] = grid_sample(Gb, NR, NZ, 0.0f, (float)zmin, Rf, zf, zt);
      ps[3] = grid_sample(Gb, NR, NZ, (float)(0.95 * R), (float)zmin, Rf, zf, zt);
    }
  }
  if (maxres > mp->tolerance) err = CLR_E_NOT_CONVERGED;

done:
  if (out->grid) memcpy(out->grid, G, total * 8);
  if (out->residual) memcpy(out->residual, res, (size_t)n * 8);
  if (out->active) memcpy(out->active, act, plane);
  if (out->coeff) memcpy(out->coeff, A, 5 * total * 8);
  free(sA); free(sB); free(sN); free(sW); free(rl); free(wr); free(act); free(G); free(A); free(res);
  return err;
}
