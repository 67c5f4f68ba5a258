/*
 * oracle.h -- CPU restatement of the CLARITAS photon-transport path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: it may be built, linked,
 * imported or executed only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product
 * (tardiis_and_claritas_amd/, include/) includes or calls it.
 *
 * The structs of the public ABI (include/claritas.h) are reused so that one
 * configuration can be handed to both the HIP library and this oracle.
 */
#ifndef CLARITAS_ORACLE_H
#define CLARITAS_ORACLE_H
#include "../include/claritas.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

/* Philox4x32-10 block function (Salmon et al., SC'11), exposed for KATs. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* First n draws of the per-history uniform stream (float, (0,1)). */
void orc_uniform_stream(uint64_t seed, uint64_t history, int n, float* out);

/* Mie tables for a PSD (FP64). Same outputs as clr_mie_tables. */
int orc_mie(const clr_particles* p, clr_mie_out* out);

/* Build the full model (Mie + transport + sediment tables). */
int orc_prepare(const clr_config* cfg, const clr_particles* p, const clr_sediment* sed,
                const float* field, size_t field_len, orc_model** out);
void orc_free(orc_model* m);
int orc_get_tables(const orc_model* m, clr_tables_out* out);
/* Replace the Mie phase table the trace samples from (used to run the
 * oracle on exactly the device's float tables in history-by-history tests). */
int orc_set_phase_table(orc_model* m, const double* theta, const double* cdf, int n_theta);

/* Track a shared FP32 field (sediment models 4, 5) with the device's
 * piecewise-constant majorant over z slabs instead of the single majorant
 * (on = 0 restores it). */
int orc_set_slab_majorant(orc_model* m, int on);

/* Trace n histories; tallies zeroed then filled.  nthreads<=0 uses all. */
int orc_run(const orc_model* m, uint64_t n, uint64_t seed, uint64_t history_offset,
            int nthreads, clr_tallies* out, clr_record* records, uint64_t n_records);

/* orc_run plus the light-transport voxel diagnostics (clr_run_light
 * contract; the oracle initialises the caller's output arrays). */
/* orc_run plus the causal provenance record of every history (crec[n],
 * detected[n]: 1 when the history reached a hardware detector) */
int orc_run_causal(const orc_model* m, uint64_t n, uint64_t seed, uint64_t offset, int nthreads,
                   clr_tallies* out, clr_causal_record* crec, uint8_t* detected);
int orc_light_heat_probe(const clr_light_config* lc, float rin, float zmin, int n, const float* segments7,
                         clr_light_out* lo);
int orc_run_light(const orc_model* m, uint64_t n, uint64_t seed, uint64_t offset, int nthreads,
                  const clr_light_config* lc, clr_tallies* out, clr_light_out* lo);
/* The reference's diagnostic test kernels (CLARITAS.py:3470-3485): segment i
 * of segments7 [n][7] (x0, y0, z0, vx, vy, vz, dist; particle order i & 1,
 * history id i) and event sites xyz3 [n][3], deposited into lo (which the
 * caller zeroes; max_rays, representative_rays and max_segments from lc). */
int orc_light_segment_probe(const clr_light_config* lc, float rin, float zmin, int n, const float* segments7,
                            clr_light_out* lo);
int orc_light_event_probe(const clr_light_config* lc, float rin, float zmin, int n, const float* xyz3,
                          clr_light_out* lo);

/* One reflecting bore section (CLARITAS.py:1296-1318) on explicit rays
 * rays6[n] = (s, q, w, vs, vq, vw) in the bore frame up to s_end, walls of
 * refractive index cfg->bore_n, at most cfg->max_bore_bounces + 1 wall hits;
 * ray i draws from stream i of `seed`.  And the SFH213 relative response. */
int orc_bore_probe(const clr_config* cfg, int n, uint64_t seed, const float* rays6, float s_end, float radius,
                   int* pass, int* inter, int* refl);
float orc_sfh213(float phi_deg);

/* Geometry probes on the reference's test-kernel semantics
 * (free_surface_geometry_test_kernel, CLARITAS.py:4438-4455). */
int orc_free_surface_probe(const clr_config* cfg, const float* rays6, int n,
                           float* zs_start, float* t_hit, float* zs_hit,
                           float* nx, float* ny, float* nz);
/* Local sediment density scale of bin `bin` at n points (CLARITAS.py:4228-4235). */
int orc_sediment_probe(const orc_model* m, int bin, const float* xyz, int n, double* out);

/* The probe modes of clr_probe (include/claritas.h: CLR_PROBE_*, item
 * layouts from clr_probe_shape's table, restated here) on a prepared model;
 * the detector acceptance is the reference's full 18-bore scan. */
int orc_probe_shape(int mode, int* in_stride, int* out_stride);
int orc_probe(const orc_model* m, int mode, int n, uint64_t seed, const float* in, double* out);

/* closed-form probes: unpolarised Fresnel R, complex-index interface,
 * Snell refraction, water-side wall exit distance */
int orc_fresnel_probe(int n, const float* ci, float n1, float n2, float* R, int* tir);
int orc_complex_probe(int n, const float* ci, float n1r, float n1k, float n2r, float n2k, float* R, float* tt,
                      float* tn, int* ok);
int orc_refract_probe(int n, const float* v, const float* nrm, float n1, float n2, float* out, int* ok);
int orc_wall_probe(int n, const float* rays6, float R, float zlo, float zhi, double* t);

/* rough-surface probes: inverse error function, Beckmann tail function
 * B(x), signed Smith Lambda, one Smith height step, alpha = 1 visible
 * slopes from two uniforms, and whole microsurface walks (history i of
 * `seed`; d is the incident direction in, the outgoing one out; status
 * 2 escape, 3 numerical failure, 4 watchdog; top = 1 when it left on the
 * top side; fc the failure subcode) */
double orc_erfinv(double y);
int orc_beckmann_B_probe(int n, const double* x, double* out);
double orc_smith_lambda(float wx, float wy, float wz, float bx, float by, float bz, float alpha);
int orc_smith_height(double U, float wx, float wy, float wz, float bx, float by, float bz, float alpha, float h,
                     float* out_h, int* fc);
int orc_beckmann_slopes_probe(int n, const double* si, const double* u1, const double* u2, double* sx, double* sy,
                              int* ok);
int orc_walk_probe(int n, uint64_t seed, float alpha, const float* b, float n_top, float k_top, float n_bot,
                   float k_bot, float* d, int* status, int* top, int* fc);

/* Per-bin sediment coefficients of cfg->sediment_model (1, 2 or 3):
 * A, B, normalisation, settling speed [n], radial log potential and radial
 * slip w_r [n][sed->n_radial] (radwr may be NULL); CLARITAS.py:3940-4038. */
int orc_sediment_bins(const clr_config* cfg, const clr_sediment* s, int n, const double* diameter, double* A,
                      double* B, double* norm, double* ws, double* radlog, double* radwr);

/* meridional.c: vortex-surface diagnostics (wall_z, center_z, h_inf, h_bar,
 * vtheta_peak; CLARITAS.py:708-792), the secondary-circulation velocity
 * (CLARITAS.py:4059-4073) and the full meridional solve
 * (CLARITAS.py:4091-4215, 6451-6725) with the clr_meridional_solve
 * contract. */
int orc_vortex_geometry(int fsm, double zmean, double R, double D, double a, double g, double out[5]);
void orc_meridional_velocity(int n, const double* r, const double* z, double zmin, double zmean, double R, int fsm,
                             double D, double a, double upeak, double* ur, double* uz);
int orc_meridional_solve(const clr_config* cfg, const clr_particles* p, const clr_sediment* sed,
                         const clr_meridional* mp, clr_meridional_out* out);

#ifdef __cplusplus
}
#endif
#endif
