/*
 * oracle.c -- CPU restatement of the CLARITAS photon-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may build, load or
 * run it, and never as the thing measured or shipped.  The product library
 * (tardiis_and_claritas_amd/csrc) does not include or link any of this.
 *
 * Every function restates the algorithm of the CLARITAS/CLARITAS.py lines it
 * cites (the reference's CUDA_SRC device code and its host driver).  Two
 * deliberate differences, shared with the HIP library:
 *   - random numbers: one counter-based Philox4x32-10 stream per history,
 *     key = 64-bit seed, counter = (block, 0, history lo, history hi),
 *     instead of the reference's three xorshift32 words per ray
 *     (CLARITAS.py:1125-1133, 4612);
 *   - 64-bit history numbers (the reference stops at 2^32, CLARITAS.py:7191).
 * Supported physics: smooth and rough (Smith microsurface walk) particle
 * surfaces, the coherent Mie and the incoherent spheroid-morphology wave
 * channels, absorbing bores, ideal photodiodes, plus the extensions named in
 * claritas.h (plain-Mie event law, collimated disk source, Rankine surface,
 * shared FP32 sediment fields, auxiliary bores).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INF_F 3.402823466e+38F
#define PI_F 3.14159265358979323846f
#define PI_D 3.141592653589793238462643383279502884
#define EPS_F 2.0e-7f
#define DEPS 2.220446049250313080847263336181640625e-16
#define FEPS 1.1920928955078125e-7
#define MIE_ORDER_CAP (1 << 22)

/* production failure codes, CLARITAS.py:1846-1866 */
enum {
  PF_OK = 0, PF_MICRO_NUMERICAL = 1, PF_MICRO_WATCHDOG = 2, PF_MACRO_GEOMETRY = 3,
  PF_MACRO_BOUNCE = 4, PF_ACRYLIC_NO_HIT = 5, PF_ACRYLIC_BOUNCE = 6, PF_BOTTOM_GEOMETRY = 7,
  PF_BOTTOM_BOUNCE = 8, PF_WATER_NO_BOUNDARY = 9, PF_ITERATION_LIMIT = 10,
  PF_NONFINITE_POSITION = 11, PF_NONFINITE_DIRECTION = 12, PF_INVALID_MEDIUM = 13,
  PF_INVALID_EVENT = 14, PF_OTHER = 15, PF_SPHEROID = 16, PF_HEADSPACE_NO_HIT = 17,
  PF_MAJORANT = 18
};
/* subcodes, CLARITAS.py:1868-1895 and micro codes CLARITAS.py:1810-1839 */
enum {
  SUB_NONE = 0, SUB_SPHERE_ENTRY_SIDE = 1001, SUB_SPHERE_CHORD = 1002, SUB_SPHERE_EXIT_SIDE = 1003,
  SUB_SPHERE_INTERNAL_SIDE = 1004, SUB_SPHERE_BOUNCE_LIMIT = 1005,
  SUB_SPHEROID_ASPECT = 1101, SUB_SPHEROID_AXES = 1102, SUB_SPHEROID_ENTRY = 1103, SUB_SPHEROID_CHORD = 1104,
  SUB_SPHEROID_EXIT_SIDE = 1105, SUB_SPHEROID_INTERNAL_SIDE = 1106, SUB_SPHEROID_BOUNCE_LIMIT = 1107,
  SUB_ACRYLIC_NO_HIT = 2001, SUB_ACRYLIC_BOUNCE = 2002, SUB_BOTTOM_TOPOLOGY = 2104,
  SUB_WATER_NO_BOUNDARY = 3001, SUB_ITERATION_LIMIT = 3002, SUB_NONFINITE_POSITION = 3003,
  SUB_NONFINITE_DIRECTION = 3004, SUB_WATER_OUTSIDE = 3005, SUB_HEADSPACE_NO_HIT = 3006,
  SUB_MAJORANT = 4001,
  MF_REFLECTION_INVALID = 31, MF_REFRACTION_INVALID = 32, MF_SIDE_STATE = 33,
  MF_NONFINITE_DIRECTION = 34, MF_FRESNEL_INVALID = 35
};

/* ======================================================================
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random
 * numbers: as easy as 1, 2, 3").  Multipliers and Weyl constants are the
 * published ones.
 * ====================================================================== */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

typedef struct {
  uint32_t key0, key1, h0, h1, blk, idx, draws;
  uint32_t buf[4];
} rng_t;

static void rng_seed(rng_t* r, uint64_t seed, uint64_t hist) {
  r->key0 = (uint32_t)seed; r->key1 = (uint32_t)(seed >> 32);
  r->h0 = (uint32_t)hist; r->h1 = (uint32_t)(hist >> 32);
  r->blk = 0; r->idx = 4; r->draws = 0;
}
static uint32_t rng_next(rng_t* r) {
  if (r->idx >= 4) {
    uint32_t c[4] = {r->blk, 0u, r->h0, r->h1}, k[2] = {r->key0, r->key1};
    orc_philox4x32_10(c, k, r->buf);
    r->blk++;
    r->idx = 0;
  }
  r->draws++;
  return r->buf[r->idx++];
}
/* float uniform on the open interval: (m + 1/2) 2^-23, m = top 23 bits */
static float U(rng_t* r) { return ((float)(rng_next(r) >> 9) + 0.5f) * 1.1920928955078125e-7f; }
/* 53-bit [0,1) roulette from two words, CLARITAS.py:4491-4495 */
static double U53(rng_t* r) {
  uint32_t a = rng_next(r), b = rng_next(r);
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 1.1102230246251565404236316680908203125e-16;
}
/* open (0,1) double from one word, CLARITAS.py:1908-1912 */
static double Uopen(rng_t* r) { return ((double)rng_next(r) + 0.5) * 2.3283064365386962890625e-10; }

void orc_uniform_stream(uint64_t seed, uint64_t history, int n, float* out) {
  rng_t r; rng_seed(&r, seed, history);
  for (int i = 0; i < n; ++i) out[i] = U(&r);
}

/* ======================================================================
 * Source samplers, CLARITAS.py:1149-1212
 * ====================================================================== */
static float normal_bm(rng_t* r) {
  float u1 = fmaxf(U(r), 1e-12f), u2 = U(r);
  return sqrtf(-2.0f * logf(u1)) * cosf(2.0f * PI_F * u2);
}
/* Marsaglia & Tsang (2000) gamma sampler with the a<1 boost, CLARITAS.py:1154-1175 */
static float gamma_draw(rng_t* r, float shape) {
  if (shape <= 0.0f) return 0.0f;
  if (fabsf(shape - 1.0f) < 1e-7f) return -logf(fmaxf(U(r), 1e-12f));
  float boost = 1.0f, a = shape;
  if (a < 1.0f) {
    float u = fmaxf(U(r), 1e-12f);
    boost = powf(u, 1.0f / a);
    a += 1.0f;
  }
  float d = a - 0.3333333333333333f, c = 1.0f / sqrtf(9.0f * d);
  for (int it = 0; it < 128; ++it) {
    float x = normal_bm(r), v = 1.0f + c * x;
    if (v <= 0.0f) continue;
    v = v * v * v;
    float u = U(r);
    if (u < 1.0f - 0.0331f * x * x * x * x) return boost * d * v;
    if (logf(fmaxf(u, 1e-12f)) < 0.5f * x * x + d * (1.0f - v + logf(v))) return boost * d * v;
  }
  return boost * d;
}
/* CLARITAS.py:1177-1191 */
static float beta_draw(rng_t* r, float a, float b) {
  if (a <= 0.0f || b <= 0.0f) return 0.0f;
  if (fabsf(a - 1.0f) < 1e-7f) {
    float u = fminf(fmaxf(U(r), 1e-12f), 1.0f - 1e-7f);
    return 1.0f - powf(1.0f - u, 1.0f / b);
  }
  if (fabsf(b - 1.0f) < 1e-7f) {
    float u = fminf(fmaxf(U(r), 1e-12f), 1.0f - 1e-7f);
    return powf(u, 1.0f / a);
  }
  float ga = gamma_draw(r, a), gb = gamma_draw(r, b), s = ga + gb;
  return s > 0.0f ? ga / s : 0.0f;
}
/* radiance law with the sin(theta) Jacobian by rejection, CLARITAS.py:1193-1212 */
static float beta_radiance_polar(rng_t* r, float a, float b) {
  const float c = 0.5f * PI_F;
  for (int it = 0; it < 128; ++it) {
    float u = fminf(fmaxf(beta_draw(r, a + 1.0f, b), 0.0f), 1.0f);
    float x = c * u, acc = (x < 1.0e-7f) ? 1.0f : sinf(x) / x;
    if (U(r) <= acc) return x;
  }
  float u = beta_draw(r, a + 1.0f, b);
  return c * fminf(fmaxf(u, 0.0f), 1.0f);
}

/* ======================================================================
 * Vector optics, CLARITAS.py:1214-1260
 * ====================================================================== */
static inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return ax * bx + ay * by + az * bz;
}
static void normalize3(float* x, float* y, float* z) {
  float q = (*x) * (*x) + (*y) * (*y) + (*z) * (*z);
  if (q <= 0.0f) { *x = 0.0f; *y = 1.0f; *z = 0.0f; return; }
  float s = 1.0f / sqrtf(q);
  *x *= s; *y *= s; *z *= s;
}
static void reflect3(float ix, float iy, float iz, float nx, float ny, float nz, float* ox,
                     float* oy, float* oz) {
  float d = dot3(ix, iy, iz, nx, ny, nz);
  *ox = ix - 2.0f * d * nx; *oy = iy - 2.0f * d * ny; *oz = iz - 2.0f * d * nz;
  normalize3(ox, oy, oz);
}
/* n points into the incident medium */
static int refract3(float ix, float iy, float iz, float nx, float ny, float nz, float n1, float n2,
                    float* ox, float* oy, float* oz) {
  float ci = fminf(fmaxf(-dot3(ix, iy, iz, nx, ny, nz), 0.0f), 1.0f);
  float eta = n1 / n2, k = 1.0f - eta * eta * (1.0f - ci * ci);
  if (k <= 0.0f) return 0;
  float ct = sqrtf(k), f = eta * ci - ct;
  *ox = eta * ix + f * nx; *oy = eta * iy + f * ny; *oz = eta * iz + f * nz;
  normalize3(ox, oy, oz);
  return 1;
}
/* unpolarised Fresnel reflectance */
static float fresnel_R(float ci, float n1, float n2, int* tir) {
  ci = fminf(fmaxf(ci, 0.0f), 1.0f);
  float eta = n1 / n2, st2 = eta * eta * fmaxf(0.0f, 1.0f - ci * ci);
  if (st2 >= 1.0f) { *tir = 1; return 1.0f; }
  *tir = 0;
  float ct = sqrtf(fmaxf(0.0f, 1.0f - st2));
  float ds = n1 * ci + n2 * ct, dp = n1 * ct + n2 * ci, rs = 1.0f, rp = 1.0f;
  if (fabsf(ds) > 1e-12f) { float a = (n1 * ci - n2 * ct) / ds; rs = a * a; }
  if (fabsf(dp) > 1e-12f) { float a = (n1 * ct - n2 * ci) / dp; rp = a * a; }
  return fminf(fmaxf(0.5f * (rs + rp), 0.0f), 1.0f);
}

/* ---- complex-index interface, CLARITAS.py:1376-1452 ---- */
typedef struct { float re, im; } cf;
static cf cmk(float a, float b) { cf z = {a, b}; return z; }
static cf cfadd(cf a, cf b) { return cmk(a.re + b.re, a.im + b.im); }
static cf cfsub(cf a, cf b) { return cmk(a.re - b.re, a.im - b.im); }
static cf cfmul(cf a, cf b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cf cfdiv(cf a, cf b) {
  float d = b.re * b.re + b.im * b.im;
  if (d < 1e-30f) d = 1e-30f;
  return cmk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
static float cfabs2(cf a) { return a.re * a.re + a.im * a.im; }
/* principal square root with non-negative real part */
static cf cfsqrt(cf z) {
  float r = sqrtf(z.re * z.re + z.im * z.im);
  float u = sqrtf(fmaxf(0.0f, 0.5f * (r + z.re))), v = sqrtf(fmaxf(0.0f, 0.5f * (r - z.re)));
  if (z.im < 0.0f) v = -v;
  if (u < 0.0f || (fabsf(u) < 1e-20f && v < 0.0f)) { u = -u; v = -v; }
  return cmk(u, v);
}
/* planar complex-index Fresnel reflectance and the s/p-averaged energy-flux
 * direction of the transmitted wave (tangential, normal components) */
static int complex_interface(float ci, float n1r, float n1k, float n2r, float n2k, float* R,
                             float* t_tan, float* t_nrm) {
  if (!(n1r > 0.f && n2r > 0.f && n1k >= 0.f && n2k >= 0.f)) return 0;
  ci = fminf(fmaxf(ci, 0.f), 1.f);
  float beta = n1r * sqrtf(fmaxf(0.f, 1.f - ci * ci));
  cf n1 = cmk(n1r, n1k), n2 = cmk(n2r, n2k);
  cf e1 = cfmul(n1, n1), e2 = cfmul(n2, n2), b2 = cmk(beta * beta, 0.f);
  cf q1 = cfsqrt(cfsub(e1, b2)), q2 = cfsqrt(cfsub(e2, b2));
  cf rs = cfdiv(cfsub(q1, q2), cfadd(q1, q2));
  cf a = cfmul(e2, q1), b = cfmul(e1, q2);
  cf rp = cfdiv(cfsub(a, b), cfadd(a, b));
  float Rv = .5f * (cfabs2(rs) + cfabs2(rp));
  if (!isfinite(Rv)) return 0;
  cf qe = cfdiv(q2, e2), be = cfdiv(cmk(beta, 0.f), e2);
  float Sx = beta, Sz = q2.re, Px = be.re, Pz = qe.re;
  float sn = hypotf(Sx, Sz), pn = hypotf(Px, Pz);
  if (!(sn > 0.f && pn > 0.f && Sz >= 0.f && Pz >= 0.f)) return 0;
  float tx = .5f * (Sx / sn + Px / pn), tz = .5f * (Sz / sn + Pz / pn);
  float nm = hypotf(tx, tz);
  if (!(nm > 0.f && isfinite(nm))) return 0;
  *R = fminf(fmaxf(Rv, 0.f), 1.f);
  *t_tan = tx / nm;
  *t_nrm = tz / nm;
  return 1;
}
static int energy_refract3(float ix, float iy, float iz, float nx, float ny, float nz, float n1r,
                           float n1k, float n2r, float n2k, float* ox, float* oy, float* oz) {
  float ci = fminf(fmaxf(-dot3(ix, iy, iz, nx, ny, nz), 0.f), 1.f);
  float R, ts, tn;
  if (!complex_interface(ci, n1r, n1k, n2r, n2k, &R, &ts, &tn)) return 0;
  float px = ix + ci * nx, py = iy + ci * ny, pz = iz + ci * nz;
  float pn = sqrtf(px * px + py * py + pz * pz);
  if (pn > 1.e-7f) { px /= pn; py /= pn; pz /= pn; }
  else { px = py = pz = 0.f; ts = 0.f; tn = 1.f; }
  *ox = ts * px - tn * nx; *oy = ts * py - tn * ny; *oz = ts * pz - tn * nz;
  normalize3(ox, oy, oz);
  return isfinite(*ox) && isfinite(*oy) && isfinite(*oz);
}

static void perpendicular_basis(float vx, float vy, float vz, float* e1x, float* e1y, float* e1z,
                                float* e2x, float* e2y, float* e2z) {
  float rx, ry, rz;
  if (fabsf(vy) < 0.9f) { rx = 0.0f; ry = 1.0f; rz = 0.0f; } else { rx = 1.0f; ry = 0.0f; rz = 0.0f; }
  *e1x = ry * vz - rz * vy; *e1y = rz * vx - rx * vz; *e1z = rx * vy - ry * vx;
  normalize3(e1x, e1y, e1z);
  *e2x = vy * (*e1z) - vz * (*e1y); *e2y = vz * (*e1x) - vx * (*e1z); *e2z = vx * (*e1y) - vy * (*e1x);
  normalize3(e2x, e2y, e2z);
}

/* ======================================================================
 * Model: float parameters as the transport kernel receives them
 * ====================================================================== */
#define CLR_ORC_MAX_SLABS 32
typedef struct {
  float n_water, n_acrylic, n_air, n_bottom_ext;
  float rin, rout, zmin, zmean, zwall_top, bdt, zlow;
  float ring_in, throat_out, ring_out, source_r, throat_r, cb_r;
  float dh, core_r, alpha1, alpha2, sigma, beam_r, legacy_acc, wavelength;
  float mu_geom, mu_wave, maj_geom, maj_wave, coherent;
  float zs_min;                 /* lowest free-surface height */
  double rk_K, rk_C0;           /* Rankine constants */
  float field_zmax, field_half;
  int fs_model, src_model, event_model, sed_model, max_pb, max_iter, cond_trials, n_aux, n_groups;
  int fn0, fn1, fn2, fnbins;
  float rough, aspect, kappa;   /* Beckmann alpha, spheroid a/b, vMF kappa */
  int orientation;              /* spheroid axis law */
  int src_bore, det_bore, face_model, pd_model, max_bb, max_fb;  /* CLARITAS.py:5558-5562 */
  float n_bore, n_face;
  /* piecewise-constant majorant over z slabs (orc_set_slab_majorant; 0 slabs:
   * the single majorant of the reference) */
  int n_slabs;
  float slab_mu[CLR_ORC_MAX_SLABS], slab_z0, slab_w, slab_inv_w, slab_mu_max;
} fparams;

struct orc_model {
  clr_config cfg;
  fparams fp;
  int n, nt, nrad;
  double *d, *w, *nr, *ki;
  double *x, *qext, *qsca, *qabs, *qback, *g, *albedo;
  int* nstop;
  double *theta, *pdf, *cdf;
  float *theta_f, *cdf_f;
  int nt_sample;
  double *radius, *nd, *mug, *mus, *mue, *muw, *muev, *gcdf, *wcdf;
  float *gcdf_f, *wcdf_f;
  double *trace_g, *trace_w;
  double sum_trace_g, sum_trace_w;
  float sum_g_f, sum_w_f, finv[3];   /* shared-field rates and inverse cell sizes */
  double mu_scalars[8];
  double *sedA, *sedB, *sednorm, *sedws, *radlog;
  double maj[4];
  float* field;
  size_t field_len;
  float det_ux[CLR_N_CHANNELS], det_uy[CLR_N_CHANNELS];
  float aux_ux[CLR_MAX_AUX], aux_uy[CLR_MAX_AUX], aux_deg[CLR_MAX_AUX];
  float n_f[512], k_f[512], r_f[512];
};

/* ======================================================================
 * Mie, CLARITAS.py:3488-3592 (FP64)
 * ====================================================================== */
typedef struct { double re, im; } cd;
static cd cdm(double a, double b) { cd z = {a, b}; return z; }
static cd cdadd(cd a, cd b) { return cdm(a.re + b.re, a.im + b.im); }
static cd cdsub(cd a, cd b) { return cdm(a.re - b.re, a.im - b.im); }
static cd cdmul(cd a, cd b) { return cdm(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cd cdscale(cd a, double s) { return cdm(a.re * s, a.im * s); }
static cd cddiv(cd a, cd b) {
  double d = b.re * b.re + b.im * b.im;
  if (d < 1e-300) d = 1e-300;
  return cdm((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

/* Lorenz-Mie coefficients a_n, b_n by downward logarithmic derivative and
 * upward Riccati-Bessel recurrences (Bohren & Huffman 1983, BHMIE);
 * efficiencies and g as CLARITAS.py:3495-3559.  a, b sized [nstop]. */
static int mie_one(double diameter, double nre, double kim, double n_med, double wl, double* xo,
                   int* nstop_o, double* q, cd** a_out, cd** b_out) {
  double x = PI_D * n_med * diameter / wl;
  *xo = x;
  if (!(x > 0.0) || !(nre > 0.0) || !(n_med > 0.0) || !(wl > 0.0)) {
    *nstop_o = 0; q[0] = q[1] = q[2] = q[3] = q[4] = 0.0; q[5] = 1.0;
    return 2;
  }
  double mr = nre / n_med, mi = kim / n_med;
  int nstop = (int)(x + 4.0 * cbrt(x) + 2.0);
  if (nstop < 1) nstop = 1;
  double mz = hypot(mr * x, mi * x);
  int nmx = (int)fmax((double)nstop, mz) + 128;
  if (nmx < nstop + 1) nmx = nstop + 1;
  if (nmx >= MIE_ORDER_CAP) { *nstop_o = 0; return 1; }
  *nstop_o = nstop;
  cd* D = (cd*)malloc(sizeof(cd) * (size_t)(nmx + 1));
  cd* A = (cd*)malloc(sizeof(cd) * (size_t)nstop);
  cd* B = (cd*)malloc(sizeof(cd) * (size_t)nstop);
  D[nmx] = cdm(0.0, 0.0);
  cd z = cdm(mr * x, mi * x), one = cdm(1.0, 0.0);
  for (int n = nmx; n >= 1; --n) {
    cd nz = cddiv(cdm((double)n, 0.0), z);
    D[n - 1] = cdsub(nz, cddiv(one, cdadd(D[n], nz)));
  }
  double psi_m1 = sin(x), psi = sin(x) / x - cos(x);
  double chi_m1 = cos(x), chi = cos(x) / x + sin(x);
  cd m = cdm(mr, mi);
  for (int j = 0; j < nstop; ++j) {
    int n = j + 1;
    cd xi_m1 = cdm(psi_m1, -chi_m1), xi = cdm(psi, -chi);
    cd ca = cdadd(cddiv(D[n], m), cdm((double)n / x, 0.0));
    cd cb = cdadd(cdmul(m, D[n]), cdm((double)n / x, 0.0));
    A[j] = cddiv(cdsub(cdscale(ca, psi), cdm(psi_m1, 0.0)), cdsub(cdmul(ca, xi), xi_m1));
    B[j] = cddiv(cdsub(cdscale(cb, psi), cdm(psi_m1, 0.0)), cdsub(cdmul(cb, xi), xi_m1));
    double psi_p1 = ((2.0 * n + 1.0) / x) * psi - psi_m1;
    double chi_p1 = ((2.0 * n + 1.0) / x) * chi - chi_m1;
    psi_m1 = psi; psi = psi_p1; chi_m1 = chi; chi = chi_p1;
  }
  free(D);
  double se = 0, ss = 0, sbr = 0, sbi = 0, t1 = 0, t2 = 0;
  for (int j = 0; j < nstop; ++j) {
    int n = j + 1;
    double cn = 2.0 * n + 1.0;
    cd aa = A[j], bb = B[j];
    se += cn * (aa.re + bb.re);
    ss += cn * (aa.re * aa.re + aa.im * aa.im + bb.re * bb.re + bb.im * bb.im);
    double sg = (n & 1) ? -1.0 : 1.0;
    sbr += cn * sg * (aa.re - bb.re);
    sbi += cn * sg * (aa.im - bb.im);
    t2 += (cn / ((double)n * (n + 1.0))) * (aa.re * bb.re + aa.im * bb.im);
    if (j + 1 < nstop) {
      cd an = A[j + 1], bn = B[j + 1];
      t1 += ((double)n * (n + 2.0) / (n + 1.0)) *
            ((aa.re * an.re + aa.im * an.im) + (bb.re * bn.re + bb.im * bn.im));
    }
  }
  double x2 = x * x, qe = 2.0 * se / x2, qs = 2.0 * ss / x2, qb = (sbr * sbr + sbi * sbi) / x2;
  if (qe < 0.0 && qe > -1e-12) qe = 0.0;
  if (qs < 0.0 && qs > -1e-12) qs = 0.0;
  double qa = qe - qs;
  if (qa < 0.0 && qa > -1e-10) qa = 0.0;
  double gg = qs > 0.0 ? 4.0 * (t1 + t2) / (x2 * qs) : 0.0;
  if (gg > 1.0) gg = 1.0;
  if (gg < -1.0) gg = -1.0;
  q[0] = qe; q[1] = qs; q[2] = qa; q[3] = qb; q[4] = gg; q[5] = qe > 0.0 ? qs / qe : 1.0;
  if (a_out) *a_out = A; else free(A);
  if (b_out) *b_out = B; else free(B);
  return 0;
}

/* theta_j = (pi/2)(1 - cos(pi j/(N-1))), CLARITAS.py:3488-3493 */
static void theta_grid(double* th, int nt) {
  for (int j = 0; j < nt; ++j) {
    if (j == 0) th[j] = 0.0;
    else if (j == nt - 1) th[j] = PI_D;
    else th[j] = 0.5 * PI_D * (1.0 - cos(PI_D * ((double)j / (double)(nt - 1))));
  }
}
/* phase PDF 0.5(|S1|^2+|S2|^2) sin(theta) by pi_n/tau_n recurrences, CLARITAS.py:3561-3581 */
static void phase_pdf(const cd* A, const cd* B, int nstop, const double* th, int nt, double* pdf) {
  for (int j = 0; j < nt; ++j) {
    double mu = cos(th[j]), s1r = 0, s1i = 0, s2r = 0, s2i = 0, pm1 = 0.0, pn = 1.0;
    for (int k = 0; k < nstop; ++k) {
      double nf = (double)(k + 1);
      double tau = nf * mu * pn - (nf + 1.0) * pm1;
      double c = (2.0 * nf + 1.0) / (nf * (nf + 1.0));
      s1r += c * (A[k].re * pn + B[k].re * tau); s1i += c * (A[k].im * pn + B[k].im * tau);
      s2r += c * (A[k].re * tau + B[k].re * pn); s2i += c * (A[k].im * tau + B[k].im * pn);
      double pp1 = ((2.0 * nf + 1.0) / nf) * mu * pn - ((nf + 1.0) / nf) * pm1;
      pm1 = pn; pn = pp1;
    }
    pdf[j] = 0.5 * (s1r * s1r + s1i * s1i + s2r * s2r + s2i * s2i) * sin(th[j]);
  }
}
/* trapezoid CDF, normalised, monotone, last = 1; CLARITAS.py:3583-3592 */
static int phase_cdf(const double* th, const double* pdf, int nt, double* cdf) {
  cdf[0] = 0.0;
  double s = 0.0;
  for (int j = 1; j < nt; ++j) {
    double area = 0.5 * (pdf[j] + pdf[j - 1]) * (th[j] - th[j - 1]);
    if (area < 0.0 && area > -1e-18) area = 0.0;
    s += area;
    cdf[j] = s;
  }
  if (!(s > 0.0) || !isfinite(s)) {
    for (int j = 0; j < nt; ++j) cdf[j] = (double)j / (double)(nt - 1);
    return 3;
  }
  double prev = 0.0;
  for (int j = 0; j < nt; ++j) {
    double v = cdf[j] / s;
    if (v < prev) v = prev;
    if (v > 1.0) v = 1.0;
    cdf[j] = v;
    prev = v;
  }
  cdf[nt - 1] = 1.0;
  return 0;
}

int orc_mie(const clr_particles* p, clr_mie_out* o) {
  if (!p || !o || p->n <= 0) return CLR_E_ARG;
  int nt = p->n_theta, err = 0;
  if (o->theta && nt >= 2) theta_grid(o->theta, nt);
  for (int i = 0; i < p->n; ++i) {
    double q[6], x;
    int nstop = 0;
    cd *A = NULL, *B = NULL;
    int e = mie_one(p->diameter[i], p->n_real[i], p->k_imag[i], p->n_medium, p->wavelength, &x,
                    &nstop, q, &A, &B);
    if (e) { err = err ? err : e; continue; }
    if (o->size_parameter) o->size_parameter[i] = x;
    if (o->qext) o->qext[i] = q[0];
    if (o->qsca) o->qsca[i] = q[1];
    if (o->qabs) o->qabs[i] = q[2];
    if (o->qback) o->qback[i] = q[3];
    if (o->asym_g) o->asym_g[i] = q[4];
    if (o->albedo) o->albedo[i] = q[5];
    if (o->nstop) o->nstop[i] = nstop;
    if (o->theta && o->phase_pdf && nt >= 2) {
      phase_pdf(A, B, nstop, o->theta, nt, o->phase_pdf + (size_t)i * nt);
      if (o->phase_cdf) {
        int ce = phase_cdf(o->theta, o->phase_pdf + (size_t)i * nt, nt, o->phase_cdf + (size_t)i * nt);
        if (ce && !err) err = ce;
      }
    }
    free(A);
    free(B);
  }
  return err;
}

/* ======================================================================
 * Transport split, CLARITAS.py:3845-3921, with the trace-channel rates
 * chosen as in CLARITAS.py:6368-6384.  The plain-Mie law (code 6) traces
 * mu_ext in one channel.
 * ====================================================================== */
static int transport_precompute(orc_model* m, const clr_particles* p) {
  int n = m->n, em = p->event_model;
  double sw = 0.0;
  for (int i = 0; i < n; ++i) sw += p->weight[i];
  if (!(sw > 0.0)) return 4;
  double* sg = (double*)calloc((size_t)n, sizeof(double));
  double* ss = (double*)calloc((size_t)n, sizeof(double));
  double* se = (double*)calloc((size_t)n, sizeof(double));
  double* swv = (double*)calloc((size_t)n, sizeof(double));
  double* mass = (double*)calloc((size_t)n, sizeof(double));
  double* srcw = (double*)calloc((size_t)n, sizeof(double));
  double avg_mass = 0.0;
  int err = 0;
  for (int i = 0; i < n; ++i) {
    double r = 0.5 * p->diameter[i];
    m->radius[i] = r;
    srcw[i] = p->weight[i] / sw;
    mass[i] = p->density * (4.0 / 3.0) * PI_D * r * r * r;
    avg_mass += srcw[i] * mass[i];
  }
  for (int i = 0; i < n && !err; ++i) {
    double nd = (p->weight_mode == 0) ? p->concentration * srcw[i] / mass[i]
                                      : (avg_mass > 0.0 ? p->concentration / avg_mass : 0.0) * srcw[i];
    m->nd[i] = nd;
    double r = m->radius[i];
    double g0 = PI_D * r * r, s0 = fmax(m->qsca[i], 0.0) * g0, e0 = fmax(m->qext[i], 0.0) * g0;
    double tol = 1e-10 * fmax(1.0, fmax(fabs(m->qext[i]), fabs(m->qsca[i])));
    double k = p->k_imag[i];
    if (!isfinite(k) || k < 0.0 || !isfinite(m->qext[i]) || !isfinite(m->qsca[i]) ||
        m->qext[i] < -tol || m->qsca[i] < -tol || m->qsca[i] > m->qext[i] + tol ||
        (k == 0.0 && fabs(m->qext[i] - m->qsca[i]) > tol)) {
      err = 5;
      break;
    }
    double w = 0.0;
    if (em == 3) w = p->hybrid_wave_scale * fmax(m->qsca[i] - 1.0, 0.0) * g0;
    else if (em == 4) {
      double rate = (k > 0.0) ? e0 : s0;
      w = (k > 0.0) ? ((s0 > 0.0) ? rate * fmax(s0 - g0, 0.0) / s0 : 0.0) : fmax(s0 - g0, 0.0);
    }
    sg[i] = g0; ss[i] = s0; se[i] = e0; swv[i] = w;
    m->mug[i] = nd * g0; m->mus[i] = nd * s0; m->mue[i] = nd * e0; m->muw[i] = nd * w;
  }
  if (!err) {
    double mg = 0, ms = 0, me = 0, mw = 0, mt = 0;
    for (int i = 0; i < n; ++i) { mg += m->mug[i]; ms += m->mus[i]; me += m->mue[i]; mw += m->muw[i]; }
    for (int i = 0; i < n; ++i) {
      double mb;
      if (em == 0) mb = m->mug[i];
      else if (em == 1) mb = m->mus[i];
      else if (em == 2 || em == 6) mb = m->mue[i];
      else if (em == 3) mb = m->mug[i] + m->muw[i];
      else mb = (p->k_imag[i] > 0.0) ? m->mue[i] : m->mus[i];
      m->muev[i] = mb;
      mt += mb;
    }
    int hyb = (em == 3 || em == 4);
    double mtg = (em == 4) ? fmax(mt - mw, 0.0) : ((em == 3) ? mg : mt);
    double mtw = hyb ? mw : 0.0;
    double cg = 0, cw = 0;
    for (int i = 0; i < n; ++i) {
      double mgi = (em == 4) ? fmax(m->muev[i] - m->muw[i], 0.0) : m->mug[i];
      double pg = hyb ? (mtg > 0.0 ? mgi / mtg : 0.0) : (mt > 0.0 ? m->muev[i] / mt : 0.0);
      double pw = (hyb && mtw > 0.0) ? m->muw[i] / mtw : 0.0;
      cg += pg; cw += pw;
      m->gcdf[i] = cg; m->wcdf[i] = cw;
    }
    if (n > 0) { if (cg > 0.0) m->gcdf[n - 1] = 1.0; if (cw > 0.0) m->wcdf[n - 1] = 1.0; }
    m->mu_scalars[0] = mtg; m->mu_scalars[1] = mtw; m->mu_scalars[2] = mg; m->mu_scalars[3] = ms;
    m->mu_scalars[4] = me; m->mu_scalars[5] = mtw; m->mu_scalars[6] = mt;
    double cwave = 1.0;
    if (hyb && p->wave_rms_height_m > 0.0 && p->wavelength > 0.0 && p->n_medium > 0.0) {
      double a = 4.0 * PI_D * p->n_medium * p->wave_rms_height_m / p->wavelength;
      cwave = exp(-a * a);
      if (cwave < 0.0) cwave = 0.0;
      if (cwave > 1.0) cwave = 1.0;
    }
    m->mu_scalars[7] = cwave;
    /* traced per-bin channel rates, CLARITAS.py:6368-6384 */
    m->sum_trace_g = 0.0; m->sum_trace_w = 0.0;
    for (int i = 0; i < n; ++i) {
      if (em == 3) { m->trace_g[i] = m->mug[i]; m->trace_w[i] = m->muw[i]; }
      else if (em == 4) { m->trace_g[i] = fmax(m->muev[i] - m->muw[i], 0.0); m->trace_w[i] = m->muw[i]; }
      else { m->trace_g[i] = m->muev[i]; m->trace_w[i] = 0.0; }
      m->sum_trace_g += m->trace_g[i];
      m->sum_trace_w += m->trace_w[i];
    }
  }
  free(sg); free(ss); free(se); free(swv); free(mass); free(srcw);
  return err;
}

/* ======================================================================
 * Sediment coefficients, CLARITAS.py:3924-4038 (device kernel) and the
 * host twin CLARITAS.py:883-1079.
 * ====================================================================== */
static double schiller_naumann_velocity(double v0, double d, double rho_w, double mu) {
  if (!(v0 > 0.0)) return 0.0;
  double v = v0, K = rho_w * d / mu;
  for (int it = 0; it < 24; ++it) {
    double Re = fmax(K * v, 0.0), corr = 0.15 * pow(Re, 0.687);
    double f = v * (1.0 + corr) - v0, df = 1.0 + 1.687 * corr;
    double vn = fmax(v - f / df, 0.0);
    if (fabs(vn - v) <= 2e-14 * fmax(1.0, fabs(v))) { v = vn; break; }
    v = vn;
  }
  return v;
}

static float field_max(const float* f, size_t n) {
  float mx = 0.0f;
  for (size_t i = 0; i < n; ++i) if (f[i] > mx) mx = f[i];
  return mx;
}

/* Per-bin coefficients of the analytic drift-diffusion fields (models 1, 2
 * and the per-bin grid model 3, which the meridional solver starts from):
 * settling speed, radial slip table w_r (radwr, may be NULL), radial log
 * potential table, A, B and the volume normalisation.  Returns 0 or the
 * reference's error code 31-34. */
static int sediment_bins(const clr_config* cfg, const clr_sediment* s, int tm, int n, const double* dia, int nrad,
                         double* Av, double* Bv, double* normv, double* wsv, double* radlog, double* radwr,
                         double* maxA_out) {
  int fsm = cfg->free_surface_model;
  double R = cfg->rin, ZMIN = cfg->zmin, ZMEAN = cfg->zmean;
  double DH = cfg->vortex_delta_h, CORE = cfg->vortex_core_radius;
  double rho_p = s->particle_density, rho_w = s->water_density, mu = s->water_viscosity;
  double grav = s->gravity, Dt = s->eddy_diffusivity;
  int NQ = s->n_quad;
  if (!(rho_p > rho_w) || !(rho_w > 0.0) || !(mu > 0.0) || !(grav > 0.0) || !(Dt > 0.0) ||
      !(R > 0.0) || NQ < 64 || nrad < 65)
    return 31;
  if (fsm < 0 || fsm > 2 || (DH > 0.0 && fsm == 0) || (fsm == 2 && !(CORE > 0.0))) return 32;
  double R2 = R * R, a2 = CORE * CORE, Hinf = 0.0, hbar = 0.0, om2 = 0.0;
  if (DH > 0.0 && fsm == 1) { Hinf = DH; hbar = 0.5 * DH; om2 = 2.0 * grav * DH / R2; }
  else if (DH > 0.0 && fsm == 2) {
    Hinf = DH * (a2 + R2) / R2; hbar = Hinf * (1.0 - (a2 / R2) * log1p(R2 / a2)); om2 = 2.0 * grav * Hinf / a2;
  }
  double fR = (fsm == 2) ? R2 / (a2 + R2) : R2;
  double Hmean = ZMEAN - ZMIN, dr = R / (double)NQ, drad = R / (double)(nrad - 1);
  if (!(Hmean > 0.0)) return 33;
  double maxA = 0.0;
  double* wr = (double*)malloc(sizeof(double) * (size_t)nrad);
  for (int i = 0; i < n; ++i) {
    double d = dia[i];
    double tau = rho_p * d * d / (18.0 * mu);
    double ws0 = (rho_p - rho_w) * grav * d * d / (18.0 * mu);
    double ws = (tm != 1) ? schiller_naumann_velocity(ws0, d, rho_w, mu) : ws0;
    wsv[i] = ws;
    double B = ws / Dt;
    Bv[i] = B;
    double* rl = radlog + (size_t)i * nrad;
    if (tm == 1) {
      double A = tau * om2 * ((fsm == 2) ? a2 : 1.0) / (2.0 * Dt);
      Av[i] = A;
      if (A > maxA) maxA = A;
      for (int j = 0; j < nrad; ++j) {
        double r = (double)j * drad, rr2 = r * r, f = (fsm == 2) ? rr2 / (a2 + rr2) : rr2;
        rl[j] = A * (f - fR);
        double acc = 0.0;
        if (r > 0.0) acc = (fsm == 1) ? om2 * r : ((fsm == 2) ? om2 * r / pow(1.0 + r * r / a2, 2.0) : 0.0);
        wr[j] = tau * acc;
      }
    } else {
      double total = 0.0, prev = 0.0;
      for (int j = 0; j < nrad; ++j) {
        double r = (double)j * drad, acc = 0.0;
        if (r > 0.0) acc = (fsm == 1) ? om2 * r : ((fsm == 2) ? om2 * r / pow(1.0 + r * r / a2, 2.0) : 0.0);
        wr[j] = schiller_naumann_velocity(tau * acc, d, rho_w, mu);
        if (j > 0) total += 0.5 * (prev + wr[j]) * drad;
        prev = wr[j];
      }
      double cum = 0.0;
      prev = 0.0;
      for (int j = 0; j < nrad; ++j) {
        if (j > 0) cum += 0.5 * (prev + wr[j]) * drad;
        prev = wr[j];
        double psi = (cum - total) / Dt;
        if (psi > 0.0) psi = 0.0;
        rl[j] = psi;
      }
      rl[nrad - 1] = 0.0;
      double Aeq = (fR > 0.0) ? -rl[0] / fR : 0.0;
      Av[i] = Aeq;
      if (Aeq > maxA) maxA = Aeq;
    }
    if (radwr) memcpy(radwr + (size_t)i * nrad, wr, sizeof(double) * (size_t)nrad);
    double integ = 0.0;
    for (int j = 0; j < NQ; ++j) {
      double r = ((double)j + 0.5) * dr, rr2 = r * r, zs = ZMEAN;
      if (fsm == 1 && DH > 0.0) zs = ZMEAN + DH * (rr2 / R2 - 0.5);
      else if (fsm == 2 && DH > 0.0) zs = ZMEAN - hbar + Hinf * rr2 / (a2 + rr2);
      double H = fmax(zs - ZMIN, 0.0), lr;
      if (tm == 1) { double f = (fsm == 2) ? rr2 / (a2 + rr2) : rr2; lr = Av[i] * (f - fR); }
      else {
        double u = r / R * (double)(nrad - 1);
        int k = (int)floor(u);
        if (k < 0) k = 0;
        if (k > nrad - 2) k = nrad - 2;
        double t = u - (double)k;
        lr = (1.0 - t) * rl[k] + t * rl[k + 1];
        if (lr > 0.0) lr = 0.0;
      }
      double rad = exp(lr), iz = (fabs(B) < 1e-14) ? rad * H : rad * (-expm1(-B * H)) / B;
      integ += r * iz;
    }
    double norm = (2.0 / (R2 * Hmean)) * integ * dr;
    if (!(norm > 0.0) || !isfinite(norm)) { free(wr); return 34; }
    normv[i] = norm;
  }
  free(wr);
  if (maxA_out) *maxA_out = maxA;
  return 0;
}

int orc_sediment_bins(const clr_config* cfg, const clr_sediment* s, int n, const double* diameter, double* A,
                      double* B, double* norm, double* ws, double* radlog, double* radwr) {
  if (!cfg || !s || n <= 0 || !diameter || !A || !B || !norm || !ws || !radlog) return -1;
  return sediment_bins(cfg, s, cfg->sediment_model, n, diameter, s->n_radial, A, B, norm, ws, radlog, radwr, NULL);
}

static int sediment_precompute(orc_model* m, const clr_sediment* s) {
  int n = m->n, tm = m->cfg.sediment_model;
  int nrad = m->nrad;
  for (int i = 0; i < n; ++i) {
    m->sedA[i] = 0.0; m->sedB[i] = 0.0; m->sednorm[i] = 1.0; m->sedws[i] = 0.0;
    for (int j = 0; j < nrad; ++j) m->radlog[(size_t)i * nrad + j] = 0.0;
  }
  if (tm == CLR_SED_UNIFORM) {
    m->maj[0] = m->sum_trace_g; m->maj[1] = m->sum_trace_w; m->maj[2] = m->maj[0] + m->maj[1]; m->maj[3] = 0.0;
    return 0;
  }
  if (tm == CLR_SED_RZ_FIELD || tm == CLR_SED_XYZ_FIELD) {
    double mx = (double)field_max(m->field, m->field_len);
    m->maj[0] = m->sum_trace_g * mx; m->maj[1] = m->sum_trace_w * mx;
    m->sum_g_f = (float)m->sum_trace_g; m->sum_w_f = (float)m->sum_trace_w;
    m->maj[2] = m->maj[0] + m->maj[1]; m->maj[3] = 0.0;
    return 0;
  }
  /* analytic drift-diffusion fields (models 1, 2) */
  double maxA = 0.0;
  int e = sediment_bins(&m->cfg, s, tm, n, m->d, nrad, m->sedA, m->sedB, m->sednorm, m->sedws, m->radlog, NULL,
                        &maxA);
  if (e) return e;
  double mjg = 0.0, mjw = 0.0;
  for (int i = 0; i < n; ++i) {
    mjg += m->trace_g[i] / m->sednorm[i];
    mjw += m->trace_w[i] / m->sednorm[i];
  }
  if (tm == CLR_SED_GRID_PER_BIN) {
    /* the caller's per-bin r-z grid replaces the analytic field */
    size_t plane = (size_t)m->fp.fn0 * (size_t)m->fp.fn1;
    mjg = 0.0; mjw = 0.0;
    for (int i = 0; i < n; ++i) {
      double mx = (double)field_max(m->field + (size_t)i * plane, plane);
      mjg += m->trace_g[i] * mx; mjw += m->trace_w[i] * mx;
    }
    maxA = 0.0;
  }
  m->maj[0] = mjg; m->maj[1] = mjw; m->maj[2] = mjg + mjw; m->maj[3] = maxA;
  if (!(m->maj[2] >= 0.0) || !isfinite(m->maj[2])) return 35;
  return 0;
}

/* ======================================================================
 * Sediment sampling at a point, CLARITAS.py:1537-1609
 * ====================================================================== */
static double rz_grid_sample(const float* G, int NR, int NZ, float x, float y, float z, float R,
                             float ZMIN, float ZMAX) {
  if (NR < 2 || NZ < 2 || !(ZMAX > ZMIN)) return 1.0;
  double rr = sqrt(fmax((double)x * (double)x + (double)y * (double)y, 0.0));
  if (rr > (double)R) rr = (double)R;
  double dr = (double)R / (double)NR, dz = ((double)ZMAX - (double)ZMIN) / (double)NZ;
  double ur = rr / dr - 0.5, uz = ((double)z - (double)ZMIN) / dz - 0.5;
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
/* trilinear, cell-centred Cartesian field over [-H,H]^2 x [zmin,zmax] */
static double xyz_grid_sample(const float* G, int NX, int NY, int NZ, float x, float y, float z,
                              float H, float ZMIN, float ZMAX) {
  double dx = 2.0 * (double)H / NX, dy = 2.0 * (double)H / NY, dz = ((double)ZMAX - (double)ZMIN) / NZ;
  double u[3] = {((double)x + H) / dx - 0.5, ((double)y + H) / dy - 0.5, ((double)z - ZMIN) / dz - 0.5};
  int nn[3] = {NX, NY, NZ}, ii[3];
  double tt[3];
  for (int a = 0; a < 3; ++a) {
    double v = u[a];
    if (v < 0.0) v = 0.0;
    if (v > (double)(nn[a] - 1)) v = (double)(nn[a] - 1);
    int i = (int)floor(v);
    double t = v - (double)i;
    if (i >= nn[a] - 1) { i = nn[a] - 2; t = 1.0; }
    ii[a] = i; tt[a] = t;
  }
  size_t sy = (size_t)NZ, sx = (size_t)NY * NZ;
  const float* b = G + (size_t)ii[0] * sx + (size_t)ii[1] * sy + ii[2];
  double c00 = (1 - tt[2]) * b[0] + tt[2] * b[1];
  double c01 = (1 - tt[2]) * b[sy] + tt[2] * b[sy + 1];
  double c10 = (1 - tt[2]) * b[sx] + tt[2] * b[sx + 1];
  double c11 = (1 - tt[2]) * b[sx + sy] + tt[2] * b[sx + sy + 1];
  double c0 = (1 - tt[1]) * c00 + tt[1] * c01, c1 = (1 - tt[1]) * c10 + tt[1] * c11;
  return fmax((1 - tt[0]) * c0 + tt[0] * c1, 0.0);
}

/* Shared fields (models 4, 5; extensions with no reference counterpart) are
 * interpolated in FP32 with float inverse cell sizes, as the device does. */
static float shared_field_scale(const orc_model* m, float x, float y, float z) {
  const fparams* P = &m->fp;
  const float* G = m->field;
  if (P->sed_model == CLR_SED_RZ_FIELD) {
    int NR = P->fn0, NZ = P->fn1;
    float rr = fminf(sqrtf(fmaxf(x * x + y * y, 0.0f)), P->rin);
    float ur = fminf(fmaxf(rr * m->finv[0] - 0.5f, 0.0f), (float)(NR - 1));
    float uz = fminf(fmaxf((z - P->zmin) * m->finv[1] - 0.5f, 0.0f), (float)(NZ - 1));
    int i = (int)floorf(ur), k = (int)floorf(uz);
    float tr = ur - (float)i, tz = uz - (float)k;
    if (i >= NR - 1) { i = NR - 2; tr = 1.0f; }
    if (k >= NZ - 1) { k = NZ - 2; tz = 1.0f; }
    const float* g = G + (size_t)i * NZ + k;
    float q0 = (1.0f - tr) * g[0] + tr * g[NZ], q1 = (1.0f - tr) * g[1] + tr * g[NZ + 1];
    return fmaxf((1.0f - tz) * q0 + tz * q1, 0.0f);
  }
  int nn[3] = {P->fn0, P->fn1, P->fn2}, ii[3];
  float u[3] = {(x + P->field_half) * m->finv[0] - 0.5f, (y + P->field_half) * m->finv[1] - 0.5f,
                (z - P->zmin) * m->finv[2] - 0.5f}, tt[3];
  for (int a = 0; a < 3; ++a) {
    float v = fminf(fmaxf(u[a], 0.0f), (float)(nn[a] - 1));
    int i = (int)floorf(v);
    float t = v - (float)i;
    if (i >= nn[a] - 1) { i = nn[a] - 2; t = 1.0f; }
    ii[a] = i; tt[a] = t;
  }
  size_t sy = (size_t)nn[2], sx = (size_t)nn[1] * nn[2];
  const float* b = G + (size_t)ii[0] * sx + (size_t)ii[1] * sy + ii[2];
  float c00 = (1.0f - tt[2]) * b[0] + tt[2] * b[1];
  float c01 = (1.0f - tt[2]) * b[sy] + tt[2] * b[sy + 1];
  float c10 = (1.0f - tt[2]) * b[sx] + tt[2] * b[sx + 1];
  float c11 = (1.0f - tt[2]) * b[sx + sy] + tt[2] * b[sx + sy + 1];
  float c0 = (1.0f - tt[1]) * c00 + tt[1] * c01, c1 = (1.0f - tt[1]) * c10 + tt[1] * c11;
  return fmaxf((1.0f - tt[0]) * c0 + tt[0] * c1, 0.0f);
}

static double sediment_local_scale(const orc_model* m, int pidx, float x, float y, float z) {
  const fparams* P = &m->fp;
  int tm = P->sed_model;
  if (tm == CLR_SED_UNIFORM) return 1.0;
  if (tm == CLR_SED_GRID_PER_BIN)
    return rz_grid_sample(m->field + (size_t)pidx * P->fn0 * P->fn1, P->fn0, P->fn1, x, y, z, P->rin,
                          P->zmin, P->field_zmax);
  if (tm == CLR_SED_RZ_FIELD)
    return rz_grid_sample(m->field, P->fn0, P->fn1, x, y, z, P->rin, P->zmin, P->field_zmax);
  if (tm == CLR_SED_XYZ_FIELD)
    return xyz_grid_sample(m->field, P->fn0, P->fn1, P->fn2, x, y, z, P->field_half, P->zmin,
                           P->field_zmax);
  double rr2 = (double)x * (double)x + (double)y * (double)y;
  double R2 = (double)P->rin * (double)P->rin;
  if (rr2 > R2) rr2 = R2;
  double rl = 0.0;
  if (tm == CLR_SED_STOKES) {
    if (P->fs_model == 1) rl = m->sedA[pidx] * (rr2 - R2);
    else {
      double a2 = (double)P->core_r * (double)P->core_r;
      rl = m->sedA[pidx] * (rr2 / (a2 + rr2) - R2 / (a2 + R2));
    }
  } else {
    int nrad = m->nrad;
    double rr = sqrt(fmax(rr2, 0.0));
    double u = (P->rin > 0.0f) ? rr / (double)P->rin * (double)(nrad - 1) : 0.0;
    int j = (int)floor(u);
    if (j < 0) j = 0;
    if (j > nrad - 2) j = nrad - 2;
    double t = u - (double)j;
    const double* L = m->radlog + (size_t)pidx * nrad;
    rl = (1.0 - t) * L[j] + t * L[j + 1];
    if (rl > 0.0) rl = 0.0;
  }
  double h = fmax((double)z - (double)P->zmin, 0.0);
  double ns = m->sednorm[pidx];
  if (!(ns > 0.0)) return 0.0;
  return exp(rl - m->sedB[pidx] * h) / ns;
}

static int sample_cdf_index(rng_t* r, const float* cdf, int n) {
  if (n <= 1) return 0;
  float u = U(r);
  int lo = 0, hi = n - 1;
  while (lo < hi) { int mid = (lo + hi) >> 1; if (u <= cdf[mid]) hi = mid; else lo = mid + 1; }
  return lo;
}
/* O(bins) walk on the local rates, CLARITAS.py:1595-1609 */
static int sample_local_bin(rng_t* r, const orc_model* m, const double* rate, float x, float y,
                            float z, double local_total) {
  if (!(local_total > 0.0)) return 0;
  double target = (double)U(r) * local_total, cum = 0.0;
  for (int i = 0; i < m->n; ++i) {
    cum += rate[i] * sediment_local_scale(m, i, x, y, z);
    if (target <= cum) return i;
  }
  return m->n > 0 ? m->n - 1 : 0;
}
/* bin selection for one channel: per-bin CDF when the density scale is
 * common to every bin (uniform, shared fields), local walk otherwise */
static int select_bin(rng_t* r, const orc_model* m, int wave, float x, float y, float z,
                      float local_channel) {
  int tm = m->fp.sed_model;
  if (tm == CLR_SED_UNIFORM || tm == CLR_SED_RZ_FIELD || tm == CLR_SED_XYZ_FIELD)
    return sample_cdf_index(r, wave ? m->wcdf_f : m->gcdf_f, m->n);
  return sample_local_bin(r, m, wave ? m->trace_w : m->trace_g, x, y, z, (double)local_channel);
}

/* ======================================================================
 * Free surface, CLARITAS.py:1503-1527, 1611-1779; Rankine is an extension
 * ====================================================================== */
static float free_surface_z(const fparams* P, float x, float y) {
  int M = P->fs_model;
  if (M == 0 || P->dh <= 0.0f) return P->zmean;
  float q = x * x + y * y, R = P->rin;
  if (M == 1) return P->zmean + P->dh * (q / (R * R) - 0.5f);
  if (M == 3) {
    double qa = (double)q, a2 = (double)P->core_r * (double)P->core_r;
    double f = (qa < a2) ? P->rk_K * qa / a2 : P->rk_K * (2.0 - a2 / qa);
    return (float)(P->rk_C0 + f);
  }
  float a2 = P->core_r * P->core_r, R2 = R * R;
  float H = P->dh * (a2 + R2) / R2;
  float hbar = H * (1.0f - (a2 / R2) * log1pf(R2 / a2));
  return P->zmean - hbar + H * q / (a2 + q);
}

static double quadratic_min_positive(double A, double B, double C) {
  if (fabs(A) < 1e-24) {
    if (fabs(B) < 1e-24) return 1.0e300;
    double t = -C / B;
    return t > 0.0 ? t : 1.0e300;
  }
  double D = B * B - 4.0 * A * C;
  if (D < 0.0) return 1.0e300;
  double sd = sqrt(fmax(D, 0.0)), q = -0.5 * (B + copysign(sd, B));
  double t1 = (fabs(q) > 1e-300) ? q / A : (-B - sd) / (2.0 * A);
  double t2 = (fabs(q) > 1e-300) ? C / q : (-B + sd) / (2.0 * A);
  double best = 1.0e300;
  if (t1 > 0.0 && t1 < best) best = t1;
  if (t2 > 0.0 && t2 < best) best = t2;
  return best;
}
static double cubic_min_positive(double A3, double A2, double A1, double A0) {
  double scale = fmax(fmax(fabs(A3), fabs(A2)), fmax(fabs(A1), fabs(A0)));
  if (scale <= 0.0) return 1.0e300;
  if (fabs(A3) <= 1e-12) return quadratic_min_positive(A2, A1, A0);
  double a = A2 / A3, b = A1 / A3, c = A0 / A3;
  double p = b - a * a / 3.0, q = 2.0 * a * a * a / 27.0 - a * b / 3.0 + c;
  double hq = 0.5 * q, tp = p / 3.0, d1 = hq * hq, d2 = tp * tp * tp, disc = d1 + d2;
  double tol = 1e-12 * (fabs(d1) + fabs(d2) + 1e-300), best = 1.0e300;
  if (disc > tol) {
    double sd = sqrt(disc), t = cbrt(-hq + sd) + cbrt(-hq - sd) - a / 3.0;
    if (t > 0.0) best = t;
  } else if (fabs(p) < 1e-24 && fabs(q) < 1e-24) {
    double t = -a / 3.0;
    if (t > 0.0) best = t;
  } else if (disc >= -tol) {
    double u = cbrt(-hq), t1 = 2.0 * u - a / 3.0, t2 = -u - a / 3.0;
    if (t1 > 0.0 && t1 < best) best = t1;
    if (t2 > 0.0 && t2 < best) best = t2;
  } else {
    double mm = 2.0 * sqrt(fmax(-tp, 0.0)), den = sqrt(fmax(-(tp * tp * tp), 1e-300));
    double ca = fmin(fmax(-hq / den, -1.0), 1.0), phi = acos(ca);
    for (int k = 0; k < 3; ++k) {
      double t = mm * cos((phi + 2.0 * PI_D * (double)k) / 3.0) - a / 3.0;
      if (t > 0.0 && t < best) best = t;
    }
  }
  return best;
}
/* monotone-interval bisection for the first sign change of a cubic on
 * [lo, hi], CLARITAS.py:1690-1722 */
static double bounded_cubic_root(double a3, double a2, double a1, double a0, double lo, double hi) {
  if (!(hi > lo) || !isfinite(hi)) return 1.0e300;
  double cuts[4];
  int n = 1;
  cuts[0] = lo;
  double A = 3.0 * a3, B = 2.0 * a2, C = a1;
  if (A != 0.0) {
    double disc = B * B - 4.0 * A * C;
    if (disc >= 0.0) {
      double q = -0.5 * (B + copysign(sqrt(disc), B));
      double r1 = (q != 0.0) ? q / A : -B / (2.0 * A), r2 = (q != 0.0) ? C / q : r1;
      if (r1 > lo && r1 < hi) cuts[n++] = r1;
      if (r2 > lo && r2 < hi && r2 != r1) cuts[n++] = r2;
    }
  } else if (B != 0.0) {
    double r = -C / B;
    if (r > lo && r < hi) cuts[n++] = r;
  }
  cuts[n++] = hi;
  for (int i = 1; i < n; ++i) {
    double v = cuts[i];
    int j = i;
    while (j > 0 && cuts[j - 1] > v) { cuts[j] = cuts[j - 1]; --j; }
    cuts[j] = v;
  }
  for (int i = 0; i < n - 1; ++i) {
    double l = cuts[i], r = cuts[i + 1];
    double fl = ((a3 * l + a2) * l + a1) * l + a0, fr = ((a3 * r + a2) * r + a1) * r + a0;
    if (l > 0.0 && fl == 0.0) return l;
    if ((fl < 0.0 && fr > 0.0) || (fl > 0.0 && fr < 0.0)) {
      for (int j = 0; j < 64; ++j) {
        double mid = l + 0.5 * (r - l);
        if (mid == l || mid == r) break;
        double fm = ((a3 * mid + a2) * mid + a1) * mid + a0;
        if (fm == 0.0) { l = r = mid; break; }
        if ((fl < 0.0 && fm < 0.0) || (fl > 0.0 && fm > 0.0)) { l = mid; fl = fm; } else r = mid;
      }
      return l + 0.5 * (r - l);
    }
    if (r > 0.0 && fr == 0.0) return r;
  }
  return 1.0e300;
}

/* Rankine surface hit: split the ray at the core cylinder r=a, solve the
 * core paraboloid (quadratic) and the outer a^2/r^2 sheet (cubic after
 * multiplying by r^2) on each piece, first root wins. */
static double rankine_hit(const fparams* P, double X, double Y, double Z, double VX, double VY,
                          double VZ) {
  double a2 = (double)P->core_r * (double)P->core_r, K = P->rk_K, C0 = P->rk_C0;
  double RR = (double)P->rin * (double)P->rin;
  double Q2 = VX * VX + VY * VY, Q1 = 2.0 * (X * VX + Y * VY), Q0 = X * X + Y * Y;
  double tend = 1.0e300;
  double tin = 1.0e300, tout = -1.0;
  if (Q2 > 0.0) {
    double dc = Q1 * Q1 - 4.0 * Q2 * (Q0 - RR);
    if (dc >= 0.0) {
      double q = -0.5 * (Q1 + copysign(sqrt(dc), Q1));
      double ta = (q != 0.0) ? q / Q2 : -Q1 / (2.0 * Q2), tb = (q != 0.0) ? (Q0 - RR) / q : ta;
      tend = fmax(ta, tb);
    }
    double dk = Q1 * Q1 - 4.0 * Q2 * (Q0 - a2);
    if (dk > 0.0) {
      double q = -0.5 * (Q1 + copysign(sqrt(dk), Q1));
      double ta = (q != 0.0) ? q / Q2 : -Q1 / (2.0 * Q2), tb = (q != 0.0) ? (Q0 - a2) / q : ta;
      tin = fmin(ta, tb);
      tout = fmax(ta, tb);
    }
  } else {
    if (Q0 < a2) { tin = -1.0e300; tout = 1.0e300; }
  }
  if (!(tend > 0.0)) return 1.0e300;
  double cuts[4] = {0.0, 0.0, 0.0, 0.0};
  int nc = 0;
  cuts[nc++] = 0.0;
  if (tin > 0.0 && tin < tend) cuts[nc++] = tin;
  if (tout > 0.0 && tout < tend) cuts[nc++] = tout;
  cuts[nc++] = (tend < 1.0e299) ? tend : 1.0;  /* vertical ray: one piece, bounded below */
  if (!(tend < 1.0e299)) {
    /* vertical ray: surface height at fixed r */
    double f = (Q0 < a2) ? K * Q0 / a2 : K * (2.0 - a2 / Q0);
    double t = (VZ != 0.0) ? (C0 + f - Z) / VZ : 1.0e300;
    return (t > 0.0) ? t : 1.0e300;
  }
  for (int i = 0; i + 1 < nc; ++i) {
    double lo = cuts[i], hi = cuts[i + 1];
    if (!(hi > lo)) continue;
    double mid = 0.5 * (lo + hi), qm = (Q2 * mid + Q1) * mid + Q0;
    double t;
    if (qm < a2) {
      double k = K / a2;
      double A = -k * Q2, B = VZ - k * Q1, C = Z - C0 - k * Q0;
      t = bounded_cubic_root(0.0, A, B, C, lo, hi);
    } else {
      double L = Z - C0 - 2.0 * K;
      t = bounded_cubic_root(VZ * Q2, L * Q2 + VZ * Q1, L * Q1 + VZ * Q0, L * Q0 + K * a2, lo, hi);
    }
    if (t < 1.0e299 && t > 0.0) return t;
  }
  return 1.0e300;
}

/* first positive intersection with the free surface, CLARITAS.py:1726-1766 */
static float free_surface_hit(const fparams* P, float x, float y, float z, float vx, float vy,
                              float vz) {
  int M = P->fs_model;
  float R = P->rin;
  if (M == 0 || P->dh <= 0.0f) {
    if (fabsf(vz) <= 1e-14f) return INF_F;
    float t = (P->zmean - z) / vz;
    return (t > 0.0f) ? t : INF_F;
  }
  if (M == 1) {
    float aa = P->dh / (R * R);
    double A = -(double)aa * ((double)vx * vx + (double)vy * vy);
    double B = (double)vz - 2.0 * (double)aa * ((double)x * vx + (double)y * vy);
    double C = (double)z - ((double)P->zmean - 0.5 * (double)P->dh) - (double)aa * ((double)x * x + (double)y * y);
    double t = quadratic_min_positive(A, B, C);
    return (t < 0.5 * (double)INF_F) ? (float)t : INF_F;
  }
  if (M == 3) {
    double t = rankine_hit(P, x, y, z, vx, vy, vz);
    return (t > 0.0 && t < 0.5 * (double)INF_F) ? (float)t : INF_F;
  }
  double X = x, Y = y, Z = z, VX = vx, VY = vy, VZ = vz;
  double RR = (double)R * (double)R, aa = (double)P->core_r * (double)P->core_r;
  double H = (double)P->dh * (aa + RR) / RR;
  double hbar = H * (1.0 - (aa / RR) * log1p(RR / aa));
  double Cs = (double)P->zmean - hbar;
  double Q2 = VX * VX + VY * VY, Q1 = 2.0 * (X * VX + Y * VY), Q0 = X * X + Y * Y, L0 = Z - Cs;
  double c3 = VZ * Q2, c2 = (L0 - H) * Q2 + VZ * Q1, c1 = (L0 - H) * Q1 + VZ * (aa + Q0);
  double c0 = L0 * aa + (L0 - H) * Q0;
  double t = cubic_min_positive(c3, c2, c1, c0);
  double xx = X + t * VX, yy = Y + t * VY, qq = xx * xx + yy * yy;
  double res = (Z + t * VZ) - (Cs + H * qq / (aa + qq));
  double tol = 32.0 * DEPS * fmax(fmax(fabs(Z), fabs(Cs) + fabs(H)), 1.0e-3);
  if (!(t < 0.5 * (double)INF_F) || qq > RR * (1.0 + 2e-6) || !isfinite(res) || fabs(res) > tol) {
    if (Q2 > 0.0) {
      double dc = Q1 * Q1 - 4.0 * Q2 * (Q0 - RR);
      if (dc < 0.0) return INF_F;
      double q = -0.5 * (Q1 + copysign(sqrt(dc), Q1));
      double ta = (q != 0.0) ? q / Q2 : -Q1 / (2.0 * Q2), tb = (q != 0.0) ? (Q0 - RR) / q : ta;
      double lo = fmax(0.0, fmin(ta, tb)), hi = fmax(ta, tb);
      t = bounded_cubic_root(c3, c2, c1, c0, lo, hi);
    } else {
      t = (VZ != 0.0) ? ((Cs + H * Q0 / (aa + Q0)) - Z) / VZ : 1.0e300;
      if (!(t > 0.0) || Q0 > RR * (1.0 + 2e-6)) return INF_F;
    }
  }
  return (t > 0.0 && t < 0.5 * (double)INF_F) ? (float)t : INF_F;
}

/* unit normal into the water, CLARITAS.py:1769-1779 */
static void free_surface_normal(const fparams* P, float x, float y, float* nx, float* ny, float* nz) {
  int M = P->fs_model;
  if (M == 0 || P->dh <= 0.0f) { *nx = 0.0f; *ny = 0.0f; *nz = -1.0f; return; }
  float R = P->rin;
  if (M == 1) {
    float aa = P->dh / (R * R);
    *nx = 2.0f * aa * x; *ny = 2.0f * aa * y; *nz = -1.0f;
    normalize3(nx, ny, nz);
    return;
  }
  float q = x * x + y * y, fac;
  if (M == 3) {
    float a2 = P->core_r * P->core_r, K = (float)P->rk_K;
    fac = (q < a2) ? 2.0f * K / a2 : 2.0f * K * a2 / (q * q);
  } else {
    float a2 = P->core_r * P->core_r, R2 = R * R;
    float H = P->dh * (a2 + R2) / R2;
    fac = 2.0f * H * a2 / ((a2 + q) * (a2 + q));
  }
  *nx = fac * x; *ny = fac * y; *nz = -1.0f;
  normalize3(nx, ny, nz);
}

/* surface height (double) at radius^2 q, used by particle accessibility */
static double free_surface_z_d(const fparams* P, double q) {
  int M = P->fs_model;
  double zm = (double)P->zmean;
  if (M == 0 || P->dh <= 0.0f) return zm;
  double R = (double)P->rin, dh = (double)P->dh, a = (double)P->core_r, R2 = R * R;
  if (M == 1) return zm + dh * (q / R2 - 0.5);
  if (M == 3) {
    double a2 = a * a;
    return P->rk_C0 + ((q < a2) ? P->rk_K * q / a2 : P->rk_K * (2.0 - a2 / q));
  }
  double a2 = a * a, H = dh * (a2 + R2) / R2, hbar = H * (1.0 - (a2 / R2) * log1p(R2 / a2));
  return zm - hbar + H * q / (a2 + q);
}

/* ======================================================================
 * Cell topology, CLARITAS.py:1473-1500, 2888-3293
 * ====================================================================== */
static void radial_normal(float x, float y, float* nx, float* ny) {
  float r = sqrtf(fmaxf(x * x + y * y, 1e-30f));
  *nx = x / r; *ny = y / r;
}
static int bottom_topology_at_top(float x, float y, float Rin, float Rout) {
  double r2 = (double)x * (double)x + (double)y * (double)y;
  double ri2 = (double)Rin * (double)Rin, ro2 = (double)Rout * (double)Rout;
  if (r2 < ri2) return 1;
  if (r2 <= ro2 * (1.0 + 8.0 * DEPS)) return 2;
  return 0;
}
static int cylinder_roots_d(float x, float y, float vx, float vy, float R, double* t0, double* t1) {
  *t0 = 1.0e300; *t1 = 1.0e300;
  double X = x, Y = y, VX = vx, VY = vy, RR = R;
  double a = VX * VX + VY * VY;
  if (!(a > 1e-30) || !isfinite(a)) return 0;
  double b = X * VX + Y * VY, c = X * X + Y * Y - RR * RR, d = b * b - a * c;
  double dtol = 16.0 * DEPS * (fabs(b * b) + fabs(a * c) + 1e-300);
  if (d < 0.0) { if (d >= -dtol) d = 0.0; else return 0; }
  double sd = sqrt(fmax(d, 0.0)), q = -b - copysign(sd, b);
  double r0 = (fabs(q) > 1e-300) ? q / a : (-b - sd) / a;
  double r1 = (fabs(q) > 1e-300) ? c / q : (-b + sd) / a;
  double a0 = 1.0e300, a1 = 1.0e300;
  if (r0 > 0.0 && isfinite(r0)) a0 = r0;
  if (r1 > 0.0 && isfinite(r1)) { if (r1 < a0) { a1 = a0; a0 = r1; } else a1 = r1; }
  *t0 = a0; *t1 = a1;
  return (a0 < 1.0e299) + (a1 < 1.0e299);
}
static double support_tol(float Rin, float Rout, float Zmin, float Zlow, float Ztop) {
  double s = fmax(fmax(fabs((double)Rout), fabs((double)Zlow)), fabs((double)Ztop));
  s = fmax(s, fmax(fabs((double)Rin), fabs((double)Zmin)));
  return 8.0 * FEPS * fmax(s, 1.0e-3);
}
static double finite_cylinder_forward_d(float x, float y, float z, float vx, float vy, float vz,
                                        float R, float zlo, float zhi) {
  double r[2];
  cylinder_roots_d(x, y, vx, vy, R, &r[0], &r[1]);
  double tol = support_tol(R, R, zlo, zlo, zhi), best = 1.0e300;
  for (int k = 0; k < 2; ++k) {
    double t = r[k];
    if (!(t > 0.0) || t >= 1.0e299) continue;
    double zh = (double)z + t * (double)vz;
    if (zh >= (double)zlo - tol && zh <= (double)zhi + tol && t < best) best = t;
  }
  return best;
}
/* water/headspace side exit through r=Rin, CLARITAS.py:2929-2950 */
static double bore_wall_hit_d(float x, float y, float z, float vx, float vy, float vz, float R,
                              float zlo, float zhi) {
  double X = x, Y = y, VX = vx, VY = vy, RR = (double)R * R;
  double rr2 = X * X + Y * Y, c = rr2 - RR, b = X * VX + Y * VY, a = VX * VX + VY * VY;
  double rtol = 8.0 * FEPS * RR, ztol = support_tol(R, R, zlo, zlo, zhi);
  if (c > rtol || !isfinite(c) || !isfinite(b) || !isfinite(a)) return -1.0;
  if (c >= 0.0) {
    if (b >= 0.0 && a > 0.0 && z >= zlo - ztol && z <= zhi + ztol) return 0.0;
    b *= sqrt(RR / rr2);
    c = 0.0;
  }
  if (!(a > 1e-30)) return 1.0e300;
  double d = b * b - a * c;
  if (d < 0.0) return 1.0e300;
  double sd = sqrt(d), q = -b - copysign(sd, b);
  double t = (b < 0.0) ? q / a : ((q != 0.0) ? c / q : 0.0);
  if (!(t >= 0.0) || !isfinite(t)) return 1.0e300;
  double zh = (double)z + t * (double)vz;
  return (zh >= zlo - ztol && zh <= zhi + ztol) ? t : 1.0e300;
}
/* CLARITAS.py:2961-2978 */
static int water_surface_offset(const fparams* P, float* x, float* y, float* z, float nx, float ny,
                                float nz) {
  double rr2 = (double)(*x) * (double)(*x) + (double)(*y) * (double)(*y);
  double RR = (double)P->rin * P->rin, tol = 8.0 * FEPS * RR;
  if (!isfinite(rr2) || rr2 > RR + tol) return 0;
  float xx = *x + EPS_F * nx, yy = *y + EPS_F * ny, zz = *z + EPS_F * nz;
  double rad = hypot((double)xx, (double)yy);
  if (rad > (double)P->rin) {
    double target = (double)P->rin * (1.0 - 2.0 * FEPS);
    xx = (float)((double)xx * target / rad);
    yy = (float)((double)yy * target / rad);
    float surf = free_surface_z(P, xx, yy);
    if (zz >= surf) zz = nextafterf(surf, -INF_F);
  }
  if (!isfinite(xx) || !isfinite(yy) || !isfinite(zz) || zz <= P->zmin || zz >= free_surface_z(P, xx, yy))
    return 0;
  *x = xx; *y = yy; *z = zz;
  return 1;
}

/* nearest physical boundary of the acrylic union, CLARITAS.py:3034-3063 */
static int acrylic_next_surface(const fparams* P, float x, float y, float z, float vx, float vy,
                                float vz, double* t_out) {
  double best = 1.0e300;
  int surf = 0;
  float Rin = P->rin, Rout = P->rout, Zmin = P->zmin, Zlow = P->zlow, Ztop = P->zwall_top;
  double tol = support_tol(Rin, Rout, Zmin, Zlow, Ztop), r[2];
  cylinder_roots_d(x, y, vx, vy, Rin, &r[0], &r[1]);
  for (int k = 0; k < 2; ++k) {
    double t = r[k];
    if (t >= 1.0e299) continue;
    double zh = (double)z + t * (double)vz;
    if (zh >= (double)Zmin - tol && zh <= (double)Ztop + tol && t > 0.0 && t < best) { best = t; surf = 1; }
  }
  cylinder_roots_d(x, y, vx, vy, Rout, &r[0], &r[1]);
  for (int k = 0; k < 2; ++k) {
    double t = r[k];
    if (t >= 1.0e299) continue;
    double zh = (double)z + t * (double)vz;
    if (zh >= (double)Zlow - tol && zh <= (double)Ztop + tol && t > 0.0 && t < best) { best = t; surf = 2; }
  }
  if (vz < 0.0f) {
    double t = ((double)Zlow - (double)z) / (double)vz;
    if (t > 0.0 && isfinite(t)) {
      double X = (double)x + t * (double)vx, Y = (double)y + t * (double)vy;
      if (hypot(X, Y) <= (double)Rout + tol && t < best) { best = t; surf = 3; }
    }
  }
  if (vz > 0.0f) {
    double t = ((double)Ztop - (double)z) / (double)vz;
    if (t > 0.0 && isfinite(t)) {
      double X = (double)x + t * (double)vx, Y = (double)y + t * (double)vy, rr = hypot(X, Y);
      if (rr >= (double)Rin - tol && rr <= (double)Rout + tol && t < best) { best = t; surf = 4; }
    }
    double ta = ((double)Zmin - (double)z) / (double)vz;
    if (ta > 0.0 && isfinite(ta)) {
      double X = (double)x + ta * (double)vx, Y = (double)y + ta * (double)vy;
      if (hypot(X, Y) < (double)Rin - tol && ta < best) { best = ta; surf = 5; }
    }
  }
  *t_out = best;
  return surf;
}

static int tir_for_ci(double ci, double n1, double n2) {
  if (!(n1 > 0.0) || !(n2 > 0.0) || n1 <= n2) return 0;
  ci = fmin(fmax(ci, 0.0), 1.0);
  return (n1 / n2) * sqrt(fmax(0.0, 1.0 - ci * ci)) >= 1.0;
}
/* conserved-orbit TIR trap classifier, CLARITAS.py:3083-3135 */
static int acrylic_trapped(const fparams* P, float x, float y, float vx, float vy, float vz) {
  double dx = vx, dy = vy, dz = vz, sp2 = dx * dx + dy * dy + dz * dz;
  if (!(sp2 > 0.0) || !isfinite(sp2)) return 0;
  double vxy2 = dx * dx + dy * dy, L = (double)x * dy - (double)y * dx, rs = sqrt(sp2);
  const double reach = 4.0 * (double)EPS_F;
  double na = P->n_acrylic, nw = P->n_water, nair = P->n_air, nb = P->n_bottom_ext;
  if (vxy2 > 0.0) {
    double tg = L / (double)P->rout, rad2 = vxy2 - tg * tg;
    if (rad2 < 0.0 && rad2 > -1.0e-12 * sp2) rad2 = 0.0;
    if (rad2 < 0.0) return 0;
    if (!tir_for_ci(sqrt(rad2) / rs, na, nair)) return 0;
    double impact = fabs(L) / sqrt(vxy2);
    if (impact <= (double)P->rin + reach) {
      double ti = L / (double)P->rin, ri2 = vxy2 - ti * ti;
      if (ri2 < 0.0 && ri2 > -1.0e-12 * sp2) ri2 = 0.0;
      if (ri2 >= 0.0 && !tir_for_ci(sqrt(ri2) / rs, na, nw)) return 0;
    }
    if (impact < (double)P->rin - reach && fabs(dz) > 0.0 && !tir_for_ci(fabs(dz) / rs, na, nw)) return 0;
  }
  if (fabs(dz) > 0.0) {
    double ca = fabs(dz) / rs;
    if (vxy2 <= 1.0e-30 * sp2 && hypot((double)x, (double)y) < (double)P->rin - reach &&
        !tir_for_ci(ca, na, nw))
      return 0;
    if (!tir_for_ci(ca, na, nair)) return 0;
    if (!tir_for_ci(ca, na, nb)) return 0;
  }
  return 1;
}

typedef struct {
  float x, y, z, vx, vy, vz;
  int cir, cor, bir, bor, ctir, btir, tr, ttir;
  uint32_t n_iface;
} photon_t;

/* propagate inside the acrylic union; returns 1 water, 2 external air,
 * 3 bottom external, 4 headspace, 5 trapped, 0 failure (sub set).
 * CLARITAS.py:3140-3198 */
static int acrylic_solid(const fparams* P, rng_t* r, photon_t* p, int* sub) {
  *sub = SUB_NONE;
  int stuck = 0, last = -1;
  for (int it = 0; it < 16384; ++it) {
    double td = 1.0e300;
    int s = acrylic_next_surface(P, p->x, p->y, p->z, p->vx, p->vy, p->vz, &td);
    if (!s || !(td > 0.0) || !isfinite(td) || td >= 1.0e299) { *sub = SUB_ACRYLIC_NO_HIT; return 0; }
    float t = (float)td, px = p->x, py = p->y, pz = p->z, pvx = p->vx, pvy = p->vy, pvz = p->vz;
    p->x += t * p->vx; p->y += t * p->vy; p->z += t * p->vz;
    float a, b, c, ci, R;
    int tir = 0;
    p->n_iface++;
    if (s == 1) {
      float nx, ny;
      radial_normal(p->x, p->y, &nx, &ny);
      float zs = free_surface_z(P, p->x, p->y);
      float n_in = (p->z <= zs + EPS_F) ? P->n_water : P->n_air;
      ci = fminf(fmaxf(-dot3(p->vx, p->vy, p->vz, nx, ny, 0.0f), 0.0f), 1.0f);
      R = fresnel_R(ci, P->n_acrylic, n_in, &tir);
      if (!tir && U(r) >= R && refract3(p->vx, p->vy, p->vz, nx, ny, 0.0f, P->n_acrylic, n_in, &a, &b, &c)) {
        p->vx = a; p->vy = b; p->vz = c; p->x -= EPS_F * nx; p->y -= EPS_F * ny;
        return (n_in == P->n_water) ? 1 : 4;
      }
      reflect3(p->vx, p->vy, p->vz, nx, ny, 0.0f, &a, &b, &c);
      p->vx = a; p->vy = b; p->vz = c; p->cir++; if (tir) p->ctir++;
      p->x += EPS_F * nx; p->y += EPS_F * ny;
    } else if (s == 2) {
      float nx, ny;
      radial_normal(p->x, p->y, &nx, &ny);
      ci = fminf(fmaxf(dot3(p->vx, p->vy, p->vz, nx, ny, 0.0f), 0.0f), 1.0f);
      R = fresnel_R(ci, P->n_acrylic, P->n_air, &tir);
      if (!tir && U(r) >= R && refract3(p->vx, p->vy, p->vz, -nx, -ny, 0.0f, P->n_acrylic, P->n_air, &a, &b, &c)) {
        p->vx = a; p->vy = b; p->vz = c; p->x += EPS_F * nx; p->y += EPS_F * ny;
        return 2;
      }
      reflect3(p->vx, p->vy, p->vz, nx, ny, 0.0f, &a, &b, &c);
      p->vx = a; p->vy = b; p->vz = c; p->cor++; if (tir) p->ctir++;
      p->x -= EPS_F * nx; p->y -= EPS_F * ny;
    } else if (s == 3) {
      p->z = P->zlow;
      ci = fminf(fmaxf(-p->vz, 0.0f), 1.0f);
      R = fresnel_R(ci, P->n_acrylic, P->n_bottom_ext, &tir);
      if (!tir && U(r) >= R && refract3(p->vx, p->vy, p->vz, 0.0f, 0.0f, 1.0f, P->n_acrylic, P->n_bottom_ext, &a, &b, &c)) {
        p->vx = a; p->vy = b; p->vz = c; p->z = P->zlow - EPS_F;
        return 3;
      }
      reflect3(p->vx, p->vy, p->vz, 0.0f, 0.0f, 1.0f, &a, &b, &c);
      p->vx = a; p->vy = b; p->vz = c; p->bor++; if (tir) p->btir++;
      p->z = P->zlow + EPS_F;
    } else if (s == 4) {
      p->z = P->zwall_top;
      ci = fminf(fmaxf(p->vz, 0.0f), 1.0f);
      R = fresnel_R(ci, P->n_acrylic, P->n_air, &tir);
      if (!tir && U(r) >= R && refract3(p->vx, p->vy, p->vz, 0.0f, 0.0f, -1.0f, P->n_acrylic, P->n_air, &a, &b, &c)) {
        p->vx = a; p->vy = b; p->vz = c; p->z = P->zwall_top + EPS_F;
        return 2;
      }
      reflect3(p->vx, p->vy, p->vz, 0.0f, 0.0f, 1.0f, &a, &b, &c);
      p->vx = a; p->vy = b; p->vz = c; p->cor++; if (tir) p->ctir++;
      p->z = P->zwall_top - EPS_F;
    } else {
      p->z = P->zmin;
      ci = fminf(fmaxf(p->vz, 0.0f), 1.0f);
      R = fresnel_R(ci, P->n_acrylic, P->n_water, &tir);
      if (!tir && U(r) >= R && refract3(p->vx, p->vy, p->vz, 0.0f, 0.0f, -1.0f, P->n_acrylic, P->n_water, &a, &b, &c)) {
        p->vx = a; p->vy = b; p->vz = c; p->z = P->zmin + EPS_F;
        return 1;
      }
      reflect3(p->vx, p->vy, p->vz, 0.0f, 0.0f, 1.0f, &a, &b, &c);
      p->vx = a; p->vy = b; p->vz = c; p->bir++; if (tir) p->btir++;
      p->z = P->zmin - EPS_F;
    }
    float pd = fabsf(p->x - px) + fabsf(p->y - py) + fabsf(p->z - pz);
    float vd = fabsf(p->vx - pvx) + fabsf(p->vy - pvy) + fabsf(p->vz - pvz);
    stuck = (s == last && pd <= 1.0e-12f && vd <= 1.0e-7f) ? stuck + 1 : 0;
    if (stuck >= 8) { *sub = SUB_ACRYLIC_BOUNCE; return 0; }
    last = s;
  }
  if (acrylic_trapped(P, p->x, p->y, p->vx, p->vy, p->vz)) return 5;
  *sub = SUB_ACRYLIC_BOUNCE;
  return 0;
}

/* water -> bottom disc, CLARITAS.py:3211-3228 */
static int bottom_disc(const fparams* P, rng_t* r, photon_t* p, int* sub) {
  *sub = SUB_NONE;
  if (bottom_topology_at_top(p->x, p->y, P->rin, P->rout) != 1) { *sub = SUB_BOTTOM_TOPOLOGY; return 0; }
  float a, b, c;
  int tir = 0;
  float ci = fminf(fmaxf(-p->vz, 0.0f), 1.0f);
  float R = fresnel_R(ci, P->n_water, P->n_acrylic, &tir);
  p->n_iface++;
  if (tir || U(r) < R || !refract3(p->vx, p->vy, p->vz, 0.0f, 0.0f, 1.0f, P->n_water, P->n_acrylic, &a, &b, &c)) {
    reflect3(p->vx, p->vy, p->vz, 0.0f, 0.0f, 1.0f, &a, &b, &c);
    p->vx = a; p->vy = b; p->vz = c; p->bir++; if (tir) p->btir++;
    p->z = P->zmin + EPS_F;
    return 1;
  }
  p->vx = a; p->vy = b; p->vz = c; p->z = P->zmin - EPS_F;
  return acrylic_solid(P, r, p, sub);
}

/* headspace inside the open tube; returns 1 water, 2 external, 3 bottom
 * external, 4 open-top escape, 5 trapped, 0 failure.  CLARITAS.py:3235-3293 */
static int headspace(const fparams* P, rng_t* r, photon_t* p, int* sub) {
  *sub = SUB_NONE;
  double tol = support_tol(P->rin, P->rout, P->zmin, P->zlow, P->zwall_top);
  for (int step = 0; step < 16384; ++step) {
    if (!isfinite(p->x) || !isfinite(p->y) || !isfinite(p->z) || !isfinite(p->vx) || !isfinite(p->vy) || !isfinite(p->vz)) {
      *sub = SUB_HEADSPACE_NO_HIT; return 0;
    }
    float zs = free_surface_z(P, p->x, p->y);
    double rr = hypot((double)p->x, (double)p->y);
    if ((double)p->z <= (double)zs - tol && rr < (double)P->rin - tol) return 1;
    if ((double)p->z > (double)P->zwall_top + tol) return 4;
    double tside = bore_wall_hit_d(p->x, p->y, p->z, p->vx, p->vy, p->vz, P->rin, P->zmin, P->zwall_top);
    if (tside < 0.0) { *sub = SUB_WATER_OUTSIDE; return 0; }
    double tsurf = (double)free_surface_hit(P, p->x, p->y, p->z, p->vx, p->vy, p->vz);
    double ttop = 1.0e300;
    if (p->vz > 1.0e-14f) {
      double t = ((double)P->zwall_top - (double)p->z) / (double)p->vz;
      if (t > 0.0) {
        double X = (double)p->x + t * (double)p->vx, Y = (double)p->y + t * (double)p->vy;
        if (hypot(X, Y) < (double)P->rin + tol) ttop = t;
      }
    }
    int which = 0;
    double best = 1.0e300;
    if (tsurf > 0.0 && tsurf < best) { best = tsurf; which = 1; }
    if (tside >= 0.0 && tside < best) { best = tside; which = 2; }
    if (ttop > 0.0 && ttop < best) { best = ttop; which = 3; }
    if (!which || !isfinite(best) || best >= 1.0e299) { *sub = SUB_HEADSPACE_NO_HIT; return 0; }
    float t = (float)best;
    p->x += t * p->vx; p->y += t * p->vy; p->z += t * p->vz;
    if (which == 3) { p->z = P->zwall_top + EPS_F; return 4; }
    float a, b, c;
    int tir = 0;
    p->n_iface++;
    if (which == 1) {
      float snx, sny, snz;
      free_surface_normal(P, p->x, p->y, &snx, &sny, &snz);
      float anx = -snx, any = -sny, anz = -snz;
      float ci = fminf(fmaxf(-dot3(p->vx, p->vy, p->vz, anx, any, anz), 0.0f), 1.0f);
      float R = fresnel_R(ci, P->n_air, P->n_water, &tir);
      if (tir || U(r) < R || !refract3(p->vx, p->vy, p->vz, anx, any, anz, P->n_air, P->n_water, &a, &b, &c)) {
        reflect3(p->vx, p->vy, p->vz, anx, any, anz, &a, &b, &c);
        p->vx = a; p->vy = b; p->vz = c; p->tr++; if (tir) p->ttir++;
        p->x += EPS_F * anx; p->y += EPS_F * any; p->z += EPS_F * anz;
        continue;
      }
      p->vx = a; p->vy = b; p->vz = c;
      if (!water_surface_offset(P, &p->x, &p->y, &p->z, snx, sny, snz)) { *sub = SUB_WATER_OUTSIDE; return 0; }
      return 1;
    }
    float nx, ny;
    radial_normal(p->x, p->y, &nx, &ny);
    float inx = -nx, iny = -ny;
    float ci = fminf(fmaxf(-dot3(p->vx, p->vy, p->vz, inx, iny, 0.0f), 0.0f), 1.0f);
    float R = fresnel_R(ci, P->n_air, P->n_acrylic, &tir);
    if (tir || U(r) < R || !refract3(p->vx, p->vy, p->vz, inx, iny, 0.0f, P->n_air, P->n_acrylic, &a, &b, &c)) {
      reflect3(p->vx, p->vy, p->vz, inx, iny, 0.0f, &a, &b, &c);
      p->vx = a; p->vy = b; p->vz = c; p->cir++; if (tir) p->ctir++;
      p->x -= EPS_F * nx; p->y -= EPS_F * ny;
      continue;
    }
    p->vx = a; p->vy = b; p->vz = c; p->x += EPS_F * nx; p->y += EPS_F * ny;
    int asub = SUB_NONE;
    int where = acrylic_solid(P, r, p, &asub);
    if (where == 4) continue;
    *sub = asub;
    return where;
  }
  *sub = SUB_HEADSPACE_NO_HIT;
  return 0;
}

/* ======================================================================
 * Particle events
 * ====================================================================== */
/* impact-parameter sampling of the struck surface point, CLARITAS.py:2436-2454 */
static void sphere_sample(rng_t* r, float vx, float vy, float vz, float* gnx, float* gny, float* gnz) {
  float rho = sqrtf(fminf(fmaxf(U(r), 0.0f), 0.99999994f));
  float ci = sqrtf(fmaxf(0.0f, 1.0f - rho * rho));
  float phi = U(r) * 2.0f * PI_F;
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(vx, vy, vz, &e1x, &e1y, &e1z, &e2x, &e2y, &e2z);
  float cp = cosf(phi), sp = sinf(phi);
  float qx = cp * e1x + sp * e2x, qy = cp * e1y + sp * e2y, qz = cp * e1z + sp * e2z;
  *gnx = -ci * vx + rho * qx; *gny = -ci * vy + rho * qy; *gnz = -ci * vz + rho * qz;
  normalize3(gnx, gny, gnz);
}
/* CLARITAS.py:2459-2495 */
static int sphere_accessible(const fparams* P, double cx, double cy, double cz, double radius) {
  double rc = hypot(cx, cy), rr = radius;
  double wall = ((double)P->rin - rc) - rr, bottom = cz - rr - (double)P->zmin;
  double rmin = fmax(0.0, rc - rr);
  double top = free_surface_z_d(P, rmin * rmin) - (cz + rr);
  double scale = fmax(fmax(fabs((double)P->rin), fabs(cz)), fmax(rr, 1.0e-6));
  double tol = 128.0 * DEPS * scale;
  return (wall >= -tol && bottom >= -tol && top >= -tol) ? 1 : 0;
}
/* ----------------------------------------------------------------------
 * Rough particle surfaces: Smith microsurface random walk over a Beckmann
 * slope distribution (Heitz, Hanika, d'Eon & Dachsbacher, ACM TOG 35(4),
 * 2016), in the signed, log-domain form of CLARITAS.py:1792-2396.
 * ---------------------------------------------------------------------- */
enum { MS_HIT = 1, MS_ESCAPE = 2, MS_NUMERIC = 3, MS_WATCHDOG = 4 };
/* walk failure subcodes, CLARITAS.py:1810-1839 */
enum {
  MF_PROJ_AREA = 10, MF_CDF = 11, MF_INV_CDF = 12, MF_VISIBILITY = 13, MF_NORMALISATION = 14,
  MF_HEIGHT_SAMPLE = 21, MF_LAMBDA = 23, MF_C0 = 24, MF_LOG_C0 = 25, MF_G1 = 26, MF_ESCAPE_PROB = 27,
  MF_RANDOM = 28, MF_LOG_CNEXT = 29, MF_NORMAL = 30, MF_CNEXT_RANGE = 36, MF_HEIGHT_INV = 37,
  MF_HEIGHT_OUT = 38, MF_WATCHDOG = 40
};
#define MICRO_WATCHDOG_LIMIT 4096
#define SQRT_PI_D 1.77245385090551602729816748334114518
#define INV_2SQRTPI_D 0.28209479177387814347403972578038629

/* inverse error function: Winitzki's closed-form estimate, then Halley
 * steps on erf (|y| <= 1/2) or on erfc with the exact 1 - |y| (|y| > 1/2) */
double orc_erfinv(double y) {
  if (!(y > -1.0 && y < 1.0)) return (y == 1.0) ? INFINITY : (y == -1.0) ? -INFINITY : NAN;
  if (y == 0.0) return 0.0;
  const double s = (y < 0.0) ? -1.0 : 1.0, a = fabs(y), c = 1.0 - a;   /* c exact for a > 1/2 */
  const double l = log(c * (1.0 + a)), k = 0.147, t = 2.0 / (PI_D * k) + 0.5 * l;
  double x = sqrt(sqrt(t * t - l / k) - t);
  for (int it = 0; it < 6; ++it) {
    const double f = (a <= 0.5) ? (erf(x) - a) : (c - erfc(x));
    const double d = 2.0 / SQRT_PI_D * exp(-x * x);           /* d/dx of erf and of -erfc */
    if (!(d > 0.0)) break;
    const double q = f / d, dx = q / (1.0 + x * q);           /* Halley: f'' = -2 x f' */
    x -= dx;
    if (fabs(dx) <= 1e-17 * fabs(x)) break;
  }
  return s * x;
}

/* B(x) = 1 - x sqrt(pi) erfcx(x), x >= 0; asymptotic series for x >= 5,
 * CLARITAS.py:1926-1942 */
static double beck_tail_B(double x) {
  if (!(x >= 0.0) || !isfinite(x)) return -1.0;
  double b;
  if (x < 5.0) {
    b = 1.0 - x * SQRT_PI_D * (exp(x * x) * erfc(x));
  } else {
    const double y = 1.0 / (x * x);
    /* sum_k (-1)^(k+1) (2k-1)!! / 2^k y^k, k = 1..7 */
    b = 0.5 * y - 0.75 * y * y + 1.875 * y * y * y - 6.5625 * y * y * y * y + 29.53125 * y * y * y * y * y -
        162.421875 * y * y * y * y * y * y + 1055.7421875 * y * y * y * y * y * y * y;
  }
  return (isfinite(b) && b > 0.0) ? b : -1.0;
}
/* visible-slope marginal CDF for slope_i >= -1, CLARITAS.py:1947-1962 */
static int beck_visible_cdf(double s, double si, double* out) {
  if (!isfinite(s) || !isfinite(si)) return 0;
  if (s >= si) { *out = 1.0; return 1; }
  const double q = INV_2SQRTPI_D * exp(-si * si) + 0.5 * si * erfc(-si);
  if (!(q > 0.0) || !isfinite(q)) return 0;
  const double n = INV_2SQRTPI_D * exp(-s * s) + 0.5 * si * erfc(-s);
  double f = n / q;
  if (!isfinite(f)) return 0;
  if (f < 0.0 && f > -1.0e-13) f = 0.0;
  if (f > 1.0 && f < 1.0 + 1.0e-13) f = 1.0;
  if (f < 0.0 || f > 1.0) return 0;
  *out = f;
  return 1;
}
/* log of the scaled tail CDF at slope -(x+u), CLARITAS.py:1986-2013 */
static int beck_tail_logcdf(double x, double u, double Bx, double* out) {
  const double y = x + u, By = beck_tail_B(y);
  if (!(By > 0.0) || !isfinite(By)) return 0;
  const double C = u / y + (x / y) * By;
  if (!(C > 0.0) || !isfinite(C)) return 0;
  const double lr = -(2.0 * x * u + u * u) + log(C / Bx);
  if (!isfinite(lr)) return 0;
  *out = lr;
  return 1;
}
/* alpha = 1 visible slope (sx, sy) for signed cot(theta_i) = si,
 * CLARITAS.py:1964-2045 */
static int beck_sample_slopes(double si, double u1, double u2, double* sxo, double* syo, int* fc) {
  *fc = 0;
  if (!isfinite(u1) || !isfinite(u2) || !(u1 > 0.0 && u1 < 1.0) || !(u2 > 0.0 && u2 < 1.0)) { *fc = MF_CDF; return 0; }
  const double COT_1E4 = 9999.999966666666;      /* cot(1e-4 rad): the visible law is the NDF */
  if (isinf(si) || si > COT_1E4) {
    if (si < 0.0) { *fc = MF_PROJ_AREA; return 0; }
    const double r = sqrt(-log(u1)), ph = 2.0 * PI_D * u2, sx = r * cos(ph), sy = r * sin(ph);
    if (!isfinite(sx) || !isfinite(sy)) { *fc = MF_INV_CDF; return 0; }
    *sxo = sx; *syo = sy;
    return 1;
  }
  if (!isfinite(si)) { *fc = MF_PROJ_AREA; return 0; }
  double sx;
  if (si < -1.0) {
    const double x = -si, Bx = beck_tail_B(x);
    if (!(Bx > 0.0) || !isfinite(Bx)) { *fc = MF_PROJ_AREA; return 0; }
    const double target = log(u1);
    double lo = 0.0, hi = fmax(1.0 / x, 1.0e-12), lr;
    int ok = 0;
    for (int it = 0; it < 64; ++it) {
      if (!beck_tail_logcdf(x, hi, Bx, &lr)) { *fc = MF_CDF; return 0; }
      if (lr <= target) { ok = 1; break; }
      hi *= 2.0;
      if (!isfinite(hi)) { *fc = MF_INV_CDF; return 0; }
    }
    if (!ok) { *fc = MF_INV_CDF; return 0; }
    for (int it = 0; it < 52; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (!beck_tail_logcdf(x, mid, Bx, &lr)) { *fc = MF_CDF; return 0; }
      if (lr > target) lo = mid; else hi = mid;
    }
    sx = -(x + 0.5 * (lo + hi));
  } else {
    double lo = fmin(-1.0, si - 1.0), hi = si, f = 1.0;
    int ok = 0;
    for (int it = 0; it < 64; ++it) {
      if (!beck_visible_cdf(lo, si, &f)) { *fc = MF_CDF; return 0; }
      if (f <= u1) { ok = 1; break; }
      lo = (lo < 0.0) ? 2.0 * lo : lo - 1.0;
      if (!isfinite(lo)) { *fc = MF_INV_CDF; return 0; }
    }
    if (!ok) { *fc = MF_INV_CDF; return 0; }
    for (int it = 0; it < 48; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (!beck_visible_cdf(mid, si, &f)) { *fc = MF_CDF; return 0; }
      if (f < u1) lo = mid; else hi = mid;
    }
    sx = 0.5 * (lo + hi);
  }
  const double ey = 2.0 * u2 - 1.0;
  if (!(ey > -1.0 && ey < 1.0) || !isfinite(ey)) { *fc = MF_INV_CDF; return 0; }
  const double sy = orc_erfinv(ey);
  if (!isfinite(sx) || !isfinite(sy)) { *fc = MF_INV_CDF; return 0; }
  *sxo = sx; *syo = sy;
  return 1;
}
/* visible microfacet normal about the TOP axis b for direction w (signed:
 * w.b may be negative), two open uniforms, CLARITAS.py:2073-2129 */
static int beck_vndf_normal(rng_t* r, float alpha, float bx, float by, float bz, float wx, float wy, float wz,
                            double* mx, double* my, double* mz, int* fc) {
  *fc = 0;
  if (!(alpha >= 0.0f) || !isfinite(alpha)) { *fc = MF_NORMALISATION; return 0; }
  normalize3(&bx, &by, &bz);
  normalize3(&wx, &wy, &wz);
  if (!isfinite(bx) || !isfinite(by) || !isfinite(bz) || !isfinite(wx) || !isfinite(wy) || !isfinite(wz)) {
    *fc = MF_NONFINITE_DIRECTION; return 0;
  }
  if (alpha <= 1.0e-8f) {
    const double vd = (double)wx * bx + (double)wy * by + (double)wz * bz;
    if (!(vd > 0.0)) { *fc = MF_VISIBILITY; return 0; }
    *mx = bx; *my = by; *mz = bz;
    return 1;
  }
  float t1x, t1y, t1z, t2x, t2y, t2z;
  perpendicular_basis(bx, by, bz, &t1x, &t1y, &t1z, &t2x, &t2y, &t2z);
  const double lx = (double)wx * t1x + (double)wy * t1y + (double)wz * t1z;
  const double ly = (double)wx * t2x + (double)wy * t2y + (double)wz * t2z;
  const double lz = (double)wx * bx + (double)wy * by + (double)wz * bz;
  double sxv = (double)alpha * lx, syv = (double)alpha * ly, szv = lz;
  const double vl = sqrt(sxv * sxv + syv * syv + szv * szv);
  if (!(vl > 0.0) || !isfinite(vl)) { *fc = MF_NORMALISATION; return 0; }
  sxv /= vl; syv /= vl; szv /= vl;
  const double radial = hypot(sxv, syv);
  double si;
  if (radial == 0.0) {
    if (szv > 0.0) si = INFINITY;
    else { *fc = MF_PROJ_AREA; return 0; }
  } else {
    si = szv / radial;
    if (!isfinite(si)) { *fc = MF_PROJ_AREA; return 0; }
  }
  const double u1 = Uopen(r), u2 = Uopen(r);
  double s0x, s0y;
  int f2 = 0;
  if (!beck_sample_slopes(si, u1, u2, &s0x, &s0y, &f2)) { *fc = f2 ? f2 : MF_INV_CDF; return 0; }
  const double ph = atan2(syv, sxv), cp = cos(ph), sp = sin(ph);
  const double tx = cp * s0x - sp * s0y, ty = sp * s0x + cp * s0y;
  const double sx = (double)alpha * tx, sy = (double)alpha * ty;
  if (!isfinite(sx) || !isfinite(sy)) { *fc = MF_NORMAL; return 0; }
  const double inv = 1.0 / sqrt(sx * sx + sy * sy + 1.0);
  if (!(inv > 0.0) || !isfinite(inv)) { *fc = MF_NORMALISATION; return 0; }
  const double nlx = -sx * inv, nly = -sy * inv, nlz = inv;
  double nx = nlx * t1x + nly * t2x + nlz * bx;
  double ny = nlx * t1y + nly * t2y + nlz * by;
  double nz = nlx * t1z + nly * t2z + nlz * bz;
  const double nn = sqrt(nx * nx + ny * ny + nz * nz);
  if (!(nn > 0.0) || !isfinite(nn)) { *fc = MF_NORMALISATION; return 0; }
  nx /= nn; ny /= nn; nz /= nn;
  if (!(nx * bx + ny * by + nz * bz > 0.0)) { *fc = MF_NORMAL; return 0; }
  if (!(nx * wx + ny * wy + nz * wz > 0.0)) { *fc = MF_VISIBILITY; return 0; }
  *mx = nx; *my = ny; *mz = nz;
  return 1;
}
/* signed Smith Lambda over the whole sphere of directions,
 * CLARITAS.py:2162-2184 */
double orc_smith_lambda(float wx, float wy, float wz, float bx, float by, float bz, float alpha) {
  double c = (double)wx * bx + (double)wy * by + (double)wz * bz;
  if (!isfinite(c) || !isfinite(alpha) || !(alpha >= 0.0f)) return NAN;
  c = fmax(-1.0, fmin(1.0, c));
  const double ss2 = fmax(0.0, 1.0 - c * c);
  if (ss2 == 0.0) return c >= 0.0 ? 0.0 : -1.0;
  if (c == 0.0) return copysign(INFINITY, c);
  if (alpha == 0.0f) return c >= 0.0 ? 0.0 : -1.0;
  const double aa = fabs(c) / ((double)alpha * sqrt(ss2));
  if (!(aa > 0.0) || !isfinite(aa)) return c >= 0.0 ? 0.0 : -1.0;
  const double B = beck_tail_B(aa);
  if (!(B > 0.0) || !isfinite(B)) return NAN;
  const double lp = exp(-(aa * aa) - log(2.0 * SQRT_PI_D * aa)) * B;
  if (!(lp >= 0.0) || !isfinite(lp)) return NAN;
  return c >= 0.0 ? lp : (-1.0 - lp);
}
static double uniform_C1(double h) { return fmin(1.0, fmax(0.0, 0.5 * (h + 1.0))); }
static double uniform_invC1(double u) { return 2.0 * fmin(1.0, fmax(0.0, u)) - 1.0; }
/* next height of the walk for direction w at height h with the open uniform
 * U, CLARITAS.py:2211-2310 */
int orc_smith_height(double U, float wx, float wy, float wz, float bx, float by, float bz, float alpha, float h,
                     float* out_h, int* fc) {
  *fc = 0;
  if (!isfinite(U) || !isfinite(wx) || !isfinite(wy) || !isfinite(wz) || !isfinite(h) || !(alpha >= 0.0f) ||
      !isfinite(alpha) || !(U > 0.0 && U < 1.0)) {
    *fc = MF_RANDOM; return MS_NUMERIC;
  }
  double c = (double)wx * bx + (double)wy * by + (double)wz * bz;
  if (!isfinite(c)) { *fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC; }
  c = fmax(-1.0, fmin(1.0, c));
  const double C0 = uniform_C1((double)h);
  if (!(C0 >= 0.0 && C0 <= 1.0) || !isfinite(C0)) { *fc = MF_C0; return MS_NUMERIC; }
  const double ss2 = fmax(0.0, 1.0 - c * c);
  if (ss2 == 0.0) {
    if (c > 0.0) return MS_ESCAPE;
    if (c < 0.0) {
      const double hn = uniform_invC1(U * C0);
      if (!isfinite(hn)) { *fc = MF_HEIGHT_INV; return MS_NUMERIC; }
      *out_h = (float)hn;
      return MS_HIT;
    }
    *out_h = h;
    return MS_HIT;
  }
  if (c == 0.0) { *out_h = h; return MS_HIT; }
  const double lam = orc_smith_lambda(wx, wy, wz, bx, by, bz, alpha);
  if (isnan(lam)) { *fc = MF_LAMBDA; return MS_NUMERIC; }
  if (c > 0.0) {
    if (!(lam >= 0.0)) { *fc = MF_LAMBDA; return MS_NUMERIC; }
    if (C0 <= 0.0 || C0 >= 1.0 || lam == 0.0) return MS_ESCAPE;
    const double lc0 = log(C0);
    if (!isfinite(lc0)) { *fc = MF_LOG_C0; return MS_NUMERIC; }
    const double lg1 = lam * lc0;
    if (!(lg1 <= 0.0) || !isfinite(lg1)) { *fc = MF_G1; return MS_NUMERIC; }
    const double p_hit = -expm1(lg1);
    if (!(p_hit >= 0.0 && p_hit <= 1.0) || !isfinite(p_hit)) { *fc = MF_ESCAPE_PROB; return MS_NUMERIC; }
    if (U > p_hit) return MS_ESCAPE;
    const double l1u = log1p(-U);
    if (!isfinite(l1u)) { *fc = MF_RANDOM; return MS_NUMERIC; }
    double lcn = lc0 - l1u / lam;
    if (!isfinite(lcn)) { *fc = MF_LOG_CNEXT; return MS_NUMERIC; }
    if (lcn > 64.0 * DEPS) { *fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
    if (lcn > 0.0) lcn = 0.0;
    const double cn = exp(lcn);
    if (!(cn >= 0.0 && cn <= 1.0) || !isfinite(cn)) { *fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
    const double hn = uniform_invC1(cn);
    if (!isfinite(hn)) { *fc = MF_HEIGHT_INV; return MS_NUMERIC; }
    *out_h = (float)hn;
    if (!isfinite(*out_h)) { *fc = MF_HEIGHT_OUT; return MS_NUMERIC; }
    return MS_HIT;
  }
  if (!(lam <= -1.0) || !isfinite(lam)) { *fc = MF_LAMBDA; return MS_NUMERIC; }
  if (C0 <= 0.0) { *out_h = -1.0f; return MS_HIT; }
  const double lc0 = log(C0);
  if (!isfinite(lc0)) { *fc = MF_LOG_C0; return MS_NUMERIC; }
  const double l1u = log1p(-U);
  if (!isfinite(l1u)) { *fc = MF_RANDOM; return MS_NUMERIC; }
  const double lcn = lc0 - l1u / lam;
  if (!isfinite(lcn)) { *fc = MF_LOG_CNEXT; return MS_NUMERIC; }
  if (lcn > 0.0) { *fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
  const double cn = exp(lcn);
  if (!(cn >= 0.0 && cn <= C0 * (1.0 + 8.0e-15)) || !isfinite(cn)) { *fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
  const double hn = uniform_invC1(cn);
  if (!isfinite(hn)) { *fc = MF_HEIGHT_INV; return MS_NUMERIC; }
  *out_h = (float)hn;
  if (!isfinite(*out_h)) { *fc = MF_HEIGHT_OUT; return MS_NUMERIC; }
  return MS_HIT;
}

/* smooth-interface limit of the microsurface walk, CLARITAS.py:2336-2356 */
static int smooth_interface(rng_t* r, float bx, float by, float bz, float n_top, float k_top,
                            float n_bot, float k_bot, float* dx, float* dy, float* dz, int* out_top,
                            int* fc) {
  float ci = fminf(fmaxf(dot3(-(*dx), -(*dy), -(*dz), bx, by, bz), 0.0f), 1.0f);
  float R, ts, tn;
  if (!complex_interface(ci, n_top, k_top, n_bot, k_bot, &R, &ts, &tn) || !isfinite(R)) {
    *fc = MF_FRESNEL_INVALID; return MS_NUMERIC;
  }
  float ox, oy, oz;
  if (U(r) < R) {
    reflect3(*dx, *dy, *dz, bx, by, bz, &ox, &oy, &oz);
    *out_top = 1;
    if (!isfinite(ox) || !isfinite(oy) || !isfinite(oz)) { *fc = MF_REFLECTION_INVALID; return MS_NUMERIC; }
  } else {
    if (!energy_refract3(*dx, *dy, *dz, bx, by, bz, n_top, k_top, n_bot, k_bot, &ox, &oy, &oz)) {
      *fc = MF_REFRACTION_INVALID; return MS_NUMERIC;
    }
    *out_top = 0;
  }
  if (!isfinite(ox) || !isfinite(oy) || !isfinite(oz)) { *fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC; }
  *dx = ox; *dy = oy; *dz = oz;
  return MS_ESCAPE;
}

/* one encounter of direction d with a (rough) dielectric interface whose
 * TOP side has normal b: returns MS_ESCAPE with out_top = 1 (left on the
 * top side) or 0 (crossed), else MS_NUMERIC / MS_WATCHDOG with fc set.
 * CLARITAS.py:2314-2430 */
static int micro_interface_ord(rng_t* r, float alpha, float bx, float by, float bz, float n_top, float k_top,
                               float n_bot, float k_bot, float* dx, float* dy, float* dz, int* out_top, int* fc,
                               int* order);
/* the walk without its scattering order */
static int micro_interface(rng_t* r, float alpha, float bx, float by, float bz, float n_top, float k_top,
                           float n_bot, float k_bot, float* dx, float* dy, float* dz, int* out_top, int* fc) {
  int order = 0;
  return micro_interface_ord(r, alpha, bx, by, bz, n_top, k_top, n_bot, k_bot, dx, dy, dz, out_top, fc, &order);
}
/* order: microfacet scatterings of the walk (the smooth limit counts 1),
 * CLARITAS.py:2328-2373 */
static int micro_interface_ord(rng_t* r, float alpha, float bx, float by, float bz, float n_top, float k_top,
                               float n_bot, float k_bot, float* dx, float* dy, float* dz, int* out_top, int* fc,
                               int* order) {
  *fc = 0;
  *order = 0;
  normalize3(&bx, &by, &bz);
  normalize3(dx, dy, dz);
  if (!isfinite(bx) || !isfinite(by) || !isfinite(bz) || !isfinite(*dx) || !isfinite(*dy) || !isfinite(*dz)) {
    *fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC;
  }
  if (dot3(*dx, *dy, *dz, bx, by, bz) >= -1.0e-7f) { *fc = MF_SIDE_STATE; return MS_NUMERIC; }
  if (alpha <= 1.0e-7f) {
    int st = smooth_interface(r, bx, by, bz, n_top, k_top, n_bot, k_bot, dx, dy, dz, out_top, fc);
    if (st == MS_ESCAPE) *order = 1;
    return st;
  }
  float h = 1.0f + (2.0f * 0.999f - 1.0f);
  int outside = 1;
  for (int it = 0; it < MICRO_WATCHDOG_LIMIT; ++it) {
    float nh = 0.0f;
    int hs, f2 = 0;
    const double Uh = Uopen(r);
    if (outside) {
      hs = orc_smith_height(Uh, *dx, *dy, *dz, bx, by, bz, alpha, h, &nh, &f2);
      if (hs == MS_ESCAPE) { *out_top = 1; return MS_ESCAPE; }
    } else {
      float t = 0.0f;
      hs = orc_smith_height(Uh, -(*dx), -(*dy), -(*dz), bx, by, bz, alpha, -h, &t, &f2);
      if (hs == MS_ESCAPE) { *out_top = 0; return MS_ESCAPE; }
      if (hs == MS_HIT) nh = -t;
    }
    if (hs != MS_HIT || !isfinite(nh)) { *fc = f2 ? f2 : MF_HEIGHT_SAMPLE; return MS_NUMERIC; }
    h = nh;
    (*order)++;
    const float wix = -(*dx), wiy = -(*dy), wiz = -(*dz);
    double mx, my, mz;
    f2 = 0;
    if (outside) {
      if (!beck_vndf_normal(r, alpha, bx, by, bz, wix, wiy, wiz, &mx, &my, &mz, &f2)) {
        *fc = f2 ? f2 : MF_INV_CDF; return MS_NUMERIC;
      }
    } else {
      double tx, ty, tz;
      if (!beck_vndf_normal(r, alpha, bx, by, bz, -wix, -wiy, -wiz, &tx, &ty, &tz, &f2)) {
        *fc = f2 ? f2 : MF_INV_CDF; return MS_NUMERIC;
      }
      mx = -tx; my = -ty; mz = -tz;
    }
    const double vis = (double)wix * mx + (double)wiy * my + (double)wiz * mz;
    if (!isfinite(mx) || !isfinite(my) || !isfinite(mz)) { *fc = MF_NORMAL; return MS_NUMERIC; }
    if (!(vis > 0.0)) { *fc = MF_VISIBILITY; return MS_NUMERIC; }
    const float ci = (float)fmin(1.0, fmax(0.0, vis));
    const float n1 = outside ? n_top : n_bot, k1 = outside ? k_top : k_bot;
    const float n2 = outside ? n_bot : n_top, k2 = outside ? k_bot : k_top;
    if (!(n1 > 0.0f && n2 > 0.0f)) { *fc = MF_REFRACTION_INVALID; return MS_NUMERIC; }
    float R, ts, tn;
    if (!complex_interface(ci, n1, k1, n2, k2, &R, &ts, &tn) || !isfinite(R)) { *fc = MF_FRESNEL_INVALID; return MS_NUMERIC; }
    double ox, oy, oz;
    if (U(r) < R) {
      const double dd = (double)(*dx) * mx + (double)(*dy) * my + (double)(*dz) * mz;
      ox = (double)(*dx) - 2.0 * dd * mx; oy = (double)(*dy) - 2.0 * dd * my; oz = (double)(*dz) - 2.0 * dd * mz;
    } else {
      float tx, ty, tz;
      if (!energy_refract3(*dx, *dy, *dz, (float)mx, (float)my, (float)mz, n1, k1, n2, k2, &tx, &ty, &tz)) {
        *fc = MF_REFRACTION_INVALID; return MS_NUMERIC;
      }
      ox = tx; oy = ty; oz = tz;
      outside = 1 - outside;
    }
    const double qn = sqrt(ox * ox + oy * oy + oz * oz);
    if (!(qn > 0.0) || !isfinite(qn) || !isfinite(ox) || !isfinite(oy) || !isfinite(oz)) {
      *fc = MF_REFRACTION_INVALID; return MS_NUMERIC;
    }
    float fx = (float)(ox / qn), fy = (float)(oy / qn), fz = (float)(oz / qn);
    if (!isfinite(fx) || !isfinite(fy) || !isfinite(fz)) { *fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC; }
    normalize3(&fx, &fy, &fz);
    if (!isfinite(fx) || !isfinite(fy) || !isfinite(fz)) { *fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC; }
    *dx = fx; *dy = fy; *dz = fz;
  }
  *fc = MF_WATCHDOG;
  return MS_WATCHDOG;
}

/* reflection / absorption counts, and the last encounter's record
 * (CLARITAS.py:2503-2599): internal path (float sum of chords), entry
 * incidence and exit angles (deg, -1 undefined), last chord index and the
 * microfacet order of the last walk */
typedef struct {
  int pr, per, pir, pabs;
  uint32_t n_abs;
  float path, inc, ext;
  int macro, micro;
} pcount_t;
static void pc_event_reset(pcount_t* pc) {
  pc->path = 0.0f; pc->inc = -1.0f; pc->ext = -1.0f; pc->macro = 0; pc->micro = 0;
}
static float angle_deg(float c) { return acosf(fminf(fmaxf(c, 0.0f), 1.0f)) * (180.0f / PI_F); }

/* Fresnel sphere (smooth or rough surface): entry, internal chords with
 * Beer-Lambert absorption 4 pi k / lambda, exit or internal reflection.
 * Returns 1 escape, 2 absorbed, 0 failure.  CLARITAS.py:2497-2599 */
static int sphere_interaction(const fparams* P, rng_t* r, float n_m, float n_p, float k_p,
                              float radius, int max_bounces, float gnx, float gny, float gnz,
                              double cx, double cy, double cz, float* x, float* y, float* z,
                              float* vx, float* vy, float* vz, pcount_t* pc, int* pf, int* pfs) {
  *pf = PF_OK; *pfs = SUB_NONE;
  pc_event_reset(pc);
  if (radius <= 0.0f) return 1;
  int top = 1, fc = 0;
  pc->inc = angle_deg(-dot3(*vx, *vy, *vz, gnx, gny, gnz));
  int ms = micro_interface_ord(r, P->rough, gnx, gny, gnz, n_m, 0.0f, n_p, k_p, vx, vy, vz, &top, &fc, &pc->micro);
  if (ms != MS_ESCAPE) { *pf = (ms == MS_WATCHDOG) ? PF_MICRO_WATCHDOG : PF_MICRO_NUMERICAL; *pfs = fc; return 0; }
  if (top) { pc->ext = angle_deg(dot3(*vx, *vy, *vz, gnx, gny, gnz)); pc->pr++; pc->per++; return 1; }
  if (dot3(*vx, *vy, *vz, gnx, gny, gnz) >= 0.0f) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_ENTRY_SIDE; return 0; }
  const double rr = (double)radius, rr2 = rr * rr;
  double qx = rr * (double)gnx, qy = rr * (double)gny, qz = rr * (double)gnz;
  double qn0 = sqrt(qx * qx + qy * qy + qz * qz);
  if (!(qn0 > 0.0) || !isfinite(qn0)) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_CHORD; return 0; }
  qx *= rr / qn0; qy *= rr / qn0; qz *= rr / qn0;
  float alpha_abs = (P->wavelength > 0.0f && k_p > 0.0f) ? (4.0f * PI_F * k_p / P->wavelength) : 0.0f;
  for (int bounce = 0; bounce <= max_bounces; ++bounce) {
    pc->macro = bounce;
    double dvx = *vx, dvy = *vy, dvz = *vz;
    double A = dvx * dvx + dvy * dvy + dvz * dvz, q2 = qx * qx + qy * qy + qz * qz;
    if (!(A > 0.0) || !isfinite(A) || !isfinite(q2)) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_CHORD; return 0; }
    double qnorm = sqrt(q2);
    if (!(qnorm > 0.0) || !isfinite(qnorm)) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_CHORD; return 0; }
    double qp = rr / qnorm;
    qx *= qp; qy *= qp; qz *= qp;
    double B = qx * dvx + qy * dvy + qz * dvz;
    if (!(B < 0.0) || !isfinite(B)) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_INTERNAL_SIDE; return 0; }
    double chord = -2.0 * B / A;
    if (!(chord > 0.0) || !isfinite(chord)) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_CHORD; return 0; }
    double tqx = qx + chord * dvx, tqy = qy + chord * dvy, tqz = qz + chord * dvz;
    double arr = tqx * dvx + tqy * dvy + tqz * dvz, tq2 = tqx * tqx + tqy * tqy + tqz * tqz;
    if (!(arr > 0.0) || !isfinite(arr) || !isfinite(tq2) || fabs(tq2 - rr2) > 512.0 * DEPS * fmax(rr2, 1.0e-30)) {
      *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_CHORD; return 0;
    }
    if (alpha_abs > 0.0f) {
      pc->n_abs++;
      double dist = -log(Uopen(r)) / (double)alpha_abs;
      if (dist < chord) {
        qx += dist * dvx; qy += dist * dvy; qz += dist * dvz;
        *x = (float)(cx + qx); *y = (float)(cy + qy); *z = (float)(cz + qz);
        pc->path += (float)dist;
        pc->pabs++;
        return 2;
      }
    }
    qx = tqx; qy = tqy; qz = tqz;
    pc->path += (float)chord;
    double qn = sqrt(qx * qx + qy * qy + qz * qz);
    if (!(qn > 0.0) || !isfinite(qn)) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_CHORD; return 0; }
    double pj = rr / qn;
    qx *= pj; qy *= pj; qz *= pj;
    *x = (float)(cx + qx); *y = (float)(cy + qy); *z = (float)(cz + qz);
    float nox = (float)(qx / rr), noy = (float)(qy / rr), noz = (float)(qz / rr);
    normalize3(&nox, &noy, &noz);
    top = 1;
    ms = micro_interface_ord(r, P->rough, -nox, -noy, -noz, n_p, k_p, n_m, 0.0f, vx, vy, vz, &top, &fc, &pc->micro);
    if (ms != MS_ESCAPE) { *pf = (ms == MS_WATCHDOG) ? PF_MICRO_WATCHDOG : PF_MICRO_NUMERICAL; *pfs = fc; return 0; }
    if (!top) {
      if (dot3(*vx, *vy, *vz, nox, noy, noz) <= 0.0f) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_EXIT_SIDE; return 0; }
      pc->ext = angle_deg(dot3(*vx, *vy, *vz, nox, noy, noz));
      return 1;
    }
    if (dot3(*vx, *vy, *vz, nox, noy, noz) >= 0.0f) { *pf = PF_MACRO_GEOMETRY; *pfs = SUB_SPHERE_INTERNAL_SIDE; return 0; }
    pc->pr++; pc->pir++;
    if (bounce == max_bounces) { *pf = PF_MACRO_BOUNCE; *pfs = SUB_SPHERE_BOUNCE_LIMIT; return 0; }
  }
  *pf = PF_OTHER; *pfs = SUB_NONE;
  return 0;
}

/* ----------------------------------------------------------------------
 * Incoherent wave channel: direction-only finite-particle event on a
 * volume-preserving spheroid a = r q^(2/3), b = r q^(-1/3) (q = a/b; a
 * sphere when q = 1), CLARITAS.py:2655-2866
 * ---------------------------------------------------------------------- */
/* CLARITAS.py:2660-2665 */
static void random_axis(rng_t* r, float* ux, float* uy, float* uz) {
  const float cz = 2.0f * U(r) - 1.0f, az = 2.0f * PI_F * U(r), sz = sqrtf(fmaxf(0.0f, 1.0f - cz * cz));
  *ux = sz * cosf(az); *uy = sz * sinf(az); *uz = cz;
}
/* isotropic, or the head-tail symmetric von Mises-Fisher law about the
 * local flow direction f (axial density ~ cosh(kappa u.f)),
 * CLARITAS.py:2675-2702 */
static void spheroid_axis(rng_t* r, int model, float kappa, float fx, float fy, float fz, float* ux, float* uy,
                          float* uz) {
  if (model == 0 || !(kappa > 1e-7f) || !isfinite(kappa)) { random_axis(r, ux, uy, uz); return; }
  normalize3(&fx, &fy, &fz);
  if (!(fx * fx + fy * fy + fz * fz > 0.5f)) { random_axis(r, ux, uy, uz); return; }
  const float u = fminf(fmaxf(U(r), 1e-12f), 1.0f - 1e-7f);
  const float em2 = expf(-2.0f * kappa);
  float c = 1.0f + logf(u + (1.0f - u) * em2) / kappa;
  c = fminf(fmaxf(c, -1.0f), 1.0f);
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(fx, fy, fz, &e1x, &e1y, &e1z, &e2x, &e2y, &e2z);
  const float ph = 2.0f * PI_F * U(r), st = sqrtf(fmaxf(0.0f, 1.0f - c * c)), cp = cosf(ph), sp = sinf(ph);
  *ux = c * fx + st * (cp * e1x + sp * e2x);
  *uy = c * fy + st * (cp * e1y + sp * e2y);
  *uz = c * fz + st * (cp * e1z + sp * e2z);
  if (U(r) < 0.5f) { *ux = -*ux; *uy = -*uy; *uz = -*uz; }
  normalize3(ux, uy, uz);
}
/* gradient of F = |p|^2/b^2 + delta (p.u)^2 - 1, CLARITAS.py:2720-2727 */
static void spheroid_normal(float x, float y, float z, float ux, float uy, float uz, float invb2, float delta,
                            float* nx, float* ny, float* nz) {
  const float pu = x * ux + y * uy + z * uz;
  *nx = invb2 * x + delta * pu * ux; *ny = invb2 * y + delta * pu * uy; *nz = invb2 * z + delta * pu * uz;
  normalize3(nx, ny, nz);
}
/* both roots of F(o + t d) = 0, ascending, CLARITAS.py:2729-2743 */
static int spheroid_roots(float ox, float oy, float oz, float dx, float dy, float dz, float ux, float uy, float uz,
                          float invb2, float delta, float* t0, float* t1) {
  const float du = dx * ux + dy * uy + dz * uz, ou = ox * ux + oy * uy + oz * uz;
  const float A = invb2 * (dx * dx + dy * dy + dz * dz) + delta * du * du;
  const float B = 2.0f * (invb2 * (ox * dx + oy * dy + oz * dz) + delta * ou * du);
  const float Cc = invb2 * (ox * ox + oy * oy + oz * oz) + delta * ou * ou - 1.0f;
  const float disc = B * B - 4.0f * A * Cc;
  if (!(A > 0.0f) || disc < 0.0f) return 0;
  const float sd = sqrtf(fmaxf(0.0f, disc));
  float r0 = (-B - sd) / (2.0f * A), r1 = (-B + sd) / (2.0f * A);
  if (r0 > r1) { const float q = r0; r0 = r1; r1 = q; }
  *t0 = r0; *t1 = r1;
  return 1;
}
/* entry point uniform over the projected silhouette (rejection from the
 * bounding disk, 64 attempts), CLARITAS.py:2745-2771 */
static int spheroid_entry(rng_t* r, float vx, float vy, float vz, float a, float b, float ux, float uy, float uz,
                          float* x, float* y, float* z, float* gnx, float* gny, float* gnz) {
  const float invb2 = 1.0f / (b * b), delta = 1.0f / (a * a) - invb2;
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(vx, vy, vz, &e1x, &e1y, &e1z, &e2x, &e2y, &e2z);
  const float du = dot3(ux, uy, uz, vx, vy, vz);
  const float rproj2 = b * b + (a * a - b * b) * fmaxf(0.0f, 1.0f - du * du);
  const float rbound = sqrtf(fmaxf(b * b, rproj2)), rmax = fmaxf(a, b);
  for (int attempt = 0; attempt < 64; ++attempt) {
    const float rr = rbound * sqrtf(U(r)), ph = 2.0f * PI_F * U(r);
    const float ox = rr * (cosf(ph) * e1x + sinf(ph) * e2x) - 2.5f * rmax * vx;
    const float oy = rr * (cosf(ph) * e1y + sinf(ph) * e2y) - 2.5f * rmax * vy;
    const float oz = rr * (cosf(ph) * e1z + sinf(ph) * e2z) - 2.5f * rmax * vz;
    float t0, t1;
    if (!spheroid_roots(ox, oy, oz, vx, vy, vz, ux, uy, uz, invb2, delta, &t0, &t1)) continue;
    const float t = (t0 > 1e-9f * rmax) ? t0 : ((t1 > 1e-9f * rmax) ? t1 : -1.0f);
    if (t <= 0.0f) continue;
    *x = ox + t * vx; *y = oy + t * vy; *z = oz + t * vz;
    spheroid_normal(*x, *y, *z, ux, uy, uz, invb2, delta, gnx, gny, gnz);
    if (dot3(vx, vy, vz, *gnx, *gny, *gnz) < -1e-6f) return 1;
  }
  return 0;
}
/* CLARITAS.py:2773-2781 */
static float spheroid_chord(float x, float y, float z, float vx, float vy, float vz, float ux, float uy, float uz,
                            float invb2, float delta, float scale) {
  float t0, t1;
  if (!spheroid_roots(x, y, z, vx, vy, vz, ux, uy, uz, invb2, delta, &t0, &t1)) return -1.0f;
  const float eps = fmaxf(1e-5f * scale, 1e-12f);
  float t = INF_F;
  if (t0 > eps) t = t0;
  if (t1 > eps && t1 < t) t = t1;
  return t < INF_F * 0.5f ? t : -1.0f;
}
/* direction-only event: 1 scattered (direction updated), 2 absorbed,
 * 0 failure (pf, pfs).  (fx, fy, fz) is the local flow direction of the
 * vMF axis law.  CLARITAS.py:2783-2841 */
static int wave_particle_event(const fparams* P, rng_t* r, float n_m, float n_p, float k_p, float r_eq, float fx,
                               float fy, float fz, float* vx, float* vy, float* vz, pcount_t* pc, int* pf,
                               int* pfs) {
  *pf = PF_OK; *pfs = SUB_NONE;
  pc_event_reset(pc);
  if (r_eq <= 0.0f) return 1;
  const float q = P->aspect;
  if (!(q > 0.0f) || !isfinite(q)) { *pf = PF_SPHEROID; *pfs = SUB_SPHEROID_ASPECT; return 0; }
  if (fabsf(q - 1.0f) < 1e-7f) {
    float gx, gy, gz, lx = 0.0f, ly = 0.0f, lz = 0.0f, lvx = *vx, lvy = *vy, lvz = *vz;
    sphere_sample(r, *vx, *vy, *vz, &gx, &gy, &gz);
    const double cx = -(double)r_eq * (double)gx, cy = -(double)r_eq * (double)gy, cz = -(double)r_eq * (double)gz;
    const int ps = sphere_interaction(P, r, n_m, n_p, k_p, r_eq, P->max_pb, gx, gy, gz, cx, cy, cz, &lx, &ly, &lz,
                                      &lvx, &lvy, &lvz, pc, pf, pfs);
    if (ps == 1) { *vx = lvx; *vy = lvy; *vz = lvz; }
    return ps;
  }
  const float q13 = powf(q, 1.0f / 3.0f), a = r_eq * q13 * q13, b = r_eq / q13;
  if (!(a > 0.0f && b > 0.0f && isfinite(a) && isfinite(b))) { *pf = PF_SPHEROID; *pfs = SUB_SPHEROID_AXES; return 0; }
  float ux, uy, uz;
  spheroid_axis(r, P->orientation, P->kappa, fx, fy, fz, &ux, &uy, &uz);
  float x, y, z, gnx, gny, gnz;
  if (!spheroid_entry(r, *vx, *vy, *vz, a, b, ux, uy, uz, &x, &y, &z, &gnx, &gny, &gnz)) {
    *pf = PF_SPHEROID; *pfs = SUB_SPHEROID_ENTRY; return 0;
  }
  const float invb2 = 1.0f / (b * b), delta = 1.0f / (a * a) - invb2, scale = fmaxf(a, b);
  int top = 1, fc = 0;
  pc->inc = angle_deg(-dot3(*vx, *vy, *vz, gnx, gny, gnz));
  int ms = micro_interface_ord(r, P->rough, gnx, gny, gnz, n_m, 0.0f, n_p, k_p, vx, vy, vz, &top, &fc, &pc->micro);
  if (ms != MS_ESCAPE) { *pf = (ms == MS_WATCHDOG) ? PF_MICRO_WATCHDOG : PF_MICRO_NUMERICAL; *pfs = fc; return 0; }
  if (top) { pc->ext = angle_deg(dot3(*vx, *vy, *vz, gnx, gny, gnz)); pc->pr++; pc->per++; return 1; }
  if (dot3(*vx, *vy, *vz, gnx, gny, gnz) >= 0.0f) { *pf = PF_SPHEROID; *pfs = SUB_SPHEROID_ENTRY; return 0; }
  const float alpha_abs = (P->wavelength > 0.0f && k_p > 0.0f) ? (4.0f * PI_F * k_p / P->wavelength) : 0.0f;
  for (int bounce = 0; bounce <= P->max_pb; ++bounce) {
    pc->macro = bounce;
    const float chord = spheroid_chord(x, y, z, *vx, *vy, *vz, ux, uy, uz, invb2, delta, scale);
    if (chord <= 0.0f) { *pf = PF_SPHEROID; *pfs = SUB_SPHEROID_CHORD; return 0; }
    if (alpha_abs > 0.0f) {
      pc->n_abs++;
      const double dist = -log(Uopen(r)) / (double)alpha_abs;
      if (dist < (double)chord) { pc->path += (float)dist; pc->pabs++; return 2; }
    }
    x += chord * (*vx); y += chord * (*vy); z += chord * (*vz);
    pc->path += chord;
    float nox, noy, noz;
    spheroid_normal(x, y, z, ux, uy, uz, invb2, delta, &nox, &noy, &noz);
    top = 1;
    ms = micro_interface_ord(r, P->rough, -nox, -noy, -noz, n_p, k_p, n_m, 0.0f, vx, vy, vz, &top, &fc, &pc->micro);
    if (ms != MS_ESCAPE) { *pf = (ms == MS_WATCHDOG) ? PF_MICRO_WATCHDOG : PF_MICRO_NUMERICAL; *pfs = fc; return 0; }
    if (!top) {
      if (dot3(*vx, *vy, *vz, nox, noy, noz) <= 0.0f) { *pf = PF_SPHEROID; *pfs = SUB_SPHEROID_EXIT_SIDE; return 0; }
      pc->ext = angle_deg(dot3(*vx, *vy, *vz, nox, noy, noz));
      return 1;
    }
    if (dot3(*vx, *vy, *vz, nox, noy, noz) >= 0.0f) { *pf = PF_SPHEROID; *pfs = SUB_SPHEROID_INTERNAL_SIDE; return 0; }
    pc->pr++; pc->pir++;
    if (bounce == P->max_pb) { *pf = PF_MACRO_BOUNCE; *pfs = SUB_SPHEROID_BOUNCE_LIMIT; return 0; }
  }
  *pf = PF_OTHER; *pfs = SUB_NONE;
  return 0;
}

/* inverse-CDF polar angle with linear interpolation on the float table,
 * CLARITAS.py:2632-2641 */
static float sample_mie_theta(rng_t* r, const orc_model* m, int pidx) {
  int nt = m->nt_sample;
  const float* cdf = m->cdf_f + (size_t)pidx * nt;
  const float* th = m->theta_f;
  float u = U(r);
  int lo = 0, hi = nt - 1;
  while (lo < hi) { int mid = (lo + hi) >> 1; if (u <= cdf[mid]) hi = mid; else lo = mid + 1; }
  int j = lo;
  if (j <= 0) return th[0];
  float c0 = cdf[j - 1], c1 = cdf[j], t0 = th[j - 1], t1 = th[j];
  float f = (c1 > c0) ? fminf(fmaxf((u - c0) / (c1 - c0), 0.0f), 1.0f) : 0.0f;
  return t0 + f * (t1 - t0);
}
/* CLARITAS.py:2643-2652 */
static void mie_scatter(rng_t* r, const orc_model* m, int pidx, float* vx, float* vy, float* vz) {
  float th = sample_mie_theta(r, m, pidx);
  float phi = 2.0f * PI_F * U(r);
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(*vx, *vy, *vz, &e1x, &e1y, &e1z, &e2x, &e2y, &e2z);
  float st = sinf(th), ct = cosf(th), cp = cosf(phi), sp = sinf(phi);
  float ox = ct * (*vx) + st * (cp * e1x + sp * e2x);
  float oy = ct * (*vy) + st * (cp * e1y + sp * e2y);
  float oz = ct * (*vz) + st * (cp * e1z + sp * e2z);
  *vx = ox; *vy = oy; *vz = oz;
  normalize3(vx, vy, vz);
}

typedef struct {
  float x, y, z, vx, vy, vz;
  int status, channel, failure, subcode;
  pcount_t pc;
  uint32_t trials, rejects, inaccessible;
} cond_t;

/* conditional finite-particle closure: the outcome (absorbed or not) is
 * drawn from the Mie albedo first; channel candidates are retried until
 * one realises it.  CLARITAS.py:4479-4536 */
static cond_t conditional_event(const orc_model* m, rng_t* r, int pidx, float x, float y, float z,
                                float vx, float vy, float vz, int max_trials) {
  const fparams* P = &m->fp;
  cond_t h;
  memset(&h, 0, sizeof(h));
  h.failure = PF_OTHER; h.subcode = 9002;
  double qs = m->qsca[pidx], qe = m->qext[pidx];
  float ki = m->k_f[pidx], nr = m->n_f[pidx], radius = m->r_f[pidx];
  double tol = 1e-10 * fmax(1.0, fmax(fabs(qs), fabs(qe)));
  if (max_trials == 0) { h.subcode = 9001; return h; }
  if (!isfinite(qs) || !isfinite(qe) || qe <= 0.0 || qs < -tol || qs > qe + tol || max_trials < 1) return h;
  int lossless = (ki == 0.0f && fabs(qe - qs) <= tol);
  double albedo = lossless ? 1.0 : fmin(fmax(qs / qe, 0.0), 1.0);
  double pabs = lossless ? 0.0 : fmin(fmax((qe - qs) / qe, 0.0), 1.0);
  if ((ki > 0.0f && (pabs <= 0.0 || albedo >= 1.0)) || (albedo > 0.0 && albedo < 2.3283064365386962890625e-10) ||
      (pabs > 0.0 && pabs < 2.3283064365386962890625e-10)) {
    h.subcode = 9003; return h;
  }
  int target_abs = (albedo < 1.0 && U53(r) >= albedo);
  double wave = qs > 0.0 ? fmax(qs - 1.0, 0.0) / qs : 0.0;
  float coherent = P->coherent;
  pcount_t acc;
  memset(&acc, 0, sizeof(acc));
  uint32_t trials = 0, rejects = 0, inacc = 0;
  for (int trial = 0; trial < max_trials; ++trial) {
    trials++;
    cond_t c;
    memset(&c, 0, sizeof(c));
    pc_event_reset(&c.pc);
    c.x = x; c.y = y; c.z = z; c.vx = vx; c.vy = vy; c.vz = vz;
    c.channel = (U53(r) < wave) ? ((coherent >= 1.0f || U53(r) < coherent) ? 1 : 2) : 0;
    if (c.channel == 1) {
      mie_scatter(r, m, pidx, &c.vx, &c.vy, &c.vz);
      c.status = 1;
    } else if (c.channel == 2) {
      float fr = sqrtf(x * x + y * y), ffx = 1.0f, ffy = 0.0f;
      if (fr > 1e-12f) { ffx = -y / fr; ffy = x / fr; }
      int pf = 0, pfs = 0;
      c.status = wave_particle_event(P, r, P->n_water, nr, ki, radius, ffx, ffy, 0.0f, &c.vx, &c.vy, &c.vz, &c.pc,
                                     &pf, &pfs);
      c.failure = pf; c.subcode = pfs;
    } else {
      float gx, gy, gz;
      sphere_sample(r, vx, vy, vz, &gx, &gy, &gz);
      double dcx = (double)x - (double)radius * gx, dcy = (double)y - (double)radius * gy,
             dcz = (double)z - (double)radius * gz;
      if (!sphere_accessible(P, dcx, dcy, dcz, radius)) { inacc++; continue; }
      int pf = 0, pfs = 0;
      c.status = sphere_interaction(P, r, P->n_water, nr, ki, radius, m->fp.max_pb, gx, gy, gz, dcx, dcy,
                                    dcz, &c.x, &c.y, &c.z, &c.vx, &c.vy, &c.vz, &c.pc, &pf, &pfs);
      c.failure = pf; c.subcode = pfs;
    }
    acc.n_abs += c.pc.n_abs;
    if (c.status != 1 && c.status != 2) {
      c.trials = trials; c.rejects = rejects; c.inaccessible = inacc; c.pc.n_abs = acc.n_abs;
      return c;
    }
    if ((c.status == 2) != target_abs) { rejects++; continue; }
    c.trials = trials; c.rejects = rejects; c.inaccessible = inacc; c.pc.n_abs = acc.n_abs;
    c.failure = PF_OK; c.subcode = 0;
    return c;
  }
  memset(&h, 0, sizeof(h));
  h.failure = PF_OTHER; h.subcode = 9001;
  h.trials = trials; h.rejects = rejects; h.inaccessible = inacc; h.pc.n_abs = acc.n_abs;
  return h;
}

/* ======================================================================
 * Detectors, CLARITAS.py:4237-4255, 4404-4435
 * ====================================================================== */
static int bore_accept(const fparams* P, float x, float y, float z, float vx, float vy, float vz,
                       float ux, float uy, float* t_out) {
  float den = vx * ux + vy * uy;
  if (den <= 1e-10f) return 0;
  float pdot = x * ux + y * uy;
  float t0 = (P->ring_in - pdot) / den, t1 = (P->throat_out - pdot) / den, t2 = (P->ring_out - pdot) / den;
  if (!(t0 > 0.0f && t1 > t0 && t2 > t1)) return 0;
  float rth2 = P->throat_r * P->throat_r, rcb2 = P->cb_r * P->cb_r;
  float q0x = x + t0 * vx - P->ring_in * ux, q0y = y + t0 * vy - P->ring_in * uy, q0z = z + t0 * vz;
  float q1x = x + t1 * vx - P->throat_out * ux, q1y = y + t1 * vy - P->throat_out * uy, q1z = z + t1 * vz;
  float q2x = x + t2 * vx - P->ring_out * ux, q2y = y + t2 * vy - P->ring_out * uy, q2z = z + t2 * vz;
  if (q0x * q0x + q0y * q0y + q0z * q0z <= rth2 && q1x * q1x + q1y * q1y + q1z * q1z <= rth2 &&
      q2x * q2x + q2y * q2y + q2z * q2z <= rcb2) {
    *t_out = t0;
    return 1;
  }
  return 0;
}
static int hardware_detector(const orc_model* m, float x, float y, float z, float vx, float vy,
                             float vz, int mirror) {
  if (mirror) { x = -x; vx = -vx; }
  float best_t = INF_F, t;
  int best = -1;
  for (int j = 0; j < CLR_N_CHANNELS; ++j)
    if (bore_accept(&m->fp, x, y, z, vx, vy, vz, m->det_ux[j], m->det_uy[j], &t) && t < best_t) { best_t = t; best = j; }
  return best;
}

/* ======================================================================
 * Reflecting collimator bores, ring-face cavity and the SFH213 photodiode,
 * CLARITAS.py:1263-1374 and 4283-4376.  Random numbers continue the
 * history's stream: source collimator at the launch; at scoring native then
 * mirror direct acceptance, then native and mirror ring-face paths.
 * ====================================================================== */
/* CLARITAS.py:1266-1279 */
static float bore_side_hit(float q, float w, float vq, float vw, float radius) {
  float A = vq * vq + vw * vw;
  if (A <= 1.0e-20f) return INF_F;
  float B = 2.0f * (q * vq + w * vw), Cc = q * q + w * w - radius * radius, D = B * B - 4.0f * A * Cc;
  if (D < 0.0f) return INF_F;
  float root = sqrtf(fmaxf(D, 0.0f));
  float t1 = (-B - root) / (2.0f * A), t2 = (-B + root) / (2.0f * A), t = INF_F;
  if (t1 > EPS_F) t = t1;
  if (t2 > EPS_F && t2 < t) t = t2;
  return t;
}

/* bore_wall_reflect + trace_bore_segment, CLARITAS.py:1281-1318 */
static int bore_section(const fparams* P, rng_t* r, float s_end, float radius, float* s, float* q, float* w,
                        float* vs, float* vq, float* vw, int* inter, int* refl) {
  for (int b = 0; b <= P->max_bb; ++b) {
    if (*vs <= 1.0e-12f) return 0;
    float tp = (s_end - *s) / (*vs);
    if (tp < -EPS_F) return 0;
    if (tp < 0.0f) tp = 0.0f;
    float twall = bore_side_hit(*q, *w, *vq, *vw, radius);
    if (tp <= twall) {
      *q += tp * (*vq); *w += tp * (*vw); *s = s_end;
      return (*q) * (*q) + (*w) * (*w) <= radius * radius * (1.0f + 2.0e-5f);
    }
    if (twall >= INF_F * 0.5f) return 0;
    *q += twall * (*vq); *w += twall * (*vw); *s += twall * (*vs);
    (*inter)++;
    float rr = sqrtf(fmaxf((*q) * (*q) + (*w) * (*w), 1.0e-30f));
    float nq = -(*q) / rr, nw = -(*w) / rr;
    float ci = fminf(fmaxf(-((*vs) * 0.0f + (*vq) * nq + (*vw) * nw), 0.0f), 1.0f);
    int tir = 0;
    float R = fresnel_R(ci, P->n_air, P->n_bore, &tir);
    if (U(r) >= R) return 0;
    float os, oq, ow;
    reflect3(*vs, *vq, *vw, 0.0f, nq, nw, &os, &oq, &ow);
    *vs = os; *vq = oq; *vw = ow;
    (*refl)++;
    *q -= EPS_F * (*q) / rr; *w -= EPS_F * (*w) / rr;
  }
  return 0;
}

/* CLARITAS.py:1321-1345 */
static int source_collimator(const fparams* P, rng_t* r, float* x, float* y, float* z, float* vx, float* vy,
                             float* vz, int* inter, int* refl) {
  if (*vy <= 1.0e-12f) return 0;
  float s = *y, q = *x, w = *z, vs = *vy, vq = *vx, vw = *vz;
  if (q * q + w * w > P->cb_r * P->cb_r) return 0;
  float shoulder = -P->throat_out;
  if (s < shoulder - EPS_F) {
    if (!bore_section(P, r, shoulder, P->cb_r, &s, &q, &w, &vs, &vq, &vw, inter, refl)) return 0;
  } else if (s > shoulder + EPS_F) {
    return 0;
  }
  if (q * q + w * w > P->throat_r * P->throat_r) return 0;
  if (!bore_section(P, r, -P->ring_in, P->throat_r, &s, &q, &w, &vs, &vq, &vw, inter, refl)) return 0;
  *x = q; *y = s; *z = w; *vx = vq; *vy = vs; *vz = vw;
  return 1;
}

/* CLARITAS.py:1347-1371 */
static int detector_collimator(const orc_model* m, rng_t* r, float x, float y, float z, float vx, float vy,
                               float vz, int mirror, int* refl_out) {
  const fparams* P = &m->fp;
  if (mirror) { x = -x; vx = -vx; }
  float best_t = INF_F;
  int best = -1;
  for (int j = 0; j < CLR_N_CHANNELS; ++j) {
    float ux = m->det_ux[j], uy = m->det_uy[j], den = vx * ux + vy * uy;
    if (den <= 1.0e-10f) continue;
    float t = (P->ring_in - (x * ux + y * uy)) / den;
    if (t <= 0.0f) continue;
    float hx = x + t * vx, hy = y + t * vy, hz = z + t * vz;
    float q = hx * uy - hy * ux, w = hz;
    if (q * q + w * w <= P->throat_r * P->throat_r && t < best_t) { best_t = t; best = j; }
  }
  if (best < 0) return -1;
  float ux = m->det_ux[best], uy = m->det_uy[best];
  float hx = x + best_t * vx, hy = y + best_t * vy;
  float s = P->ring_in, q = hx * uy - hy * ux, w = z + best_t * vz;
  float vs = vx * ux + vy * uy, vq = vx * uy - vy * ux, vw = vz;
  int inter = 0, refl = 0;
  if (!bore_section(P, r, P->throat_out, P->throat_r, &s, &q, &w, &vs, &vq, &vw, &inter, &refl)) return -1;
  if (!bore_section(P, r, P->ring_out, P->cb_r, &s, &q, &w, &vs, &vq, &vw, &inter, &refl)) return -1;
  *refl_out = refl;
  return best;
}

/* CLARITAS.py:4304-4316 (the source throat faces -y) */
static int throat_opening(const orc_model* m, float x, float y, float z, float vx, float vy, float vz) {
  const fparams* P = &m->fp;
  float best_t = INF_F, r2 = P->throat_r * P->throat_r;
  int best = -1;
  for (int j = 0; j <= CLR_N_CHANNELS; ++j) {
    float ux = (j < CLR_N_CHANNELS) ? m->det_ux[j] : 0.0f, uy = (j < CLR_N_CHANNELS) ? m->det_uy[j] : -1.0f;
    float den = vx * ux + vy * uy;
    if (den <= 1.0e-10f) continue;
    float t = (P->ring_in - (x * ux + y * uy)) / den;
    if (t <= 0.0f) continue;
    float qx = x + t * vx - P->ring_in * ux, qy = y + t * vy - P->ring_in * uy, qz = z + t * vz;
    if (qx * qx + qy * qy + qz * qz <= r2 && t < best_t) { best_t = t; best = j; }
  }
  return best;
}

/* cylinder_hit of the reference: nearest x-y circle hit beyond EPS_F */
static float circle_hit(float x, float y, float vx, float vy, float R) {
  float a = vx * vx + vy * vy;
  if (a < 1e-18f) return INF_F;
  float b = x * vx + y * vy, c = x * x + y * y - R * R, d = b * b - a * c;
  if (d < 0.0f) return INF_F;
  float sd = sqrtf(fmaxf(d, 0.0f)), t1 = (-b - sd) / a, t2 = (-b + sd) / a, t = INF_F;
  if (t1 > EPS_F) t = t1;
  if (t2 > EPS_F && t2 < t) t = t2;
  return t;
}

/* CLARITAS.py:4318-4328 */
static int detector_from_air(const orc_model* m, rng_t* r, float x, float y, float z, float vx, float vy, float vz,
                             int* bounces) {
  *bounces = 0;
  if (m->fp.det_bore == 0) return hardware_detector(m, x, y, z, vx, vy, vz, 0);
  return detector_collimator(m, r, x, y, z, vx, vy, vz, 0, bounces);
}

/* CLARITAS.py:4330-4376 */
static int ring_face_path(const orc_model* m, rng_t* r, float x, float y, float z, float vx, float vy, float vz,
                          int mirror, int* bounces, int* face_refl) {
  const fparams* P = &m->fp;
  if (mirror) { x = -x; vx = -vx; }
  *bounces = 0; *face_refl = 0;
  if (P->face_model == 0) return -1;
  if (throat_opening(m, x, y, z, vx, vy, vz) >= 0) return -1;
  for (int b = 0; b < P->max_fb; ++b) {
    float tr = circle_hit(x, y, vx, vy, P->ring_in);
    if (tr >= INF_F * 0.5f) return -1;
    x += tr * vx; y += tr * vy; z += tr * vz;
    float nx, ny, a, c, d;
    radial_normal(x, y, &nx, &ny);
    float ci = fminf(fmaxf(vx * nx + vy * ny, 0.0f), 1.0f);
    int tir = 0;
    float R = fresnel_R(ci, P->n_air, P->n_face, &tir);
    if (U(r) >= R) return -1;
    reflect3(vx, vy, vz, -nx, -ny, 0.0f, &a, &c, &d);
    vx = a; vy = c; vz = d;
    (*face_refl)++;
    x -= EPS_F * nx; y -= EPS_F * ny;
    float ta = circle_hit(x, y, vx, vy, P->rout);
    if (ta >= INF_F * 0.5f) return -1;
    x += ta * vx; y += ta * vy; z += ta * vz;
    radial_normal(x, y, &nx, &ny);
    ci = fminf(fmaxf(-dot3(vx, vy, vz, nx, ny, 0.0f), 0.0f), 1.0f);
    R = fresnel_R(ci, P->n_air, P->n_acrylic, &tir);
    if (U(r) >= R) return -1;
    reflect3(vx, vy, vz, nx, ny, 0.0f, &a, &c, &d);
    vx = a; vy = c; vz = d;
    x += EPS_F * nx; y += EPS_F * ny;
    int db = 0;
    int det = detector_from_air(m, r, x, y, z, vx, vy, vz, &db);
    if (det >= 0) { *bounces = db; return det; }
    if (throat_opening(m, x, y, z, vx, vy, vz) >= 0) return -1;
  }
  return -1;
}

/* bore_section on explicit rays (s, q, w, vs, vq, vw), ray i with stream i of
 * `seed`, walls of cfg->bore_n, cfg->max_bore_bounces (test probe) */
int orc_bore_probe(const clr_config* cfg, int n, uint64_t seed, const float* rays6, float s_end, float radius,
                   int* pass, int* inter, int* refl) {
  fparams P;
  memset(&P, 0, sizeof(P));
  P.n_air = (float)cfg->n_air; P.n_bore = (float)cfg->bore_n; P.max_bb = cfg->max_bore_bounces;
  for (int i = 0; i < n; ++i) {
    rng_t r;
    rng_seed(&r, seed, (uint64_t)i);
    const float* q = rays6 + 6 * (size_t)i;
    float s = q[0], qq = q[1], w = q[2], vs = q[3], vq = q[4], vw = q[5];
    inter[i] = 0; refl[i] = 0;
    pass[i] = bore_section(&P, &r, s_end, radius, &s, &qq, &w, &vs, &vq, &vw, &inter[i], &refl[i]);
  }
  return 0;
}

/* CLARITAS.py:4283-4290 */
static float sfh213_response(float phi_deg) {
  float a = fabsf(phi_deg);
  if (a >= 90.0f) return 0.0f;
  return 1.0f / (1.0f + powf(a / 10.0f, 3.169925001442312f));
}
float orc_sfh213(float phi_deg) { return sfh213_response(phi_deg); }

/* ======================================================================
 * One history, CLARITAS.py:4606-5068
 * ====================================================================== */
typedef struct {
  int st, pf, pfs, entered;
  int ic, gic, wic, wcoh;
  pcount_t pc;
  photon_t ph;
  uint32_t n_flight, n_null, trials, rejects, inacc, steps;
  uint32_t draws;
  /* conditional-law sampler outcomes, CLARITAS.py:4491-4536, 4854-4933:
     accepted geometric / coherent / morphology, accepted absorbing /
     scattering, numerical failures, trial exhaustions */
  uint32_t cs[7];
  int native, mirror;
  int aux_native[CLR_MAX_AUX], aux_mirror[CLR_MAX_AUX];
  int auxw_native[CLR_MAX_AUX], auxw_mirror[CLR_MAX_AUX];
  int legacy_native[CLR_N_CHANNELS], legacy_mirror[CLR_N_CHANNELS];
  float inc_native, inc_mirror;
  int sbi, sbr, nb, mb, nfb, mfb;  /* collimator / ring-face interactions and reflections */
  float pd_native, pd_mirror;      /* SFH213 weights of the accepted hits */
  /* causal path provenance (CLARITAS.py:4629-4647), observational */
  struct {
    float wp, first_defl, last_defl, ipath, iopl, wp_last, cum, cum2, maxd;
    float site_max[3], site_first[3], site_last[3];
    float sum_r, sum_z, min_r, max_r, min_z, max_z, wp_first;
    float m_path, m_pod, m_inc, m_exit;
    int ret, after_ret, first_type, last_type, last_order, first_back, first_hemi;
    int bins[6];  /* first / last particle, geometric, wave bin */
    int btrans, etrans, nsmall, nmed, nlarge, prev_branch, prev_type;
    int m_type, m_order, m_bin, m_macro, m_micro;
  } pv;
} hist_t;

/* a closure call's outcome in the sampler counts (every trial's own
   outcome is in h->trials / rejects / inaccessible) */
static void count_closure(hist_t* o, const cond_t* h) {
  if (h->status == 1 || h->status == 2) {
    o->cs[h->channel]++;
    o->cs[h->status == 2 ? 3 : 4]++;
  } else {
    o->cs[h->subcode == 9001 ? 6 : 5]++;
  }
}

static void prov_reset(hist_t* o) {
  memset(&o->pv, 0, sizeof(o->pv));
  o->pv.first_defl = o->pv.last_defl = o->pv.wp_last = o->pv.maxd = o->pv.wp_first = -1.0f;
  o->pv.min_r = o->pv.min_z = INF_F;
  o->pv.max_r = o->pv.max_z = -INF_F;
  o->pv.m_inc = o->pv.m_exit = -1.0f;
  for (int k = 0; k < 6; ++k) o->pv.bins[k] = -1;
  o->pv.m_bin = -1;
}

/* a surviving particle event of code `type` on bin `pidx` at site s, incoming
 * direction vin, outgoing direction p->v; e holds the encounter's record.
 * CLARITAS.py:4936-4980 */
static void prov_event(const orc_model* m, hist_t* o, const float vin[3], const float s[3], int type, int pidx,
                       const pcount_t* e) {
  const photon_t* p = &o->ph;
  float c = dot3(vin[0], vin[1], vin[2], p->vx, p->vy, p->vz);
  c = fminf(fmaxf(c, -1.0f), 1.0f);
  const float defl = acosf(c) * (180.0f / PI_F);
  const float sr = sqrtf(fmaxf(s[0] * s[0] + s[1] * s[1], 0.0f));
  o->pv.ipath += e->path;
  if (pidx >= 0) o->pv.iopl += e->path * m->n_f[pidx];
  o->pv.cum += defl;
  o->pv.cum2 += defl * defl;
  if (defl > o->pv.maxd) {
    o->pv.maxd = defl;
    memcpy(o->pv.site_max, s, sizeof(float) * 3);
    o->pv.m_type = type; o->pv.m_order = o->ic; o->pv.m_bin = pidx;
    o->pv.m_macro = e->macro; o->pv.m_micro = e->micro; o->pv.m_path = e->path;
    o->pv.m_pod = (pidx >= 0 && m->r_f[pidx] > 0.0f) ? e->path / (2.0f * m->r_f[pidx]) : 0.0f;
    o->pv.m_inc = e->inc; o->pv.m_exit = e->ext;
  }
  if (defl < 30.0f) o->pv.nsmall++;
  else if (defl < 90.0f) o->pv.nmed++;
  else o->pv.nlarge++;
  int branch = (type >= 1 && type <= 3) ? 1 : 2;
  if (o->pv.prev_branch > 0 && branch != o->pv.prev_branch) o->pv.btrans++;
  if (o->pv.prev_type > 0 && type != o->pv.prev_type) o->pv.etrans++;
  o->pv.prev_branch = branch; o->pv.prev_type = type;
  o->pv.sum_r += sr; o->pv.sum_z += s[2];
  o->pv.min_r = fminf(o->pv.min_r, sr); o->pv.max_r = fmaxf(o->pv.max_r, sr);
  o->pv.min_z = fminf(o->pv.min_z, s[2]); o->pv.max_z = fmaxf(o->pv.max_z, s[2]);
  if (o->ic == 1) { memcpy(o->pv.site_first, s, sizeof(float) * 3); o->pv.wp_first = o->pv.wp; }
  memcpy(o->pv.site_last, s, sizeof(float) * 3);
  o->pv.wp_last = o->pv.wp;
  if (pidx >= 0) {
    int geo = (type >= 1 && type <= 3), wav = (type >= 4 && type <= 7);
    if (o->pv.bins[0] < 0) o->pv.bins[0] = pidx;
    o->pv.bins[1] = pidx;
    if (geo) { if (o->pv.bins[2] < 0) o->pv.bins[2] = pidx; o->pv.bins[3] = pidx; }
    if (wav) { if (o->pv.bins[4] < 0) o->pv.bins[4] = pidx; o->pv.bins[5] = pidx; }
    if (o->pv.ret > 0) o->pv.after_ret++;
  }
  o->pv.last_type = type; o->pv.last_order = o->ic; o->pv.last_defl = defl;
  if (o->ic == 1) {
    o->pv.first_type = type; o->pv.first_defl = defl;
    o->pv.first_back = (c < 0.0f); o->pv.first_hemi = (p->vy < 0.0f);
  }
}

/* the record of a history, with the write-out sentinels of
 * CLARITAS.py:5076-5103 */
static void prov_record(const hist_t* o, uint64_t hist, clr_causal_record* q) {
  const int any = o->ic > 0;
  const photon_t* p = &o->ph;
  memset(q, 0, sizeof(*q));
  q->history = hist; q->native_id = o->native; q->mirror_id = o->mirror;
  q->ic = o->ic; q->gic = o->gic; q->wic = o->wic; q->wcoh = o->wcoh;
  q->per = o->pc.per; q->pir = o->pc.pir; q->pabs = o->pc.pabs > 0;
  q->acrylic_reflections = p->cir + p->cor + p->bir + p->bor;
  q->acrylic_returns = o->pv.ret; q->events_after_return = o->pv.after_ret;
  q->first_type = o->pv.first_type; q->last_type = o->pv.last_type; q->last_order = o->pv.last_order;
  q->first_backscatter = o->pv.first_back; q->first_source_hemisphere = o->pv.first_hemi;
  q->first_bin = o->pv.bins[0]; q->last_bin = o->pv.bins[1];
  q->first_geometric_bin = o->pv.bins[2]; q->last_geometric_bin = o->pv.bins[3];
  q->first_wave_bin = o->pv.bins[4]; q->last_wave_bin = o->pv.bins[5];
  q->branch_transitions = o->pv.btrans; q->type_transitions = o->pv.etrans;
  q->small_deflections = o->pv.nsmall; q->medium_deflections = o->pv.nmed; q->large_deflections = o->pv.nlarge;
  q->max_type = any ? o->pv.m_type : 0; q->max_order = any ? o->pv.m_order : 0;
  q->max_bin = any ? o->pv.m_bin : -1;
  q->max_macro_bounces = any ? o->pv.m_macro : 0; q->max_micro_order = any ? o->pv.m_micro : 0;
  q->first_deflection = o->pv.first_defl; q->last_deflection = o->pv.last_defl;
  q->internal_path = o->pv.ipath; q->internal_opl = o->pv.iopl;
  q->water_path_after_last =
      (o->pv.last_order > 0 && o->pv.wp_last >= 0.0f) ? fmaxf(o->pv.wp - o->pv.wp_last, 0.0f) : -1.0f;
  q->cumulative_deflection = any ? o->pv.cum : -1.0f;
  q->cumulative_deflection_sq = any ? o->pv.cum2 : -1.0f;
  q->max_deflection = any ? o->pv.maxd : -1.0f;
  if (any) {
    q->turn_x = o->pv.site_max[0]; q->turn_y = o->pv.site_max[1]; q->turn_z = o->pv.site_max[2];
    q->first_x = o->pv.site_first[0]; q->first_y = o->pv.site_first[1]; q->first_z = o->pv.site_first[2];
    q->last_x = o->pv.site_last[0]; q->last_y = o->pv.site_last[1]; q->last_z = o->pv.site_last[2];
  }
  q->mean_r = any ? o->pv.sum_r / (float)o->ic : -1.0f;
  q->mean_z = any ? o->pv.sum_z / (float)o->ic : -1.0f;
  q->min_r = any ? o->pv.min_r : -1.0f; q->max_r = any ? o->pv.max_r : -1.0f;
  q->min_z = any ? o->pv.min_z : -1.0f; q->max_z = any ? o->pv.max_z : -1.0f;
  q->water_path_first = any ? o->pv.wp_first : -1.0f; q->water_path_last = any ? o->pv.wp_last : -1.0f;
  q->max_internal_path = any ? o->pv.m_path : 0.0f; q->max_path_over_diameter = any ? o->pv.m_pod : 0.0f;
  q->max_incidence = any ? o->pv.m_inc : -1.0f; q->max_exit = any ? o->pv.m_exit : -1.0f;
}

/* ======================================================================
 * Light-transport voxel diagnostics, CLARITAS.py:3312-3467 (observational:
 * no random numbers, no change to the photon).  Histories run in parallel,
 * so grid and total updates are atomic; a history's trajectory slots are
 * written only by its own thread.
 * ====================================================================== */
typedef struct {
  const clr_light_config* c;
  clr_light_out* o;
  float rin, zmin, zmax;
} light_t;

static void light_count_u32(uint32_t* p, uint32_t v) {
#pragma omp atomic
  *p += v;
}

/* clip [t0, t1] of p + t d to [lo, hi], CLARITAS.py:3312-3316 */
static int light_clip(double p, double d, double lo, double hi, double* t0, double* t1) {
  if (fabs(d) < 1.0e-30f) return p >= lo && p <= hi;
  double a = (lo - p) / d, b = (hi - p) / d;
  if (a > b) { double q = a; a = b; b = q; }
  *t0 = fmax(*t0, a);
  *t1 = fmin(*t1, b);
  return *t1 >= *t0;
}

static double light_face(double lo, int i, int st, double h, double a, double e) {
  return (st == 0) ? (double)INF_F : (lo + (double)(st > 0 ? i + 1 : i) * h - a) / e;
}

/* exact track-length deposition by 3-D voxel traversal, CLARITAS.py:3318-3400 */
static void light_segment(const light_t* L, uint64_t gid, float x0, float y0, float z0, float vx, float vy, float vz,
                          float dist, int order) {
  const clr_light_config* c = L->c;
  clr_light_out* o = L->o;
  if (gid >= c->max_rays || !(dist > 0.0f) || !isfinite(dist)) return;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const float Rin = L->rin, zmin = L->zmin, zmax = L->zmax;
  if (nx <= 0 || ny <= 0 || nz <= 0 || !(zmax > zmin)) { light_count_u32(&o->error_count, 1); return; }
  double dd = (double)dist, sx = (double)vx * dd, sy = (double)vy * dd, sz = (double)vz * dd;
  double c0 = 0.0, c1 = 1.0;
  if (!light_clip(x0, sx, -Rin, Rin, &c0, &c1) || !light_clip(y0, sy, -Rin, Rin, &c0, &c1) ||
      !light_clip(z0, sz, zmin, zmax, &c0, &c1))
    return;
  double ax = x0 + sx * c0, ay = y0 + sy * c0, az = z0 + sz * c0;
  double bx = x0 + sx * c1, by = y0 + sy * c1, bz = z0 + sz * c1;
  double clipped = dist * (c1 - c0);
  if (!(clipped > 0.0f)) return;
#pragma omp atomic
  o->segment_total += clipped;
#pragma omp atomic
  o->segment_count += 1;
  if (gid < (uint64_t)c->representative_rays && c->max_segments > 0) {
    int g = (int)gid, slot = o->traj_count ? o->traj_count[g] : 0;
    if (slot < c->max_segments) {
      size_t q = (size_t)g * c->max_segments + slot;
      if (o->traj_x0) { o->traj_x0[q] = (float)ax; o->traj_y0[q] = (float)ay; o->traj_z0[q] = (float)az; }
      if (o->traj_x1) { o->traj_x1[q] = (float)bx; o->traj_y1[q] = (float)by; o->traj_z1[q] = (float)bz; }
      if (o->traj_order) o->traj_order[q] = order;
      if (o->traj_event) o->traj_event[q] = 0;
      if (o->traj_count) o->traj_count[g] = slot + 1;
    } else if (o->traj_overflow) {
      o->traj_overflow[g] = 1;
    }
  }
  double xlo = -(double)Rin, ylo = -(double)Rin, zlo = (double)zmin;
  double hx = (2.0 * (double)Rin) / nx, hy = (2.0 * (double)Rin) / ny, hz = ((double)zmax - zlo) / nz;
  int ix = (int)floor((ax - xlo) / hx), iy = (int)floor((ay - ylo) / hy), iz = (int)floor((az - zlo) / hz);
  if (ix < 0) ix = 0;
  if (ix >= nx) ix = nx - 1;
  if (iy < 0) iy = 0;
  if (iy >= ny) iy = ny - 1;
  if (iz < 0) iz = 0;
  if (iz >= nz) iz = nz - 1;
  double ex = bx - ax, ey = by - ay, ez = bz - az;
  int stx = (ex > 0.0f) ? 1 : ((ex < 0.0f) ? -1 : 0);
  int sty = (ey > 0.0f) ? 1 : ((ey < 0.0f) ? -1 : 0);
  int stz = (ez > 0.0f) ? 1 : ((ez < 0.0f) ? -1 : 0);
  double tmx = light_face(xlo, ix, stx, hx, ax, ex), tmy = light_face(ylo, iy, sty, hy, ay, ey);
  double tmz = light_face(zlo, iz, stz, hz, az, ez);
  if (tmx < 0.0f) tmx = 0.0f;
  if (tmy < 0.0f) tmy = 0.0f;
  if (tmz < 0.0f) tmz = 0.0f;
  double u = 0.0, local = 0.0;
  const float tol = 2.0e-6f;
  int max_steps = nx + ny + nz + 8;
  for (int k = 0; k < max_steps && u < 1.0f; ++k) {
    if (ix < 0 || ix >= nx || iy < 0 || iy >= ny || iz < 0 || iz >= nz) {
      if (u >= 1.0f - tol) { light_count_u32(&o->boundary_adjustment_count, 1); u = 1.0f; break; }
      light_count_u32(&o->error_count, 1);
      break;
    }
    double un = fmin(1.0, fmin(tmx, fmin(tmy, tmz)));
    if (un < 1.0f && un >= 1.0f - tol) { un = 1.0f; light_count_u32(&o->boundary_adjustment_count, 1); }
    if (un < u) {
      if (u - un <= tol) { un = u; light_count_u32(&o->boundary_adjustment_count, 1); }
      else { light_count_u32(&o->error_count, 1); break; }
    }
    double len = (un - u) * clipped;
    if (len > 0.0f) {
      size_t q = ((size_t)iz * ny + iy) * nx + ix;
      float lf = (float)len;
      if (o->track_length) {
#pragma omp atomic
        o->track_length[q] += lf;
      }
      if (order > 0 && o->post_scatter_track_length) {
#pragma omp atomic
        o->post_scatter_track_length[q] += lf;
      }
      local += len;
    }
    if (un >= 1.0f) { u = 1.0f; break; }
    double eps = 8.0 * 2.2204460492503131e-16 * fmax(1.0, fabs(un));
    if (tmx <= un + eps) { ix += stx; tmx = light_face(xlo, ix, stx, hx, ax, ex); }
    if (tmy <= un + eps) { iy += sty; tmy = light_face(ylo, iy, sty, hy, ay, ey); }
    if (tmz <= un + eps) { iz += stz; tmz = light_face(zlo, iz, stz, hz, az, ez); }
    u = un;
  }
  if (u < 1.0f - tol) light_count_u32(&o->error_count, 1);
#pragma omp atomic
  o->deposited_total += local;
}

/* CLARITAS.py:3402-3407 */
/* water-path heat maps: a sample every 10 um along the segment (at most
 * 20000) into the x-y and z-y projections, z up to zvis.
 * CLARITAS.py:3296-3309 */
static void heat_raster(const light_t* L, float zvis, float x0, float y0, float z0, float vx, float vy, float vz,
                        float dist) {
  const int vis = L->c->heatmap_size;
  if (vis <= 0 || !(dist > 0.0f) || !L->o->heatmap_xy || !L->o->heatmap_zy) return;
  const float spacing = 1.0e-5f, Rin = L->rin, zmin = L->zmin;
  int n = (int)ceilf(dist / spacing);
  n = n < 1 ? 1 : (n > 20000 ? 20000 : n);
  for (int k = 1; k <= n; ++k) {
    float f = (float)k / (float)n;
    float x = x0 + vx * dist * f, y = y0 + vy * dist * f, z = z0 + vz * dist * f;
    int ix = (int)(((x + Rin) / (2.0f * Rin)) * vis);
    int iy = vis - 1 - (int)(((y + Rin) / (2.0f * Rin)) * vis);
    int iz = (int)(((z - zmin) / (zvis - zmin)) * vis);
#ifdef _OPENMP
#pragma omp atomic
#endif
    L->o->heatmap_xy[(ix >= 0 && ix < vis && iy >= 0 && iy < vis) ? iy * vis + ix : 0] +=
        (ix >= 0 && ix < vis && iy >= 0 && iy < vis) ? 1.0f : 0.0f;
#ifdef _OPENMP
#pragma omp atomic
#endif
    L->o->heatmap_zy[(iz >= 0 && iz < vis && iy >= 0 && iy < vis) ? iy * vis + iz : 0] +=
        (iz >= 0 && iz < vis && iy >= 0 && iy < vis) ? 1.0f : 0.0f;
  }
}

static void light_mark(const light_t* L, uint64_t gid, int code) {
  const clr_light_config* c = L->c;
  if (gid >= c->max_rays || gid >= (uint64_t)c->representative_rays || c->max_segments <= 0) return;
  if (!L->o->traj_count || !L->o->traj_event) return;
  int g = (int)gid, n = L->o->traj_count[g];
  if (n > 0 && n <= c->max_segments) L->o->traj_event[(size_t)g * c->max_segments + n - 1] = code;
}

/* one event site into a count grid (and the weight grid when given),
 * CLARITAS.py:3409-3467 */
static void light_site(const light_t* L, uint64_t gid, float x, float y, float z, float* weight, uint32_t* count,
                       uint64_t* total) {
  const clr_light_config* c = L->c;
  clr_light_out* o = L->o;
  if (gid >= c->max_rays) return;
#pragma omp atomic
  *total += 1;
  int nx = c->nx, ny = c->ny, nz = c->nz;
  float Rin = L->rin, zmin = L->zmin, zmax = L->zmax;
  float ctol = 2.0e-6f * fmaxf(2.0f * Rin, zmax - zmin);
  if (!isfinite(x) || !isfinite(y) || !isfinite(z) || x < -Rin - ctol || x > Rin + ctol || y < -Rin - ctol ||
      y > Rin + ctol || z < zmin - ctol || z > zmax + ctol) {
#pragma omp critical(light_outside)
    {
      if (o->outside[0]++ == 0) {
        memcpy(&o->outside[1], &x, 4); memcpy(&o->outside[2], &y, 4); memcpy(&o->outside[3], &z, 4);
        o->outside[4] = (uint32_t)gid; memcpy(&o->outside[5], &ctol, 4); o->outside[6] = 1;
      }
    }
    light_count_u32(&o->error_count, 1);
    return;
  }
  int ix = (int)floorf(((x + Rin) / (2.0f * Rin)) * nx);
  int iy = (int)floorf(((y + Rin) / (2.0f * Rin)) * ny);
  int iz = (int)floorf(((z - zmin) / (zmax - zmin)) * nz);
  int snapped = 0;
  if (ix < 0) { ix = 0; snapped = 1; } else if (ix >= nx) { ix = nx - 1; snapped = 1; }
  if (iy < 0) { iy = 0; snapped = 1; } else if (iy >= ny) { iy = ny - 1; snapped = 1; }
  if (iz < 0) { iz = 0; snapped = 1; } else if (iz >= nz) { iz = nz - 1; snapped = 1; }
  if (snapped) light_count_u32(&o->boundary_adjustment_count, 1);
  size_t q = ((size_t)iz * ny + iy) * nx + ix;
  if (weight) {
#pragma omp atomic
    weight[q] += 1.0f;
  }
  if (count) light_count_u32(&count[q], 1);
}
static void light_event(const light_t* L, uint64_t gid, float x, float y, float z, int code) {
  light_site(L, gid, x, y, z, L->o->scatter_event_weight, L->o->scatter_event_count, &L->o->scatter_event_total);
  light_mark(L, gid, code);
}
static void light_absorb(const light_t* L, uint64_t gid, float x, float y, float z) {
  light_site(L, gid, x, y, z, NULL, L->o->absorption_event_count, &L->o->absorption_event_total);
  light_mark(L, gid, 9);
}
/* Fresnel outcome code of a finite-particle event: 1 external reflection,
 * 3 internal reflection, 2 transmission (+4 for the morphology channel) */
static int light_code(int per, int pir, int morph) { return (per ? 1 : (pir ? 3 : 2)) + (morph ? 4 : 0); }

/* ----------------------------------------------------------------------
 * Woodcock tracking with a majorant that is constant on z slabs: the
 * device's tracking of shared FP32 fields (no reference counterpart; the
 * reference tracks with one majorant, CLARITAS.py:4757-4799).  The optical
 * depth -log u is spent slab by slab (the exponential is memoryless at the
 * slab faces) and a collision is kept with probability mu_local / mu_slab.
 * ---------------------------------------------------------------------- */
static int slab_of(const fparams* P, float z) {
  int j = (int)floorf((z - P->slab_z0) * P->slab_inv_w);
  return j < 0 ? 0 : (j > P->n_slabs - 1 ? P->n_slabs - 1 : j);
}
static float slab_free_path(const fparams* P, float z, float vz, float u, float limit) {
  float tau = -logf(u), t = 0.0f;
  /* no slab majorant exceeds slab_mu_max: no collision before limit */
  if (!(tau < P->slab_mu_max * limit)) return INF_F;
  int j = slab_of(P, z);
  const int step = (vz > 0.0f) ? 1 : -1;
  const float ivz = 1.0f / vz;
  for (int n = 0; n < P->n_slabs; ++n) {
    float t_exit = INF_F;
    if (vz > 1e-14f && j < P->n_slabs - 1) t_exit = (P->slab_z0 + (float)(j + 1) * P->slab_w - z) * ivz;
    else if (vz < -1e-14f && j > 0) t_exit = (P->slab_z0 + (float)j * P->slab_w - z) * ivz;
    t_exit = fmaxf(t_exit, t);
    const float mu = P->slab_mu[j];
    if (mu > 0.0f && tau < mu * (t_exit - t)) return t + tau / mu;
    if (t_exit >= limit) return INF_F;
    tau -= mu * (t_exit - t);
    t = t_exit;
    j += step;
  }
  return INF_F;
}

/* source position on the launch plane y = -source_r and direction before
 * the collimator, CLARITAS.py:4650-4680 (the collimated disk has no
 * reference counterpart) */
static void source_sample(const fparams* P, rng_t* rp, float* xo, float* yo, float* zo, float* vxo, float* vyo,
                          float* vzo) {
  rng_t r = *rp;
  float x, y, z, vx = 0.0f, vy = 1.0f, vz = 0.0f;
  if (P->src_model == CLR_SOURCE_COLLIMATED_DISK) {
    float u1 = U(&r), u2 = U(&r);
    float rho = P->beam_r * sqrtf(u1), ang = 2.0f * PI_F * u2;
    x = rho * cosf(ang);
    z = rho * sinf(ang);
  } else {
    float u1 = fmaxf(U(&r), 1e-12f), u2 = U(&r);
    float mag = sqrtf(-2.0f * logf(u1));
    x = P->sigma * mag * cosf(2.0f * PI_F * u2);
    z = P->sigma * mag * sinf(2.0f * PI_F * u2);
  }
  y = -P->source_r;
  if (P->src_model == CLR_SOURCE_LEGACY_POLAR || P->src_model == CLR_SOURCE_BETA_RADIANCE) {
    float polar = (P->src_model == 0) ? beta_draw(&r, P->alpha1, P->alpha2) * (PI_F / 2.0f)
                                      : beta_radiance_polar(&r, P->alpha1, P->alpha2);
    float az = 2.0f * PI_F * U(&r), sp = sinf(polar);
    vx = sp * cosf(az); vy = cosf(polar); vz = sp * sinf(az);
  } else if (P->src_model == CLR_SOURCE_THESIS_PLANAR) {
    float h = beta_draw(&r, P->alpha1, P->alpha2) * (PI_F / 2.0f);
    if (U(&r) < 0.5f) h = -h;
    vx = sinf(h); vy = cosf(h); vz = 0.0f;
  }
  normalize3(&vx, &vy, &vz);
  *rp = r;
  *xo = x; *yo = y; *zo = z; *vxo = vx; *vyo = vy; *vzo = vz;
}

static void trace_history(const orc_model* m, uint64_t seed, uint64_t hist, hist_t* o, const light_t* Ld) {
  const fparams* P = &m->fp;
  rng_t r;
  rng_seed(&r, seed, hist);
  memset(o, 0, sizeof(*o));
  prov_reset(o);
  o->native = o->mirror = -1;
  for (int a = 0; a < CLR_MAX_AUX; ++a) o->aux_native[a] = o->aux_mirror[a] = -1;
  photon_t* p = &o->ph;
  int st = 0, pf = PF_OK, pfs = SUB_NONE;
  /* ---- launch, CLARITAS.py:4650-4697 ---- */
  float x, y, z, vx, vy, vz;
  source_sample(P, &r, &x, &y, &z, &vx, &vy, &vz);
  p->x = x; p->y = y; p->z = z; p->vx = vx; p->vy = vy; p->vz = vz;
  if (vy <= 1e-12f) { st = 3; goto WRITE_OUT; }
  if (P->src_bore) {
    /* reflecting source collimator, CLARITAS.py:4693-4697 */
    if (!source_collimator(P, &r, &x, &y, &z, &vx, &vy, &vz, &o->sbi, &o->sbr)) { st = 3; goto WRITE_OUT; }
  } else {
    /* absorbing bore: hard apertures, CLARITAS.py:4681-4692 */
    if (x * x + z * z > P->cb_r * P->cb_r) { st = 3; goto WRITE_OUT; }
    float t = (-P->throat_out - y) / vy;
    if (t < -EPS_F) { st = 3; goto WRITE_OUT; }
    if (t < 0.0f) t = 0.0f;
    float xt = x + t * vx, zt = z + t * vz;
    if (xt * xt + zt * zt > P->throat_r * P->throat_r) { st = 3; goto WRITE_OUT; }
    t = (-P->ring_in - y) / vy;
    if (t <= 0.0f) { st = 3; goto WRITE_OUT; }
    x += t * vx; y += t * vy; z += t * vz;
    if (x * x + z * z > P->throat_r * P->throat_r) { st = 3; goto WRITE_OUT; }
  }
  /* outer acrylic wall, CLARITAS.py:4700-4735 */
  {
    double td = finite_cylinder_forward_d(x, y, z, vx, vy, vz, P->rout, P->zlow, P->zwall_top);
    float t = (td < 0.5 * (double)INF_F) ? (float)td : INF_F;
    if (t >= INF_F * 0.5f) { st = 3; goto WRITE_OUT; }
    x += t * vx; y += t * vy; z += t * vz;
    float nx, ny, a, b, c;
    radial_normal(x, y, &nx, &ny);
    float ci = fminf(fmaxf(-dot3(vx, vy, vz, nx, ny, 0.0f), 0.0f), 1.0f);
    int tir = 0;
    float R = fresnel_R(ci, P->n_air, P->n_acrylic, &tir);
    p->n_iface++;
    p->x = x; p->y = y; p->z = z;
    if (tir || U(&r) < R || !refract3(vx, vy, vz, nx, ny, 0.0f, P->n_air, P->n_acrylic, &a, &b, &c)) {
      reflect3(vx, vy, vz, nx, ny, 0.0f, &a, &b, &c);
      p->vx = a; p->vy = b; p->vz = c; p->cor++; if (tir) p->ctir++;
      p->x = x + EPS_F * nx; p->y = y + EPS_F * ny;
      st = 1;
      goto AIR_EXIT;
    }
    p->vx = a; p->vy = b; p->vz = c; p->x = x - EPS_F * nx; p->y = y - EPS_F * ny;
    int sub = 0;
    int where = acrylic_solid(P, &r, p, &sub);
    if (where == 2) { st = 1; goto AIR_EXIT; }
    if (where == 3) { st = 7; goto WRITE_OUT; }
    if (where == 5) { st = 8; goto WRITE_OUT; }
    if (where == 4) {
      int hsub = 0, hw = headspace(P, &r, p, &hsub);
      if (hw == 1) o->entered = 1;
      else if (hw == 2) { st = 1; goto AIR_EXIT; }
      else if (hw == 3) { st = 7; goto WRITE_OUT; }
      else if (hw == 4) { st = 6; goto WRITE_OUT; }
      else if (hw == 5) { st = 8; goto WRITE_OUT; }
      else { pf = PF_HEADSPACE_NO_HIT; pfs = hsub; st = 4; goto WRITE_OUT; }
    } else if (where != 1) {
      pf = (sub == SUB_ACRYLIC_BOUNCE) ? PF_ACRYLIC_BOUNCE : PF_ACRYLIC_NO_HIT; pfs = sub; st = 4;
      goto WRITE_OUT;
    }
    o->entered = 1;
  }
  /* ---- water loop, CLARITAS.py:4734-5066 ---- */
  {
    const float MU_TOTAL = P->mu_geom + P->mu_wave;
    const float MU_TRACK = (P->sed_model == CLR_SED_UNIFORM) ? MU_TOTAL : (P->maj_geom + P->maj_wave);
    const double ri2w = (double)P->rin * (double)P->rin, wtol = 8.0 * FEPS * ri2w;
    int step;
    for (step = 0; step < P->max_iter; ++step) {
      o->steps++;
      if (!isfinite(p->x) || !isfinite(p->y) || !isfinite(p->z)) { pf = PF_NONFINITE_POSITION; pfs = SUB_NONFINITE_POSITION; st = 4; break; }
      if (!isfinite(p->vx) || !isfinite(p->vy) || !isfinite(p->vz)) { pf = PF_NONFINITE_DIRECTION; pfs = SUB_NONFINITE_DIRECTION; st = 4; break; }
      float zsh = free_surface_z(P, p->x, p->y);
      if (p->z <= P->zmin || p->z >= zsh) { st = 2; break; }
      double rr2 = (double)p->x * (double)p->x + (double)p->y * (double)p->y;
      if (rr2 > ri2w + wtol) { pf = PF_INVALID_MEDIUM; pfs = SUB_WATER_OUTSIDE; st = 4; break; }
      double tsd = bore_wall_hit_d(p->x, p->y, p->z, p->vx, p->vy, p->vz, P->rin, P->zmin, P->zwall_top);
      if (tsd < 0.0) { pf = PF_INVALID_MEDIUM; pfs = SUB_WATER_OUTSIDE; st = 4; break; }
      float tside = (tsd < 0.5 * (double)INF_F) ? (float)tsd : INF_F;
      float ttop = INF_F, tbot = INF_F;
      if (p->vz < -1e-14f) tbot = (P->zmin - p->z) / p->vz;
      {
        /* exact skip: the segment up to the side/bottom stays below the
         * lowest point of the surface, so no surface hit can come first */
        float tc = fminf(tside, (tbot > 0.0f) ? tbot : INF_F);
        int skip = 0;
        if (P->fs_model != 0 && tc < 0.5f * INF_F) {
          float zend = p->z + tc * p->vz;
          skip = (fmaxf(p->z, zend) < P->zs_min);
        }
        if (!skip) ttop = free_surface_hit(P, p->x, p->y, p->z, p->vx, p->vy, p->vz);
      }
      if (!(ttop > 0.0f)) ttop = INF_F;
      if (!(tbot > 0.0f)) tbot = INF_F;
      int axial_top = (ttop <= tbot && ttop < tside);
      int axial_bot = (tbot < ttop && tbot < tside);
      float tbound = axial_top ? ttop : (axial_bot ? tbot : tside);
      if (tbound >= INF_F * 0.5f) { pf = PF_WATER_NO_BOUNDARY; pfs = SUB_WATER_NO_BOUNDARY; st = 4; break; }
      float freep = INF_F;
      if (tbound > 0.0f && MU_TRACK > 0.0f) {
        float u = fmaxf(U(&r), 1e-12f);
        freep = (P->n_slabs > 0) ? slab_free_path(P, p->z, p->vz, u, tbound) : -logf(u) / MU_TRACK;
      }
      int do_particle = freep < tbound;
      float travel = do_particle ? freep : tbound;
      if (Ld) {
        if (Ld->c->heatmap_size > 0)
          heat_raster(Ld, free_surface_z(P, P->rin, 0.0f), p->x, p->y, p->z, p->vx, p->vy, p->vz, travel);
        light_segment(Ld, hist, p->x, p->y, p->z, p->vx, p->vy, p->vz, travel, o->ic);
      }
      o->n_flight++;
      p->x += travel * p->vx; p->y += travel * p->vy; p->z += travel * p->vz;
      o->pv.wp += travel;
      if (do_particle) {
        double er2 = (double)p->x * p->x + (double)p->y * p->y;
        if (er2 > ri2w + wtol || p->z < P->zmin - EPS_F || p->z > free_surface_z(P, p->x, p->y) + EPS_F) {
          pf = PF_INVALID_MEDIUM; pfs = SUB_WATER_OUTSIDE; st = 4; break;
        }
        float LG = P->mu_geom, LW = P->mu_wave, LT = MU_TOTAL;
        if (P->sed_model != CLR_SED_UNIFORM) {
          if (P->sed_model == CLR_SED_RZ_FIELD || P->sed_model == CLR_SED_XYZ_FIELD) {
            float s = shared_field_scale(m, p->x, p->y, p->z);
            LG = m->sum_g_f * s; LW = m->sum_w_f * s;
          } else {
            double lg = 0.0, lw = 0.0;
            for (int i = 0; i < m->n; ++i) {
              double sc = sediment_local_scale(m, i, p->x, p->y, p->z);
              lg += m->trace_g[i] * sc; lw += m->trace_w[i] * sc;
            }
            LG = (float)lg; LW = (float)lw;
          }
          LT = LG + LW;
          if (!(LT > 0.0f)) { o->n_null++; continue; }
          const float MU_LOC = (P->n_slabs > 0) ? P->slab_mu[slab_of(P, p->z)] : MU_TRACK;
          float slack = 64.0f * 1.1920928955078125e-7f * fmaxf(fabsf(MU_LOC), 1.0f);
          if (LT > MU_LOC + slack) { pf = PF_MAJORANT; pfs = SUB_MAJORANT; st = 4; break; }
          float pacc = LT / fmaxf(MU_LOC, 1e-30f);
          if (pacc > 1.0f) pacc = 1.0f;
          if (U(&r) >= pacc) { o->n_null++; continue; }
        }
        o->ic++;
        const float sx0 = p->x, sy0 = p->y, sz0 = p->z;  /* event site */
        const float site[3] = {sx0, sy0, sz0}, vin[3] = {p->vx, p->vy, p->vz};
        pcount_t none;
        memset(&none, 0, sizeof(none));
        pc_event_reset(&none);
        int use_wave = (LW > 0.0f && U(&r) * LT >= LG);
        int em = P->event_model;
        if (em == CLR_EVENT_PLAIN_MIE) {
          int b = select_bin(&r, m, 0, p->x, p->y, p->z, LG);
          o->pc.n_abs++;
          if (U(&r) >= (float)m->albedo[b]) {
            o->pc.pabs++;
            if (Ld) light_absorb(Ld, hist, sx0, sy0, sz0);
            st = 5;
            break;
          }
          o->wic++; o->wcoh++;
          mie_scatter(&r, m, b, &p->vx, &p->vy, &p->vz);
          if (Ld) light_event(Ld, hist, sx0, sy0, sz0, 4);
          prov_event(m, o, vin, site, 4, b, &none);
          continue;
        }
        int cond = (em == CLR_EVENT_HYBRID_EXACT_RATE);
        int sel = -1;
        if (cond) sel = select_bin(&r, m, use_wave, p->x, p->y, p->z, use_wave ? LW : LG);
        if (cond && m->ki[sel] > 0.0) {
          cond_t h = conditional_event(m, &r, sel, p->x, p->y, p->z, p->vx, p->vy, p->vz, P->cond_trials);
          o->trials += h.trials; o->rejects += h.rejects; o->inacc += h.inaccessible; o->pc.n_abs += h.pc.n_abs;
          count_closure(o, &h);
          if (h.status != 1 && h.status != 2) { pf = h.failure ? h.failure : PF_OTHER; pfs = h.subcode; st = 4; break; }
          p->x = h.x; p->y = h.y; p->z = h.z; p->vx = h.vx; p->vy = h.vy; p->vz = h.vz;
          o->pc.pr += h.pc.pr; o->pc.per += h.pc.per; o->pc.pir += h.pc.pir; o->pc.pabs += h.pc.pabs;
          if (h.channel == 0) o->gic++; else { o->wic++; if (h.channel == 1) o->wcoh++; }
          if (h.status == 2) {
            if (Ld) light_absorb(Ld, hist, sx0, sy0, sz0);
            st = 5;
            break;
          }
          if (Ld) light_event(Ld, hist, sx0, sy0, sz0, h.channel == 1 ? 4 : light_code(h.pc.per, h.pc.pir, h.channel == 2));
          prov_event(m, o, vin, site, h.channel == 1 ? 4 : light_code(h.pc.per, h.pc.pir, h.channel == 2), sel, &h.pc);
          continue;
        }
        if (use_wave) {
          o->wic++;
          int pidx = sel >= 0 ? sel : select_bin(&r, m, 1, p->x, p->y, p->z, LW);
          int coherent = (P->coherent >= 1.0f) || (U(&r) < P->coherent);
          if (cond) o->trials++;
          if (!coherent) {
            /* morphology-wave event about the local azimuthal flow direction */
            float fr = sqrtf(p->x * p->x + p->y * p->y), ffx = 1.0f, ffy = 0.0f;
            if (fr > 1e-12f) { ffx = -p->y / fr; ffy = p->x / fr; }
            int ppf = 0, ppfs = 0, per_b = o->pc.per, pir_b = o->pc.pir;
            int ps = wave_particle_event(P, &r, P->n_water, m->n_f[pidx], m->k_f[pidx], m->r_f[pidx], ffx, ffy, 0.0f,
                                         &p->vx, &p->vy, &p->vz, &o->pc, &ppf, &ppfs);
            if (cond && ps == 1) { o->cs[2]++; o->cs[4]++; }
            if (cond && ps != 1 && ps != 2) o->cs[5]++;
            if (ps == 2) { if (Ld) light_mark(Ld, hist, 9); st = 5; break; }
            if (ps != 1) { if (Ld) light_mark(Ld, hist, 10); pf = ppf ? ppf : PF_OTHER; pfs = ppfs; st = 4; break; }
            if (Ld) light_event(Ld, hist, sx0, sy0, sz0, light_code(o->pc.per > per_b, o->pc.pir > pir_b, 1));
            prov_event(m, o, vin, site, light_code(o->pc.per > per_b, o->pc.pir > pir_b, 1), pidx, &o->pc);
            continue;
          }
          o->wcoh++;
          if (cond) { o->cs[1]++; o->cs[4]++; }
          mie_scatter(&r, m, pidx, &p->vx, &p->vy, &p->vz);
          if (Ld) light_event(Ld, hist, sx0, sy0, sz0, 4);
          prov_event(m, o, vin, site, 4, pidx, &none);
          continue;
        }
        o->gic++;
        if (cond) o->trials++;
        int pidx = sel >= 0 ? sel : select_bin(&r, m, 0, p->x, p->y, p->z, LG);
        float rad = m->r_f[pidx], gx, gy, gz;
        sphere_sample(&r, p->vx, p->vy, p->vz, &gx, &gy, &gz);
        double dcx = (double)p->x - (double)rad * (double)gx;
        double dcy = (double)p->y - (double)rad * (double)gy;
        double dcz = (double)p->z - (double)rad * (double)gz;
        int accessible = sphere_accessible(P, dcx, dcy, dcz, rad);
        if (!accessible && cond) {
          o->inacc++;
          cond_t h = conditional_event(m, &r, pidx, p->x, p->y, p->z, p->vx, p->vy, p->vz, P->cond_trials - 1);
          o->trials += h.trials; o->rejects += h.rejects; o->inacc += h.inaccessible; o->pc.n_abs += h.pc.n_abs;
          count_closure(o, &h);
          if (h.status != 1) { pf = h.failure ? h.failure : PF_OTHER; pfs = h.subcode; st = 4; break; }
          p->x = h.x; p->y = h.y; p->z = h.z; p->vx = h.vx; p->vy = h.vy; p->vz = h.vz;
          o->pc.pr += h.pc.pr; o->pc.per += h.pc.per; o->pc.pir += h.pc.pir; o->pc.pabs += h.pc.pabs;
          o->gic--;
          if (h.channel == 0) o->gic++; else { o->wic++; if (h.channel == 1) o->wcoh++; }
          if (Ld) light_event(Ld, hist, sx0, sy0, sz0, h.channel == 1 ? 4 : light_code(h.pc.per, h.pc.pir, h.channel == 2));
          prov_event(m, o, vin, site, h.channel == 1 ? 4 : light_code(h.pc.per, h.pc.pir, h.channel == 2), pidx, &h.pc);
          continue;
        }
        if (!accessible) { o->ic--; o->gic--; o->n_null++; if (Ld) light_mark(Ld, hist, 8); continue; }
        int ppf = 0, ppfs = 0, per_b = o->pc.per, pir_b = o->pc.pir;
        int ps = sphere_interaction(P, &r, P->n_water, m->n_f[pidx], m->k_f[pidx], rad, P->max_pb, gx, gy, gz,
                                    dcx, dcy, dcz, &p->x, &p->y, &p->z, &p->vx, &p->vy, &p->vz, &o->pc, &ppf, &ppfs);
        if (cond && ps == 1) { o->cs[0]++; o->cs[4]++; }
        if (cond && ps != 1 && ps != 2) o->cs[5]++;
        if (ps == 2) { if (Ld) light_mark(Ld, hist, 9); st = 5; break; }
        if (ps != 1) { if (Ld) light_mark(Ld, hist, 10); pf = ppf ? ppf : PF_OTHER; pfs = ppfs; st = 4; break; }
        if (Ld) light_event(Ld, hist, sx0, sy0, sz0, light_code(o->pc.per > per_b, o->pc.pir > pir_b, 0));
        prov_event(m, o, vin, site, light_code(o->pc.per > per_b, o->pc.pir > pir_b, 0), pidx, &o->pc);
        continue;
      }
      if (axial_top || axial_bot) {
        float a, b, c;
        int tir = 0;
        if (axial_top) {
          float snx = 0.0f, sny = 0.0f, snz = -1.0f;
          free_surface_normal(P, p->x, p->y, &snx, &sny, &snz);
          float ci = fminf(fmaxf(-dot3(p->vx, p->vy, p->vz, snx, sny, snz), 0.0f), 1.0f);
          float R = fresnel_R(ci, P->n_water, P->n_air, &tir);
          p->n_iface++;
          if (tir || U(&r) < R || !refract3(p->vx, p->vy, p->vz, snx, sny, snz, P->n_water, P->n_air, &a, &b, &c)) {
            reflect3(p->vx, p->vy, p->vz, snx, sny, snz, &a, &b, &c);
            p->vx = a; p->vy = b; p->vz = c; p->tr++; if (tir) p->ttir++;
            if (!water_surface_offset(P, &p->x, &p->y, &p->z, snx, sny, snz)) { pf = PF_INVALID_MEDIUM; pfs = SUB_WATER_OUTSIDE; st = 4; break; }
            continue;
          }
          p->vx = a; p->vy = b; p->vz = c;
          p->x -= EPS_F * snx; p->y -= EPS_F * sny; p->z -= EPS_F * snz;
          int hsub = 0, hw = headspace(P, &r, p, &hsub);
          if (hw == 1) { if (o->ic > 0) o->pv.ret++; continue; }
          if (hw == 2) { st = 1; goto AIR_EXIT; }
          if (hw == 3) { st = 7; goto WRITE_OUT; }
          if (hw == 4) { st = 6; goto WRITE_OUT; }
          pf = PF_HEADSPACE_NO_HIT; pfs = hsub; st = 4; goto WRITE_OUT;
        } else {
          int bsub = 0, where = bottom_disc(P, &r, p, &bsub);
          if (where == 1) { if (o->ic > 0) o->pv.ret++; continue; }
          if (where == 2) { st = 1; goto AIR_EXIT; }
          if (where == 3) { st = 7; break; }
          if (where == 5) { st = 8; break; }
          if (where == 4) {
            int hsub = 0, hw = headspace(P, &r, p, &hsub);
            if (hw == 1) { if (o->ic > 0) o->pv.ret++; continue; }
            if (hw == 2) { st = 1; goto AIR_EXIT; }
            if (hw == 3) { st = 7; goto WRITE_OUT; }
            if (hw == 4) { st = 6; goto WRITE_OUT; }
            pf = PF_HEADSPACE_NO_HIT; pfs = hsub; st = 4; goto WRITE_OUT;
          }
          pf = (bsub == SUB_ACRYLIC_BOUNCE) ? PF_ACRYLIC_BOUNCE : PF_BOTTOM_GEOMETRY; pfs = bsub; st = 4;
          break;
        }
      }
      /* water -> acrylic at the side wall, CLARITAS.py:5038-5065 */
      {
        float nx, ny, a, b, c;
        radial_normal(p->x, p->y, &nx, &ny);
        float ci = fminf(fmaxf(dot3(p->vx, p->vy, p->vz, nx, ny, 0.0f), 0.0f), 1.0f);
        int tir = 0;
        float R = fresnel_R(ci, P->n_water, P->n_acrylic, &tir);
        p->n_iface++;
        if (tir || U(&r) < R || !refract3(p->vx, p->vy, p->vz, -nx, -ny, 0.0f, P->n_water, P->n_acrylic, &a, &b, &c)) {
          reflect3(p->vx, p->vy, p->vz, nx, ny, 0.0f, &a, &b, &c);
          p->vx = a; p->vy = b; p->vz = c; p->cir++; if (tir) p->ctir++;
          p->x -= EPS_F * nx; p->y -= EPS_F * ny;
          continue;
        }
        p->vx = a; p->vy = b; p->vz = c; p->x += EPS_F * nx; p->y += EPS_F * ny;
        int sub = 0, where = acrylic_solid(P, &r, p, &sub);
        if (where == 1) { if (o->ic > 0) o->pv.ret++; continue; }
        if (where == 2) { st = 1; goto AIR_EXIT; }
        if (where == 3) { st = 7; break; }
        if (where == 5) { st = 8; break; }
        if (where == 4) {
          int hsub = 0, hw = headspace(P, &r, p, &hsub);
          if (hw == 1) { if (o->ic > 0) o->pv.ret++; continue; }
          if (hw == 2) { st = 1; goto AIR_EXIT; }
          if (hw == 3) { st = 7; goto WRITE_OUT; }
          if (hw == 4) { st = 6; goto WRITE_OUT; }
          if (hw == 5) { st = 8; goto WRITE_OUT; }
          pf = PF_HEADSPACE_NO_HIT; pfs = hsub; st = 4; goto WRITE_OUT;
        }
        pf = (sub == SUB_ACRYLIC_BOUNCE) ? PF_ACRYLIC_BOUNCE : PF_ACRYLIC_NO_HIT; pfs = sub; st = 4;
        break;
      }
    }
    if (st == 0) { pf = PF_ITERATION_LIMIT; pfs = SUB_ITERATION_LIMIT; st = 4; }
    goto WRITE_OUT;
  }
AIR_EXIT:
  /* detector scoring, CLARITAS.py:4397-4435 (absorbing bores, ideal photodiode) */
  {
    float ex = p->x, ey = p->y, ez = p->z, evx = p->vx, evy = p->vy, evz = p->vz;
    if (isfinite(ex) && isfinite(ey) && isfinite(ez) && isfinite(evx) && isfinite(evy) && isfinite(evz)) {
      if (P->det_bore || P->face_model) {
        /* direct acceptance first, then the ring-face cavity, CLARITAS.py:4398-4413 */
        if (P->det_bore) {
          o->native = detector_collimator(m, &r, ex, ey, ez, evx, evy, evz, 0, &o->nb);
          o->mirror = detector_collimator(m, &r, ex, ey, ez, evx, evy, evz, 1, &o->mb);
        } else {
          o->native = hardware_detector(m, ex, ey, ez, evx, evy, evz, 0);
          o->mirror = hardware_detector(m, ex, ey, ez, evx, evy, evz, 1);
        }
        if (o->native < 0 && P->face_model)
          o->native = ring_face_path(m, &r, ex, ey, ez, evx, evy, evz, 0, &o->nb, &o->nfb);
        if (o->mirror < 0 && P->face_model)
          o->mirror = ring_face_path(m, &r, ex, ey, ez, evx, evy, evz, 1, &o->mb, &o->mfb);
      } else {
        o->native = hardware_detector(m, ex, ey, ez, evx, evy, evz, 0);
        o->mirror = hardware_detector(m, ex, ey, ez, evx, evy, evz, 1);
      }
      if (o->native >= 0) {
        float c = fminf(fmaxf(evx * m->det_ux[o->native] + evy * m->det_uy[o->native], -1.0f), 1.0f);
        o->inc_native = acosf(c) * (180.0f / PI_F);
        o->pd_native = P->pd_model ? sfh213_response(o->inc_native) : 1.0f;
      }
      if (o->mirror >= 0) {
        float c = fminf(fmaxf(-evx * m->det_ux[o->mirror] + evy * m->det_uy[o->mirror], -1.0f), 1.0f);
        o->inc_mirror = acosf(c) * (180.0f / PI_F);
        o->pd_mirror = P->pd_model ? sfh213_response(o->inc_mirror) : 1.0f;
      }
      for (int a = 0; a < P->n_aux; ++a) {
        float t;
        o->aux_native[a] = bore_accept(P, ex, ey, ez, evx, evy, evz, m->aux_ux[a], m->aux_uy[a], &t) ? 1 : 0;
        o->aux_mirror[a] = bore_accept(P, -ex, ey, ez, -evx, evy, evz, m->aux_ux[a], m->aux_uy[a], &t) ? 1 : 0;
      }
      for (int mir = 0; mir < 2; ++mir) {
        float xx = mir ? -ex : ex;
        float ang = atan2f(xx, ey) * (180.0f / PI_F);
        if (ang < 0.0f) ang += 360.0f;
        if (ang < 0.0f || ang > 180.0f) continue;
        for (int j = 0; j < CLR_N_CHANNELS; ++j) {
          float deg = 10.0f * j;
          if (o->ic <= 0 && deg >= 90.0f) continue;
          if (fabsf(ang - deg) <= P->legacy_acc) {
            if (mir) o->legacy_mirror[j] = 1; else o->legacy_native[j] = 1;
          }
        }
        /* the same window rule at the auxiliary angles (extension) */
        for (int a = 0; a < P->n_aux; ++a) {
          float deg = m->aux_deg[a];
          if (o->ic <= 0 && deg >= 90.0f) continue;
          if (fabsf(ang - deg) <= P->legacy_acc) {
            if (mir) o->auxw_mirror[a] = 1; else o->auxw_native[a] = 1;
          }
        }
      }
    }
  }
WRITE_OUT:
  o->st = st; o->pf = pf; o->pfs = pfs;
  o->draws = r.draws;
}

/* ======================================================================
 * Tally one history
 * ====================================================================== */
/* tally slot of a failure subcode (CLR_SUBCODE_VALUES; the last slot: any other) */
static int subcode_slot(int sub) {
  static const int vals[CLR_N_SUBCODE_KNOWN] = CLR_SUBCODE_VALUES;
  for (int k = 0; k < CLR_N_SUBCODE_KNOWN; ++k)
    if (vals[k] == sub) return k;
  return CLR_N_SUBCODE - 1;
}

static void tally(clr_tallies* T, const hist_t* h, int g) {
  T->counters[CLR_C_HISTORIES]++;
  T->group_histories[g]++;
  if (h->st >= 0 && h->st < CLR_N_STATUS) T->status[h->st]++;
  if (h->st == 4 && h->pf >= 0 && h->pf < CLR_N_FAILURE) T->failures[h->pf]++;
  if (h->st == 4) T->failure_subcodes[subcode_slot(h->pfs)]++;
  const photon_t* p = &h->ph;
  if (h->st == 1 && isfinite(p->x) && isfinite(p->vx)) T->counters[CLR_C_VALID_AIR]++;
  if (h->st == 1) T->counters[CLR_C_N_EXIT]++;
  if (h->entered) {
    T->counters[CLR_C_ENTERED]++;
    T->counters[CLR_C_SUM_IC] += (uint64_t)h->ic;
    T->counters[CLR_C_SUM_GIC] += (uint64_t)h->gic;
    T->counters[CLR_C_SUM_WIC] += (uint64_t)h->wic;
    T->counters[CLR_C_SUM_WCOH] += (uint64_t)h->wcoh;
    if (h->ic == 0) T->counters[CLR_C_BALLISTIC]++;
    T->ic_hist[h->ic < CLR_N_ICHIST - 1 ? h->ic : CLR_N_ICHIST - 1]++;
    T->counters[CLR_C_SUM_PR] += (uint64_t)h->pc.pr;
    if (h->pc.per > 0) T->counters[CLR_C_RAYS_PER]++;
    if (h->pc.pir > 0) T->counters[CLR_C_RAYS_PIR]++;
    if (h->pc.pabs > 0) T->counters[CLR_C_RAYS_PABS]++;
    T->counters[CLR_C_SUM_CIR] += (uint64_t)p->cir;
    T->counters[CLR_C_SUM_COR] += (uint64_t)p->cor;
    if (p->ctir > 0) T->counters[CLR_C_RAYS_CTIR]++;
    T->counters[CLR_C_SUM_TOPR] += (uint64_t)p->tr;
    if (p->ttir > 0) T->counters[CLR_C_RAYS_TOPTIR]++;
    T->counters[CLR_C_SUM_BIR] += (uint64_t)p->bir;
    T->counters[CLR_C_SUM_BOR] += (uint64_t)p->bor;
    if (p->btir > 0) T->counters[CLR_C_RAYS_BTIR]++;
  }
  T->counters[CLR_C_N_FLIGHT] += h->n_flight;
  T->counters[CLR_C_N_NULL] += h->n_null;
  T->counters[CLR_C_N_SCAT] += (uint64_t)(h->ic > 0 ? h->ic : 0);
  T->counters[CLR_C_N_IFACE] += p->n_iface;
  T->counters[CLR_C_N_ABS] += h->pc.n_abs;
  T->counters[CLR_C_COND_TRIALS] += h->trials;
  T->counters[CLR_C_COND_REJECT] += h->rejects;
  T->counters[CLR_C_COND_INACCESSIBLE] += h->inacc;
  for (int k = 0; k < 7; ++k) T->counters[CLR_C_COND_ACC_GEOM + k] += h->cs[k];
  T->counters[CLR_C_RNG_DRAWS] += h->draws;
  T->counters[CLR_C_N_STEPS] += h->steps;
  if (h->native >= 0) {
    T->det[g][0][h->native]++;
    T->incidence_sum[0][h->native] += h->inc_native;
    T->photodiode[g][0][h->native] += (double)h->pd_native;
    if (h->nb > 0 || h->nfb > 0) T->stray[0][h->native]++;
    if (h->nfb > 0) T->face_stray[0][h->native]++;
  }
  if (h->mirror >= 0) {
    T->det[g][1][h->mirror]++;
    T->incidence_sum[1][h->mirror] += h->inc_mirror;
    T->photodiode[g][1][h->mirror] += (double)h->pd_mirror;
    if (h->mb > 0 || h->mfb > 0) T->stray[1][h->mirror]++;
    if (h->mfb > 0) T->face_stray[1][h->mirror]++;
  }
  T->counters[CLR_C_SRC_BORE_INTER] += (uint64_t)h->sbi;
  T->counters[CLR_C_SRC_BORE_REFL] += (uint64_t)h->sbr;
  if (h->st != 3 && h->sbr > 0) T->counters[CLR_C_SRC_REFL_ACCEPT]++;
  for (int j = 0; j < CLR_N_CHANNELS; ++j) {
    if (h->legacy_native[j]) T->det[g][2][j]++;
    if (h->legacy_mirror[j]) T->det[g][3][j]++;
  }
  for (int a = 0; a < CLR_MAX_AUX; ++a) {
    if (h->aux_native[a] > 0) T->aux[g][0][a]++;
    if (h->aux_mirror[a] > 0) T->aux[g][1][a]++;
    if (h->auxw_native[a] > 0) T->aux_window[g][0][a]++;
    if (h->auxw_mirror[a] > 0) T->aux_window[g][1][a]++;
  }
}

static void merge(clr_tallies* D, const clr_tallies* S) {
  const uint64_t* s = (const uint64_t*)S->det;
  uint64_t* d = (uint64_t*)D->det;
  for (size_t i = 0; i < sizeof(S->det) / 8; ++i) d[i] += s[i];
  s = (const uint64_t*)S->aux; d = (uint64_t*)D->aux;
  for (size_t i = 0; i < sizeof(S->aux) / 8; ++i) d[i] += s[i];
  s = (const uint64_t*)S->aux_window; d = (uint64_t*)D->aux_window;
  for (size_t i = 0; i < sizeof(S->aux_window) / 8; ++i) d[i] += s[i];
  for (int g = 0; g < CLR_MAX_GROUPS; ++g) D->group_histories[g] += S->group_histories[g];
  for (int k = 0; k < 2; ++k)
    for (int c = 0; c < CLR_N_CHANNELS; ++c) {
      D->incidence_sum[k][c] += S->incidence_sum[k][c];
      D->stray[k][c] += S->stray[k][c];
      D->face_stray[k][c] += S->face_stray[k][c];
    }
  for (int g = 0; g < CLR_MAX_GROUPS; ++g)
    for (int k = 0; k < 2; ++k)
      for (int c = 0; c < CLR_N_CHANNELS; ++c) D->photodiode[g][k][c] += S->photodiode[g][k][c];
  for (int i = 0; i < CLR_N_STATUS; ++i) D->status[i] += S->status[i];
  for (int i = 0; i < CLR_N_FAILURE; ++i) D->failures[i] += S->failures[i];
  for (int i = 0; i < CLR_N_SUBCODE; ++i) D->failure_subcodes[i] += S->failure_subcodes[i];
  for (int i = 0; i < CLR_N_COUNTERS; ++i) D->counters[i] += S->counters[i];
  for (int i = 0; i < CLR_N_ICHIST; ++i) D->ic_hist[i] += S->ic_hist[i];
}

static int run_impl(const orc_model* m, uint64_t n, uint64_t seed, uint64_t offset, int nthreads,
                    clr_tallies* out, clr_record* rec, uint64_t n_rec, const light_t* Ld,
                    clr_causal_record* crec, uint8_t* cdet) {
  if (!m || !out) return CLR_E_ARG;
  memset(out, 0, sizeof(*out));
  int G = m->fp.n_groups;
  out->n_groups = G;
  if (n == 0) return 0;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  int nt = 1;
  (void)nthreads;
#endif
  clr_tallies* part = (clr_tallies*)calloc((size_t)nt, sizeof(clr_tallies));
  if (!part) return CLR_E_NOMEM;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    clr_tallies* T = &part[tid];
    hist_t h;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      uint64_t hist = offset + (uint64_t)i;
      trace_history(m, seed, hist, &h, Ld);
      int g = (int)(((unsigned __int128)(uint64_t)i * (unsigned)G) / n);
      tally(T, &h, g);
      if (rec && (uint64_t)i < n_rec) {
        clr_record* q = &rec[i];
        q->x = h.ph.x; q->y = h.ph.y; q->z = h.ph.z; q->vx = h.ph.vx; q->vy = h.ph.vy; q->vz = h.ph.vz;
        q->status = h.st; q->native_id = h.native; q->mirror_id = h.mirror; q->interactions = h.ic;
        q->failure = h.pf; q->draws = (int32_t)h.draws;
      }
      if (crec) {
        prov_record(&h, hist, &crec[i]);
        cdet[i] = (h.native >= 0 || h.mirror >= 0);
      }
    }
  }
  for (int t = 0; t < nt; ++t) merge(out, &part[t]);
  free(part);
  return 0;
}

int orc_run(const orc_model* m, uint64_t n, uint64_t seed, uint64_t offset, int nthreads,
            clr_tallies* out, clr_record* rec, uint64_t n_rec) {
  return run_impl(m, n, seed, offset, nthreads, out, rec, n_rec, NULL, NULL, NULL);
}

/* orc_run with the causal provenance record of every history (crec[n]);
 * detected[i] = 1 when history i reached a hardware detector, the histories
 * clr_run_causal records */
int orc_run_causal(const orc_model* m, uint64_t n, uint64_t seed, uint64_t offset, int nthreads,
                   clr_tallies* out, clr_causal_record* crec, uint8_t* detected) {
  if (!crec || !detected) return CLR_E_ARG;
  return run_impl(m, n, seed, offset, nthreads, out, NULL, 0, NULL, crec, detected);
}

/* orc_run plus the light-transport diagnostics (clr_run_light contract):
 * the caller's output arrays are initialised here (grids 0, trajectory
 * coordinates NaN, orders and codes -1) and filled. */
int orc_run_light(const orc_model* m, uint64_t n, uint64_t seed, uint64_t offset, int nthreads,
                  const clr_light_config* lc, clr_tallies* out, clr_light_out* lo) {
  if (!m || !lc || !lo || !out) return CLR_E_ARG;
  if ((lc->max_rays < 1 && lc->heatmap_size < 1) || lc->heatmap_size < 0 || lc->nx < 4 || lc->ny < 4 || lc->nz < 4 ||
      lc->representative_rays < 0 ||
      (uint64_t)lc->representative_rays > lc->max_rays || lc->max_segments < 1 || !(lc->zmax > m->cfg.zmin))
    return CLR_E_ARG;
  size_t vox = (size_t)lc->nx * lc->ny * lc->nz, rep = (size_t)lc->representative_rays;
  size_t tb = rep * (size_t)lc->max_segments;
  float* g4[] = {lo->track_length, lo->post_scatter_track_length, lo->scatter_event_weight};
  for (int k = 0; k < 3; ++k) if (g4[k]) memset(g4[k], 0, vox * 4);
  if (lo->scatter_event_count) memset(lo->scatter_event_count, 0, vox * 4);
  if (lo->absorption_event_count) memset(lo->absorption_event_count, 0, vox * 4);
  float* tr[] = {lo->traj_x0, lo->traj_y0, lo->traj_z0, lo->traj_x1, lo->traj_y1, lo->traj_z1};
  for (int k = 0; k < 6; ++k) if (tr[k]) for (size_t i = 0; i < tb; ++i) tr[k][i] = NAN;
  if (lo->traj_order) for (size_t i = 0; i < tb; ++i) lo->traj_order[i] = -1;
  if (lo->traj_event) for (size_t i = 0; i < tb; ++i) lo->traj_event[i] = -1;
  if (lo->traj_count) memset(lo->traj_count, 0, rep * 4);
  size_t vis2 = (size_t)lc->heatmap_size * (size_t)lc->heatmap_size;
  if (lo->heatmap_xy) memset(lo->heatmap_xy, 0, vis2 * 4);
  if (lo->heatmap_zy) memset(lo->heatmap_zy, 0, vis2 * 4);
  if (lo->traj_overflow) memset(lo->traj_overflow, 0, rep * 4);
  lo->segment_total = lo->deposited_total = 0.0;
  lo->segment_count = lo->scatter_event_total = lo->absorption_event_total = 0;
  lo->error_count = lo->boundary_adjustment_count = 0;
  memset(lo->outside, 0, sizeof(lo->outside));
  light_t L = {lc, lo, (float)m->cfg.rin, (float)m->cfg.zmin, (float)lc->zmax};
  return run_impl(m, n, seed, offset, nthreads, out, NULL, 0, &L, NULL, NULL);
}

/* the reference's segment and event test kernels (CLARITAS.py:3470-3485)
 * on explicit inputs: segments7 [n][7] = x0, y0, z0, vx, vy, vz, dist with
 * particle order i & 1, xyz3 [n][3] event sites */
/* heat-map raster of segments7 [n][7] with the z axis up to lc->zmax (the
 * caller zeroes the maps) */
int orc_light_heat_probe(const clr_light_config* lc, float rin, float zmin, int n, const float* segments7,
                         clr_light_out* lo) {
  light_t L = {lc, lo, rin, zmin, (float)lc->zmax};
  for (int i = 0; i < n; ++i) {
    const float* q = segments7 + 7 * (size_t)i;
    heat_raster(&L, (float)lc->zmax, q[0], q[1], q[2], q[3], q[4], q[5], q[6]);
  }
  return 0;
}

int orc_light_segment_probe(const clr_light_config* lc, float rin, float zmin, int n, const float* segments7,
                            clr_light_out* lo) {
  light_t L = {lc, lo, rin, zmin, (float)lc->zmax};
  for (int i = 0; i < n; ++i) {
    const float* q = segments7 + 7 * (size_t)i;
    light_segment(&L, (uint64_t)i, q[0], q[1], q[2], q[3], q[4], q[5], q[6], i & 1);
  }
  return 0;
}
int orc_light_event_probe(const clr_light_config* lc, float rin, float zmin, int n, const float* xyz3,
                          clr_light_out* lo) {
  light_t L = {lc, lo, rin, zmin, (float)lc->zmax};
  for (int i = 0; i < n; ++i) {
    const float* q = xyz3 + 3 * (size_t)i;
    light_site(&L, (uint64_t)i, q[0], q[1], q[2], lo->scatter_event_weight, lo->scatter_event_count,
               &lo->scatter_event_total);
  }
  return 0;
}

/* ======================================================================
 * Model construction
 * ====================================================================== */
static void fill_fparams(orc_model* m) {
  const clr_config* c = &m->cfg;
  fparams* P = &m->fp;
  P->n_water = (float)c->n_water; P->n_acrylic = (float)c->n_acrylic; P->n_air = (float)c->n_air;
  P->n_bottom_ext = (float)c->bottom_external_n;
  P->rin = (float)c->rin; P->rout = (float)c->rout; P->zmin = (float)c->zmin; P->zmean = (float)c->zmean;
  P->zwall_top = (float)c->zwall_top; P->bdt = (float)c->bottom_disc_thickness; P->zlow = P->zmin - P->bdt;
  P->ring_in = (float)c->ring_in; P->throat_out = (float)c->throat_out; P->ring_out = (float)c->ring_out;
  P->source_r = (float)c->source_r; P->throat_r = (float)c->throat_r; P->cb_r = (float)c->counterbore_r;
  P->dh = (float)c->vortex_delta_h; P->core_r = (float)c->vortex_core_radius;
  P->alpha1 = (float)c->alpha1; P->alpha2 = (float)c->alpha2; P->sigma = (float)c->beam_sigma;
  P->beam_r = (float)c->beam_radius; P->legacy_acc = (float)c->legacy_accept_deg;
  P->wavelength = (float)c->wavelength;
  P->fs_model = c->free_surface_model; P->src_model = c->source_angular_model;
  P->event_model = c->event_model; P->sed_model = c->sediment_model;
  P->max_pb = c->max_particle_bounces; P->max_iter = c->max_iterations;
  P->rough = (float)c->rough_alpha; P->aspect = (float)c->spheroid_aspect; P->kappa = (float)c->spheroid_kappa;
  P->orientation = c->spheroid_orientation;
  P->src_bore = c->source_bore_model; P->det_bore = c->detector_bore_model; P->face_model = c->ring_face_model;
  P->pd_model = c->photodiode_model; P->max_bb = c->max_bore_bounces; P->max_fb = c->max_ring_face_bounces;
  P->n_bore = (float)c->bore_n; P->n_face = (float)c->ring_face_n;
  P->cond_trials = c->conditional_max_trials; P->n_aux = c->n_aux;
  P->n_groups = (c->n_groups >= 1 && c->n_groups <= CLR_MAX_GROUPS) ? c->n_groups : CLR_MAX_GROUPS;
  P->fn0 = c->field_n0; P->fn1 = c->field_n1; P->fn2 = c->field_n2; P->fnbins = c->field_nbins;
  P->field_zmax = (float)c->field_zmax; P->field_half = (float)c->field_xy_half;
  m->finv[0] = m->finv[1] = m->finv[2] = 0.0f;
  if (c->sediment_model == CLR_SED_RZ_FIELD && c->field_n0 > 0 && c->field_n1 > 0) {
    m->finv[0] = (float)c->field_n0 / P->rin;
    m->finv[1] = (float)c->field_n1 / (P->field_zmax - P->zmin);
  } else if (c->sediment_model == CLR_SED_XYZ_FIELD && c->field_n0 > 0) {
    m->finv[0] = (float)c->field_n0 / (2.0f * P->field_half);
    m->finv[1] = (float)c->field_n1 / (2.0f * P->field_half);
    m->finv[2] = (float)c->field_n2 / (P->field_zmax - P->zmin);
  }
  /* Rankine constants (K = wall-minus-centre drop / (2 - a^2/R^2)) */
  if (c->free_surface_model == CLR_SURFACE_RANKINE) {
    double R = c->rin, a = c->vortex_core_radius, R2 = R * R, a2 = a * a;
    double K = c->vortex_delta_h / (2.0 - a2 / R2);
    double hbar = (2.0 * K / R2) * (R2 - 0.75 * a2 - a2 * log(R / a));
    P->rk_K = K; P->rk_C0 = c->zmean - hbar;
  } else { P->rk_K = 0.0; P->rk_C0 = c->zmean; }
  P->zs_min = (P->fs_model == 0 || P->dh <= 0.0f) ? P->zmean : free_surface_z(P, 0.0f, 0.0f);
  P->zs_min = P->zs_min - 1e-6f * fmaxf(fabsf(P->zs_min), 1e-3f);
  for (int j = 0; j < CLR_N_CHANNELS; ++j) {
    float th = (10.0f * (float)j) * (PI_F / 180.0f);
    m->det_ux[j] = (float)sin((double)th);
    m->det_uy[j] = (float)cos((double)th);
  }
  for (int a = 0; a < CLR_MAX_AUX; ++a) {
    float th = (float)c->aux_angles_deg[a] * (PI_F / 180.0f);
    m->aux_ux[a] = (float)sin((double)th);
    m->aux_uy[a] = (float)cos((double)th);
    m->aux_deg[a] = (float)c->aux_angles_deg[a];
  }
}

void orc_free(orc_model* m) {
  if (!m) return;
  double* arrs[] = {m->d, m->w, m->nr, m->ki, m->x, m->qext, m->qsca, m->qabs, m->qback, m->g, m->albedo,
                    m->theta, m->pdf, m->cdf, m->radius, m->nd, m->mug, m->mus, m->mue, m->muw, m->muev,
                    m->gcdf, m->wcdf, m->trace_g, m->trace_w, m->sedA, m->sedB, m->sednorm, m->sedws, m->radlog};
  for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); ++i) free(arrs[i]);
  free(m->nstop); free(m->theta_f); free(m->cdf_f); free(m->gcdf_f); free(m->wcdf_f); free(m->field);
  free(m);
}

int orc_prepare(const clr_config* cfg, const clr_particles* p, const clr_sediment* sed,
                const float* field, size_t field_len, orc_model** out) {
  if (!cfg || !p || !out || p->n <= 0 || p->n > 512 || p->n_theta < 2) return CLR_E_ARG;
  /* CLARITAS.py:5797-5826 */
  if (!(cfg->rough_alpha >= 0.0) || !isfinite(cfg->rough_alpha) || !(cfg->spheroid_aspect >= 0.2 && cfg->spheroid_aspect <= 5.0) ||
      !(cfg->spheroid_kappa >= 0.0 && cfg->spheroid_kappa <= 100.0) ||
      (cfg->spheroid_orientation != 0 && cfg->spheroid_orientation != 1))
    return CLR_E_ARG;
  orc_model* m = (orc_model*)calloc(1, sizeof(orc_model));
  if (!m) return CLR_E_NOMEM;
  m->cfg = *cfg;
  int n = p->n, nt = p->n_theta;
  m->n = n; m->nt = nt;
  m->nrad = (sed && sed->n_radial >= 2) ? sed->n_radial : 2049;
#define ALLOC(f, cnt) m->f = (double*)calloc((size_t)(cnt), sizeof(double))
  ALLOC(d, n); ALLOC(w, n); ALLOC(nr, n); ALLOC(ki, n); ALLOC(x, n); ALLOC(qext, n); ALLOC(qsca, n);
  ALLOC(qabs, n); ALLOC(qback, n); ALLOC(g, n); ALLOC(albedo, n); ALLOC(theta, nt);
  ALLOC(pdf, (size_t)n * nt); ALLOC(cdf, (size_t)n * nt); ALLOC(radius, n); ALLOC(nd, n); ALLOC(mug, n);
  ALLOC(mus, n); ALLOC(mue, n); ALLOC(muw, n); ALLOC(muev, n); ALLOC(gcdf, n); ALLOC(wcdf, n);
  ALLOC(trace_g, n); ALLOC(trace_w, n); ALLOC(sedA, n); ALLOC(sedB, n); ALLOC(sednorm, n); ALLOC(sedws, n);
  ALLOC(radlog, (size_t)n * m->nrad);
#undef ALLOC
  m->nstop = (int*)calloc((size_t)n, sizeof(int));
  memcpy(m->d, p->diameter, sizeof(double) * n);
  memcpy(m->w, p->weight, sizeof(double) * n);
  memcpy(m->nr, p->n_real, sizeof(double) * n);
  memcpy(m->ki, p->k_imag, sizeof(double) * n);
  if (field && field_len) {
    m->field = (float*)malloc(sizeof(float) * field_len);
    memcpy(m->field, field, sizeof(float) * field_len);
    m->field_len = field_len;
  }
  fill_fparams(m);
  int err = 0;
  clr_mie_out mo = {m->x, m->qext, m->qsca, m->qabs, m->qback, m->g, m->albedo, m->nstop, m->theta, m->pdf, m->cdf};
  err = orc_mie(p, &mo);
  if (!err) err = transport_precompute(m, p);
  if (!err) {
    int tm = cfg->sediment_model;
    if ((tm == CLR_SED_GRID_PER_BIN && (!m->field || field_len < (size_t)n * cfg->field_n0 * cfg->field_n1)) ||
        (tm == CLR_SED_RZ_FIELD && (!m->field || field_len < (size_t)cfg->field_n0 * cfg->field_n1)) ||
        (tm == CLR_SED_XYZ_FIELD && (!m->field || field_len < (size_t)cfg->field_n0 * cfg->field_n1 * cfg->field_n2)))
      err = CLR_E_ARG;
  }
  if (!err) err = sediment_precompute(m, sed);
  if (err) { orc_free(m); return err; }
  m->fp.mu_geom = (float)m->mu_scalars[0];
  m->fp.mu_wave = (float)m->mu_scalars[1];
  m->fp.maj_geom = (float)m->maj[0];
  m->fp.maj_wave = (float)m->maj[1];
  m->fp.coherent = fminf(fmaxf((float)m->mu_scalars[7], 0.0f), 1.0f);
  m->gcdf_f = (float*)malloc(sizeof(float) * n);
  m->wcdf_f = (float*)malloc(sizeof(float) * n);
  for (int i = 0; i < n; ++i) {
    m->gcdf_f[i] = (float)m->gcdf[i]; m->wcdf_f[i] = (float)m->wcdf[i];
    m->n_f[i] = (float)m->nr[i]; m->k_f[i] = (float)m->ki[i]; m->r_f[i] = (float)m->radius[i];
  }
  m->nt_sample = nt;
  m->theta_f = (float*)malloc(sizeof(float) * nt);
  m->cdf_f = (float*)malloc(sizeof(float) * (size_t)n * nt);
  for (int j = 0; j < nt; ++j) m->theta_f[j] = (float)m->theta[j];
  for (size_t j = 0; j < (size_t)n * nt; ++j) m->cdf_f[j] = (float)m->cdf[j];
  *out = m;
  return 0;
}

int orc_set_phase_table(orc_model* m, const double* theta, const double* cdf, int nt) {
  if (!m || !theta || !cdf || nt < 2) return CLR_E_ARG;
  free(m->theta_f); free(m->cdf_f);
  m->nt_sample = nt;
  m->theta_f = (float*)malloc(sizeof(float) * nt);
  m->cdf_f = (float*)malloc(sizeof(float) * (size_t)m->n * nt);
  for (int j = 0; j < nt; ++j) m->theta_f[j] = (float)theta[j];
  for (size_t j = 0; j < (size_t)m->n * nt; ++j) m->cdf_f[j] = (float)cdf[j];
  return 0;
}

int orc_get_tables(const orc_model* m, clr_tables_out* o) {
  if (!m || !o) return CLR_E_ARG;
  size_t n = (size_t)m->n;
#define CP(dst, src) if (o->dst) memcpy(o->dst, m->src, sizeof(double) * n)
  CP(radius, radius); CP(number_density, nd); CP(mu_geom_bin, mug); CP(mu_sca_bin, mus);
  CP(mu_ext_bin, mue); CP(trace_geom_bin, trace_g); CP(trace_wave_bin, trace_w); CP(geom_cdf, gcdf);
  CP(wave_cdf, wcdf); CP(sed_A, sedA); CP(sed_B, sedB); CP(sed_norm, sednorm); CP(sed_settling, sedws);
#undef CP
  if (o->sed_radial_log) memcpy(o->sed_radial_log, m->radlog, sizeof(double) * n * (size_t)m->nrad);
  memcpy(o->mu_scalars, m->mu_scalars, sizeof(m->mu_scalars));
  memcpy(o->majorants, m->maj, sizeof(m->maj));
  return 0;
}

int orc_free_surface_probe(const clr_config* cfg, const float* rays6, int n, float* zs0, float* th,
                           float* zsh, float* nx, float* ny, float* nz) {
  orc_model m;
  memset(&m, 0, sizeof(m));
  m.cfg = *cfg;
  fill_fparams(&m);
  const fparams* P = &m.fp;
  for (int i = 0; i < n; ++i) {
    const float* q = rays6 + 6 * i;
    zs0[i] = free_surface_z(P, q[0], q[1]);
    float t = free_surface_hit(P, q[0], q[1], q[2], q[3], q[4], q[5]);
    th[i] = t;
    if (t < INF_F * 0.5f) {
      float hx = q[0] + t * q[3], hy = q[1] + t * q[4];
      zsh[i] = free_surface_z(P, hx, hy);
      free_surface_normal(P, hx, hy, &nx[i], &ny[i], &nz[i]);
    } else {
      zsh[i] = INF_F; nx[i] = ny[i] = nz[i] = 0.0f;
    }
  }
  return 0;
}

int orc_sediment_probe(const orc_model* m, int bin, const float* xyz, int n, double* out) {
  if (!m || bin < 0 || bin >= m->n) return CLR_E_ARG;
  for (int i = 0; i < n; ++i) out[i] = sediment_local_scale(m, bin, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  return 0;
}

/* ---- probes for closed-form tests ---- */
int orc_fresnel_probe(int n, const float* ci, float n1, float n2, float* R, int* tir) {
  for (int i = 0; i < n; ++i) R[i] = fresnel_R(ci[i], n1, n2, &tir[i]);
  return 0;
}
int orc_complex_probe(int n, const float* ci, float n1r, float n1k, float n2r, float n2k, float* R, float* tt,
                      float* tn, int* ok) {
  for (int i = 0; i < n; ++i) ok[i] = complex_interface(ci[i], n1r, n1k, n2r, n2k, &R[i], &tt[i], &tn[i]);
  return 0;
}
int orc_refract_probe(int n, const float* v, const float* nrm, float n1, float n2, float* out, int* ok) {
  for (int i = 0; i < n; ++i)
    ok[i] = refract3(v[3 * i], v[3 * i + 1], v[3 * i + 2], nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], n1, n2,
                     &out[3 * i], &out[3 * i + 1], &out[3 * i + 2]);
  return 0;
}
/* water-side exit distance through r = R on [zlo, zhi] (CLARITAS.py:2929-2950) */
int orc_wall_probe(int n, const float* rays6, float R, float zlo, float zhi, double* t) {
  orc_model m;
  memset(&m, 0, sizeof(m));
  for (int i = 0; i < n; ++i) {
    const float* q = rays6 + 6 * i;
    t[i] = bore_wall_hit_d(q[0], q[1], q[2], q[3], q[4], q[5], R, zlo, zhi);
  }
  return 0;
}

int orc_beckmann_B_probe(int n, const double* x, double* out) {
  for (int i = 0; i < n; ++i) out[i] = beck_tail_B(x[i]);
  return 0;
}
int orc_beckmann_slopes_probe(int n, const double* si, const double* u1, const double* u2, double* sx, double* sy,
                              int* ok) {
  for (int i = 0; i < n; ++i) {
    int fc = 0;
    ok[i] = beck_sample_slopes(si[i], u1[i], u2[i], &sx[i], &sy[i], &fc) ? 1 : -fc;
  }
  return 0;
}
int orc_walk_probe(int n, uint64_t seed, float alpha, const float* b, float n_top, float k_top, float n_bot,
                   float k_bot, float* d, int* status, int* top, int* fc) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    rng_t r;
    rng_seed(&r, seed, (uint64_t)i);
    float* v = d + 3 * i;
    const float* bb = b + 3 * i;
    top[i] = -1; fc[i] = 0;
    status[i] = micro_interface(&r, alpha, bb[0], bb[1], bb[2], n_top, k_top, n_bot, k_bot, &v[0], &v[1], &v[2],
                                &top[i], &fc[i]);
  }
  return 0;
}

/* ======================================================================
 * Probes of the building blocks on explicit inputs (clr_probe modes)
 * ====================================================================== */
int orc_probe_shape(int mode, int* in_stride, int* out_stride) {
  static const int kIn[CLR_PROBE_N_MODES] = {6, 6, 6, 4, 6, 3, 1, 1, 7};
  static const int kOut[CLR_PROBE_N_MODES] = {6, 4, 4, 2, 2, 2, 6, 1, 9};
  if (mode < 0 || mode >= CLR_PROBE_N_MODES) return CLR_E_ARG;
  *in_stride = kIn[mode];
  *out_stride = kOut[mode];
  return 0;
}

int orc_probe(const orc_model* m, int mode, int n, uint64_t seed, const float* in, double* out) {
  int si, so;
  if (!m || orc_probe_shape(mode, &si, &so)) return CLR_E_ARG;
  const fparams* P = &m->fp;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float* q = in + (size_t)i * si;
    double* o = out + (size_t)i * so;
    rng_t r;
    rng_seed(&r, seed, (uint64_t)i);
    switch (mode) {
      case CLR_PROBE_SURFACE: {
        o[0] = free_surface_z(P, q[0], q[1]);
        float t = free_surface_hit(P, q[0], q[1], q[2], q[3], q[4], q[5]), nx = 0.0f, ny = 0.0f, nz = 0.0f,
              zh = INF_F;
        o[1] = t;
        if (t < 0.5f * INF_F) {
          float hx = q[0] + t * q[3], hy = q[1] + t * q[4];
          zh = free_surface_z(P, hx, hy);
          free_surface_normal(P, hx, hy, &nx, &ny, &nz);
        }
        o[2] = zh; o[3] = nx; o[4] = ny; o[5] = nz;
        break;
      }
      case CLR_PROBE_WALL: {
        o[0] = bore_wall_hit_d(q[0], q[1], q[2], q[3], q[4], q[5], P->rin, P->zmin, P->zwall_top);
        o[1] = finite_cylinder_forward_d(q[0], q[1], q[2], q[3], q[4], q[5], P->rout, P->zlow, P->zwall_top);
        double ta = 0.0;
        int surf = acrylic_next_surface(P, q[0], q[1], q[2], q[3], q[4], q[5], &ta);
        o[2] = ta; o[3] = surf;
        break;
      }
      case CLR_PROBE_OFFSET: {
        float x = q[0], y = q[1], z = q[2];
        int ok = water_surface_offset(P, &x, &y, &z, q[3], q[4], q[5]);
        o[0] = ok; o[1] = x; o[2] = y; o[3] = z;
        break;
      }
      case CLR_PROBE_SEDIMENT: {
        int b = (int)q[0];
        o[0] = (b >= 0 && b < m->n) ? sediment_local_scale(m, b, q[1], q[2], q[3]) : -1.0;
        o[1] = (P->sed_model == CLR_SED_RZ_FIELD || P->sed_model == CLR_SED_XYZ_FIELD)
                   ? shared_field_scale(m, q[1], q[2], q[3]) : 1.0;
        break;
      }
      case CLR_PROBE_DETECTOR:
        o[0] = hardware_detector(m, q[0], q[1], q[2], q[3], q[4], q[5], 0);
        o[1] = hardware_detector(m, q[0], q[1], q[2], q[3], q[4], q[5], 1);
        break;
      case CLR_PROBE_SAMPLER: {
        int k = (int)q[0];
        o[0] = k == 0 ? gamma_draw(&r, q[1]) : k == 1 ? beta_draw(&r, q[1], q[2])
                                           : k == 2 ? beta_radiance_polar(&r, q[1], q[2]) : normal_bm(&r);
        o[1] = r.draws;
        break;
      }
      case CLR_PROBE_SOURCE: {
        float x, y, z, vx, vy, vz;
        source_sample(P, &r, &x, &y, &z, &vx, &vy, &vz);
        o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
        break;
      }
      case CLR_PROBE_MIE_THETA: {
        int b = (int)q[0];
        o[0] = (b >= 0 && b < m->n) ? sample_mie_theta(&r, m, b) : -1.0;
        break;
      }
      case CLR_PROBE_CONDITIONAL: {
        int b = (int)q[0];
        if (b < 0 || b >= m->n) { o[0] = -1.0; break; }
        cond_t c = conditional_event(m, &r, b, q[1], q[2], q[3], q[4], q[5], q[6], P->cond_trials);
        o[0] = c.status; o[1] = c.channel; o[2] = c.trials; o[3] = c.rejects; o[4] = c.inaccessible;
        o[5] = c.failure; o[6] = c.vx; o[7] = c.vy; o[8] = c.vz;
        break;
      }
      default:
        break;
    }
  }
  return 0;
}

/* Slab j covers [zmin + j w, zmin + (j+1) w), w = (field_zmax - zmin) / S,
 * S = min(32, nz); its majorant is the trace rate times the largest field
 * node over all r (x, y) and the z nodes whose interpolation weight can
 * reach the slab (floor(u) and floor(u) + 1 for u = (z - zmin) nz/(zmax -
 * zmin) - 1/2), widened by one node each side against rounding; the end
 * slabs reach the clamped edges. */
int orc_set_slab_majorant(orc_model* m, int on) {
  if (!m) return CLR_E_ARG;
  fparams* P = &m->fp;
  P->n_slabs = 0;
  if (!on) return 0;
  const int rz = P->sed_model == CLR_SED_RZ_FIELD, xyz = P->sed_model == CLR_SED_XYZ_FIELD;
  if (!(rz || xyz) || !m->field) return CLR_E_ARG;
  const int nz = rz ? P->fn1 : P->fn2;
  const size_t planes = rz ? (size_t)P->fn0 : (size_t)P->fn0 * P->fn1;
  const double zmin = m->cfg.zmin, zmax = m->cfg.field_zmax;
  if (nz < 2 || planes == 0 || !(zmax > zmin)) return CLR_E_ARG;
  const int S = nz < CLR_ORC_MAX_SLABS ? nz : CLR_ORC_MAX_SLABS;
  for (int j = 0; j < S; ++j) {
    int k0 = (int)floor((double)j * nz / S - 0.5) - 1, k1 = (int)floor((double)(j + 1) * nz / S - 0.5) + 2;
    if (j == 0) k0 = 0;
    if (j == S - 1) k1 = nz - 1;
    if (k0 < 0) k0 = 0;
    if (k1 > nz - 1) k1 = nz - 1;
    float mx = 0.0f;
    for (size_t pl = 0; pl < planes; ++pl)
      for (int k = k0; k <= k1; ++k) mx = fmaxf(mx, m->field[pl * nz + k]);
    P->slab_mu[j] = (float)(m->sum_trace_g * (double)mx) + (float)(m->sum_trace_w * (double)mx);
  }
  P->n_slabs = S;
  P->slab_mu_max = 0.0f;
  for (int j = 0; j < S; ++j) P->slab_mu_max = fmaxf(P->slab_mu_max, P->slab_mu[j]);
  P->slab_z0 = P->zmin;
  P->slab_w = (float)((zmax - zmin) / S);
  P->slab_inv_w = (float)(S / (zmax - zmin));
  return 0;
}
