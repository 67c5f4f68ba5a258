// transport.hip -- persistent photon-transport kernel for gfx950.
//
// Replaces the reference's trace_kernel + score_detectors_kernel pair
// (CLARITAS/CLARITAS.py:4378-4436, 4551-5104) and the host-side reduction of
// its per-ray arrays (CLARITAS.py:7849-7864, 8650-8733).  The physics follows
// the reference routine by routine (cited at each function); the execution
// model is MI355X-first:
//
//  * one photon history per lane, advanced as a state machine in short steps
//    (one water segment + event, one acrylic surface, or one headspace
//    segment), so a lane whose history ends takes the next history from the
//    workgroup's counter at once: live histories stay packed in every
//    wavefront instead of waiting for the longest history of a launch;
//  * a persistent grid pulls work items (configuration, history range) from
//    a device queue, so one launch serves one or hundreds of configurations;
//  * Philox4x32-10 counter-based streams (key = 64-bit seed, counter =
//    (block, 0, history lo, history hi)) replace per-ray xorshift seeds;
//    64-bit history numbers replace the reference's uint32 ray ids;
//  * tallies (detector channels x jackknife groups, status, failure codes,
//    diagnostic sums, interaction histogram) accumulate in LDS with LDS
//    atomics and are flushed to the per-configuration FP64/u64 tallies with
//    one global atomic per non-zero LDS word; no per-ray arrays are written.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "precompute.h"
#include "transport.h"

namespace {

constexpr float kInfF = 3.402823466e+38F;
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kEpsF = 2.0e-7f;
constexpr double kDEps = 2.220446049250313080847263336181640625e-16;
constexpr double kPiD = 3.141592653589793238462643383279502884;
constexpr double kFEps = 1.1920928955078125e-7;
constexpr int kEmergencyCap = 16384;  // acrylic / headspace loop caps, CLARITAS.py:3147,3244
constexpr int kLongChain = 128;       // acrylic bounces after which a visit is stepped to its end

// production failure codes and subcodes, CLARITAS.py:1846-1895
enum : int { MS_HIT = 1, MS_ESCAPE = 2, MS_NUMERIC = 3, MS_WATCHDOG = 4 };  // walk outcomes
enum : int {
  PF_OK = 0, PF_MICRO_NUMERICAL = 1, PF_MICRO_WATCHDOG = 2, PF_MACRO_GEOMETRY = 3, PF_MACRO_BOUNCE = 4,
  PF_ACRYLIC_NO_HIT = 5, PF_ACRYLIC_BOUNCE = 6, PF_BOTTOM_GEOMETRY = 7,
  PF_WATER_NO_BOUNDARY = 9, PF_ITERATION_LIMIT = 10, PF_NONFINITE_POSITION = 11,
  PF_NONFINITE_DIRECTION = 12, PF_INVALID_MEDIUM = 13, PF_OTHER = 15, PF_SPHEROID = 16,
  PF_HEADSPACE_NO_HIT = 17, PF_MAJORANT = 18
};
enum : int {
  SUB_NONE = 0, SUB_SPHERE_ENTRY_SIDE = 1001, SUB_SPHERE_CHORD = 1002, SUB_SPHERE_EXIT_SIDE = 1003,
  SUB_SPHERE_INTERNAL_SIDE = 1004, SUB_SPHERE_BOUNCE_LIMIT = 1005, SUB_ACRYLIC_NO_HIT = 2001,
  SUB_ACRYLIC_BOUNCE = 2002, SUB_BOTTOM_TOPOLOGY = 2104, SUB_WATER_NO_BOUNDARY = 3001,
  SUB_ITERATION_LIMIT = 3002, SUB_NONFINITE_POSITION = 3003, SUB_NONFINITE_DIRECTION = 3004,
  SUB_WATER_OUTSIDE = 3005, SUB_HEADSPACE_NO_HIT = 3006, SUB_MAJORANT = 4001,
  MF_REFLECTION_INVALID = 31, MF_REFRACTION_INVALID = 32, MF_SIDE_STATE = 33,
  MF_NONFINITE_DIRECTION = 34, MF_FRESNEL_INVALID = 35,
  // microsurface walk subcodes, CLARITAS.py:1810-1839
  MF_PROJ_AREA = 10, MF_CDF = 11, MF_INV_CDF = 12, MF_VISIBILITY = 13, MF_NORMALISATION = 14,
  MF_HEIGHT_SAMPLE = 21, MF_LAMBDA = 23, MF_C0 = 24, MF_LOG_C0 = 25, MF_G1 = 26, MF_ESCAPE_PROB = 27,
  MF_RANDOM = 28, MF_LOG_CNEXT = 29, MF_NORMAL = 30, MF_CNEXT_RANGE = 36, MF_HEIGHT_INV = 37,
  MF_HEIGHT_OUT = 38, MF_WATCHDOG = 40,
  SUB_SPHEROID_ASPECT = 1101, SUB_SPHEROID_AXES = 1102, SUB_SPHEROID_ENTRY = 1103, SUB_SPHEROID_CHORD = 1104,
  SUB_SPHEROID_EXIT_SIDE = 1105, SUB_SPHEROID_INTERNAL_SIDE = 1106, SUB_SPHEROID_BOUNCE_LIMIT = 1107
};

// ------------------------------------------------------------------ RNG --
// Philox4x32-10 (Salmon et al., SC'11) with the published constants.
// The key (the run's seed) is the same for every lane of a work item, so it
// lives once per workgroup in LDS rather than in two VGPRs per lane.
__shared__ uint32_t s_key[2];
// the z-slab majorants of the configuration a workgroup is tracing (read per
// lane with a lane-dependent index, so from LDS rather than the parameter block)
__shared__ float s_slab_mu[CLR_MAJ_SLABS];
// conditional-closure state of a lane in R_WCLOSE (one slot per thread of the
// workgroup; blocks have at most 256 threads): bin | fallback << 16 |
// target-absorbed << 17, trial budget, trials, rejects, inaccessible, absorptions
__shared__ uint32_t s_clo[6][256];
// The lane keeps an 8-word queue of its stream: words are consumed from b0
// and the queue shifts down (a dynamically indexed pick would make the
// compiler keep it in scratch memory).  rng_topup appends the next block when
// at most 4 words remain, at the start of a pass where the whole wave takes
// the same path, so the draws inside the pass seldom need a block of their
// own: a refill inside divergent code costs the whole wave a Philox block.
struct Rng {
  uint32_t h0, h1;  // 64-bit stream counter (history number + launch offset)
  uint32_t blk;     // next Philox block of the stream
  uint32_t n;       // words left in the queue, 0..8
  uint32_t b0, b1, b2, b3, b4, b5, b6, b7;
};
__device__ __forceinline__ void rng_seed(Rng& r, uint64_t hist) {
  r.h0 = (uint32_t)hist; r.h1 = (uint32_t)(hist >> 32);
  r.blk = 0; r.n = 0;
  r.b0 = r.b1 = r.b2 = r.b3 = r.b4 = r.b5 = r.b6 = r.b7 = 0;
}
// words consumed so far
__device__ __forceinline__ uint32_t rng_draws(const Rng& r) { return 4u * r.blk - r.n; }
__device__ __forceinline__ uint64_t rng_counter(const Rng& r) { return ((uint64_t)r.h1 << 32) | r.h0; }
__device__ __forceinline__ void philox_block(Rng& r, uint32_t& w0, uint32_t& w1, uint32_t& w2, uint32_t& w3) {
  uint32_t c0 = r.blk, c1 = 0u, c2 = r.h0, c3 = r.h1, k0 = s_key[0], k1 = s_key[1];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    // one 32x32->64 product per multiplier (a single v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  w0 = c0; w1 = c1; w2 = c2; w3 = c3;
  r.blk++;
}
__device__ __forceinline__ uint32_t rng_next(Rng& r) {
  if (r.n == 0) {
    philox_block(r, r.b0, r.b1, r.b2, r.b3);
    r.n = 4;
  }
  const uint32_t v = r.b0;
  r.b0 = r.b1; r.b1 = r.b2; r.b2 = r.b3; r.b3 = r.b4; r.b4 = r.b5; r.b5 = r.b6; r.b6 = r.b7;
  r.n--;
  return v;
}
// append the next block after the n <= 4 queued words
__device__ __forceinline__ void rng_topup(Rng& r) {
  if (r.n > 4) return;
  uint32_t w0, w1, w2, w3;
  philox_block(r, w0, w1, w2, w3);
  const uint32_t n = r.n;
  r.b7 = (n == 4) ? w3 : r.b7;
  r.b6 = (n == 4) ? w2 : (n == 3) ? w3 : r.b6;
  r.b5 = (n == 4) ? w1 : (n == 3) ? w2 : (n == 2) ? w3 : r.b5;
  r.b4 = (n == 4) ? w0 : (n == 3) ? w1 : (n == 2) ? w2 : (n == 1) ? w3 : r.b4;
  r.b3 = (n == 3) ? w0 : (n == 2) ? w1 : (n == 1) ? w2 : (n == 0) ? w3 : r.b3;
  r.b2 = (n == 2) ? w0 : (n == 1) ? w1 : (n == 0) ? w2 : r.b2;
  r.b1 = (n == 1) ? w0 : (n == 0) ? w1 : r.b1;
  r.b0 = (n == 0) ? w0 : r.b0;
  r.n = n + 4;
}
// (m + 1/2) 2^-23 with m the top 23 bits: exact, strictly inside (0,1)
__device__ __forceinline__ float u_of(uint32_t m) { return ((float)m + 0.5f) * 1.1920928955078125e-7f; }
__device__ __forceinline__ float U(Rng& r) { return u_of(rng_next(r) >> 9); }
__device__ __forceinline__ double U53(Rng& r) {
  const uint32_t a = rng_next(r), b = rng_next(r);
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) *
         1.1102230246251565404236316680908203125e-16;
}
__device__ __forceinline__ double Uopen(Rng& r) {
  return ((double)rng_next(r) + 0.5) * 2.3283064365386962890625e-10;
}

// ------------------------------------------------------ source samplers --
// CLARITAS.py:1149-1212
__device__ __forceinline__ float normal_bm(Rng& r) {
  const float u1 = fmaxf(U(r), 1e-12f), u2 = U(r);
  return sqrtf(-2.0f * logf(u1)) * cosf(2.0f * kPiF * u2);
}
// Marsaglia-Tsang gamma with the shape<1 boost
__device__ __forceinline__ float gamma_draw(Rng& r, float shape) {
  if (shape <= 0.0f) return 0.0f;
  if (fabsf(shape - 1.0f) < 1e-7f) return -logf(fmaxf(U(r), 1e-12f));
  float boost = 1.0f, a = shape;
  if (a < 1.0f) {
    boost = powf(fmaxf(U(r), 1e-12f), 1.0f / a);
    a += 1.0f;
  }
  const float d = a - 0.3333333333333333f, c = 1.0f / sqrtf(9.0f * d);
  for (int it = 0; it < 128; ++it) {
    const float x = normal_bm(r);
    float v = 1.0f + c * x;
    if (v <= 0.0f) continue;
    v = v * v * v;
    const float u = U(r);
    if (u < 1.0f - 0.0331f * x * x * x * x) return boost * d * v;
    if (logf(fmaxf(u, 1e-12f)) < 0.5f * x * x + d * (1.0f - v + logf(v))) return boost * d * v;
  }
  return boost * d;
}
__device__ __forceinline__ float beta_draw(Rng& r, float a, float b) {
  if (a <= 0.0f || b <= 0.0f) return 0.0f;
  if (fabsf(a - 1.0f) < 1e-7f) {
    const float u = fminf(fmaxf(U(r), 1e-12f), 1.0f - 1e-7f);
    return 1.0f - powf(1.0f - u, 1.0f / b);
  }
  if (fabsf(b - 1.0f) < 1e-7f) {
    const float u = fminf(fmaxf(U(r), 1e-12f), 1.0f - 1e-7f);
    return powf(u, 1.0f / a);
  }
  const float ga = gamma_draw(r, a), gb = gamma_draw(r, b), s = ga + gb;
  return s > 0.0f ? ga / s : 0.0f;
}
__device__ __forceinline__ float beta_radiance_polar(Rng& r, float a, float b) {
  const float c = 0.5f * kPiF;
  for (int it = 0; it < 128; ++it) {
    const float u = fminf(fmaxf(beta_draw(r, a + 1.0f, b), 0.0f), 1.0f);
    const float x = c * u, acc = (x < 1.0e-7f) ? 1.0f : sinf(x) / x;
    if (U(r) <= acc) return x;
  }
  return c * fminf(fmaxf(beta_draw(r, a + 1.0f, b), 0.0f), 1.0f);
}

// --------------------------------------------------------- vector optics --
// CLARITAS.py:1214-1260
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return ax * bx + ay * by + az * bz;
}
__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
  const float q = x * x + y * y + z * z;
  if (q <= 0.0f) { x = 0.0f; y = 1.0f; z = 0.0f; return; }
  // the reference's rsqrtf is the hardware reciprocal square root; v_rsq_f32
  // is its gfx950 counterpart (<= 1 ulp)
  const float s = __builtin_amdgcn_rsqf(q);
  x *= s; y *= s; z *= s;
}
__device__ __forceinline__ void reflect3(float& vx, float& vy, float& vz, float nx, float ny,
                                         float nz) {
  const float d = dot3(vx, vy, vz, nx, ny, nz);
  vx = vx - 2.0f * d * nx; vy = vy - 2.0f * d * ny; vz = vz - 2.0f * d * nz;
  normalize3(vx, vy, vz);
}
// n points into the incident medium; writes the refracted direction to o*
__device__ __forceinline__ bool refract3(float ix, float iy, float iz, float nx, float ny, float nz,
                                         float n1, float n2, float& ox, float& oy, float& oz) {
  const float ci = fminf(fmaxf(-dot3(ix, iy, iz, nx, ny, nz), 0.0f), 1.0f);
  const float eta = n1 / n2, k = 1.0f - eta * eta * (1.0f - ci * ci);
  if (k <= 0.0f) return false;
  const float f = eta * ci - sqrtf(k);
  ox = eta * ix + f * nx; oy = eta * iy + f * ny; oz = eta * iz + f * nz;
  normalize3(ox, oy, oz);
  return true;
}
__device__ __forceinline__ float fresnel_R(float ci, float n1, float n2, bool& tir) {
  ci = fminf(fmaxf(ci, 0.0f), 1.0f);
  const float eta = n1 / n2, st2 = eta * eta * fmaxf(0.0f, 1.0f - ci * ci);
  if (st2 >= 1.0f) { tir = true; return 1.0f; }
  tir = false;
  const float ct = sqrtf(fmaxf(0.0f, 1.0f - st2));
  const float ds = n1 * ci + n2 * ct, dp = n1 * ct + n2 * ci;
  float rs = 1.0f, rp = 1.0f;
  if (fabsf(ds) > 1e-12f) { const float a = (n1 * ci - n2 * ct) / ds; rs = a * a; }
  if (fabsf(dp) > 1e-12f) { const float a = (n1 * ct - n2 * ci) / dp; rp = a * a; }
  return fminf(fmaxf(0.5f * (rs + rp), 0.0f), 1.0f);
}
// Fresnel decision at an interface with normal n into the incident medium:
// returns true (and updates v) for transmission; reflects v otherwise.
// The uniform is drawn only when the interface is not TIR, matching the
// short-circuit order `tir || u < R || !refract` of CLARITAS.py:4709,4986.
__device__ __forceinline__ bool fresnel_transmit(Rng& r, float& vx, float& vy, float& vz, float nx,
                                                 float ny, float nz, float n1, float n2,
                                                 bool& tir) {
  const float ci = fminf(fmaxf(-dot3(vx, vy, vz, nx, ny, nz), 0.0f), 1.0f);
  const float R = fresnel_R(ci, n1, n2, tir);
  float a, b, c;
  if (!tir && !(U(r) < R) && refract3(vx, vy, vz, nx, ny, nz, n1, n2, a, b, c)) {
    vx = a; vy = b; vz = c;
    return true;
  }
  reflect3(vx, vy, vz, nx, ny, nz);
  return false;
}

// ---------------------------------------------- complex-index interface --
// CLARITAS.py:1376-1452 (float complex arithmetic)
struct cf { float re, im; };
__device__ __forceinline__ cf cmk(float a, float b) { return cf{a, b}; }
__device__ __forceinline__ cf cadd(cf a, cf b) { return cmk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cf csub(cf a, cf b) { return cmk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cf cmul(cf a, cf b) {
  return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cf cdivf(cf a, cf b) {
  float d = b.re * b.re + b.im * b.im;
  if (d < 1e-30f) d = 1e-30f;
  return cmk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
__device__ __forceinline__ float cabs2(cf a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ cf csqrtp(cf z) {
  const float r = sqrtf(z.re * z.re + z.im * z.im);
  float u = sqrtf(fmaxf(0.0f, 0.5f * (r + z.re))), v = sqrtf(fmaxf(0.0f, 0.5f * (r - z.re)));
  if (z.im < 0.0f) v = -v;
  if (u < 0.0f || (fabsf(u) < 1e-20f && v < 0.0f)) { u = -u; v = -v; }
  return cmk(u, v);
}
__device__ __forceinline__ bool complex_interface(float ci, float n1r, float n1k, float n2r, float n2k, float& R,
                                  float& t_tan, float& t_nrm) {
  if (!(n1r > 0.f && n2r > 0.f && n1k >= 0.f && n2k >= 0.f)) return false;
  ci = fminf(fmaxf(ci, 0.f), 1.f);
  const float beta = n1r * sqrtf(fmaxf(0.f, 1.f - ci * ci));
  const cf n1 = cmk(n1r, n1k), n2 = cmk(n2r, n2k);
  const cf e1 = cmul(n1, n1), e2 = cmul(n2, n2), b2 = cmk(beta * beta, 0.f);
  const cf q1 = csqrtp(csub(e1, b2)), q2 = csqrtp(csub(e2, b2));
  const cf rs = cdivf(csub(q1, q2), cadd(q1, q2));
  const cf pa = cmul(e2, q1), pb = cmul(e1, q2);
  const cf rp = cdivf(csub(pa, pb), cadd(pa, pb));
  const float Rv = .5f * (cabs2(rs) + cabs2(rp));
  if (!isfinite(Rv)) return false;
  const cf qe = cdivf(q2, e2), be = cdivf(cmk(beta, 0.f), e2);
  const float Sx = beta, Sz = q2.re, Px = be.re, Pz = qe.re;
  const float sn = hypotf(Sx, Sz), pn = hypotf(Px, Pz);
  if (!(sn > 0.f && pn > 0.f && Sz >= 0.f && Pz >= 0.f)) return false;
  const float tx = .5f * (Sx / sn + Px / pn), tz = .5f * (Sz / sn + Pz / pn);
  const float nm = hypotf(tx, tz);
  if (!(nm > 0.f && isfinite(nm))) return false;
  R = fminf(fmaxf(Rv, 0.f), 1.f);
  t_tan = tx / nm;
  t_nrm = tz / nm;
  return true;
}
__device__ __forceinline__ bool energy_refract3(float ix, float iy, float iz, float nx, float ny, float nz,
                                float n1r, float n1k, float n2r, float n2k, float& ox, float& oy,
                                float& oz) {
  const float ci = fminf(fmaxf(-dot3(ix, iy, iz, nx, ny, nz), 0.f), 1.f);
  float R, ts, tn;
  if (!complex_interface(ci, n1r, n1k, n2r, n2k, R, ts, tn)) return false;
  float px = ix + ci * nx, py = iy + ci * ny, pz = iz + ci * nz;
  const float pn = sqrtf(px * px + py * py + pz * pz);
  if (pn > 1.e-7f) { px /= pn; py /= pn; pz /= pn; }
  else { px = py = pz = 0.f; ts = 0.f; tn = 1.f; }
  ox = ts * px - tn * nx; oy = ts * py - tn * ny; oz = ts * pz - tn * nz;
  normalize3(ox, oy, oz);
  return isfinite(ox) && isfinite(oy) && isfinite(oz);
}
__device__ __forceinline__ void perpendicular_basis(float vx, float vy, float vz, float& e1x,
                                                    float& e1y, float& e1z, float& e2x, float& e2y,
                                                    float& e2z) {
  float rx, ry, rz;
  if (fabsf(vy) < 0.9f) { rx = 0.0f; ry = 1.0f; rz = 0.0f; } else { rx = 1.0f; ry = 0.0f; rz = 0.0f; }
  e1x = ry * vz - rz * vy; e1y = rz * vx - rx * vz; e1z = rx * vy - ry * vx;
  normalize3(e1x, e1y, e1z);
  e2x = vy * e1z - vz * e1y; e2y = vz * e1x - vx * e1z; e2z = vx * e1y - vy * e1x;
  normalize3(e2x, e2y, e2z);
}

// ----------------------------------------------------------- free surface --
// CLARITAS.py:1503-1527, 1611-1779; the Rankine sheet has no reference twin
__device__ __forceinline__ float free_surface_z(const DevParams& P, float x, float y) {
  const int M = P.fs_model;
  if (M == 0 || P.dh <= 0.0f) return P.zmean;
  const float q = x * x + y * y, R = P.rin;
  if (M == 1) return P.zmean + P.dh * (q / (R * R) - 0.5f);
  if (M == 3) {
    const double qa = (double)q, a2 = (double)P.core_r * (double)P.core_r;
    const double f = (qa < a2) ? P.rk_K * qa / a2 : P.rk_K * (2.0 - a2 / qa);
    return (float)(P.rk_C0 + f);
  }
  const float a2 = P.core_r * P.core_r, R2 = R * R;
  const float H = P.dh * (a2 + R2) / R2;
  const float hbar = H * (1.0f - (a2 / R2) * log1pf(R2 / a2));
  return P.zmean - hbar + H * q / (a2 + q);
}
// Free-surface height for a comparison with z: when z is below the surface's
// lowest point (zs_min, a lower bound of every device-evaluated height) the
// comparison's outcome is known and zs_min is returned without evaluating
// the surface (an FP64 division for the Rankine vortex).
__device__ __forceinline__ float surface_z_vs(const DevParams& P, float x, float y, float z) {
  float zs = P.zs_min;
  if (!(z < P.zs_min)) zs = free_surface_z(P, x, y);
  return zs;
}
// A lower bound of the free-surface height over the horizontal track of the
// segment p + t v, 0 <= t <= d: every surface model here rises with
// q = x^2 + y^2, so its lowest point over the track is at the track's least q
// (the end points, or the closest approach to the axis when that lies inside
// the segment).  q is rounded down by a relative and an absolute pad and the
// height by a relative pad, which cover the float rounding of both.
__device__ __forceinline__ float surface_floor(const DevParams& P, float x, float y, float vx, float vy, float d) {
  const float q0 = x * x + y * y, q1 = x * vx + y * vy, q2 = vx * vx + vy * vy;
  float q = fminf(q0, q0 + d * (2.0f * q1 + d * q2));
  if (q1 < 0.0f && -q1 < q2 * d) {
    const float cr = x * vy - y * vx;
    q = cr * cr * __builtin_amdgcn_rcpf(q2);
  }
  q = fmaxf(q * (1.0f - 1e-5f) - P.fs_qpad, 0.0f);
  float z;
  if (P.fs_model == CLR_SURFACE_RANKINE)
    z = (q < P.fs_a2) ? P.fs_c0 + P.fs_k * q * __builtin_amdgcn_rcpf(P.fs_a2)
                      : P.fs_c0 + P.fs_k * (2.0f - P.fs_a2 * __builtin_amdgcn_rcpf(q));
  else if (P.fs_model == CLR_SURFACE_SCULLY)
    z = P.fs_c0 + P.fs_k * q * __builtin_amdgcn_rcpf(P.fs_a2 + q);
  else
    z = P.fs_c0 + P.fs_k * q;
  return z - 1e-6f * fmaxf(fabsf(z), 1e-3f);
}
__device__ double quadratic_min_positive(double A, double B, double C) {
  if (fabs(A) < 1e-24) {
    if (fabs(B) < 1e-24) return 1.0e300;
    const double t = -C / B;
    return t > 0.0 ? t : 1.0e300;
  }
  const double D = B * B - 4.0 * A * C;
  if (D < 0.0) return 1.0e300;
  const double sd = sqrt(fmax(D, 0.0)), q = -0.5 * (B + copysign(sd, B));
  const double t1 = (fabs(q) > 1e-300) ? q / A : (-B - sd) / (2.0 * A);
  const double t2 = (fabs(q) > 1e-300) ? C / q : (-B + sd) / (2.0 * A);
  double best = 1.0e300;
  if (t1 > 0.0 && t1 < best) best = t1;
  if (t2 > 0.0 && t2 < best) best = t2;
  return best;
}
__device__ __forceinline__ double cubic_min_positive(double A3, double A2, double A1, double A0) {
  const double scale = fmax(fmax(fabs(A3), fabs(A2)), fmax(fabs(A1), fabs(A0)));
  if (scale <= 0.0) return 1.0e300;
  if (fabs(A3) <= 1e-12) return quadratic_min_positive(A2, A1, A0);
  const double a = A2 / A3, b = A1 / A3, c = A0 / A3;
  const double p = b - a * a / 3.0, q = 2.0 * a * a * a / 27.0 - a * b / 3.0 + c;
  const double hq = 0.5 * q, tp = p / 3.0, d1 = hq * hq, d2 = tp * tp * tp, disc = d1 + d2;
  const double tol = 1e-12 * (fabs(d1) + fabs(d2) + 1e-300);
  double best = 1.0e300;
  if (disc > tol) {
    const double sd = sqrt(disc), t = cbrt(-hq + sd) + cbrt(-hq - sd) - a / 3.0;
    if (t > 0.0) best = t;
  } else if (fabs(p) < 1e-24 && fabs(q) < 1e-24) {
    const double t = -a / 3.0;
    if (t > 0.0) best = t;
  } else if (disc >= -tol) {
    const double u = cbrt(-hq), t1 = 2.0 * u - a / 3.0, t2 = -u - a / 3.0;
    if (t1 > 0.0 && t1 < best) best = t1;
    if (t2 > 0.0 && t2 < best) best = t2;
  } else {
    const double mm = 2.0 * sqrt(fmax(-tp, 0.0)), den = sqrt(fmax(-(tp * tp * tp), 1e-300));
    const double phi = acos(fmin(fmax(-hq / den, -1.0), 1.0));
    for (int k = 0; k < 3; ++k) {
      const double t = mm * cos((phi + 2.0 * 3.141592653589793238462643383279502884 * (double)k) / 3.0) - a / 3.0;
      if (t > 0.0 && t < best) best = t;
    }
  }
  return best;
}
// first sign change of a cubic on [lo,hi] by monotone-piece bisection,
// CLARITAS.py:1690-1722
__device__ __forceinline__ double bounded_cubic_root(double a3, double a2, double a1, double a0, double lo,
                                     double hi) {
  if (!(hi > lo) || !isfinite(hi)) return 1.0e300;
  double c0 = lo, c1 = hi, c2 = hi, c3 = hi;
  int n = 1;
  const double A = 3.0 * a3, B = 2.0 * a2, C = a1;
  double r1 = 2.0 * hi, r2 = 2.0 * hi;
  if (A != 0.0) {
    const double disc = B * B - 4.0 * A * C;
    if (disc >= 0.0) {
      const double q = -0.5 * (B + copysign(sqrt(disc), B));
      r1 = (q != 0.0) ? q / A : -B / (2.0 * A);
      r2 = (q != 0.0) ? C / q : r1;
      if (r2 == r1) r2 = 2.0 * hi;
    }
  } else if (B != 0.0) {
    r1 = -C / B;
  }
  // sorted cut list lo < (r1, r2 inside) < hi
  const double sa = fmin(r1, r2), sb = fmax(r1, r2);
  if (sa > lo && sa < hi) { c1 = sa; n = 2; }
  if (sb > lo && sb < hi && sb != sa) { if (n == 2) { c2 = sb; n = 3; } else { c1 = sb; n = 2; } }
  if (n == 1) c1 = hi; else if (n == 2) c2 = hi; else c3 = hi;
  const int ncut = n + 1;
  for (int i = 0; i < ncut - 1; ++i) {
    double l = (i == 0) ? c0 : (i == 1) ? c1 : c2;
    double r = (i == 0) ? c1 : (i == 1) ? c2 : c3;
    double fl = ((a3 * l + a2) * l + a1) * l + a0, fr = ((a3 * r + a2) * r + a1) * r + a0;
    if (l > 0.0 && fl == 0.0) return l;
    if ((fl < 0.0 && fr > 0.0) || (fl > 0.0 && fr < 0.0)) {
      for (int j = 0; j < 64; ++j) {
        const double mid = l + 0.5 * (r - l);
        if (mid == l || mid == r) break;
        const double fm = ((a3 * mid + a2) * mid + a1) * mid + a0;
        if (fm == 0.0) { l = r = mid; break; }
        if ((fl < 0.0 && fm < 0.0) || (fl > 0.0 && fm > 0.0)) { l = mid; fl = fm; } else r = mid;
      }
      return l + 0.5 * (r - l);
    }
    if (r > 0.0 && fr == 0.0) return r;
  }
  return 1.0e300;
}
__device__ __forceinline__ double rankine_hit(const DevParams& P, double X, double Y, double Z, double VX,
                              double VY, double VZ) {
  const double a2 = (double)P.core_r * (double)P.core_r, K = P.rk_K, C0 = P.rk_C0;
  const double RR = (double)P.rin * (double)P.rin;
  const double Q2 = VX * VX + VY * VY, Q1 = 2.0 * (X * VX + Y * VY), Q0 = X * X + Y * Y;
  if (!(Q2 > 0.0)) {
    const double f = (Q0 < a2) ? K * Q0 / a2 : K * (2.0 - a2 / Q0);
    const double t = (VZ != 0.0) ? (C0 + f - Z) / VZ : 1.0e300;
    return (t > 0.0) ? t : 1.0e300;
  }
  double tend = 1.0e300, tin = 1.0e300, tout = -1.0;
  const double dc = Q1 * Q1 - 4.0 * Q2 * (Q0 - RR);
  if (dc >= 0.0) {
    const double q = -0.5 * (Q1 + copysign(sqrt(dc), Q1));
    const double ta = (q != 0.0) ? q / Q2 : -Q1 / (2.0 * Q2), tb = (q != 0.0) ? (Q0 - RR) / q : ta;
    tend = fmax(ta, tb);
  }
  const double dk = Q1 * Q1 - 4.0 * Q2 * (Q0 - a2);
  if (dk > 0.0) {
    const double q = -0.5 * (Q1 + copysign(sqrt(dk), Q1));
    const double ta = (q != 0.0) ? q / Q2 : -Q1 / (2.0 * Q2), tb = (q != 0.0) ? (Q0 - a2) / q : ta;
    tin = fmin(ta, tb);
    tout = fmax(ta, tb);
  }
  if (!(tend > 0.0) || !(tend < 1.0e299)) return 1.0e300;
  // up to three pieces [0,b1) [b1,b2) [b2,tend) split where the ray crosses
  // r = a (a ray that starts inside the core has only the exit crossing)
  const bool vin = (tin > 0.0 && tin < tend), vout = (tout > 0.0 && tout < tend);
  double b1 = tend, b2 = tend;
  if (vin && vout) { b1 = tin; b2 = tout; }
  else if (vin) b1 = tin;
  else if (vout) b1 = tout;
  double lo = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double hi = (i == 0) ? b1 : (i == 1) ? b2 : tend;
    if (hi > lo) {
      const double mid = 0.5 * (lo + hi), qm = (Q2 * mid + Q1) * mid + Q0;
      double t;
      if (qm < a2) {
        const double k = K / a2;
        t = bounded_cubic_root(0.0, -k * Q2, VZ - k * Q1, Z - C0 - k * Q0, lo, hi);
      } else {
        const double L = Z - C0 - 2.0 * K;
        t = bounded_cubic_root(VZ * Q2, L * Q2 + VZ * Q1, L * Q1 + VZ * Q0, L * Q0 + K * a2, lo, hi);
      }
      if (t < 1.0e299 && t > 0.0) return t;
    }
    lo = fmax(lo, hi);
  }
  return 1.0e300;
}
__device__ __forceinline__ float free_surface_hit_curved(const DevParams& P, float x, float y, float z, float vx,
                                                  float vy, float vz);
__device__ __forceinline__ float free_surface_hit(const DevParams& P, float x, float y, float z, float vx,
                                                  float vy, float vz) {
  if (P.fs_model == 0 || P.dh <= 0.0f) {
    if (fabsf(vz) <= 1e-14f) return kInfF;
    const float t = (P.zmean - z) / vz;
    return (t > 0.0f) ? t : kInfF;
  }
  return free_surface_hit_curved(P, x, y, z, vx, vy, vz);
}
__device__ __forceinline__ float free_surface_hit_curved(const DevParams& P, float x, float y, float z, float vx,
                                                  float vy, float vz) {
  const int M = P.fs_model;
  const float R = P.rin;
  if (M == 1) {
    const float aa = P.dh / (R * R);
    const double A = -(double)aa * ((double)vx * vx + (double)vy * vy);
    const double B = (double)vz - 2.0 * (double)aa * ((double)x * vx + (double)y * vy);
    const double C = (double)z - ((double)P.zmean - 0.5 * (double)P.dh) -
                     (double)aa * ((double)x * x + (double)y * y);
    const double t = quadratic_min_positive(A, B, C);
    return (t < 0.5 * (double)kInfF) ? (float)t : kInfF;
  }
  if (M == 3) {
    const double t = rankine_hit(P, x, y, z, vx, vy, vz);
    return (t > 0.0 && t < 0.5 * (double)kInfF) ? (float)t : kInfF;
  }
  const double X = x, Y = y, Z = z, VX = vx, VY = vy, VZ = vz;
  const double RR = (double)R * (double)R, aa = (double)P.core_r * (double)P.core_r;
  const double H = (double)P.dh * (aa + RR) / RR;
  const double hbar = H * (1.0 - (aa / RR) * log1p(RR / aa));
  const double Cs = (double)P.zmean - hbar;
  const double Q2 = VX * VX + VY * VY, Q1 = 2.0 * (X * VX + Y * VY), Q0 = X * X + Y * Y, L0 = Z - Cs;
  const double c3 = VZ * Q2, c2 = (L0 - H) * Q2 + VZ * Q1, c1 = (L0 - H) * Q1 + VZ * (aa + Q0);
  const double c0 = L0 * aa + (L0 - H) * Q0;
  double t = cubic_min_positive(c3, c2, c1, c0);
  const double xx = X + t * VX, yy = Y + t * VY, qq = xx * xx + yy * yy;
  const double res = (Z + t * VZ) - (Cs + H * qq / (aa + qq));
  const double tol = 32.0 * kDEps * fmax(fmax(fabs(Z), fabs(Cs) + fabs(H)), 1.0e-3);
  if (!(t < 0.5 * (double)kInfF) || qq > RR * (1.0 + 2e-6) || !isfinite(res) || fabs(res) > tol) {
    if (Q2 > 0.0) {
      const double dc = Q1 * Q1 - 4.0 * Q2 * (Q0 - RR);
      if (dc < 0.0) return kInfF;
      const double q = -0.5 * (Q1 + copysign(sqrt(dc), Q1));
      const double ta = (q != 0.0) ? q / Q2 : -Q1 / (2.0 * Q2), tb = (q != 0.0) ? (Q0 - RR) / q : ta;
      t = bounded_cubic_root(c3, c2, c1, c0, fmax(0.0, fmin(ta, tb)), fmax(ta, tb));
    } else {
      t = (VZ != 0.0) ? ((Cs + H * Q0 / (aa + Q0)) - Z) / VZ : 1.0e300;
      if (!(t > 0.0) || Q0 > RR * (1.0 + 2e-6)) return kInfF;
    }
  }
  return (t > 0.0 && t < 0.5 * (double)kInfF) ? (float)t : kInfF;
}
__device__ __forceinline__ void free_surface_normal(const DevParams& P, float x, float y, float& nx, float& ny,
                                    float& nz) {
  const int M = P.fs_model;
  if (M == 0 || P.dh <= 0.0f) { nx = 0.0f; ny = 0.0f; nz = -1.0f; return; }
  const float R = P.rin;
  if (M == 1) {
    const float aa = P.dh / (R * R);
    nx = 2.0f * aa * x; ny = 2.0f * aa * y; nz = -1.0f;
    normalize3(nx, ny, nz);
    return;
  }
  const float q = x * x + y * y;
  float fac;
  if (M == 3) {
    const float a2 = P.core_r * P.core_r, K = (float)P.rk_K;
    fac = (q < a2) ? 2.0f * K / a2 : 2.0f * K * a2 / (q * q);
  } else {
    const float a2 = P.core_r * P.core_r, R2 = R * R, H = P.dh * (a2 + R2) / R2;
    fac = 2.0f * H * a2 / ((a2 + q) * (a2 + q));
  }
  nx = fac * x; ny = fac * y; nz = -1.0f;
  normalize3(nx, ny, nz);
}
__device__ double free_surface_z_d(const DevParams& P, double q) {
  const int M = P.fs_model;
  const double zm = (double)P.zmean;
  if (M == 0 || P.dh <= 0.0f) return zm;
  const double R = (double)P.rin, dh = (double)P.dh, a = (double)P.core_r, R2 = R * R, a2 = a * a;
  if (M == 1) return zm + dh * (q / R2 - 0.5);
  if (M == 3) return P.rk_C0 + ((q < a2) ? P.rk_K * q / a2 : P.rk_K * (2.0 - a2 / q));
  const double H = dh * (a2 + R2) / R2, hbar = H * (1.0 - (a2 / R2) * log1p(R2 / a2));
  return zm - hbar + H * q / (a2 + q);
}

// ----------------------------------------------------------- cell geometry --
// CLARITAS.py:1473-1500, 2888-3063
__device__ __forceinline__ void radial_normal(float x, float y, float& nx, float& ny) {
  const float r = sqrtf(fmaxf(x * x + y * y, 1e-30f));
  nx = x / r; ny = y / r;
}
__device__ __forceinline__ void cylinder_roots_d(float x, float y, float vx, float vy, float R,
                                                 double& t0, double& t1) {
  t0 = 1.0e300; t1 = 1.0e300;
  const double X = x, Y = y, VX = vx, VY = vy, RR = R;
  const double a = VX * VX + VY * VY;
  if (!(a > 1e-30) || !isfinite(a)) return;
  const double b = X * VX + Y * VY, c = X * X + Y * Y - RR * RR;
  double d = b * b - a * c;
  const double dtol = 16.0 * kDEps * (fabs(b * b) + fabs(a * c) + 1e-300);
  if (d < 0.0) { if (d >= -dtol) d = 0.0; else return; }
  const double sd = sqrt(fmax(d, 0.0)), q = -b - copysign(sd, b);
  const double r0 = (fabs(q) > 1e-300) ? q / a : (-b - sd) / a;
  const double r1 = (fabs(q) > 1e-300) ? c / q : (-b + sd) / a;
  double a0 = 1.0e300, a1 = 1.0e300;
  if (r0 > 0.0 && isfinite(r0)) a0 = r0;
  if (r1 > 0.0 && isfinite(r1)) { if (r1 < a0) { a1 = a0; a0 = r1; } else a1 = r1; }
  t0 = a0; t1 = a1;
}
__device__ __forceinline__ double support_tol(const DevParams& P, float Rin, float Rout, float Zmin,
                                              float Zlow, float Ztop) {
  double s = fmax(fmax(fabs((double)Rout), fabs((double)Zlow)), fabs((double)Ztop));
  s = fmax(s, fmax(fabs((double)Rin), fabs((double)Zmin)));
  return 8.0 * kFEps * fmax(s, 1.0e-3);
}
__device__ double finite_cylinder_forward_d(const DevParams& P, float x, float y, float z, float vx,
                                            float vy, float vz, float R, float zlo, float zhi) {
  double r0, r1;
  cylinder_roots_d(x, y, vx, vy, R, r0, r1);
  const double tol = support_tol(P, R, R, zlo, zlo, zhi);
  double best = 1.0e300;
  if (r0 > 0.0 && r0 < 1.0e299) {
    const double zh = (double)z + r0 * (double)vz;
    if (zh >= (double)zlo - tol && zh <= (double)zhi + tol) best = r0;
  }
  if (r1 > 0.0 && r1 < 1.0e299 && r1 < best) {
    const double zh = (double)z + r1 * (double)vz;
    if (zh >= (double)zlo - tol && zh <= (double)zhi + tol) best = r1;
  }
  return best;
}
__device__ double bore_wall_hit_d(const DevParams& P, float x, float y, float z, float vx, float vy,
                                  float vz, float R, float zlo, float zhi) {
  const double X = x, Y = y, VX = vx, VY = vy, RR = (double)R * R;
  const double rr2 = X * X + Y * Y, a = VX * VX + VY * VY;
  double c = rr2 - RR, b = X * VX + Y * VY;
  const double rtol = 8.0 * kFEps * RR, ztol = support_tol(P, R, R, zlo, zlo, zhi);
  if (c > rtol || !isfinite(c) || !isfinite(b) || !isfinite(a)) return -1.0;
  if (c >= 0.0) {
    if (b >= 0.0 && a > 0.0 && z >= zlo - ztol && z <= zhi + ztol) return 0.0;
    b *= sqrt(RR / rr2);
    c = 0.0;
  }
  if (!(a > 1e-30)) return 1.0e300;
  const double d = b * b - a * c;
  if (d < 0.0) return 1.0e300;
  const double sd = sqrt(d), q = -b - copysign(sd, b);
  const double t = (b < 0.0) ? q / a : ((q != 0.0) ? c / q : 0.0);
  if (!(t >= 0.0) || !isfinite(t)) return 1.0e300;
  const double zh = (double)z + t * (double)vz;
  return (zh >= zlo - ztol && zh <= zhi + ztol) ? t : 1.0e300;
}
__device__ __forceinline__ bool water_surface_offset(const DevParams& P, float& x, float& y, float& z, float nx,
                                     float ny, float nz) {
  const double rr2 = (double)x * (double)x + (double)y * (double)y;
  const double RR = (double)P.rin * P.rin, tol = 8.0 * kFEps * RR;
  if (!isfinite(rr2) || rr2 > RR + tol) return false;
  float xx = x + kEpsF * nx, yy = y + kEpsF * ny, zz = z + kEpsF * nz;
  const double rad = hypot((double)xx, (double)yy);
  if (rad > (double)P.rin) {
    const double target = (double)P.rin * (1.0 - 2.0 * kFEps);
    xx = (float)((double)xx * target / rad);
    yy = (float)((double)yy * target / rad);
    const float surf = free_surface_z(P, xx, yy);
    if (zz >= surf) zz = nextafterf(surf, -kInfF);
  }
  if (!isfinite(xx) || !isfinite(yy) || !isfinite(zz) || zz <= P.zmin ||
      zz >= free_surface_z(P, xx, yy))
    return false;
  x = xx; y = yy; z = zz;
  return true;
}
__device__ __forceinline__ int acrylic_next_surface(const DevParams& P, float x, float y, float z, float vx,
                                    float vy, float vz, double& t_out) {
  double best = 1.0e300;
  int surf = 0;
  const double tol = support_tol(P, P.rin, P.rout, P.zmin, P.zlow, P.zwall_top);
  double r0, r1;
  cylinder_roots_d(x, y, vx, vy, P.rin, r0, r1);
  if (r0 < 1.0e299) {
    const double zh = (double)z + r0 * (double)vz;
    if (zh >= (double)P.zmin - tol && zh <= (double)P.zwall_top + tol && r0 > 0.0 && r0 < best) { best = r0; surf = 1; }
  }
  if (r1 < 1.0e299) {
    const double zh = (double)z + r1 * (double)vz;
    if (zh >= (double)P.zmin - tol && zh <= (double)P.zwall_top + tol && r1 > 0.0 && r1 < best) { best = r1; surf = 1; }
  }
  cylinder_roots_d(x, y, vx, vy, P.rout, r0, r1);
  if (r0 < 1.0e299) {
    const double zh = (double)z + r0 * (double)vz;
    if (zh >= (double)P.zlow - tol && zh <= (double)P.zwall_top + tol && r0 > 0.0 && r0 < best) { best = r0; surf = 2; }
  }
  if (r1 < 1.0e299) {
    const double zh = (double)z + r1 * (double)vz;
    if (zh >= (double)P.zlow - tol && zh <= (double)P.zwall_top + tol && r1 > 0.0 && r1 < best) { best = r1; surf = 2; }
  }
  // planes: the radius of the crossing is needed only for a nearer hit
  if (vz < 0.0f) {
    const double t = ((double)P.zlow - (double)z) / (double)vz;
    if (t > 0.0 && isfinite(t) && t < best) {
      const double X = (double)x + t * (double)vx, Y = (double)y + t * (double)vy;
      if (hypot(X, Y) <= (double)P.rout + tol) { best = t; surf = 3; }
    }
  }
  if (vz > 0.0f) {
    const double iv = 1.0 / (double)vz;
    const double t = ((double)P.zwall_top - (double)z) * iv;
    if (t > 0.0 && isfinite(t) && t < best) {
      const double X = (double)x + t * (double)vx, Y = (double)y + t * (double)vy, rr = hypot(X, Y);
      if (rr >= (double)P.rin - tol && rr <= (double)P.rout + tol) { best = t; surf = 4; }
    }
    const double ta = ((double)P.zmin - (double)z) * iv;
    if (ta > 0.0 && isfinite(ta) && ta < best) {
      const double X = (double)x + ta * (double)vx, Y = (double)y + ta * (double)vy;
      if (hypot(X, Y) < (double)P.rin - tol) { best = ta; surf = 5; }
    }
  }
  t_out = best;
  return surf;
}
__device__ __forceinline__ bool tir_for_ci(double ci, double n1, double n2) {
  if (!(n1 > 0.0) || !(n2 > 0.0) || n1 <= n2) return false;
  ci = fmin(fmax(ci, 0.0), 1.0);
  return (n1 / n2) * sqrt(fmax(0.0, 1.0 - ci * ci)) >= 1.0;
}
// conserved-orbit trap classifier at the emergency cap, CLARITAS.py:3083-3135
__device__ __forceinline__ bool acrylic_trapped(const DevParams& P, float x, float y, float vx, float vy, float vz) {
  const double dx = vx, dy = vy, dz = vz, sp2 = dx * dx + dy * dy + dz * dz;
  if (!(sp2 > 0.0) || !isfinite(sp2)) return false;
  const double vxy2 = dx * dx + dy * dy, L = (double)x * dy - (double)y * dx, rs = sqrt(sp2);
  const double reach = 4.0 * (double)kEpsF;
  const double na = P.n_acrylic, nw = P.n_water, nair = P.n_air, nb = P.n_bottom_ext;
  if (vxy2 > 0.0) {
    const double tg = L / (double)P.rout;
    double rad2 = vxy2 - tg * tg;
    if (rad2 < 0.0 && rad2 > -1.0e-12 * sp2) rad2 = 0.0;
    if (rad2 < 0.0) return false;
    if (!tir_for_ci(sqrt(rad2) / rs, na, nair)) return false;
    const double impact = fabs(L) / sqrt(vxy2);
    if (impact <= (double)P.rin + reach) {
      const double ti = L / (double)P.rin;
      double ri2 = vxy2 - ti * ti;
      if (ri2 < 0.0 && ri2 > -1.0e-12 * sp2) ri2 = 0.0;
      if (ri2 >= 0.0 && !tir_for_ci(sqrt(ri2) / rs, na, nw)) return false;
    }
    if (impact < (double)P.rin - reach && fabs(dz) > 0.0 && !tir_for_ci(fabs(dz) / rs, na, nw)) return false;
  }
  if (fabs(dz) > 0.0) {
    const double ca = fabs(dz) / rs;
    if (vxy2 <= 1.0e-30 * sp2 && hypot((double)x, (double)y) < (double)P.rin - reach &&
        !tir_for_ci(ca, na, nw))
      return false;
    if (!tir_for_ci(ca, na, nair)) return false;
    if (!tir_for_ci(ca, na, nb)) return false;
  }
  return true;
}

// -------------------------------------------------------- sediment field --
// CLARITAS.py:1537-1609; shared r-z / x-y-z fields are extensions
__device__ __forceinline__ double rz_sample(const float* __restrict__ G, int NR, int NZ, float x, float y, float z,
                            float R, float ZMIN, float ZMAX) {
  if (NR < 2 || NZ < 2 || !(ZMAX > ZMIN)) return 1.0;
  double rr = sqrt(fmax((double)x * (double)x + (double)y * (double)y, 0.0));
  if (rr > (double)R) rr = (double)R;
  const double dr = (double)R / (double)NR, dz = ((double)ZMAX - (double)ZMIN) / (double)NZ;
  double ur = rr / dr - 0.5, uz = ((double)z - (double)ZMIN) / dz - 0.5;
  ur = fmin(fmax(ur, 0.0), (double)(NR - 1));
  uz = fmin(fmax(uz, 0.0), (double)(NZ - 1));
  int i = (int)floor(ur), k = (int)floor(uz);
  double tr = ur - (double)i, tz = uz - (double)k;
  if (i >= NR - 1) { i = NR - 2; tr = 1.0; }
  if (k >= NZ - 1) { k = NZ - 2; tz = 1.0; }
  const float* g = G + (size_t)i * NZ + k;
  const double q00 = g[0], q01 = g[1], q10 = g[NZ], q11 = g[NZ + 1];
  const double q0 = (1.0 - tr) * q00 + tr * q10, q1 = (1.0 - tr) * q01 + tr * q11;
  return fmax((1.0 - tz) * q0 + tz * q1, 0.0);
}
// Shared FP32 fields (sediment models 4 and 5, extensions with no reference
// counterpart): cell-centred bilinear (r, z) / trilinear (x, y, z)
// interpolation in FP32 with the inverse cell sizes precomputed on the host.
__device__ __forceinline__ float rz_field_f32(const float* __restrict__ G, int NR, int NZ, float x, float y, float z,
                                              float R, float ZMIN, float inv_dr, float inv_dz) {
  const float rr = fminf(sqrtf(fmaxf(x * x + y * y, 0.0f)), R);
  const float ur = fminf(fmaxf(rr * inv_dr - 0.5f, 0.0f), (float)(NR - 1));
  const float uz = fminf(fmaxf((z - ZMIN) * inv_dz - 0.5f, 0.0f), (float)(NZ - 1));
  int i = (int)floorf(ur), k = (int)floorf(uz);
  float tr = ur - (float)i, tz = uz - (float)k;
  if (i >= NR - 1) { i = NR - 2; tr = 1.0f; }
  if (k >= NZ - 1) { k = NZ - 2; tz = 1.0f; }
  const float* g = G + (size_t)i * NZ + k;
  const float q0 = (1.0f - tr) * g[0] + tr * g[NZ], q1 = (1.0f - tr) * g[1] + tr * g[NZ + 1];
  return fmaxf((1.0f - tz) * q0 + tz * q1, 0.0f);
}
__device__ __forceinline__ float xyz_field_f32(const float* __restrict__ G, int NX, int NY, int NZ, float x, float y,
                                               float z, float H, float ZMIN, float inv_dx, float inv_dy, float inv_dz) {
  const float u0 = fminf(fmaxf((x + H) * inv_dx - 0.5f, 0.0f), (float)(NX - 1));
  const float u1 = fminf(fmaxf((y + H) * inv_dy - 0.5f, 0.0f), (float)(NY - 1));
  const float u2 = fminf(fmaxf((z - ZMIN) * inv_dz - 0.5f, 0.0f), (float)(NZ - 1));
  int i0 = (int)floorf(u0), i1 = (int)floorf(u1), i2 = (int)floorf(u2);
  float t0 = u0 - (float)i0, t1 = u1 - (float)i1, t2 = u2 - (float)i2;
  if (i0 >= NX - 1) { i0 = NX - 2; t0 = 1.0f; }
  if (i1 >= NY - 1) { i1 = NY - 2; t1 = 1.0f; }
  if (i2 >= NZ - 1) { i2 = NZ - 2; t2 = 1.0f; }
  const size_t sy = (size_t)NZ, sx = (size_t)NY * NZ;
  const float* b = G + (size_t)i0 * sx + (size_t)i1 * sy + i2;
  const float c00 = (1.0f - t2) * b[0] + t2 * b[1];
  const float c01 = (1.0f - t2) * b[sy] + t2 * b[sy + 1];
  const float c10 = (1.0f - t2) * b[sx] + t2 * b[sx + 1];
  const float c11 = (1.0f - t2) * b[sx + sy] + t2 * b[sx + sy + 1];
  const float c0 = (1.0f - t1) * c00 + t1 * c01, c1 = (1.0f - t1) * c10 + t1 * c11;
  return fmaxf((1.0f - t0) * c0 + t0 * c1, 0.0f);
}
__device__ double xyz_sample(const float* __restrict__ G, int NX, int NY, int NZ, float x, float y,
                             float z, float H, float ZMIN, float ZMAX) {
  const double dx = 2.0 * (double)H / NX, dy = 2.0 * (double)H / NY, dz = ((double)ZMAX - (double)ZMIN) / NZ;
  double u0 = ((double)x + H) / dx - 0.5, u1 = ((double)y + H) / dy - 0.5, u2 = ((double)z - ZMIN) / dz - 0.5;
  u0 = fmin(fmax(u0, 0.0), (double)(NX - 1));
  u1 = fmin(fmax(u1, 0.0), (double)(NY - 1));
  u2 = fmin(fmax(u2, 0.0), (double)(NZ - 1));
  int i0 = (int)floor(u0), i1 = (int)floor(u1), i2 = (int)floor(u2);
  double t0 = u0 - i0, t1 = u1 - i1, t2 = u2 - i2;
  if (i0 >= NX - 1) { i0 = NX - 2; t0 = 1.0; }
  if (i1 >= NY - 1) { i1 = NY - 2; t1 = 1.0; }
  if (i2 >= NZ - 1) { i2 = NZ - 2; t2 = 1.0; }
  const size_t sy = (size_t)NZ, sx = (size_t)NY * NZ;
  const float* b = G + (size_t)i0 * sx + (size_t)i1 * sy + i2;
  const double c00 = (1 - t2) * b[0] + t2 * b[1];
  const double c01 = (1 - t2) * b[sy] + t2 * b[sy + 1];
  const double c10 = (1 - t2) * b[sx] + t2 * b[sx + 1];
  const double c11 = (1 - t2) * b[sx + sy] + t2 * b[sx + sy + 1];
  const double c0 = (1 - t1) * c00 + t1 * c01, c1 = (1 - t1) * c10 + t1 * c11;
  return fmax((1 - t0) * c0 + t0 * c1, 0.0);
}
__device__ __forceinline__ double sediment_scale(const DevParams& P, int pidx, float x, float y, float z) {
  const int tm = P.sed_model;
  if (tm == CLR_SED_UNIFORM) return 1.0;
  if (tm == CLR_SED_GRID_PER_BIN)
    return rz_sample(P.field + (size_t)pidx * P.fn0 * P.fn1, P.fn0, P.fn1, x, y, z, P.rin, P.zmin, P.field_zmax);
  if (tm == CLR_SED_RZ_FIELD) return rz_sample(P.field, P.fn0, P.fn1, x, y, z, P.rin, P.zmin, P.field_zmax);
  if (tm == CLR_SED_XYZ_FIELD)
    return xyz_sample(P.field, P.fn0, P.fn1, P.fn2, x, y, z, P.field_half, P.zmin, P.field_zmax);
  double rr2 = (double)x * (double)x + (double)y * (double)y;
  const double R2 = (double)P.rin * (double)P.rin;
  if (rr2 > R2) rr2 = R2;
  double rl;
  if (tm == CLR_SED_STOKES) {
    if (P.fs_model == 1) rl = P.sedA[pidx] * (rr2 - R2);
    else {
      const double a2 = (double)P.core_r * (double)P.core_r;
      rl = P.sedA[pidx] * (rr2 / (a2 + rr2) - R2 / (a2 + R2));
    }
  } else {
    const int nr = P.n_radial;
    const double u = (P.rin > 0.0f) ? sqrt(fmax(rr2, 0.0)) / (double)P.rin * (double)(nr - 1) : 0.0;
    int j = (int)floor(u);
    j = min(max(j, 0), nr - 2);
    const double t = u - (double)j;
    const double* L = P.radlog + (size_t)pidx * nr;
    rl = fmin((1.0 - t) * L[j] + t * L[j + 1], 0.0);
  }
  const double h = fmax((double)z - (double)P.zmin, 0.0), ns = P.sednorm[pidx];
  if (!(ns > 0.0)) return 0.0;
  return exp(rl - P.sedB[pidx] * h) / ns;
}
// Lower bound of u = u_of(m) in a CDF row: the smallest i with u <= cdf[i]
// (or n-1).  The guide entry of m's bucket brackets the answer, so the
// bisection runs over a few entries instead of the whole row and returns
// the same index as a full binary search.
__device__ __forceinline__ int guided_lower_bound(const float* __restrict__ cdf,
                                                  const unsigned int* __restrict__ guide, uint32_t m, float u) {
  const unsigned int g = guide[m >> (23 - kGuideBits)];
  int lo = (int)(g & 0xffffu), hi = (int)(g >> 16);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (u <= cdf[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}
__device__ __forceinline__ int sample_cdf_index(Rng& r, const float* __restrict__ cdf,
                                                const unsigned int* __restrict__ guide, int n) {
  if (n <= 1) return 0;
  const uint32_t m = rng_next(r) >> 9;
  return guided_lower_bound(cdf, guide, m, u_of(m));
}
__device__ __forceinline__ int select_bin(Rng& r, const DevParams& P, bool wave, float x, float y, float z,
                          float local_channel) {
  const int tm = P.sed_model;
  if (tm == CLR_SED_UNIFORM || tm == CLR_SED_RZ_FIELD || tm == CLR_SED_XYZ_FIELD)
    return sample_cdf_index(r, wave ? P.wcdf : P.gcdf, wave ? P.wguide : P.gguide, P.n_bins);
  if (!((double)local_channel > 0.0)) return 0;
  const double target = (double)U(r) * (double)local_channel;
  const double* rate = wave ? P.trace_w : P.trace_g;
  double cum = 0.0;
  for (int i = 0; i < P.n_bins; ++i) {
    cum += rate[i] * sediment_scale(P, i, x, y, z);
    if (target <= cum) return i;
  }
  return P.n_bins > 0 ? P.n_bins - 1 : 0;
}

// --------------------------------------------------------- particle events --
__device__ __forceinline__ void sphere_sample(Rng& r, float vx, float vy, float vz, float& gx, float& gy, float& gz) {
  const float rho = sqrtf(fminf(fmaxf(U(r), 0.0f), 0.99999994f));
  const float ci = sqrtf(fmaxf(0.0f, 1.0f - rho * rho));
  const float phi = U(r) * 2.0f * kPiF;
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(vx, vy, vz, e1x, e1y, e1z, e2x, e2y, e2z);
  const float cp = cosf(phi), sp = sinf(phi);
  gx = -ci * vx + rho * (cp * e1x + sp * e2x);
  gy = -ci * vy + rho * (cp * e1y + sp * e2y);
  gz = -ci * vz + rho * (cp * e1z + sp * e2z);
  normalize3(gx, gy, gz);
}
__device__ __forceinline__ bool sphere_accessible(const DevParams& P, double cx, double cy, double cz, double rr) {
  const double rc = hypot(cx, cy);
  const double wall = ((double)P.rin - rc) - rr, bottom = cz - rr - (double)P.zmin;
  const double rmin = fmax(0.0, rc - rr);
  const double top = free_surface_z_d(P, rmin * rmin) - (cz + rr);
  const double scale = fmax(fmax(fabs((double)P.rin), fabs(cz)), fmax(rr, 1.0e-6));
  const double tol = 128.0 * kDEps * scale;
  return wall >= -tol && bottom >= -tol && top >= -tol;
}
// ------------------------------------------------ rough particle surfaces --
// Smith microsurface random walk over a Beckmann slope distribution (Heitz,
// Hanika, d'Eon & Dachsbacher, ACM TOG 35(4), 2016), in the signed,
// log-domain form of CLARITAS.py:1792-2430, FP64 where the reference uses it.
// The per-event helpers are force-inlined: a call inside the pass loop makes
// the compiler save the whole live photon state around it, and those spills
// and reloads land on the common path of every pass.  The only calls left in
// the persistent kernels are the LDS tally clear and flush at configuration
// changes and, in the ROUGH/LIGHT/CAUSAL variants, the ring-face throat test
// and the LIGHT variant's voxel traversal.  The finite-particle event
// (particle_event) holds the only micro_interface site of each caller, so
// the walk is inlined twice per kernel.
constexpr int kMicroWatchdog = 4096;
constexpr double kSqrtPi = 1.77245385090551602729816748334114518;
constexpr double kInv2SqrtPi = 0.28209479177387814347403972578038629;

// B(x) = 1 - x sqrt(pi) erfcx(x), x >= 0; asymptotic series for x >= 5,
// CLARITAS.py:1926-1942
__device__ double beck_tail_B(double x) {
  if (!(x >= 0.0) || !isfinite(x)) return -1.0;
  double b;
  if (x < 5.0) {
    b = 1.0 - x * kSqrtPi * (exp(x * x) * erfc(x));
  } else {
    const double y = 1.0 / (x * x);
    // sum_k (-1)^(k+1) (2k-1)!! / 2^k y^k, k = 1..7
    b = 0.5 * y - 0.75 * y * y + 1.875 * y * y * y - 6.5625 * y * y * y * y + 29.53125 * y * y * y * y * y -
        162.421875 * y * y * y * y * y * y + 1055.7421875 * y * y * y * y * y * y * y;
  }
  return (isfinite(b) && b > 0.0) ? b : -1.0;
}
// visible-slope marginal CDF for slope_i >= -1, CLARITAS.py:1947-1962
__device__ bool beck_visible_cdf(double s, double si, double& out) {
  if (!isfinite(s) || !isfinite(si)) return false;
  if (s >= si) { out = 1.0; return true; }
  const double q = kInv2SqrtPi * exp(-si * si) + 0.5 * si * erfc(-si);
  if (!(q > 0.0) || !isfinite(q)) return false;
  const double n = kInv2SqrtPi * exp(-s * s) + 0.5 * si * erfc(-s);
  double f = n / q;
  if (!isfinite(f)) return false;
  if (f < 0.0 && f > -1.0e-13) f = 0.0;
  if (f > 1.0 && f < 1.0 + 1.0e-13) f = 1.0;
  if (f < 0.0 || f > 1.0) return false;
  out = f;
  return true;
}
// log of the scaled tail CDF at slope -(x+u), CLARITAS.py:1986-2013
__device__ bool beck_tail_logcdf(double x, double u, double Bx, double& out) {
  const double y = x + u, By = beck_tail_B(y);
  if (!(By > 0.0) || !isfinite(By)) return false;
  const double C = u / y + (x / y) * By;
  if (!(C > 0.0) || !isfinite(C)) return false;
  const double lr = -(2.0 * x * u + u * u) + log(C / Bx);
  if (!isfinite(lr)) return false;
  out = lr;
  return true;
}
// alpha = 1 visible slopes for signed cot(theta_i) = si, CLARITAS.py:1964-2045
__device__ bool beck_sample_slopes(double si, double u1, double u2, double& sxo, double& syo, int& fc) {
  fc = 0;
  if (!isfinite(u1) || !isfinite(u2) || !(u1 > 0.0 && u1 < 1.0) || !(u2 > 0.0 && u2 < 1.0)) { fc = MF_CDF; return false; }
  const double kCot1e4 = 9999.999966666666;  // cot(1e-4 rad): the visible law is the NDF
  if (isinf(si) || si > kCot1e4) {
    if (si < 0.0) { fc = MF_PROJ_AREA; return false; }
    const double r = sqrt(-log(u1)), ph = 2.0 * kPiD * u2, sx = r * cos(ph), sy = r * sin(ph);
    if (!isfinite(sx) || !isfinite(sy)) { fc = MF_INV_CDF; return false; }
    sxo = sx; syo = sy;
    return true;
  }
  if (!isfinite(si)) { fc = MF_PROJ_AREA; return false; }
  double sx;
  if (si < -1.0) {
    const double x = -si, Bx = beck_tail_B(x);
    if (!(Bx > 0.0) || !isfinite(Bx)) { fc = MF_PROJ_AREA; return false; }
    const double target = log(u1);
    double lo = 0.0, hi = fmax(1.0 / x, 1.0e-12), lr = 0.0;
    bool ok = false;
    for (int it = 0; it < 64; ++it) {
      if (!beck_tail_logcdf(x, hi, Bx, lr)) { fc = MF_CDF; return false; }
      if (lr <= target) { ok = true; break; }
      hi *= 2.0;
      if (!isfinite(hi)) { fc = MF_INV_CDF; return false; }
    }
    if (!ok) { fc = MF_INV_CDF; return false; }
    for (int it = 0; it < 52; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (!beck_tail_logcdf(x, mid, Bx, lr)) { fc = MF_CDF; return false; }
      if (lr > target) lo = mid; else hi = mid;
    }
    sx = -(x + 0.5 * (lo + hi));
  } else {
    double lo = fmin(-1.0, si - 1.0), hi = si, f = 1.0;
    bool ok = false;
    for (int it = 0; it < 64; ++it) {
      if (!beck_visible_cdf(lo, si, f)) { fc = MF_CDF; return false; }
      if (f <= u1) { ok = true; break; }
      lo = (lo < 0.0) ? 2.0 * lo : lo - 1.0;
      if (!isfinite(lo)) { fc = MF_INV_CDF; return false; }
    }
    if (!ok) { fc = MF_INV_CDF; return false; }
    for (int it = 0; it < 48; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (!beck_visible_cdf(mid, si, f)) { fc = MF_CDF; return false; }
      if (f < u1) lo = mid; else hi = mid;
    }
    sx = 0.5 * (lo + hi);
  }
  const double ey = 2.0 * u2 - 1.0;
  if (!(ey > -1.0 && ey < 1.0) || !isfinite(ey)) { fc = MF_INV_CDF; return false; }
  const double sy = erfinv(ey);
  if (!isfinite(sx) || !isfinite(sy)) { fc = MF_INV_CDF; return false; }
  sxo = sx; syo = sy;
  return true;
}
// visible microfacet normal about the TOP axis b for direction w (signed),
// two open uniforms, CLARITAS.py:2073-2129
__device__ bool beck_vndf_normal(Rng& r, float alpha, float bx, float by, float bz, float wx, float wy, float wz,
                                 double& mx, double& my, double& mz, int& fc) {
  fc = 0;
  if (!(alpha >= 0.0f) || !isfinite(alpha)) { fc = MF_NORMALISATION; return false; }
  normalize3(bx, by, bz);
  normalize3(wx, wy, wz);
  if (!isfinite(bx) || !isfinite(by) || !isfinite(bz) || !isfinite(wx) || !isfinite(wy) || !isfinite(wz)) {
    fc = MF_NONFINITE_DIRECTION; return false;
  }
  if (alpha <= 1.0e-8f) {
    const double vd = (double)wx * bx + (double)wy * by + (double)wz * bz;
    if (!(vd > 0.0)) { fc = MF_VISIBILITY; return false; }
    mx = bx; my = by; mz = bz;
    return true;
  }
  float t1x, t1y, t1z, t2x, t2y, t2z;
  perpendicular_basis(bx, by, bz, t1x, t1y, t1z, t2x, t2y, t2z);
  const double lx = (double)wx * t1x + (double)wy * t1y + (double)wz * t1z;
  const double ly = (double)wx * t2x + (double)wy * t2y + (double)wz * t2z;
  const double lz = (double)wx * bx + (double)wy * by + (double)wz * bz;
  double sxv = (double)alpha * lx, syv = (double)alpha * ly, szv = lz;
  const double vl = sqrt(sxv * sxv + syv * syv + szv * szv);
  if (!(vl > 0.0) || !isfinite(vl)) { fc = MF_NORMALISATION; return false; }
  sxv /= vl; syv /= vl; szv /= vl;
  const double radial = hypot(sxv, syv);
  double si;
  if (radial == 0.0) {
    if (szv > 0.0) si = __builtin_inf();
    else { fc = MF_PROJ_AREA; return false; }
  } else {
    si = szv / radial;
    if (!isfinite(si)) { fc = MF_PROJ_AREA; return false; }
  }
  const double u1 = Uopen(r), u2 = Uopen(r);
  double s0x = 0.0, s0y = 0.0;
  int f2 = 0;
  if (!beck_sample_slopes(si, u1, u2, s0x, s0y, f2)) { fc = f2 ? f2 : MF_INV_CDF; return false; }
  const double ph = atan2(syv, sxv), cp = cos(ph), sp = sin(ph);
  const double tx = cp * s0x - sp * s0y, ty = sp * s0x + cp * s0y;
  const double sx = (double)alpha * tx, sy = (double)alpha * ty;
  if (!isfinite(sx) || !isfinite(sy)) { fc = MF_NORMAL; return false; }
  const double inv = 1.0 / sqrt(sx * sx + sy * sy + 1.0);
  if (!(inv > 0.0) || !isfinite(inv)) { fc = MF_NORMALISATION; return false; }
  const double nlx = -sx * inv, nly = -sy * inv, nlz = inv;
  double nx = nlx * t1x + nly * t2x + nlz * bx;
  double ny = nlx * t1y + nly * t2y + nlz * by;
  double nz = nlx * t1z + nly * t2z + nlz * bz;
  const double nn = sqrt(nx * nx + ny * ny + nz * nz);
  if (!(nn > 0.0) || !isfinite(nn)) { fc = MF_NORMALISATION; return false; }
  nx /= nn; ny /= nn; nz /= nn;
  if (!(nx * bx + ny * by + nz * bz > 0.0)) { fc = MF_NORMAL; return false; }
  if (!(nx * wx + ny * wy + nz * wz > 0.0)) { fc = MF_VISIBILITY; return false; }
  mx = nx; my = ny; mz = nz;
  return true;
}
// signed Smith Lambda over the whole sphere of directions, CLARITAS.py:2162-2184
__device__ double smith_lambda(float wx, float wy, float wz, float bx, float by, float bz, float alpha) {
  double c = (double)wx * bx + (double)wy * by + (double)wz * bz;
  if (!isfinite(c) || !isfinite(alpha) || !(alpha >= 0.0f)) return __builtin_nan("");
  c = fmax(-1.0, fmin(1.0, c));
  const double ss2 = fmax(0.0, 1.0 - c * c);
  if (ss2 == 0.0) return c >= 0.0 ? 0.0 : -1.0;
  if (c == 0.0) return copysign(__builtin_inf(), c);
  if (alpha == 0.0f) return c >= 0.0 ? 0.0 : -1.0;
  const double aa = fabs(c) / ((double)alpha * sqrt(ss2));
  if (!(aa > 0.0) || !isfinite(aa)) return c >= 0.0 ? 0.0 : -1.0;
  const double B = beck_tail_B(aa);
  if (!(B > 0.0) || !isfinite(B)) return __builtin_nan("");
  const double lp = exp(-(aa * aa) - log(2.0 * kSqrtPi * aa)) * B;
  if (!(lp >= 0.0) || !isfinite(lp)) return __builtin_nan("");
  return c >= 0.0 ? lp : (-1.0 - lp);
}
__device__ __forceinline__ double uniform_C1(double h) { return fmin(1.0, fmax(0.0, 0.5 * (h + 1.0))); }
__device__ __forceinline__ double uniform_invC1(double u) { return 2.0 * fmin(1.0, fmax(0.0, u)) - 1.0; }
// next height of the walk, CLARITAS.py:2211-2310
__device__ int smith_height(double U, float wx, float wy, float wz, float bx, float by, float bz, float alpha,
                            float h, float& out_h, int& fc) {
  fc = 0;
  if (!isfinite(U) || !isfinite(wx) || !isfinite(wy) || !isfinite(wz) || !isfinite(h) || !(alpha >= 0.0f) ||
      !isfinite(alpha) || !(U > 0.0 && U < 1.0)) {
    fc = MF_RANDOM; return MS_NUMERIC;
  }
  double c = (double)wx * bx + (double)wy * by + (double)wz * bz;
  if (!isfinite(c)) { fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC; }
  c = fmax(-1.0, fmin(1.0, c));
  const double C0 = uniform_C1((double)h);
  if (!(C0 >= 0.0 && C0 <= 1.0) || !isfinite(C0)) { fc = MF_C0; return MS_NUMERIC; }
  const double ss2 = fmax(0.0, 1.0 - c * c);
  if (ss2 == 0.0) {
    if (c > 0.0) return MS_ESCAPE;
    if (c < 0.0) {
      const double hn = uniform_invC1(U * C0);
      if (!isfinite(hn)) { fc = MF_HEIGHT_INV; return MS_NUMERIC; }
      out_h = (float)hn;
      return MS_HIT;
    }
    out_h = h;
    return MS_HIT;
  }
  if (c == 0.0) { out_h = h; return MS_HIT; }
  const double lam = smith_lambda(wx, wy, wz, bx, by, bz, alpha);
  if (isnan(lam)) { fc = MF_LAMBDA; return MS_NUMERIC; }
  if (c > 0.0) {
    if (!(lam >= 0.0)) { fc = MF_LAMBDA; return MS_NUMERIC; }
    if (C0 <= 0.0 || C0 >= 1.0 || lam == 0.0) return MS_ESCAPE;
    const double lc0 = log(C0);
    if (!isfinite(lc0)) { fc = MF_LOG_C0; return MS_NUMERIC; }
    const double lg1 = lam * lc0;
    if (!(lg1 <= 0.0) || !isfinite(lg1)) { fc = MF_G1; return MS_NUMERIC; }
    const double p_hit = -expm1(lg1);
    if (!(p_hit >= 0.0 && p_hit <= 1.0) || !isfinite(p_hit)) { fc = MF_ESCAPE_PROB; return MS_NUMERIC; }
    if (U > p_hit) return MS_ESCAPE;
    const double l1u = log1p(-U);
    if (!isfinite(l1u)) { fc = MF_RANDOM; return MS_NUMERIC; }
    double lcn = lc0 - l1u / lam;
    if (!isfinite(lcn)) { fc = MF_LOG_CNEXT; return MS_NUMERIC; }
    if (lcn > 64.0 * kDEps) { fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
    if (lcn > 0.0) lcn = 0.0;
    const double cn = exp(lcn);
    if (!(cn >= 0.0 && cn <= 1.0) || !isfinite(cn)) { fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
    const double hn = uniform_invC1(cn);
    if (!isfinite(hn)) { fc = MF_HEIGHT_INV; return MS_NUMERIC; }
    out_h = (float)hn;
    if (!isfinite(out_h)) { fc = MF_HEIGHT_OUT; return MS_NUMERIC; }
    return MS_HIT;
  }
  if (!(lam <= -1.0) || !isfinite(lam)) { fc = MF_LAMBDA; return MS_NUMERIC; }
  if (C0 <= 0.0) { out_h = -1.0f; return MS_HIT; }
  const double lc0 = log(C0);
  if (!isfinite(lc0)) { fc = MF_LOG_C0; return MS_NUMERIC; }
  const double l1u = log1p(-U);
  if (!isfinite(l1u)) { fc = MF_RANDOM; return MS_NUMERIC; }
  const double lcn = lc0 - l1u / lam;
  if (!isfinite(lcn)) { fc = MF_LOG_CNEXT; return MS_NUMERIC; }
  if (lcn > 0.0) { fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
  const double cn = exp(lcn);
  if (!(cn >= 0.0 && cn <= C0 * (1.0 + 8.0e-15)) || !isfinite(cn)) { fc = MF_CNEXT_RANGE; return MS_NUMERIC; }
  const double hn = uniform_invC1(cn);
  if (!isfinite(hn)) { fc = MF_HEIGHT_INV; return MS_NUMERIC; }
  out_h = (float)hn;
  if (!isfinite(out_h)) { fc = MF_HEIGHT_OUT; return MS_NUMERIC; }
  return MS_HIT;
}

// direction in/out and the outcome of one interface encounter
struct MicroIO {
  float dx, dy, dz;
  int top, fc, order;  // order: microfacet scatterings of the walk
};
// rough-surface walk (alpha > 1e-7) from an incident direction that points
// into the TOP side, CLARITAS.py:2358-2430
__device__ __forceinline__ int rough_walk(Rng* __restrict__ rp, float alpha, float bx, float by, float bz,
                                       float n_top, float k_top, float n_bot, float k_bot,
                                       MicroIO* __restrict__ io) {
  Rng r = *rp;
  float dx = io->dx, dy = io->dy, dz = io->dz;
  int status = MS_WATCHDOG, fc = MF_WATCHDOG, top = 0, order = 0;
  float h = 1.0f + (2.0f * 0.999f - 1.0f);
  bool outside = true;
  for (int it = 0; it < kMicroWatchdog; ++it) {
    float nh = 0.0f;
    int hs, f2 = 0;
    const double Uh = Uopen(r);
    if (outside) {
      hs = smith_height(Uh, dx, dy, dz, bx, by, bz, alpha, h, nh, f2);
      if (hs == MS_ESCAPE) { status = MS_ESCAPE; fc = 0; top = 1; break; }
    } else {
      float t = 0.0f;
      hs = smith_height(Uh, -dx, -dy, -dz, bx, by, bz, alpha, -h, t, f2);
      if (hs == MS_ESCAPE) { status = MS_ESCAPE; fc = 0; top = 0; break; }
      if (hs == MS_HIT) nh = -t;
    }
    if (hs != MS_HIT || !isfinite(nh)) { status = MS_NUMERIC; fc = f2 ? f2 : MF_HEIGHT_SAMPLE; break; }
    h = nh;
    ++order;
    const float wix = -dx, wiy = -dy, wiz = -dz;
    double mx = 0.0, my = 0.0, mz = 0.0;
    f2 = 0;
    bool ok;
    if (outside) {
      ok = beck_vndf_normal(r, alpha, bx, by, bz, wix, wiy, wiz, mx, my, mz, f2);
    } else {
      ok = beck_vndf_normal(r, alpha, bx, by, bz, -wix, -wiy, -wiz, mx, my, mz, f2);
      mx = -mx; my = -my; mz = -mz;
    }
    if (!ok) { status = MS_NUMERIC; fc = f2 ? f2 : MF_INV_CDF; break; }
    const double vis = (double)wix * mx + (double)wiy * my + (double)wiz * mz;
    if (!isfinite(mx) || !isfinite(my) || !isfinite(mz)) { status = MS_NUMERIC; fc = MF_NORMAL; break; }
    if (!(vis > 0.0)) { status = MS_NUMERIC; fc = MF_VISIBILITY; break; }
    const float ci = (float)fmin(1.0, fmax(0.0, vis));
    const float n1 = outside ? n_top : n_bot, k1 = outside ? k_top : k_bot;
    const float n2 = outside ? n_bot : n_top, k2 = outside ? k_bot : k_top;
    if (!(n1 > 0.0f && n2 > 0.0f)) { status = MS_NUMERIC; fc = MF_REFRACTION_INVALID; break; }
    float R, ts, tn;
    if (!complex_interface(ci, n1, k1, n2, k2, R, ts, tn) || !isfinite(R)) {
      status = MS_NUMERIC; fc = MF_FRESNEL_INVALID; break;
    }
    double ox, oy, oz;
    if (U(r) < R) {
      const double dd = (double)dx * mx + (double)dy * my + (double)dz * mz;
      ox = (double)dx - 2.0 * dd * mx; oy = (double)dy - 2.0 * dd * my; oz = (double)dz - 2.0 * dd * mz;
    } else {
      float tx, ty, tz;
      if (!energy_refract3(dx, dy, dz, (float)mx, (float)my, (float)mz, n1, k1, n2, k2, tx, ty, tz)) {
        status = MS_NUMERIC; fc = MF_REFRACTION_INVALID; break;
      }
      ox = tx; oy = ty; oz = tz;
      outside = !outside;
    }
    const double qn = sqrt(ox * ox + oy * oy + oz * oz);
    if (!(qn > 0.0) || !isfinite(qn) || !isfinite(ox) || !isfinite(oy) || !isfinite(oz)) {
      status = MS_NUMERIC; fc = MF_REFRACTION_INVALID; break;
    }
    float fx = (float)(ox / qn), fy = (float)(oy / qn), fz = (float)(oz / qn);
    if (!isfinite(fx) || !isfinite(fy) || !isfinite(fz)) { status = MS_NUMERIC; fc = MF_NONFINITE_DIRECTION; break; }
    normalize3(fx, fy, fz);
    if (!isfinite(fx) || !isfinite(fy) || !isfinite(fz)) { status = MS_NUMERIC; fc = MF_NONFINITE_DIRECTION; break; }
    dx = fx; dy = fy; dz = fz;
  }
  *rp = r;
  io->dx = dx; io->dy = dy; io->dz = dz; io->top = top; io->fc = fc; io->order = order;
  return status;
}
// smooth limit, CLARITAS.py:2336-2356
__device__ __forceinline__ int smooth_interface(Rng& r, float bx, float by, float bz, float n_top, float k_top,
                                                float n_bot, float k_bot, float& dx, float& dy, float& dz,
                                                bool& out_top, int& fc) {
  const float ci = fminf(fmaxf(dot3(-dx, -dy, -dz, bx, by, bz), 0.0f), 1.0f);
  float R, ts, tn;
  if (!complex_interface(ci, n_top, k_top, n_bot, k_bot, R, ts, tn) || !isfinite(R)) {
    fc = MF_FRESNEL_INVALID; return MS_NUMERIC;
  }
  float ox = dx, oy = dy, oz = dz;
  if (U(r) < R) {
    reflect3(ox, oy, oz, bx, by, bz);
    out_top = true;
    if (!isfinite(ox) || !isfinite(oy) || !isfinite(oz)) { fc = MF_REFLECTION_INVALID; return MS_NUMERIC; }
  } else {
    if (!energy_refract3(dx, dy, dz, bx, by, bz, n_top, k_top, n_bot, k_bot, ox, oy, oz)) {
      fc = MF_REFRACTION_INVALID; return MS_NUMERIC;
    }
    out_top = false;
  }
  if (!isfinite(ox) || !isfinite(oy) || !isfinite(oz)) { fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC; }
  dx = ox; dy = oy; dz = oz;
  return MS_ESCAPE;
}
// one encounter of direction d with a (rough) dielectric interface whose TOP
// side has normal b: MS_ESCAPE with out_top (left on the top side) or not
// (crossed), else MS_NUMERIC / MS_WATCHDOG with fc.  CLARITAS.py:2314-2430
template <bool kRough>
__device__ __forceinline__ int micro_interface(Rng& r, float alpha, float bx, float by, float bz, float n_top,
                                               float k_top, float n_bot, float k_bot, float& dx, float& dy,
                                               float& dz, bool& out_top, int& fc, int& order) {
  fc = 0;
  order = 0;
  normalize3(bx, by, bz);
  normalize3(dx, dy, dz);
  if (!isfinite(bx) || !isfinite(by) || !isfinite(bz) || !isfinite(dx) || !isfinite(dy) || !isfinite(dz)) {
    fc = MF_NONFINITE_DIRECTION; return MS_NUMERIC;
  }
  if (dot3(dx, dy, dz, bx, by, bz) >= -1.0e-7f) { fc = MF_SIDE_STATE; return MS_NUMERIC; }
  if (!kRough || alpha <= 1.0e-7f) {
    const int st = smooth_interface(r, bx, by, bz, n_top, k_top, n_bot, k_bot, dx, dy, dz, out_top, fc);
    order = (st == MS_ESCAPE) ? 1 : 0;
    return st;
  }
  Rng t = r;
  MicroIO io{dx, dy, dz, 0, 0, 0};
  const int st = rough_walk(&t, alpha, bx, by, bz, n_top, k_top, n_bot, k_bot, &io);
  r = t;
  dx = io.dx; dy = io.dy; dz = io.dz; out_top = io.top != 0; fc = io.fc; order = io.order;
  return st;
}

// inverse-CDF polar angle on the float table, CLARITAS.py:2632-2641
__device__ __forceinline__ float sample_mie_theta(Rng& r, const DevParams& P, int pidx) {
  const int nt = P.n_theta;
  const float* __restrict__ cdf = P.cdf + (size_t)pidx * nt;
  const uint32_t m = rng_next(r) >> 9;
  const float u = u_of(m);
  const int lo = guided_lower_bound(cdf, P.theta_guide + (size_t)pidx * kGuideN, m, u);
  if (lo <= 0) return P.theta[0];
  const float c0 = cdf[lo - 1], c1 = cdf[lo], t0 = P.theta[lo - 1], t1 = P.theta[lo];
  const float f = (c1 > c0) ? fminf(fmaxf((u - c0) / (c1 - c0), 0.0f), 1.0f) : 0.0f;
  return t0 + f * (t1 - t0);
}
__device__ __forceinline__ void mie_scatter(Rng& r, const DevParams& P, int pidx, float& vx, float& vy, float& vz) {
  const float th = sample_mie_theta(r, P, pidx);
  const float phi = 2.0f * kPiF * U(r);
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(vx, vy, vz, e1x, e1y, e1z, e2x, e2y, e2z);
  const float st = sinf(th), ct = cosf(th), cp = cosf(phi), sp = sinf(phi);
  const float ox = ct * vx + st * (cp * e1x + sp * e2x);
  const float oy = ct * vy + st * (cp * e1y + sp * e2y);
  const float oz = ct * vz + st * (cp * e1z + sp * e2z);
  vx = ox; vy = oy; vz = oz;
  normalize3(vx, vy, vz);
}

// ------------------------------------------- incoherent wave channel event --
// direction-only finite-particle event on a volume-preserving spheroid
// a = r q^(2/3), b = r q^(-1/3) (q = a/b), CLARITAS.py:2655-2866
__device__ __forceinline__ void random_axis(Rng& r, float& ux, float& uy, float& uz) {
  const float cz = 2.0f * U(r) - 1.0f, az = 2.0f * kPiF * U(r), sz = sqrtf(fmaxf(0.0f, 1.0f - cz * cz));
  ux = sz * cosf(az); uy = sz * sinf(az); uz = cz;
}
// isotropic, or the head-tail symmetric von Mises-Fisher law about the local
// flow direction f, CLARITAS.py:2675-2702
__device__ void spheroid_axis(Rng& r, int model, float kappa, float fx, float fy, float fz, float& ux, float& uy,
                              float& uz) {
  if (model == 0 || !(kappa > 1e-7f) || !isfinite(kappa)) { random_axis(r, ux, uy, uz); return; }
  normalize3(fx, fy, fz);
  if (!(fx * fx + fy * fy + fz * fz > 0.5f)) { random_axis(r, ux, uy, uz); return; }
  const float u = fminf(fmaxf(U(r), 1e-12f), 1.0f - 1e-7f);
  const float em2 = expf(-2.0f * kappa);
  float c = 1.0f + logf(u + (1.0f - u) * em2) / kappa;
  c = fminf(fmaxf(c, -1.0f), 1.0f);
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(fx, fy, fz, e1x, e1y, e1z, e2x, e2y, e2z);
  const float ph = 2.0f * kPiF * U(r), st = sqrtf(fmaxf(0.0f, 1.0f - c * c)), cp = cosf(ph), sp = sinf(ph);
  ux = c * fx + st * (cp * e1x + sp * e2x);
  uy = c * fy + st * (cp * e1y + sp * e2y);
  uz = c * fz + st * (cp * e1z + sp * e2z);
  if (U(r) < 0.5f) { ux = -ux; uy = -uy; uz = -uz; }
  normalize3(ux, uy, uz);
}
__device__ __forceinline__ void spheroid_normal(float x, float y, float z, float ux, float uy, float uz, float invb2,
                                                float delta, float& nx, float& ny, float& nz) {
  const float pu = x * ux + y * uy + z * uz;
  nx = invb2 * x + delta * pu * ux; ny = invb2 * y + delta * pu * uy; nz = invb2 * z + delta * pu * uz;
  normalize3(nx, ny, nz);
}
__device__ __forceinline__ bool spheroid_roots(float ox, float oy, float oz, float dx, float dy, float dz, float ux,
                                               float uy, float uz, float invb2, float delta, float& t0, float& t1) {
  const float du = dx * ux + dy * uy + dz * uz, ou = ox * ux + oy * uy + oz * uz;
  const float A = invb2 * (dx * dx + dy * dy + dz * dz) + delta * du * du;
  const float B = 2.0f * (invb2 * (ox * dx + oy * dy + oz * dz) + delta * ou * du);
  const float Cc = invb2 * (ox * ox + oy * oy + oz * oz) + delta * ou * ou - 1.0f;
  const float disc = B * B - 4.0f * A * Cc;
  if (!(A > 0.0f) || disc < 0.0f) return false;
  const float sd = sqrtf(fmaxf(0.0f, disc));
  float r0 = (-B - sd) / (2.0f * A), r1 = (-B + sd) / (2.0f * A);
  if (r0 > r1) { const float q = r0; r0 = r1; r1 = q; }
  t0 = r0; t1 = r1;
  return true;
}
// entry point uniform over the projected silhouette, CLARITAS.py:2745-2771
__device__ bool spheroid_entry(Rng& r, float vx, float vy, float vz, float a, float b, float ux, float uy, float uz,
                               float& x, float& y, float& z, float& gnx, float& gny, float& gnz) {
  const float invb2 = 1.0f / (b * b), delta = 1.0f / (a * a) - invb2;
  float e1x, e1y, e1z, e2x, e2y, e2z;
  perpendicular_basis(vx, vy, vz, e1x, e1y, e1z, e2x, e2y, e2z);
  const float du = dot3(ux, uy, uz, vx, vy, vz);
  const float rproj2 = b * b + (a * a - b * b) * fmaxf(0.0f, 1.0f - du * du);
  const float rbound = sqrtf(fmaxf(b * b, rproj2)), rmax = fmaxf(a, b);
  for (int attempt = 0; attempt < 64; ++attempt) {
    const float rr = rbound * sqrtf(U(r)), ph = 2.0f * kPiF * U(r);
    const float ox = rr * (cosf(ph) * e1x + sinf(ph) * e2x) - 2.5f * rmax * vx;
    const float oy = rr * (cosf(ph) * e1y + sinf(ph) * e2y) - 2.5f * rmax * vy;
    const float oz = rr * (cosf(ph) * e1z + sinf(ph) * e2z) - 2.5f * rmax * vz;
    float t0, t1;
    if (!spheroid_roots(ox, oy, oz, vx, vy, vz, ux, uy, uz, invb2, delta, t0, t1)) continue;
    const float t = (t0 > 1e-9f * rmax) ? t0 : ((t1 > 1e-9f * rmax) ? t1 : -1.0f);
    if (t <= 0.0f) continue;
    x = ox + t * vx; y = oy + t * vy; z = oz + t * vz;
    spheroid_normal(x, y, z, ux, uy, uz, invb2, delta, gnx, gny, gnz);
    if (dot3(vx, vy, vz, gnx, gny, gnz) < -1e-6f) return true;
  }
  return false;
}
__device__ __forceinline__ float spheroid_chord(float x, float y, float z, float vx, float vy, float vz, float ux,
                                                float uy, float uz, float invb2, float delta, float scale) {
  float t0, t1;
  if (!spheroid_roots(x, y, z, vx, vy, vz, ux, uy, uz, invb2, delta, t0, t1)) return -1.0f;
  const float eps = fmaxf(1e-5f * scale, 1e-12f);
  float t = kInfF;
  if (t0 > eps) t = t0;
  if (t1 > eps && t1 < t) t = t1;
  return t < kInfF * 0.5f ? t : -1.0f;
}
struct PCount {
  int pr, per, pir, pabs;
  uint32_t n_abs;
  // the encounter's observational record (CLARITAS.py:2503-2599): internal
  // path (float sum of chords), entry incidence and exit angles (deg, -1
  // undefined), last chord index, microfacet order of the last walk
  float path, inc, ext;
  int macro, micro;
};
// One finite-particle encounter with smooth or rough surfaces: a Fresnel
// sphere of the given radius whose centre is c (the entry normal g is the
// sampled surface point's), or a volume-preserving spheroid a = r q^(2/3),
// b = r q^(-1/3) about the local origin (direction only; the axis, the entry
// point and its normal are sampled here).  The entry and every internal
// encounter share one micro_interface site, so the walk is inlined once per
// caller.  Returns 1 escaped, 2 absorbed, 0 failure (pf, pfs).
// CLARITAS.py:2497-2599 (sphere), 2783-2841 (spheroid)
template <bool kRough>
__device__ __forceinline__ int particle_event(const DevParams& P, Rng& r, bool spheroid, float n_m, float n_p,
                                              float k_p, float radius, float gx, float gy, float gz, double cx,
                                              double cy, double cz, float fx, float fy, float fz, float& x, float& y,
                                              float& z, float& vx, float& vy, float& vz, PCount& pc, int& pf,
                                              int& pfs) {
  pf = PF_OK; pfs = SUB_NONE;
  pc.path = 0.0f; pc.inc = -1.0f; pc.ext = -1.0f; pc.macro = 0; pc.micro = 0;
  if (radius <= 0.0f) return 1;
  float ux = 0.0f, uy = 0.0f, uz = 0.0f, invb2 = 0.0f, delta = 0.0f, scale = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
  if (!kRough) spheroid = false;  // the morphology channel is reached only in the ROUGH variant
  if (spheroid) {
    const float q13 = powf(P.aspect, 1.0f / 3.0f), a = radius * q13 * q13, b = radius / q13;
    if (!(a > 0.0f && b > 0.0f && isfinite(a) && isfinite(b))) { pf = PF_SPHEROID; pfs = SUB_SPHEROID_AXES; return 0; }
    spheroid_axis(r, P.orientation, P.kappa, fx, fy, fz, ux, uy, uz);
    if (!spheroid_entry(r, vx, vy, vz, a, b, ux, uy, uz, sx, sy, sz, gx, gy, gz)) {
      pf = PF_SPHEROID; pfs = SUB_SPHEROID_ENTRY; return 0;
    }
    invb2 = 1.0f / (b * b); delta = 1.0f / (a * a) - invb2; scale = fmaxf(a, b);
  }
  const int pf_geo = spheroid ? PF_SPHEROID : PF_MACRO_GEOMETRY;
  const double rr = (double)radius, rr2 = rr * rr;
  double qx = 0.0, qy = 0.0, qz = 0.0;  // sphere: position relative to the centre (FP64)
  const float alpha_abs = (P.wavelength > 0.0f && k_p > 0.0f) ? (4.0f * kPiF * k_p / P.wavelength) : 0.0f;
  float bx = gx, by = gy, bz = gz;       // TOP normal of the current encounter
  float nox = 0.0f, noy = 0.0f, noz = 0.0f;  // outward normal at an internal encounter
  float n_top = n_m, k_top = 0.0f, n_bot = n_p, k_bot = k_p;
  for (int stage = 0;; ++stage) {
    bool top = true;
    int fc = 0, order = 0;
    if (stage == 0) pc.inc = acosf(fminf(fmaxf(-dot3(vx, vy, vz, gx, gy, gz), 0.0f), 1.0f)) * (180.0f / kPiF);
    const int ms = micro_interface<kRough>(r, P.rough, bx, by, bz, n_top, k_top, n_bot, k_bot, vx, vy, vz, top, fc,
                                           order);
    pc.micro = order;
    if (ms != MS_ESCAPE) { pf = (ms == MS_WATCHDOG) ? PF_MICRO_WATCHDOG : PF_MICRO_NUMERICAL; pfs = fc; return 0; }
    if (stage == 0) {
      if (top) {
        pc.ext = acosf(fminf(fmaxf(dot3(vx, vy, vz, gx, gy, gz), 0.0f), 1.0f)) * (180.0f / kPiF);
        pc.pr++; pc.per++;
        return 1;
      }
      if (dot3(vx, vy, vz, gx, gy, gz) >= 0.0f) {
        pf = pf_geo; pfs = spheroid ? SUB_SPHEROID_ENTRY : SUB_SPHERE_ENTRY_SIDE; return 0;
      }
      if (!spheroid) {
        qx = rr * (double)gx; qy = rr * (double)gy; qz = rr * (double)gz;
        const double qn0 = sqrt(qx * qx + qy * qy + qz * qz);
        if (!(qn0 > 0.0) || !isfinite(qn0)) { pf = PF_MACRO_GEOMETRY; pfs = SUB_SPHERE_CHORD; return 0; }
        qx *= rr / qn0; qy *= rr / qn0; qz *= rr / qn0;
      }
      n_top = n_p; k_top = k_p; n_bot = n_m; k_bot = 0.0f;
    } else {
      if (!top) {
        if (dot3(vx, vy, vz, nox, noy, noz) <= 0.0f) {
          pf = pf_geo; pfs = spheroid ? SUB_SPHEROID_EXIT_SIDE : SUB_SPHERE_EXIT_SIDE; return 0;
        }
        pc.ext = acosf(fminf(fmaxf(dot3(vx, vy, vz, nox, noy, noz), 0.0f), 1.0f)) * (180.0f / kPiF);
        return 1;
      }
      if (dot3(vx, vy, vz, nox, noy, noz) >= 0.0f) {
        pf = pf_geo; pfs = spheroid ? SUB_SPHEROID_INTERNAL_SIDE : SUB_SPHERE_INTERNAL_SIDE; return 0;
      }
      pc.pr++; pc.pir++;
      if (stage - 1 == P.max_pb) {
        pf = PF_MACRO_BOUNCE; pfs = spheroid ? SUB_SPHEROID_BOUNCE_LIMIT : SUB_SPHERE_BOUNCE_LIMIT; return 0;
      }
    }
    if (stage > P.max_pb) { pf = PF_OTHER; pfs = SUB_NONE; return 0; }
    // chord of internal bounce `stage` to the next encounter
    pc.macro = stage;
    if (kRough && spheroid) {
      const float chord = spheroid_chord(sx, sy, sz, vx, vy, vz, ux, uy, uz, invb2, delta, scale);
      if (chord <= 0.0f) { pf = PF_SPHEROID; pfs = SUB_SPHEROID_CHORD; return 0; }
      if (alpha_abs > 0.0f) {
        pc.n_abs++;
        const double dist = -log(Uopen(r)) / (double)alpha_abs;
        if (dist < (double)chord) { pc.path += (float)dist; pc.pabs++; return 2; }
      }
      sx += chord * vx; sy += chord * vy; sz += chord * vz;
      pc.path += chord;
      spheroid_normal(sx, sy, sz, ux, uy, uz, invb2, delta, nox, noy, noz);
    } else {
      const double dvx = vx, dvy = vy, dvz = vz;
      const double A = dvx * dvx + dvy * dvy + dvz * dvz, q2 = qx * qx + qy * qy + qz * qz;
      if (!(A > 0.0) || !isfinite(A) || !isfinite(q2)) { pf = PF_MACRO_GEOMETRY; pfs = SUB_SPHERE_CHORD; return 0; }
      const double qnorm = sqrt(q2);
      if (!(qnorm > 0.0) || !isfinite(qnorm)) { pf = PF_MACRO_GEOMETRY; pfs = SUB_SPHERE_CHORD; return 0; }
      const double qp = rr / qnorm;
      qx *= qp; qy *= qp; qz *= qp;
      const double B = qx * dvx + qy * dvy + qz * dvz;
      if (!(B < 0.0) || !isfinite(B)) { pf = PF_MACRO_GEOMETRY; pfs = SUB_SPHERE_INTERNAL_SIDE; return 0; }
      const double chord = -2.0 * B / A;
      if (!(chord > 0.0) || !isfinite(chord)) { pf = PF_MACRO_GEOMETRY; pfs = SUB_SPHERE_CHORD; return 0; }
      const double tqx = qx + chord * dvx, tqy = qy + chord * dvy, tqz = qz + chord * dvz;
      const double arr = tqx * dvx + tqy * dvy + tqz * dvz, tq2 = tqx * tqx + tqy * tqy + tqz * tqz;
      if (!(arr > 0.0) || !isfinite(arr) || !isfinite(tq2) || fabs(tq2 - rr2) > 512.0 * kDEps * fmax(rr2, 1.0e-30)) {
        pf = PF_MACRO_GEOMETRY; pfs = SUB_SPHERE_CHORD; return 0;
      }
      if (alpha_abs > 0.0f) {
        pc.n_abs++;
        const double dist = -log(Uopen(r)) / (double)alpha_abs;
        if (dist < chord) {
          qx += dist * dvx; qy += dist * dvy; qz += dist * dvz;
          x = (float)(cx + qx); y = (float)(cy + qy); z = (float)(cz + qz);
          pc.path += (float)dist;
          pc.pabs++;
          return 2;
        }
      }
      qx = tqx; qy = tqy; qz = tqz;
      pc.path += (float)chord;
      const double qn = sqrt(qx * qx + qy * qy + qz * qz);
      if (!(qn > 0.0) || !isfinite(qn)) { pf = PF_MACRO_GEOMETRY; pfs = SUB_SPHERE_CHORD; return 0; }
      const double pj = rr / qn;
      qx *= pj; qy *= pj; qz *= pj;
      x = (float)(cx + qx); y = (float)(cy + qy); z = (float)(cz + qz);
      nox = (float)(qx / rr); noy = (float)(qy / rr); noz = (float)(qz / rr);
      normalize3(nox, noy, noz);
    }
    bx = -nox; by = -noy; bz = -noz;
  }
}

struct CondOut {
  float x, y, z, vx, vy, vz;
  int status, channel, failure, subcode;
  PCount pc;
  uint32_t trials, rejects, inacc;
};
// conditional finite-particle closure, CLARITAS.py:4479-4536, in two parts:
// cond_begin checks the bin and draws the target outcome (absorbed with
// probability 1 - albedo); cond_trial runs one trial.  A trial samples the
// channel (coherent Mie, Fresnel sphere at an accessible centre, or the
// morphology wave) and the event; it is kept when its outcome (scattered /
// absorbed) matches the target.  The transport kernel runs one trial per
// CLOSE pass for every lane in the closure, so a lane that needs many trials
// (an absorbing target, accepted with the event's small absorption
// probability) no longer holds the rest of its wave; conditional_event runs
// the whole loop at once (device probes).
struct CondState {
  uint32_t trials, rejects, inacc, nabs;
  bool target_abs;
};
// 0: trials follow (st filled); otherwise the closure has ended with h
template <bool kRough>
__device__ __forceinline__ int cond_begin(const DevParams& P, Rng& r, int pidx, float x, float y, float z, float vx,
                                          float vy, float vz, int max_trials, CondState& st, CondOut& h) {
  h.x = x; h.y = y; h.z = z; h.vx = vx; h.vy = vy; h.vz = vz;
  h.status = 0; h.channel = 0; h.failure = PF_OTHER; h.subcode = 9002;
  h.pc = PCount{0, 0, 0, 0, 0u, 0.0f, -1.0f, -1.0f, 0, 0};
  h.trials = h.rejects = h.inacc = 0;
  st.trials = st.rejects = st.inacc = st.nabs = 0;
  st.target_abs = false;
  const double qs = P.qsca[pidx], qe = P.qext[pidx];
  const float ki = P.k_imag[pidx];
  const double tol = 1e-10 * fmax(1.0, fmax(fabs(qs), fabs(qe)));
  if (max_trials == 0) { h.subcode = 9001; return 1; }
  if (!isfinite(qs) || !isfinite(qe) || qe <= 0.0 || qs < -tol || qs > qe + tol || max_trials < 1) return 1;
  const bool lossless = (ki == 0.0f && fabs(qe - qs) <= tol);
  const double albedo = lossless ? 1.0 : fmin(fmax(qs / qe, 0.0), 1.0);
  const double pabs = lossless ? 0.0 : fmin(fmax((qe - qs) / qe, 0.0), 1.0);
  if ((ki > 0.0f && (pabs <= 0.0 || albedo >= 1.0)) || (albedo > 0.0 && albedo < 2.3283064365386962890625e-10) ||
      (pabs > 0.0 && pabs < 2.3283064365386962890625e-10)) {
    h.subcode = 9003; return 1;
  }
  st.target_abs = (albedo < 1.0 && U53(r) >= albedo);
  return 0;
}
// One trial from the closure's entry state (x, v).  Returns 0 to go on
// (rejected or inaccessible), 1 when the closure has ended with c (accepted,
// failed, or out of trials).
template <bool kRough>
__device__ __forceinline__ int cond_trial(const DevParams& P, Rng& r, int pidx, float x, float y, float z, float vx,
                                          float vy, float vz, int max_trials, CondState& st, CondOut& c) {
  const double qs = P.qsca[pidx];
  const float ki = P.k_imag[pidx], nr = P.n_real[pidx], radius = P.radius[pidx];
  const double wave = qs > 0.0 ? fmax(qs - 1.0, 0.0) / qs : 0.0;
  st.trials++;
  c.x = x; c.y = y; c.z = z; c.vx = vx; c.vy = vy; c.vz = vz;
  c.pc = PCount{0, 0, 0, 0, 0u, 0.0f, -1.0f, -1.0f, 0, 0};
  c.failure = PF_OK; c.subcode = 0; c.status = 0;
  c.channel = (U53(r) < wave) ? ((P.coherent >= 1.0f || U53(r) < (double)P.coherent) ? 1 : 2) : 0;
  bool go_on = false;
  if (c.channel == 1) {
    mie_scatter(r, P, pidx, c.vx, c.vy, c.vz);
    c.status = 1;
  } else {
    // channel 0: Fresnel sphere at a sampled accessible centre; channel 2:
    // morphology-wave event (direction only), CLARITAS.py:4512-4531
    const bool dir_only = (c.channel == 2);
    bool sph = false;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f, fx = 1.0f, fy = 0.0f;
    double dcx = 0.0, dcy = 0.0, dcz = 0.0;
    int pre = -1;  // outcome settled before the event
    if (dir_only) {
      const float fr = sqrtf(x * x + y * y);
      if (fr > 1e-12f) { fx = -y / fr; fy = x / fr; }
      if (radius <= 0.0f) {
        pre = 1;
      } else if (!(P.aspect > 0.0f) || !isfinite(P.aspect)) {
        pre = 0; c.failure = PF_SPHEROID; c.subcode = SUB_SPHEROID_ASPECT;
      } else if (fabsf(P.aspect - 1.0f) < 1e-7f) {
        sphere_sample(r, vx, vy, vz, gx, gy, gz);
        dcx = -(double)radius * (double)gx; dcy = -(double)radius * (double)gy; dcz = -(double)radius * (double)gz;
      } else {
        sph = true;
      }
    } else {
      sphere_sample(r, vx, vy, vz, gx, gy, gz);
      dcx = (double)x - (double)radius * gx; dcy = (double)y - (double)radius * gy;
      dcz = (double)z - (double)radius * gz;
      if (!sphere_accessible(P, dcx, dcy, dcz, radius)) { st.inacc++; go_on = true; }
    }
    if (!go_on) {
      if (pre >= 0) {
        c.status = pre;
      } else {
        float px = dir_only ? 0.0f : c.x, py = dir_only ? 0.0f : c.y, pz = dir_only ? 0.0f : c.z;
        float pvx = c.vx, pvy = c.vy, pvz = c.vz;
        int pf = 0, pfs = 0;
        c.status = particle_event<kRough>(P, r, sph, P.n_water, nr, ki, radius, gx, gy, gz, dcx, dcy, dcz, fx, fy,
                                          0.0f, px, py, pz, pvx, pvy, pvz, c.pc, pf, pfs);
        if (!dir_only) { c.x = px; c.y = py; c.z = pz; }
        if (!dir_only || sph || c.status == 1) { c.vx = pvx; c.vy = pvy; c.vz = pvz; }
        c.failure = pf; c.subcode = pfs;
      }
    }
  }
  if (!go_on) {
    st.nabs += c.pc.n_abs;
    if (c.status != 1 && c.status != 2) {
      c.trials = st.trials; c.rejects = st.rejects; c.inacc = st.inacc; c.pc.n_abs = st.nabs;
      return 1;
    }
    if ((c.status == 2) != st.target_abs) {
      st.rejects++;
    } else {
      c.trials = st.trials; c.rejects = st.rejects; c.inacc = st.inacc; c.pc.n_abs = st.nabs;
      c.failure = PF_OK; c.subcode = 0;
      return 1;
    }
  }
  if ((int)st.trials >= max_trials) {
    c.x = x; c.y = y; c.z = z; c.vx = vx; c.vy = vy; c.vz = vz;
    c.status = 0; c.channel = 0; c.failure = PF_OTHER; c.subcode = 9001;
    c.pc = PCount{0, 0, 0, 0, 0u, 0.0f, -1.0f, -1.0f, 0, 0};
    c.trials = st.trials; c.rejects = st.rejects; c.inacc = st.inacc; c.pc.n_abs = st.nabs;
    return 1;
  }
  return 0;
}
template <bool kRough>
__device__ __forceinline__ CondOut conditional_event(const DevParams& P, Rng& r, int pidx, float x, float y, float z,
                                                     float vx, float vy, float vz, int max_trials) {
  CondOut h;
  CondState st;
  if (cond_begin<kRough>(P, r, pidx, x, y, z, vx, vy, vz, max_trials, st, h)) return h;
  while (!cond_trial<kRough>(P, r, pidx, x, y, z, vx, vy, vz, max_trials, st, h)) {}
  return h;
}

// --------------------------------------------------------------- detectors --
// three-plane hard aperture of one bore, CLARITAS.py:4237-4255.  The ray
// meets plane k (u.p = a_k: ring_in, throat_out, ring_out) at
// t_k = (a_k - u.p) / (u.v); with e_k = a_k - u.p the offset from the axis
// point there, times u.v > 0, is (p - a_k u) (u.v) + e_k v, so each aperture
// test compares squared lengths scaled by (u.v)^2 and needs no division
// (the order t2 > t1 > t0 > 0 reduces to e_0 > 0 since ring_out > throat_out
// > ring_in).  A ray through both throats of radius r_t, a distance D apart
// along the axis, makes an angle with the axis of at most atan(2 r_t / D):
// rays less aligned than that (P.det_cos_min, with a margin) are rejected
// first.  Only an accepted ray's t_0 is divided out.
__device__ __forceinline__ bool bore_accept(const DevParams& P, float x, float y, float z, float vx,
                                            float vy, float vz, float ux, float uy, float& t_out) {
  const float den = vx * ux + vy * uy;
  if (den <= 1e-10f || den < P.det_cos_min) return false;
  const float pdot = x * ux + y * uy;
  const float e0 = P.ring_in - pdot;
  if (!(e0 > 0.0f)) return false;
  const float rth2 = P.throat_r * P.throat_r * (den * den), rcb2 = P.cb_r * P.cb_r * (den * den);
  const float xd = x * den, yd = y * den, zd = z * den;
  const float ax = ux * den, ay = uy * den;
  const float q0x = xd - P.ring_in * ax + e0 * vx, q0y = yd - P.ring_in * ay + e0 * vy, q0z = zd + e0 * vz;
  if (!(q0x * q0x + q0y * q0y + q0z * q0z <= rth2)) return false;
  const float e1 = P.throat_out - pdot;
  const float q1x = xd - P.throat_out * ax + e1 * vx, q1y = yd - P.throat_out * ay + e1 * vy, q1z = zd + e1 * vz;
  if (!(q1x * q1x + q1y * q1y + q1z * q1z <= rth2)) return false;
  const float e2 = P.ring_out - pdot;
  const float q2x = xd - P.ring_out * ax + e2 * vx, q2y = yd - P.ring_out * ay + e2 * vy, q2z = zd + e2 * vz;
  if (!(q2x * q2x + q2y * q2y + q2z * q2z <= rcb2)) return false;
  t_out = e0 / den;
  return true;
}
// First bore (smallest t0) whose three apertures the exit ray passes.  Only
// the bores within P.det_window of the bore nearest to the azimuth where the
// ray crosses r = ring_in are tested.  The host derives the window from the
// cell (window_for_cell in api.cpp): an accepted ray passes both throat
// apertures, which bounds its crossing azimuth's distance from the bore axis;
// the tested set then contains every bore the full 18-bore scan of
// CLARITAS.py:4241-4253 could accept, in the same order.
// bores tested around window centre jc for the ray (x, y, z, v)
__device__ __forceinline__ int bore_scan(const DevParams& P, float x, float y, float z, float vx, float vy, float vz,
                                         float deg) {
  if (deg < -90.0f) deg += 360.0f;  // bores lie at 0..170 deg: keep the azimuth continuous past 180
  const int jc = (int)rintf(deg * 0.1f), W = P.det_window;
  float best_t = kInfF, tt;
  int best = -1;
  for (int k = -W; k <= W; ++k) {
    const int j = jc + k;
    if (j < 0 || j >= CLR_N_CHANNELS) continue;
    if (bore_accept(P, x, y, z, vx, vy, vz, P.det_ux[j], P.det_uy[j], tt) && tt < best_t) { best_t = tt; best = j; }
  }
  return best;
}
__device__ __forceinline__ int hardware_detector(const DevParams& P, float x, float y, float z, float vx, float vy, float vz,
                                 bool mirror) {
  if (mirror) { x = -x; vx = -vx; }
  const float a = vx * vx + vy * vy;
  if (!(a > 0.0f)) return -1;
  const float b = x * vx + y * vy, c = x * x + y * y - P.ring_in * P.ring_in;
  const float disc = b * b - a * c;
  if (!(disc >= 0.0f)) return -1;
  const float t = (-b + sqrtf(disc)) / a;
  return bore_scan(P, x, y, z, vx, vy, vz, atan2f(x + t * vx, y + t * vy) * (180.0f / kPiF));
}
// Native and mirrored detection of one exit ray.  The mirror (x, vx negated)
// meets r = ring_in at the same t, at the reflected point, whose azimuth is
// the negated one, so the crossing is solved once.
__device__ __forceinline__ void hardware_detectors(const DevParams& P, float x, float y, float z, float vx, float vy,
                                                   float vz, int& nat, int& mir) {
  nat = -1; mir = -1;
  const float a = vx * vx + vy * vy;
  if (!(a > 0.0f)) return;
  const float b = x * vx + y * vy, c = x * x + y * y - P.ring_in * P.ring_in;
  const float disc = b * b - a * c;
  if (!(disc >= 0.0f)) return;
  const float t = (-b + sqrtf(disc)) / a;
  const float deg = atan2f(x + t * vx, y + t * vy) * (180.0f / kPiF);
  nat = bore_scan(P, x, y, z, vx, vy, vz, deg);
  mir = bore_scan(P, -x, y, z, -vx, vy, vz, -deg);
}

// ---- reflective collimator bores, ring-face cavity, SFH213 photodiode ----
// (ROUGH and LIGHT variants only; CLARITAS.py:1263-1360, 4283-4376).  A bore
// is a cylinder of radius `radius` along its local axis s; (q, w) are the
// transverse coordinates.  A smooth opaque nylon wall reflects with the
// unpolarised Fresnel probability (one uniform per wall hit) and absorbs
// otherwise.

// nearest wall hit beyond kEpsF of the transverse circle, CLARITAS.py:1266-1279
__device__ __forceinline__ float bore_side_hit(float q, float w, float vq, float vw, float radius) {
  const float A = vq * vq + vw * vw;
  if (A <= 1.0e-20f) return kInfF;
  const float B = 2.0f * (q * vq + w * vw), Cc = q * q + w * w - radius * radius;
  const float D = B * B - 4.0f * A * Cc;
  if (D < 0.0f) return kInfF;
  const float root = sqrtf(fmaxf(D, 0.0f));
  const float t1 = (-B - root) / (2.0f * A), t2 = (-B + root) / (2.0f * A);
  float t = kInfF;
  if (t1 > kEpsF) t = t1;
  if (t2 > kEpsF && t2 < t) t = t2;
  return t;
}

// one bore section from s to s_end (CLARITAS.py:1296-1318): reflections off
// the wall until the exit plane; true when the ray leaves inside the radius
__device__ bool bore_section(Rng& r, const DevParams& P, float s_end, float radius, float& s, float& q, float& w,
                             float& vs, float& vq, float& vw, int& interactions, int& reflections) {
  for (int b = 0; b <= P.max_bore_bounces; ++b) {
    if (vs <= 1.0e-12f) return false;
    float tp = (s_end - s) / vs;
    if (tp < -kEpsF) return false;
    if (tp < 0.0f) tp = 0.0f;
    const float twall = bore_side_hit(q, w, vq, vw, radius);
    if (tp <= twall) {
      q += tp * vq; w += tp * vw; s = s_end;
      return q * q + w * w <= radius * radius * (1.0f + 2.0e-5f);
    }
    if (twall >= kInfF * 0.5f) return false;
    q += twall * vq; w += twall * vw; s += twall * vs;
    // wall normal from the solid into the air cavity, CLARITAS.py:1281-1294
    interactions++;  // callers reach here only with a reflecting bore model
    const float rr = sqrtf(fmaxf(q * q + w * w, 1.0e-30f));
    const float nq = -q / rr, nw = -w / rr;
    const float ci = fminf(fmaxf(-(vq * nq + vw * nw), 0.0f), 1.0f);
    bool tir;
    const float R = fresnel_R(ci, P.n_air, P.n_bore, tir);
    if (U(r) >= R) return false;  // transmitted into the nylon: lost
    reflect3(vs, vq, vw, 0.0f, nq, nw);
    reflections++;
    q -= kEpsF * q / rr; w -= kEpsF * w / rr;
  }
  return false;
}

// source collimator, local axis +y (q = x, w = z): the counterbore from the
// luminous plane to the shoulder, the opaque shoulder, then the throat to
// the ring face, CLARITAS.py:1321-1345
__device__ bool source_collimator(Rng& r, const DevParams& P, float& x, float& y, float& z, float& vx, float& vy,
                                  float& vz, int& interactions, int& reflections) {
  if (vy <= 1.0e-12f) return false;
  float s = y, q = x, w = z, vs = vy, vq = vx, vw = vz;
  if (q * q + w * w > P.cb_r * P.cb_r) return false;
  const float shoulder = -P.throat_out;
  if (s < shoulder - kEpsF) {
    if (!bore_section(r, P, shoulder, P.cb_r, s, q, w, vs, vq, vw, interactions, reflections)) return false;
  } else if (s > shoulder + kEpsF) {
    return false;
  }
  if (q * q + w * w > P.throat_r * P.throat_r) return false;
  if (!bore_section(r, P, -P.ring_in, P.throat_r, s, q, w, vs, vq, vw, interactions, reflections)) return false;
  x = q; y = s; z = w; vx = vq; vy = vs; vz = vw;
  return true;
}

// first detector throat entrance the air ray reaches, then the throat and
// the counterbore with reflecting walls (CLARITAS.py:1347-1371); -1 if lost
__device__ int detector_collimator(Rng& r, const DevParams& P, float x, float y, float z, float vx, float vy,
                                   float vz, bool mirror, int& reflections) {
  if (mirror) { x = -x; vx = -vx; }
  float best_t = kInfF;
  int best = -1;
  for (int j = 0; j < CLR_N_CHANNELS; ++j) {
    const float ux = P.det_ux[j], uy = P.det_uy[j];
    const float den = vx * ux + vy * uy;
    if (den <= 1.0e-10f) continue;
    const float t = (P.ring_in - (x * ux + y * uy)) / den;
    if (t <= 0.0f) continue;
    const float hx = x + t * vx, hy = y + t * vy, hz = z + t * vz;
    const float q = hx * uy - hy * ux, w = hz;
    if (q * q + w * w <= P.throat_r * P.throat_r && t < best_t) { best_t = t; best = j; }
  }
  if (best < 0) return -1;
  const float ux = P.det_ux[best], uy = P.det_uy[best];
  const float hx = x + best_t * vx, hy = y + best_t * vy;
  float s = P.ring_in, q = hx * uy - hy * ux, w = z + best_t * vz;
  float vs = vx * ux + vy * uy, vq = vx * uy - vy * ux, vw = vz;
  int inter = 0, refl = 0;
  if (!bore_section(r, P, P.throat_out, P.throat_r, s, q, w, vs, vq, vw, inter, refl)) return -1;
  if (!bore_section(r, P, P.ring_out, P.cb_r, s, q, w, vs, vq, vw, inter, refl)) return -1;
  reflections = refl;
  return best;
}

// throat opening (0..17 detectors, 18 the source) the ray enters first at
// the ring face, or -1 for the solid face, CLARITAS.py:4304-4316
__device__ int throat_opening(const DevParams& P, float x, float y, float z, float vx, float vy, float vz) {
  float best_t = kInfF;
  int best = -1;
  const float r2 = P.throat_r * P.throat_r;
  for (int j = 0; j <= CLR_N_CHANNELS; ++j) {
    const float ux = (j < CLR_N_CHANNELS) ? P.det_ux[j] : 0.0f, uy = (j < CLR_N_CHANNELS) ? P.det_uy[j] : -1.0f;
    const float den = vx * ux + vy * uy;
    if (den <= 1.0e-10f) continue;
    const float t = (P.ring_in - (x * ux + y * uy)) / den;
    if (t <= 0.0f) continue;
    const float qx = x + t * vx - P.ring_in * ux, qy = y + t * vy - P.ring_in * uy, qz = z + t * vz;
    if (qx * qx + qy * qy + qz * qz <= r2 && t < best_t) { best_t = t; best = j; }
  }
  return best;
}

// nearest circle hit beyond kEpsF in x-y, CLARITAS.py cylinder_hit
__device__ __forceinline__ float circle_hit(float x, float y, float vx, float vy, float R) {
  const float a = vx * vx + vy * vy;
  if (a < 1e-18f) return kInfF;
  const float b = x * vx + y * vy, c = x * x + y * y - R * R, d = b * b - a * c;
  if (d < 0.0f) return kInfF;
  const float sd = sqrtf(fmaxf(d, 0.0f));
  const float t1 = (-b - sd) / a, t2 = (-b + sd) / a;
  float t = kInfF;
  if (t1 > kEpsF) t = t1;
  if (t2 > kEpsF && t2 < t) t = t2;
  return t;
}

// a detector reached from the air gap once: hard apertures or the
// reflecting collimator, CLARITAS.py:4318-4328
__device__ __forceinline__ int detector_from_air(Rng& r, const DevParams& P, float x, float y, float z, float vx,
                                                 float vy, float vz, int& bounces) {
  bounces = 0;
  if (P.det_bore == 0) return hardware_detector(P, x, y, z, vx, vy, vz, false);
  return detector_collimator(r, P, x, y, z, vx, vy, vz, false, bounces);
}

// stray path between the nylon ring face and the acrylic outer wall: the
// ray reflects off the ring face (Fresnel, else lost), off the acrylic
// outer surface (Fresnel, else it re-enters the cell and is not counted),
// and is offered to the detectors after each round trip,
// CLARITAS.py:4330-4376
__device__ int ring_face_path(Rng& r, const DevParams& P, float x, float y, float z, float vx, float vy, float vz,
                              bool mirror, int& bounces, int& face_reflections) {
  if (mirror) { x = -x; vx = -vx; }
  bounces = 0;
  face_reflections = 0;
  if (P.face_model == 0) return -1;
  if (throat_opening(P, x, y, z, vx, vy, vz) >= 0) return -1;
  for (int b = 0; b < P.max_face_bounces; ++b) {
    const float tr = circle_hit(x, y, vx, vy, P.ring_in);
    if (tr >= kInfF * 0.5f) return -1;
    x += tr * vx; y += tr * vy; z += tr * vz;
    float nx, ny;
    radial_normal(x, y, nx, ny);
    bool tir;
    float ci = fminf(fmaxf(vx * nx + vy * ny, 0.0f), 1.0f);
    float R = fresnel_R(ci, P.n_air, P.n_face, tir);
    if (U(r) >= R) return -1;
    reflect3(vx, vy, vz, -nx, -ny, 0.0f);
    face_reflections++;
    x -= kEpsF * nx; y -= kEpsF * ny;
    const float ta = circle_hit(x, y, vx, vy, P.rout);
    if (ta >= kInfF * 0.5f) return -1;
    x += ta * vx; y += ta * vy; z += ta * vz;
    radial_normal(x, y, nx, ny);
    ci = fminf(fmaxf(-(vx * nx + vy * ny), 0.0f), 1.0f);
    R = fresnel_R(ci, P.n_air, P.n_acrylic, tir);
    if (U(r) >= R) return -1;
    reflect3(vx, vy, vz, nx, ny, 0.0f);
    x += kEpsF * nx; y += kEpsF * ny;
    int db = 0;
    const int det = detector_from_air(r, P, x, y, z, vx, vy, vz, db);
    if (det >= 0) { bounces = db; return det; }
    if (throat_opening(P, x, y, z, vx, vy, vz) >= 0) return -1;
  }
  return -1;
}

// SFH213 relative response 1 / (1 + (phi / 10 deg)^P), P = ln 9 / ln 2 (the
// response is 0.1 at 20 deg), zero from 90 deg, CLARITAS.py:4283-4290
__device__ __forceinline__ float sfh213_response(float phi_deg) {
  const float a = fabsf(phi_deg);
  if (a >= 90.0f) return 0.0f;
  return 1.0f / (1.0f + powf(a / 10.0f, 3.169925001442312f));
}

// ------------------------------------------------------------ lane state --
// R_WATER: next free flight; R_WPART / R_WIFACE: the flight ended at a
// particle event / at a boundary of the water (event handled by the next pass)
// R_WCLOSE: a particle event whose conditional closure is running its trials (CLOSE passes)
enum : int { R_WATER = 1, R_ACRYLIC = 2, R_HEADSPACE = 3, R_DONE = 4, R_WPART = 5, R_WIFACE = 6, R_WCLOSE = 7 };
enum : int { EV_TOP = 1, EV_BOTTOM = 2, EV_SIDE = 3 };
// kinds of scheduler pass
enum : int { K_RECYCLE = 0, K_WATER = 1, K_IFACE = 2, K_ACRYLIC = 3, K_HEAD = 4, K_CLOSE = 5, K_N = 6 };
// who entered the acrylic / headspace (maps their return codes)
enum : int { A_LAUNCH = 0, A_WATER = 1, A_BOTTOM = 2, A_HEAD = 3 };
enum : int { H_LAUNCH = 0, H_WATERWALL = 1, H_TOP = 2, H_BOTTOM = 3 };
enum : unsigned { F_ENTERED = 1u, F_PER = 2u, F_PIR = 4u, F_PABS = 8u, F_CTIR = 16u, F_TTIR = 32u,
                  F_BTIR = 64u };

struct Lane {
  float x, y, z, vx, vy, vz;
  Rng r;
  int region, actx, hctx, it, hit, stuck, last_surf, wstep, ev;
  int st, pf;
  int ic, gic, wic, wcoh, pr, cir, cor, tr, bir, bor;
  unsigned flags;
};

// Causal path provenance of one history, kept by the LIGHT and CAUSAL
// variants (CLARITAS.py:4629-4647).  Observational: it draws no random
// numbers and feeds no transport decision.
struct Prov {
  float wp;                                   // water path travelled so far
  float first_defl, last_defl, ipath, iopl, wp_last;
  float cum, cum2, maxd;
  float mx, my, mz, fx, fy, fz, lx, ly, lz;   // largest-turn, first and last event sites
  float sum_r, sum_z, min_r, max_r, min_z, max_z, wp_first;
  float m_path, m_pod, m_inc, m_exit;         // the largest turn's particle record
  int per, pir, ret, after_ret;  // ret: returns to the water from the acrylic or headspace after an event
  int first_type, last_type, last_order, first_back, first_hemi;
  int fpbin, lpbin, fgbin, lgbin, fwbin, lwbin;
  int btrans, etrans, nsmall, nmed, nlarge, prev_branch, prev_type;
  int m_type, m_order, m_bin, m_macro, m_micro;
};
__device__ __forceinline__ void prov_init(Prov& V) {
  V.wp = 0.0f;
  V.first_defl = V.last_defl = -1.0f; V.ipath = V.iopl = 0.0f; V.wp_last = -1.0f;
  V.cum = V.cum2 = 0.0f; V.maxd = -1.0f;
  V.mx = V.my = V.mz = V.fx = V.fy = V.fz = V.lx = V.ly = V.lz = 0.0f;
  V.sum_r = V.sum_z = 0.0f;
  V.min_r = kInfF; V.max_r = -kInfF; V.min_z = kInfF; V.max_z = -kInfF; V.wp_first = -1.0f;
  V.m_path = V.m_pod = 0.0f; V.m_inc = V.m_exit = -1.0f;
  V.per = V.pir = V.ret = V.after_ret = 0;
  V.first_type = V.last_type = V.last_order = V.first_back = V.first_hemi = 0;
  V.fpbin = V.lpbin = V.fgbin = V.lgbin = V.fwbin = V.lwbin = -1;
  V.btrans = V.etrans = V.nsmall = V.nmed = V.nlarge = V.prev_branch = V.prev_type = 0;
  V.m_type = V.m_order = 0; V.m_bin = -1; V.m_macro = V.m_micro = 0;
}
// One surviving particle event of type `type` (1-7) on bin `pidx` at site s
// with incoming direction vin (the lane holds the outgoing one),
// CLARITAS.py:4936-4980.
__device__ __forceinline__ void prov_commit(Prov& V, const Lane& L, const DevParams& P, float vix, float viy,
                                            float viz, float sx, float sy, float sz, int type, int pidx,
                                            const PCount& pc) {
  const float dcos = fminf(fmaxf(dot3(vix, viy, viz, L.vx, L.vy, L.vz), -1.0f), 1.0f);
  const float defl = acosf(dcos) * (180.0f / kPiF);
  const float sr = sqrtf(fmaxf(sx * sx + sy * sy, 0.0f));
  V.ipath += pc.path;
  if (pidx >= 0) V.iopl += pc.path * P.n_real[pidx];
  V.cum += defl;
  V.cum2 += defl * defl;
  if (defl > V.maxd) {  // strict: the first of equal turns is kept
    V.maxd = defl; V.mx = sx; V.my = sy; V.mz = sz;
    V.m_type = type; V.m_order = L.ic; V.m_bin = pidx; V.m_macro = pc.macro; V.m_micro = pc.micro;
    V.m_path = pc.path;
    V.m_pod = (pidx >= 0 && P.radius[pidx] > 0.0f) ? pc.path / (2.0f * P.radius[pidx]) : 0.0f;
    V.m_inc = pc.inc; V.m_exit = pc.ext;
  }
  if (defl < 30.0f) V.nsmall++; else if (defl < 90.0f) V.nmed++; else V.nlarge++;
  const int branch = (type >= 1 && type <= 3) ? 1 : 2;
  if (V.prev_branch > 0 && branch != V.prev_branch) V.btrans++;
  if (V.prev_type > 0 && type != V.prev_type) V.etrans++;
  V.prev_branch = branch; V.prev_type = type;
  V.sum_r += sr; V.sum_z += sz;
  if (sr < V.min_r) V.min_r = sr;
  if (sr > V.max_r) V.max_r = sr;
  if (sz < V.min_z) V.min_z = sz;
  if (sz > V.max_z) V.max_z = sz;
  if (L.ic == 1) { V.fx = sx; V.fy = sy; V.fz = sz; V.wp_first = V.wp; }
  V.lx = sx; V.ly = sy; V.lz = sz; V.wp_last = V.wp;
  if (pidx >= 0) {
    if (V.fpbin < 0) V.fpbin = pidx;
    V.lpbin = pidx;
    if (type >= 1 && type <= 3) { if (V.fgbin < 0) V.fgbin = pidx; V.lgbin = pidx; }
    else if (type >= 4 && type <= 7) { if (V.fwbin < 0) V.fwbin = pidx; V.lwbin = pidx; }
    if (V.ret > 0) V.after_ret++;
  }
  V.last_type = type; V.last_order = L.ic; V.last_defl = defl;
  if (L.ic == 1) { V.first_type = type; V.first_defl = defl; V.first_back = dcos < 0.0f; V.first_hemi = L.vy < 0.0f; }
}
// the record of a detected history (sentinels of CLARITAS.py:5076-5103)
__device__ void prov_record(clr_causal_record* __restrict__ q, const Lane& L, const Prov& V, uint64_t hist, int nat,
                            int mir) {
  const bool any = L.ic > 0;
  clr_causal_record c;
  c.history = hist; c.native_id = nat; c.mirror_id = mir;
  c.ic = L.ic; c.gic = L.gic; c.wic = L.wic; c.wcoh = L.wcoh;
  c.per = V.per; c.pir = V.pir; c.pabs = (L.flags & F_PABS) ? 1 : 0;
  c.acrylic_reflections = L.cir + L.cor + L.bir + L.bor;
  c.acrylic_returns = V.ret; c.events_after_return = V.after_ret;
  c.first_type = V.first_type; c.last_type = V.last_type; c.last_order = V.last_order;
  c.first_backscatter = V.first_back; c.first_source_hemisphere = V.first_hemi;
  c.first_bin = V.fpbin; c.last_bin = V.lpbin; c.first_geometric_bin = V.fgbin; c.last_geometric_bin = V.lgbin;
  c.first_wave_bin = V.fwbin; c.last_wave_bin = V.lwbin;
  c.branch_transitions = V.btrans; c.type_transitions = V.etrans;
  c.small_deflections = V.nsmall; c.medium_deflections = V.nmed; c.large_deflections = V.nlarge;
  c.max_type = any ? V.m_type : 0; c.max_order = any ? V.m_order : 0; c.max_bin = any ? V.m_bin : -1;
  c.max_macro_bounces = any ? V.m_macro : 0; c.max_micro_order = any ? V.m_micro : 0;
  c.first_deflection = V.first_defl; c.last_deflection = V.last_defl;
  c.internal_path = V.ipath; c.internal_opl = V.iopl;
  c.water_path_after_last = (V.last_order > 0 && V.wp_last >= 0.0f) ? fmaxf(V.wp - V.wp_last, 0.0f) : -1.0f;
  c.cumulative_deflection = any ? V.cum : -1.0f; c.cumulative_deflection_sq = any ? V.cum2 : -1.0f;
  c.max_deflection = any ? V.maxd : -1.0f;
  c.turn_x = any ? V.mx : 0.0f; c.turn_y = any ? V.my : 0.0f; c.turn_z = any ? V.mz : 0.0f;
  c.first_x = any ? V.fx : 0.0f; c.first_y = any ? V.fy : 0.0f; c.first_z = any ? V.fz : 0.0f;
  c.last_x = any ? V.lx : 0.0f; c.last_y = any ? V.ly : 0.0f; c.last_z = any ? V.lz : 0.0f;
  c.mean_r = any ? V.sum_r / (float)L.ic : -1.0f; c.mean_z = any ? V.sum_z / (float)L.ic : -1.0f;
  c.min_r = any ? V.min_r : -1.0f; c.max_r = any ? V.max_r : -1.0f;
  c.min_z = any ? V.min_z : -1.0f; c.max_z = any ? V.max_z : -1.0f;
  c.water_path_first = any ? V.wp_first : -1.0f; c.water_path_last = any ? V.wp_last : -1.0f;
  c.max_internal_path = any ? V.m_path : 0.0f; c.max_path_over_diameter = any ? V.m_pod : 0.0f;
  c.max_incidence = any ? V.m_inc : -1.0f; c.max_exit = any ? V.m_exit : -1.0f;
  *q = c;
}
// per-lane work counters accumulated over all histories of a launch
struct Acc {
  uint32_t n_flight, n_null, n_iface, n_abs, steps, draws, trials, rejects, inacc;
  uint32_t entered, n_exit, valid_air;
};

// L.pf keeps the failure code in its low 8 bits and the subcode above them
__device__ __forceinline__ void terminate(Lane& L, int st, int pf = PF_OK, int pfs = SUB_NONE) {
  L.st = st;
  L.pf = pf | (pfs << 8);
  L.region = R_DONE;
}
// tally slot of a failure subcode: the position of the subcode in
// CLR_SUBCODE_VALUES (runs of consecutive codes), the last slot for any other
__host__ __device__ constexpr int subcode_slot(int s) {
  return s == 0                    ? 0
         : (s >= 10 && s <= 14)     ? 1 + (s - 10)
         : (s >= 20 && s <= 41)     ? 6 + (s - 20)
         : (s >= 1001 && s <= 1005) ? 28 + (s - 1001)
         : (s >= 1101 && s <= 1107) ? 33 + (s - 1101)
         : (s >= 2001 && s <= 2002) ? 40 + (s - 2001)
         : (s >= 2101 && s <= 2104) ? 42 + (s - 2101)
         : (s >= 3001 && s <= 3006) ? 46 + (s - 3001)
         : (s == 4001)              ? 52
         : (s >= 9001 && s <= 9003) ? 53 + (s - 9001)
                                    : CLR_N_SUBCODE - 1;
}
constexpr bool subcode_slots_match_table() {
  constexpr int vals[CLR_N_SUBCODE_KNOWN] = CLR_SUBCODE_VALUES;
  for (int k = 0; k < CLR_N_SUBCODE_KNOWN; ++k)
    if (subcode_slot(vals[k]) != k) return false;
  return subcode_slot(5) == CLR_N_SUBCODE - 1 && subcode_slot(9004) == CLR_N_SUBCODE - 1;
}
static_assert(subcode_slots_match_table(), "subcode slots disagree with CLR_SUBCODE_VALUES");

// launch + collimator + air->acrylic entry, CLARITAS.py:4650-4735.  The
// reflecting source collimator (ROUGH and LIGHT variants) reports its wall
// interactions and reflections in sbi / sbr.
// variants with the rough-particle physics and the detector-side models
template <int kKind>
constexpr bool kind_rough() { return kKind == CLR_KERNEL_ROUGH || kKind == CLR_KERNEL_LIGHT || kKind == CLR_KERNEL_CAUSAL; }
// variants that keep the causal provenance
template <int kKind>
constexpr bool kind_diag() { return kKind == CLR_KERNEL_LIGHT || kKind == CLR_KERNEL_CAUSAL; }

// Source position on the launch plane y = -source_r and direction, before
// the collimator, CLARITAS.py:4650-4680 (the collimated disk has no reference
// counterpart).
__device__ __forceinline__ void source_sample(Rng& r, const DevParams& P, float& x, float& y, float& z, float& vx,
                                              float& vy, float& vz) {
  vx = 0.0f; vy = 1.0f; vz = 0.0f;
  if (P.src_model == CLR_SOURCE_COLLIMATED_DISK) {
    const float u1 = U(r), u2 = U(r);
    const float rho = P.beam_r * sqrtf(u1), ang = 2.0f * kPiF * u2;
    x = rho * cosf(ang);
    z = rho * sinf(ang);
  } else {
    const float u1 = fmaxf(U(r), 1e-12f), u2 = U(r);
    const float mag = sqrtf(-2.0f * logf(u1));
    x = P.sigma * mag * cosf(2.0f * kPiF * u2);
    z = P.sigma * mag * sinf(2.0f * kPiF * u2);
  }
  y = -P.source_r;
  if (P.src_model == CLR_SOURCE_LEGACY_POLAR || P.src_model == CLR_SOURCE_BETA_RADIANCE) {
    const float polar = (P.src_model == CLR_SOURCE_LEGACY_POLAR)
                            ? beta_draw(r, P.alpha1, P.alpha2) * (kPiF / 2.0f)
                            : beta_radiance_polar(r, P.alpha1, P.alpha2);
    const float az = 2.0f * kPiF * U(r), sp = sinf(polar);
    vx = sp * cosf(az); vy = cosf(polar); vz = sp * sinf(az);
  } else if (P.src_model == CLR_SOURCE_THESIS_PLANAR) {
    float h = beta_draw(r, P.alpha1, P.alpha2) * (kPiF / 2.0f);
    if (U(r) < 0.5f) h = -h;
    vx = sinf(h); vy = cosf(h); vz = 0.0f;
  }
  normalize3(vx, vy, vz);
}

template <int kKind>
__device__ __forceinline__ void launch_history(Lane& L, const DevParams& P, Acc& A, int& sbi, int& sbr) {
  constexpr bool kBores = kind_rough<kKind>();
  Rng& r = L.r;
  float x, y, z, vx, vy, vz;
  source_sample(r, P, x, y, z, vx, vy, vz);
  L.x = x; L.y = y; L.z = z; L.vx = vx; L.vy = vy; L.vz = vz;
  if (vy <= 1e-12f) { terminate(L, 3); return; }
  bool reflecting = false;
  if constexpr (kBores) reflecting = P.src_bore != 0;
  if (reflecting) {
    if (!source_collimator(r, P, x, y, z, vx, vy, vz, sbi, sbr)) { terminate(L, 3); return; }
  } else {
    if (x * x + z * z > P.cb_r * P.cb_r) { terminate(L, 3); return; }
    float t = (-P.throat_out - y) / vy;
    if (t < -kEpsF) { terminate(L, 3); return; }
    if (t < 0.0f) t = 0.0f;
    const float xt = x + t * vx, zt = z + t * vz;
    if (xt * xt + zt * zt > P.throat_r * P.throat_r) { terminate(L, 3); return; }
    t = (-P.ring_in - y) / vy;
    if (t <= 0.0f) { terminate(L, 3); return; }
    x += t * vx; y += t * vy; z += t * vz;
    if (x * x + z * z > P.throat_r * P.throat_r) { L.x = x; L.y = y; L.z = z; terminate(L, 3); return; }
  }
  const double td = finite_cylinder_forward_d(P, x, y, z, vx, vy, vz, P.rout, P.zlow, P.zwall_top);
  const float tw = (td < 0.5 * (double)kInfF) ? (float)td : kInfF;
  if (tw >= kInfF * 0.5f) { L.x = x; L.y = y; L.z = z; terminate(L, 3); return; }
  x += tw * vx; y += tw * vy; z += tw * vz;
  float nx, ny;
  radial_normal(x, y, nx, ny);
  bool tir;
  A.n_iface++;
  L.z = z;
  if (!fresnel_transmit(r, vx, vy, vz, nx, ny, 0.0f, P.n_air, P.n_acrylic, tir)) {
    L.cor++;
    if (tir) L.flags |= F_CTIR;
    L.x = x + kEpsF * nx; L.y = y + kEpsF * ny;
    L.vx = vx; L.vy = vy; L.vz = vz;
    terminate(L, 1);
    return;
  }
  L.vx = vx; L.vy = vy; L.vz = vz;
  L.x = x - kEpsF * nx; L.y = y - kEpsF * ny;
  L.region = R_ACRYLIC; L.actx = A_LAUNCH; L.it = 0; L.stuck = 0; L.last_surf = -1;
}

// headspace return code -> next state (callers differ, CLARITAS.py:4720-4731,
// 4995-5002, 5024-5031, 5051-5059)
__device__ __forceinline__ void headspace_return(Lane& L, int hw, int sub = SUB_HEADSPACE_NO_HIT) {
  if (hw == 1) { L.region = R_WATER; L.flags |= F_ENTERED; return; }
  if (hw == 2) { terminate(L, 1); return; }
  if (hw == 3) { terminate(L, 7); return; }
  if (hw == 4) { terminate(L, 6); return; }
  if (hw == 5 && (L.hctx == H_LAUNCH || L.hctx == H_WATERWALL)) { terminate(L, 8); return; }
  terminate(L, 4, PF_HEADSPACE_NO_HIT, sub);
}
// acrylic return code -> next state
__device__ __forceinline__ void acrylic_return(Lane& L, int where, int sub) {
  if (L.actx == A_HEAD) {
    if (where == 4) { L.region = R_HEADSPACE; return; }  // same headspace walk continues
    headspace_return(L, where, sub);
    return;
  }
  if (where == 1) { L.region = R_WATER; L.flags |= F_ENTERED; return; }
  if (where == 2) { terminate(L, 1); return; }
  if (where == 3) { terminate(L, 7); return; }
  if (where == 5) { terminate(L, 8); return; }
  if (where == 4) {
    L.region = R_HEADSPACE;
    L.hit = 0;
    L.hctx = (L.actx == A_LAUNCH) ? H_LAUNCH : (L.actx == A_WATER) ? H_WATERWALL : H_BOTTOM;
    return;
  }
  int pf;
  if (L.actx == A_BOTTOM) pf = (sub == SUB_ACRYLIC_BOUNCE) ? PF_ACRYLIC_BOUNCE : PF_BOTTOM_GEOMETRY;
  else pf = (sub == SUB_ACRYLIC_BOUNCE) ? PF_ACRYLIC_BOUNCE : PF_ACRYLIC_NO_HIT;
  terminate(L, 4, pf, sub);
}
__device__ __forceinline__ void enter_acrylic(Lane& L, int ctx) {
  L.region = R_ACRYLIC; L.actx = ctx; L.it = 0; L.stuck = 0; L.last_surf = -1;
}

// one surface interaction inside the acrylic union, CLARITAS.py:3140-3198
__device__ __forceinline__ void acrylic_step(Lane& L, const DevParams& P, Acc& A) {
  if (L.it >= kEmergencyCap) {
    if (acrylic_trapped(P, L.x, L.y, L.vx, L.vy, L.vz)) acrylic_return(L, 5, SUB_NONE);
    else acrylic_return(L, 0, SUB_ACRYLIC_BOUNCE);
    return;
  }
  L.it++;
  double td = 1.0e300;
  const int s = acrylic_next_surface(P, L.x, L.y, L.z, L.vx, L.vy, L.vz, td);
  if (!s || !(td > 0.0) || !isfinite(td) || td >= 1.0e299) { acrylic_return(L, 0, SUB_ACRYLIC_NO_HIT); return; }
  const float t = (float)td;
  const float px = L.x, py = L.y, pz = L.z, pvx = L.vx, pvy = L.vy, pvz = L.vz;
  L.x += t * L.vx; L.y += t * L.vy; L.z += t * L.vz;
  bool tir = false;
  A.n_iface++;
  if (s == 1 || s == 2) {
    float nx, ny;
    radial_normal(L.x, L.y, nx, ny);
    if (s == 1) {
      const float zs = surface_z_vs(P, L.x, L.y, L.z);
      const float n_in = (L.z <= zs + kEpsF) ? P.n_water : P.n_air;
      if (fresnel_transmit(L.r, L.vx, L.vy, L.vz, nx, ny, 0.0f, P.n_acrylic, n_in, tir)) {
        // normal convention of the reference: +radial incident side here
        L.x -= kEpsF * nx; L.y -= kEpsF * ny;
        acrylic_return(L, (n_in == P.n_water) ? 1 : 4, SUB_NONE);
        return;
      }
      L.cir++;
      if (tir) L.flags |= F_CTIR;
      L.x += kEpsF * nx; L.y += kEpsF * ny;
    } else {
      if (fresnel_transmit(L.r, L.vx, L.vy, L.vz, -nx, -ny, 0.0f, P.n_acrylic, P.n_air, tir)) {
        L.x += kEpsF * nx; L.y += kEpsF * ny;
        acrylic_return(L, 2, SUB_NONE);
        return;
      }
      L.cor++;
      if (tir) L.flags |= F_CTIR;
      L.x -= kEpsF * nx; L.y -= kEpsF * ny;
    }
  } else {
    // planes: 3 bottom external face, 4 annular top face, 5 water aperture
    const float zplane = (s == 3) ? P.zlow : (s == 4) ? P.zwall_top : P.zmin;
    const float nz = (s == 3) ? 1.0f : -1.0f;  // normal into the incident acrylic
    const float n2 = (s == 3) ? P.n_bottom_ext : (s == 4) ? P.n_air : P.n_water;
    L.z = zplane;
    // the reference reflects about +z for all three planes (CLARITAS.py:3175-3189)
    const float ci = fminf(fmaxf((s == 3) ? -L.vz : L.vz, 0.0f), 1.0f);
    const float R = fresnel_R(ci, P.n_acrylic, n2, tir);
    float a, b, c;
    if (!tir && !(U(L.r) < R) && refract3(L.vx, L.vy, L.vz, 0.0f, 0.0f, nz, P.n_acrylic, n2, a, b, c)) {
      L.vx = a; L.vy = b; L.vz = c;
      L.z = (s == 3) ? P.zlow - kEpsF : (s == 4) ? P.zwall_top + kEpsF : P.zmin + kEpsF;
      acrylic_return(L, (s == 3) ? 3 : (s == 4) ? 2 : 1, SUB_NONE);
      return;
    }
    reflect3(L.vx, L.vy, L.vz, 0.0f, 0.0f, 1.0f);
    if (s == 4) { L.cor++; if (tir) L.flags |= F_CTIR; }
    else if (s == 3) { L.bor++; if (tir) L.flags |= F_BTIR; }
    else { L.bir++; if (tir) L.flags |= F_BTIR; }
    L.z = (s == 3) ? P.zlow + kEpsF : (s == 4) ? P.zwall_top - kEpsF : P.zmin - kEpsF;
  }
  const float pd = fabsf(L.x - px) + fabsf(L.y - py) + fabsf(L.z - pz);
  const float vd = fabsf(L.vx - pvx) + fabsf(L.vy - pvy) + fabsf(L.vz - pvz);
  L.stuck = (s == L.last_surf && pd <= 1.0e-12f && vd <= 1.0e-7f) ? L.stuck + 1 : 0;
  if (L.stuck >= 8) { acrylic_return(L, 0, SUB_ACRYLIC_BOUNCE); return; }
  L.last_surf = s;
}

// one headspace segment, CLARITAS.py:3235-3293
__device__ __forceinline__ void headspace_step(Lane& L, const DevParams& P, Acc& A) {
  if (L.hit >= kEmergencyCap) { headspace_return(L, 0); return; }
  L.hit++;
  if (!isfinite(L.x) || !isfinite(L.y) || !isfinite(L.z) || !isfinite(L.vx) || !isfinite(L.vy) || !isfinite(L.vz)) {
    headspace_return(L, 0); return;
  }
  const double tol = support_tol(P, P.rin, P.rout, P.zmin, P.zlow, P.zwall_top);
  const float zs = free_surface_z(P, L.x, L.y);
  const double rr = hypot((double)L.x, (double)L.y);
  if ((double)L.z <= (double)zs - tol && rr < (double)P.rin - tol) { headspace_return(L, 1); return; }
  if ((double)L.z > (double)P.zwall_top + tol) { headspace_return(L, 4); return; }
  const double tside = bore_wall_hit_d(P, L.x, L.y, L.z, L.vx, L.vy, L.vz, P.rin, P.zmin, P.zwall_top);
  if (tside < 0.0) { headspace_return(L, 0, SUB_WATER_OUTSIDE); return; }
  const double tsurf = (double)free_surface_hit(P, L.x, L.y, L.z, L.vx, L.vy, L.vz);
  double ttop = 1.0e300;
  if (L.vz > 1.0e-14f) {
    const double t = ((double)P.zwall_top - (double)L.z) / (double)L.vz;
    if (t > 0.0) {
      const double X = (double)L.x + t * (double)L.vx, Y = (double)L.y + t * (double)L.vy;
      if (hypot(X, Y) < (double)P.rin + tol) ttop = t;
    }
  }
  int which = 0;
  double best = 1.0e300;
  if (tsurf > 0.0 && tsurf < best) { best = tsurf; which = 1; }
  if (tside >= 0.0 && tside < best) { best = tside; which = 2; }
  if (ttop > 0.0 && ttop < best) { best = ttop; which = 3; }
  if (!which || !isfinite(best) || best >= 1.0e299) { headspace_return(L, 0); return; }
  const float t = (float)best;
  L.x += t * L.vx; L.y += t * L.vy; L.z += t * L.vz;
  if (which == 3) { L.z = P.zwall_top + kEpsF; headspace_return(L, 4); return; }
  bool tir = false;
  A.n_iface++;
  if (which == 1) {
    float snx, sny, snz;
    free_surface_normal(P, L.x, L.y, snx, sny, snz);
    if (!fresnel_transmit(L.r, L.vx, L.vy, L.vz, -snx, -sny, -snz, P.n_air, P.n_water, tir)) {
      L.tr++;
      if (tir) L.flags |= F_TTIR;
      L.x -= kEpsF * snx; L.y -= kEpsF * sny; L.z -= kEpsF * snz;
      return;
    }
    if (!water_surface_offset(P, L.x, L.y, L.z, snx, sny, snz)) { headspace_return(L, 0, SUB_WATER_OUTSIDE); return; }
    headspace_return(L, 1);
    return;
  }
  float nx, ny;
  radial_normal(L.x, L.y, nx, ny);
  if (!fresnel_transmit(L.r, L.vx, L.vy, L.vz, -nx, -ny, 0.0f, P.n_air, P.n_acrylic, tir)) {
    L.cir++;
    if (tir) L.flags |= F_CTIR;
    L.x -= kEpsF * nx; L.y -= kEpsF * ny;
    return;
  }
  L.x += kEpsF * nx; L.y += kEpsF * ny;
  enter_acrylic(L, A_HEAD);
}

// Water loop, CLARITAS.py:4734-5066, in three passes: the free flight to the
// next particle event or boundary (water_flight), the particle event
// (water_particle) and the boundary interaction (water_iface).  The wave
// schedules each kind separately so divergent lanes do not serialise them.
// majorant slab of height z (end slabs open)
// ---- light-transport voxel diagnostics (LIGHT variant only) ----
// Observational: no random numbers, no change to the photon; writes only
// the diagnostic buffers.  The arithmetic types (FP64 traversal of FP32
// segments, FP32 tolerances) follow CLARITAS.py:3312-3467 so the same
// segments deposit the same lengths and count the same snaps and errors.

// clip the parameter interval [t0, t1] of p + t d to lo <= . <= hi
__device__ __forceinline__ bool light_clip(double p, double d, double lo, double hi, double& t0, double& t1) {
  if (fabs(d) < 1.0e-30f) return p >= lo && p <= hi;
  double a = (lo - p) / d, b = (hi - p) / d;
  if (a > b) { const double q = a; a = b; b = q; }
  t0 = fmax(t0, a);
  t1 = fmin(t1, b);
  return t1 >= t0;
}

// Exact track-length deposition of the water segment (x0 + t v, 0 < t <
// dist) into the voxel grid by a 3-D DDA over its clipped part; `order` is
// the number of particle events before it (post-scatter grid when > 0).
// CLARITAS.py:3318-3400.
__device__ void light_segment(const LightDiag& D, uint64_t gid, float x0, float y0, float z0, float vx, float vy,
                              float vz, float dist, int order) {
  if (gid >= D.max_rays || !(dist > 0.0f) || !isfinite(dist)) return;
  const int nx = D.nx, ny = D.ny, nz = D.nz;
  const float Rin = D.rin, zmin = D.zmin, zmax = D.zmax;
  if (nx <= 0 || ny <= 0 || nz <= 0 || !(zmax > zmin)) { atomicAdd(D.error_count, 1u); return; }
  const double ddist = (double)dist;
  const double sx = (double)vx * ddist, sy = (double)vy * ddist, sz = (double)vz * ddist;
  double c0 = 0.0, c1 = 1.0;
  if (!light_clip(x0, sx, -Rin, Rin, c0, c1) || !light_clip(y0, sy, -Rin, Rin, c0, c1) ||
      !light_clip(z0, sz, zmin, zmax, c0, c1))
    return;
  const double ax = x0 + sx * c0, ay = y0 + sy * c0, az = z0 + sz * c0;
  const double bx = x0 + sx * c1, by = y0 + sy * c1, bz = z0 + sz * c1;
  const double clipped = dist * (c1 - c0);
  if (!(clipped > 0.0f)) return;
  atomicAdd(D.segment_total, clipped);
  atomicAdd(D.segment_count, 1ull);
  if (gid < (uint64_t)D.rep_rays && D.max_segments > 0) {
    const int g = (int)gid, slot = D.tcount[g];
    if (slot < D.max_segments) {
      const size_t q = (size_t)g * D.max_segments + slot;
      D.tx0[q] = (float)ax; D.ty0[q] = (float)ay; D.tz0[q] = (float)az;
      D.tx1[q] = (float)bx; D.ty1[q] = (float)by; D.tz1[q] = (float)bz;
      D.torder[q] = order; D.tevent[q] = 0; D.tcount[g] = slot + 1;
    } else {
      D.toverflow[g] = 1;
    }
  }
  const double xlo = -(double)Rin, ylo = -(double)Rin, zlo = (double)zmin;
  const double hx = (2.0 * (double)Rin) / nx, hy = (2.0 * (double)Rin) / ny, hz = ((double)zmax - zlo) / nz;
  int ix = min(max((int)floor((ax - xlo) / hx), 0), nx - 1);
  int iy = min(max((int)floor((ay - ylo) / hy), 0), ny - 1);
  int iz = min(max((int)floor((az - zlo) / hz), 0), nz - 1);
  const double ex = bx - ax, ey = by - ay, ez = bz - az;
  const int stx = (ex > 0.0f) ? 1 : ((ex < 0.0f) ? -1 : 0);
  const int sty = (ey > 0.0f) ? 1 : ((ey < 0.0f) ? -1 : 0);
  const int stz = (ez > 0.0f) ? 1 : ((ez < 0.0f) ? -1 : 0);
  // parameter (fraction of the clipped segment) of the next face crossing per axis
  auto face = [](double lo, int i, int st, double h, double a, double e) -> double {
    return (st == 0) ? (double)kInfF : (lo + (double)(st > 0 ? i + 1 : i) * h - a) / e;
  };
  double tmx = fmax(face(xlo, ix, stx, hx, ax, ex), 0.0);
  double tmy = fmax(face(ylo, iy, sty, hy, ay, ey), 0.0);
  double tmz = fmax(face(zlo, iz, stz, hz, az, ez), 0.0);
  double u = 0.0, local = 0.0;
  const float tol = 2.0e-6f;
  const int max_steps = nx + ny + nz + 8;
  for (int k = 0; k < max_steps && u < 1.0f; ++k) {
    if (ix < 0 || ix >= nx || iy < 0 || iy >= ny || iz < 0 || iz >= nz) {
      if (u >= 1.0f - tol) { atomicAdd(D.boundary_adjust, 1u); u = 1.0f; break; }
      atomicAdd(D.error_count, 1u);
      break;
    }
    double un = fmin(1.0, fmin(tmx, fmin(tmy, tmz)));
    if (un < 1.0f && un >= 1.0f - tol) { un = 1.0f; atomicAdd(D.boundary_adjust, 1u); }
    if (un < u) {
      if (u - un <= tol) { un = u; atomicAdd(D.boundary_adjust, 1u); }
      else { atomicAdd(D.error_count, 1u); break; }
    }
    const double len = (un - u) * clipped;
    if (len > 0.0f) {
      const int q = (iz * ny + iy) * nx + ix;
      atomicAdd(D.total + q, (float)len);
      if (order > 0) atomicAdd(D.post + q, (float)len);
      local += len;
    }
    if (un >= 1.0f) { u = 1.0f; break; }
    const double eps = 8.0 * 2.2204460492503131e-16 * fmax(1.0, fabs(un));
    if (tmx <= un + eps) { ix += stx; tmx = face(xlo, ix, stx, hx, ax, ex); }
    if (tmy <= un + eps) { iy += sty; tmy = face(ylo, iy, sty, hy, ay, ey); }
    if (tmz <= un + eps) { iz += stz; tmz = face(zlo, iz, stz, hz, az, ez); }
    u = un;
  }
  if (u < 1.0f - tol) atomicAdd(D.error_count, 1u);
  atomicAdd(D.deposited_total, local);
}

// terminal event code of the history's last stored segment,
// CLARITAS.py:3402-3407 (0 propagation/boundary, 1-3 geometric external
// reflection / transmission / internal reflection, 4 coherent Mie, 5-7 the
// same for morphology, 8 inaccessible placement, 9 absorption, 10 failure)
// Water-path heat maps: a sample every 10 um along the segment (at most
// 20000), binned into the x-y and z-y projections; z runs from zmin to the
// surface height at the wall.  CLARITAS.py:3296-3309, 4764-4765
__device__ void heat_raster(const LightDiag& D, float zvis, float x0, float y0, float z0, float vx, float vy,
                            float vz, float dist) {
  const int vis = D.vis;
  if (vis <= 0 || !(dist > 0.0f)) return;
  const float spacing = 1.0e-5f, Rin = D.rin, zmin = D.zmin;
  int n = (int)ceilf(dist / spacing);
  n = n < 1 ? 1 : (n > 20000 ? 20000 : n);
  for (int k = 1; k <= n; ++k) {
    const float f = (float)k / (float)n;
    const float x = x0 + vx * dist * f, y = y0 + vy * dist * f, z = z0 + vz * dist * f;
    const int ix = (int)(((x + Rin) / (2.0f * Rin)) * vis);
    const int iy = vis - 1 - (int)(((y + Rin) / (2.0f * Rin)) * vis);
    const int iz = (int)(((z - zmin) / (zvis - zmin)) * vis);
    if (ix >= 0 && ix < vis && iy >= 0 && iy < vis) atomicAdd(&D.heat_xy[iy * vis + ix], 1.0f);
    if (iz >= 0 && iz < vis && iy >= 0 && iy < vis) atomicAdd(&D.heat_zy[iy * vis + iz], 1.0f);
  }
}

__device__ __forceinline__ void light_mark(const LightDiag& D, uint64_t gid, int code) {
  if (gid >= D.max_rays || gid >= (uint64_t)D.rep_rays || D.max_segments <= 0) return;
  const int g = (int)gid, n = D.tcount[g];
  if (n > 0 && n <= D.max_segments) D.tevent[(size_t)g * D.max_segments + n - 1] = code;
}

// one event site into a count grid (and the weight grid when given),
// CLARITAS.py:3409-3467: sites beyond the box by more than the coordinate
// tolerance are errors (the first is recorded), sites within it are
// snapped to the edge voxel
__device__ void light_site(const LightDiag& D, uint64_t gid, float x, float y, float z, float* weight,
                           unsigned int* count, unsigned long long* total) {
  if (gid >= D.max_rays) return;
  atomicAdd(total, 1ull);
  const int nx = D.nx, ny = D.ny, nz = D.nz;
  const float Rin = D.rin, zmin = D.zmin, zmax = D.zmax;
  const float ctol = 2.0e-6f * fmaxf(2.0f * Rin, zmax - zmin);
  if (!isfinite(x) || !isfinite(y) || !isfinite(z) || x < -Rin - ctol || x > Rin + ctol || y < -Rin - ctol ||
      y > Rin + ctol || z < zmin - ctol || z > zmax + ctol) {
    const unsigned int prev = atomicAdd(D.outside, 1u);
    if (prev == 0u) {
      D.outside[1] = __float_as_uint(x); D.outside[2] = __float_as_uint(y); D.outside[3] = __float_as_uint(z);
      D.outside[4] = (unsigned int)gid; D.outside[5] = __float_as_uint(ctol); D.outside[6] = 1u;
    }
    atomicAdd(D.error_count, 1u);
    return;
  }
  int ix = (int)floorf(((x + Rin) / (2.0f * Rin)) * nx);
  int iy = (int)floorf(((y + Rin) / (2.0f * Rin)) * ny);
  int iz = (int)floorf(((z - zmin) / (zmax - zmin)) * nz);
  bool snapped = false;
  if (ix < 0) { ix = 0; snapped = true; } else if (ix >= nx) { ix = nx - 1; snapped = true; }
  if (iy < 0) { iy = 0; snapped = true; } else if (iy >= ny) { iy = ny - 1; snapped = true; }
  if (iz < 0) { iz = 0; snapped = true; } else if (iz >= nz) { iz = nz - 1; snapped = true; }
  if (snapped) atomicAdd(D.boundary_adjust, 1u);
  const int q = (iz * ny + iy) * nx + ix;
  if (weight) atomicAdd(weight + q, 1.0f);
  atomicAdd(count + q, 1u);
}
// a surviving particle event (scatter-event grids) and its code
__device__ __forceinline__ void light_event(const LightDiag& D, uint64_t gid, float x, float y, float z, int code) {
  light_site(D, gid, x, y, z, D.event_weight, D.event_count, D.event_total);
  light_mark(D, gid, code);
}
// an accepted extinction encounter that ends in absorption, code 9
__device__ __forceinline__ void light_absorb(const LightDiag& D, uint64_t gid, float x, float y, float z) {
  light_site(D, gid, x, y, z, nullptr, D.absorption_count, D.absorption_total);
  light_mark(D, gid, 9);
}

__device__ __forceinline__ int slab_of(const DevParams& P, float z) {
  const int j = (int)floorf((z - P.slab_z0) * P.slab_inv_w);
  return min(max(j, 0), P.n_slabs - 1);
}
// Free path from height z along direction cosine vz under the piecewise-
// constant majorant of the z slabs: the optical depth -log u is spent slab by
// slab (delta tracking with a spatially varying majorant; the exponential is
// memoryless at slab faces, so the collision density is exact).  Returns
// kInfF when the path reaches `limit` first.
__device__ __forceinline__ float slab_free_path(const DevParams& P, float z, float vz, float u, float limit) {
  float tau = -logf(fmaxf(u, 1e-12f));
  // every slab's majorant is at most slab_mu_max: a depth that the path
  // cannot spend before `limit` even at that rate means no collision first
  if (!(tau < P.slab_mu_max * limit)) return kInfF;
  int j = slab_of(P, z);
  const int step = (vz > 0.0f) ? 1 : -1;
  const float ivz = 1.0f / vz;
  float t = 0.0f;
  for (int n = 0; n < P.n_slabs; ++n) {
    float t_exit = kInfF;
    if (vz > 1e-14f && j < P.n_slabs - 1) t_exit = (P.slab_z0 + (float)(j + 1) * P.slab_w - z) * ivz;
    else if (vz < -1e-14f && j > 0) t_exit = (P.slab_z0 + (float)j * P.slab_w - z) * ivz;
    t_exit = fmaxf(t_exit, t);
    const float mu = s_slab_mu[j];
    if (mu > 0.0f && tau < mu * (t_exit - t)) return t + tau / mu;
    if (t_exit >= limit) return kInfF;
    tau -= mu * (t_exit - t);
    t = t_exit;
    j += step;
  }
  return kInfF;
}

template <int kKind>
__device__ __forceinline__ void water_flight(Lane& L, const DevParams& P, Acc& A, const LightDiag* __restrict__ D,
                                             Prov& V) {
  constexpr bool kPlain = (kKind == CLR_KERNEL_PLAIN);
  if (L.wstep >= P.max_iter) { terminate(L, 4, PF_ITERATION_LIMIT, SUB_ITERATION_LIMIT); return; }
  L.wstep++;
  A.steps++;
  if (!isfinite(L.x) || !isfinite(L.y) || !isfinite(L.z)) {
    terminate(L, 4, PF_NONFINITE_POSITION, SUB_NONFINITE_POSITION); return;
  }
  if (!isfinite(L.vx) || !isfinite(L.vy) || !isfinite(L.vz)) {
    terminate(L, 4, PF_NONFINITE_DIRECTION, SUB_NONFINITE_DIRECTION); return;
  }
  const float zsh = surface_z_vs(P, L.x, L.y, L.z);
  if (L.z <= P.zmin || L.z >= zsh) { terminate(L, 2); return; }
  const double ri2w = (double)P.rin * (double)P.rin, wtol = 8.0 * kFEps * ri2w;
  const double rr2 = (double)L.x * (double)L.x + (double)L.y * (double)L.y;
  if (rr2 > ri2w + wtol) { terminate(L, 4, PF_INVALID_MEDIUM, SUB_WATER_OUTSIDE); return; }
  const double tsd = bore_wall_hit_d(P, L.x, L.y, L.z, L.vx, L.vy, L.vz, P.rin, P.zmin, P.zwall_top);
  if (tsd < 0.0) { terminate(L, 4, PF_INVALID_MEDIUM, SUB_WATER_OUTSIDE); return; }
  const float tside = (tsd < 0.5 * (double)kInfF) ? (float)tsd : kInfF;
  float ttop = kInfF, tbot = kInfF;
  if (L.vz < -1e-14f) tbot = (P.zmin - L.z) / L.vz;
  if (!(tbot > 0.0f)) tbot = kInfF;
  const float tc = fminf(tside, tbot);  // nearer of side wall and bottom
  const float MU_TRACK = P.mu_track;
  // The boundary distance is min(ttop, tbot, tside) with ttop, tbot in
  // (0, inf], so when tside > 0 and tc is finite the free path is drawn
  // whatever the surface gives.  Drawing it first (the surface intersection
  // uses no random numbers) lets the surface be skipped exactly when the
  // segment up to min(free path, tc) stays below its lowest point: then the
  // surface can neither come first nor change the outcome of the flight.
  // With tside > 0 the flight always draws its free path (the boundary is
  // at a positive distance): a vertical ray (tc infinite) has the surface
  // ahead, and the draw's outcome against the boundary does not depend on
  // which of the two limits it was measured with.  With tside = 0 the ray is
  // leaving through the wall here and draws nothing.
  const bool early = (tside > 0.0f && MU_TRACK > 0.0f);
  const bool slabs = P.n_slabs > 0;
  float freep = kInfF;
  if (early) freep = slabs ? slab_free_path(P, L.z, L.vz, U(L.r), tc) : -logf(fmaxf(U(L.r), 1e-12f)) / MU_TRACK;
  {
    // the surface may be skipped only when the side wall or the bottom
    // bounds the flight (tc finite): otherwise it is the only boundary
    const float d = fminf(freep, tc);
    bool skip = false;
    if (P.fs_model != 0 && tc < 0.5f * kInfF) {
      // below the surface's lowest point, or below its lowest point over
      // this segment's track
      const float ztop = fmaxf(L.z, L.z + d * L.vz);
      skip = ztop < P.zs_min || ztop < surface_floor(P, L.x, L.y, L.vx, L.vy, d);
    }
    if (!skip) ttop = free_surface_hit(P, L.x, L.y, L.z, L.vx, L.vy, L.vz);
  }
  if (!(ttop > 0.0f)) ttop = kInfF;
  const bool axial_top = (ttop <= tbot && ttop < tside);
  const bool axial_bot = (tbot < ttop && tbot < tside);
  const float tbound = axial_top ? ttop : (axial_bot ? tbot : tside);
  if (tbound >= kInfF * 0.5f) { terminate(L, 4, PF_WATER_NO_BOUNDARY, SUB_WATER_NO_BOUNDARY); return; }
  const bool do_particle = freep < tbound;
  const float travel = do_particle ? freep : tbound;
  if constexpr (kKind == CLR_KERNEL_LIGHT) {
    if (D->vis > 0) heat_raster(*D, free_surface_z(P, P.rin, 0.0f), L.x, L.y, L.z, L.vx, L.vy, L.vz, travel);
    light_segment(*D, rng_counter(L.r), L.x, L.y, L.z, L.vx, L.vy, L.vz, travel, L.ic);
  }
  A.n_flight++;
  L.x += travel * L.vx; L.y += travel * L.vy; L.z += travel * L.vz;
  if constexpr (kind_diag<kKind>()) V.wp += travel;
  if (do_particle) { L.region = R_WPART; return; }
  L.ev = axial_top ? EV_TOP : (axial_bot ? EV_BOTTOM : EV_SIDE);
  L.region = R_WIFACE;
}

// Adds the number of active lanes with `pred` to the workgroup's LDS counter
// cc[idx], with one atomic per wave (the conditional law's rarer sampler
// counts, which keep no per-lane register).
__device__ __forceinline__ void wave_count(unsigned long long* cc, int idx, bool pred) {
  const uint64_t m = __ballot(pred);
  const uint64_t act = __ballot(1);
  if (m && (int)__lane_id() == __ffsll((unsigned long long)act) - 1) atomicAdd(cc + idx, (unsigned long long)__popcll(m));
}
// a closure call's outcome in the sampler counts (CLARITAS.py:4498-4535)
__device__ __forceinline__ void count_closure(unsigned long long* cc, const CondOut& h) {
  const bool ok = (h.status == 1 || h.status == 2);
  wave_count(cc, CLR_C_COND_ACC_GEOM, ok && h.channel == 0);
  wave_count(cc, CLR_C_COND_ACC_COHERENT, ok && h.channel == 1);
  wave_count(cc, CLR_C_COND_ACC_MORPH, ok && h.channel == 2);
  wave_count(cc, CLR_C_COND_ACC_ABS, h.status == 2);
  wave_count(cc, CLR_C_COND_ACC_SCAT, h.status == 1);
  wave_count(cc, CLR_C_COND_EXHAUSTED, !ok && h.subcode == 9001);
  wave_count(cc, CLR_C_COND_NUMERICAL, !ok && h.subcode != 9001);
}

// The end of a conditional closure (CLARITAS.py:4854-4933): the sampler
// counts, a failure, or the accepted event applied to the lane.  The lane is
// still at the event site with its incoming direction.
template <int kKind>
__device__ __forceinline__ void closure_end(Lane& L, const DevParams& P, Acc& A, const LightDiag* __restrict__ D,
                                            Prov& V, unsigned long long* cc, const CondOut& h, bool fallback,
                                            int cbin) {
  constexpr bool kLight = (kKind == CLR_KERNEL_LIGHT), kDiag = kind_diag<kKind>();
  const float sx0 = L.x, sy0 = L.y, sz0 = L.z, vix = L.vx, viy = L.vy, viz = L.vz;
  const uint64_t gid = kLight ? rng_counter(L.r) : 0ull;
  A.trials += h.trials; A.rejects += h.rejects; A.inacc += h.inacc; A.n_abs += h.pc.n_abs;
  count_closure(cc, h);
  if (fallback ? (h.status != 1) : (h.status != 1 && h.status != 2)) {
    terminate(L, 4, h.failure ? h.failure : PF_OTHER, h.subcode);
    return;
  }
  L.x = h.x; L.y = h.y; L.z = h.z; L.vx = h.vx; L.vy = h.vy; L.vz = h.vz;
  L.pr += h.pc.pr;
  if constexpr (kDiag) { V.per += h.pc.per; V.pir += h.pc.pir; }
  if (h.pc.per) L.flags |= F_PER;
  if (h.pc.pir) L.flags |= F_PIR;
  if (h.pc.pabs) L.flags |= F_PABS;
  if (fallback) L.gic--;
  if (h.channel == 0) L.gic++; else { L.wic++; if (h.channel == 1) L.wcoh++; }
  if (!fallback && h.status == 2) {
    if constexpr (kLight) light_absorb(*D, gid, sx0, sy0, sz0);
    terminate(L, 5);
    return;
  }
  {
    const int fres = h.pc.per ? 1 : (h.pc.pir ? 3 : 2);  // external reflection / internal reflection / transmission
    const int type = (h.channel == 1) ? 4 : (h.channel == 2 ? fres + 4 : fres);
    if constexpr (kLight) light_event(*D, gid, sx0, sy0, sz0, type);
    if constexpr (kDiag) prov_commit(V, L, P, vix, viy, viz, sx0, sy0, sz0, type, cbin, h.pc);
  }
}
// CLOSE pass: one trial of the lane's conditional closure (state in s_clo)
template <int kKind>
__device__ __forceinline__ void water_closure(Lane& L, const DevParams& P, Acc& A, const LightDiag* __restrict__ D,
                                              Prov& V, unsigned long long* cc) {
  if constexpr (kKind != CLR_KERNEL_PLAIN) {
    constexpr bool kRough = kind_rough<kKind>();
    const int tid = threadIdx.x;
    const uint32_t w0 = s_clo[0][tid];
    const int cbin = (int)(w0 & 0xFFFFu), ctrials = (int)s_clo[1][tid];
    const bool fallback = (w0 & 0x10000u) != 0;
    CondState st;
    st.target_abs = (w0 & 0x20000u) != 0;
    st.trials = s_clo[2][tid]; st.rejects = s_clo[3][tid]; st.inacc = s_clo[4][tid]; st.nabs = s_clo[5][tid];
    CondOut h;
    if (cond_trial<kRough>(P, L.r, cbin, L.x, L.y, L.z, L.vx, L.vy, L.vz, ctrials, st, h)) {
      L.region = R_WATER;
      closure_end<kKind>(L, P, A, D, V, cc, h, fallback, cbin);
    } else {
      s_clo[2][tid] = st.trials; s_clo[3][tid] = st.rejects; s_clo[4][tid] = st.inacc; s_clo[5][tid] = st.nabs;
    }
  }
}

// the particle event at the end of a flight (CLARITAS.py:4774-4981).  Under
// the conditional law the sampler counts of CLARITAS.py:4854-4933 go to cc
// (the workgroup's LDS counters).
template <int kKind>
__device__ __forceinline__ void water_particle(Lane& L, const DevParams& P, Acc& A, const LightDiag* __restrict__ D,
                                               Prov& V, unsigned long long* cc) {
  constexpr bool kPlain = (kKind == CLR_KERNEL_PLAIN);
  constexpr bool kLight = (kKind == CLR_KERNEL_LIGHT), kRough = kind_rough<kKind>();
  constexpr bool kDiag = kind_diag<kKind>();
  L.region = R_WATER;  // a terminate() below overrides
  // event site and incoming direction (diagnostics), history id (light grids)
  const float sx0 = L.x, sy0 = L.y, sz0 = L.z;
  const float vix = L.vx, viy = L.vy, viz = L.vz;
  const uint64_t gid = kLight ? rng_counter(L.r) : 0ull;
  const double ri2w = (double)P.rin * (double)P.rin, wtol = 8.0 * kFEps * ri2w;
  const float MU_TOTAL = P.mu_geom + P.mu_wave;
  const float MU_TRACK = P.mu_track;
  const double er2 = (double)L.x * L.x + (double)L.y * L.y;
  if (er2 > ri2w + wtol || L.z < P.zmin - kEpsF || L.z > surface_z_vs(P, L.x, L.y, L.z) + kEpsF) {
    terminate(L, 4, PF_INVALID_MEDIUM, SUB_WATER_OUTSIDE);
    return;
  }
  float LG = P.mu_geom, LW = P.mu_wave, LT = MU_TOTAL;
  if (P.sed_model != CLR_SED_UNIFORM) {
    if (kPlain || P.sed_model == CLR_SED_RZ_FIELD || P.sed_model == CLR_SED_XYZ_FIELD) {
      const float s = (P.sed_model == CLR_SED_RZ_FIELD)
                          ? rz_field_f32(P.field, P.fn0, P.fn1, L.x, L.y, L.z, P.rin, P.zmin, P.field_inv0,
                                         P.field_inv1)
                          : xyz_field_f32(P.field, P.fn0, P.fn1, P.fn2, L.x, L.y, L.z, P.field_half, P.zmin,
                                          P.field_inv0, P.field_inv1, P.field_inv2);
      LG = P.sum_g_f * s;
      LW = P.sum_w_f * s;
    } else {
      double lg = 0.0, lw = 0.0;
      for (int i = 0; i < P.n_bins; ++i) {
        const double sc = sediment_scale(P, i, L.x, L.y, L.z);
        lg += P.trace_g[i] * sc;
        lw += P.trace_w[i] * sc;
      }
      LG = (float)lg; LW = (float)lw;
    }
    LT = LG + LW;
    if (!(LT > 0.0f)) { A.n_null++; return; }
    // the majorant the collision was sampled with: the slab's, or mu_track
    const float MU_LOC = (P.n_slabs > 0) ? s_slab_mu[slab_of(P, L.z)] : MU_TRACK;
    const float slack = 64.0f * 1.1920928955078125e-7f * fmaxf(fabsf(MU_LOC), 1.0f);
    if (LT > MU_LOC + slack) { terminate(L, 4, PF_MAJORANT, SUB_MAJORANT); return; }
    const float pacc = fminf(LT / fmaxf(MU_LOC, 1e-30f), 1.0f);
    if (U(L.r) >= pacc) { A.n_null++; return; }
  }
  L.ic++;
  const bool use_wave = (LW > 0.0f && U(L.r) * LT >= LG);
  const int em = P.event_model;
  if (kPlain || em == CLR_EVENT_PLAIN_MIE) {
    const int b = select_bin(L.r, P, false, L.x, L.y, L.z, LG);
    A.n_abs++;
    if (U(L.r) >= P.albedo[b]) {
      L.flags |= F_PABS;
      if constexpr (kLight) light_absorb(*D, gid, sx0, sy0, sz0);
      terminate(L, 5);
      return;
    }
    L.wic++; L.wcoh++;
    mie_scatter(L.r, P, b, L.vx, L.vy, L.vz);
    if constexpr (kLight) light_event(*D, gid, sx0, sy0, sz0, 4);
    if constexpr (kDiag) prov_commit(V, L, P, vix, viy, viz, sx0, sy0, sz0, 4, b, PCount{0, 0, 0, 0, 0u, 0.0f, -1.0f, -1.0f, 0, 0});
    return;
  }
  if constexpr (!kPlain) {
  // One conditional-closure call site and one finite-particle call site serve
  // every branch below (each inlines the microsurface walk once).
  const bool cond = (em == CLR_EVENT_HYBRID_EXACT_RATE);
  int sel = -1;
  if (cond) sel = select_bin(L.r, P, use_wave, L.x, L.y, L.z, use_wave ? LW : LG);
  int cbin = -1, ctrials = 0;
  bool fallback = false;  // closure reached from an inaccessible geometric candidate
  int pidx = -1;
  bool sph = false, dir_only = false;
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, fx = 1.0f, fy = 0.0f, rad = 0.0f;
  double dcx = 0.0, dcy = 0.0, dcz = 0.0;
  if (cond && P.k_imag[sel] > 0.0f) {
    cbin = sel; ctrials = P.cond_trials;
  } else if (use_wave) {
    L.wic++;
    pidx = sel >= 0 ? sel : select_bin(L.r, P, true, L.x, L.y, L.z, LW);
    const bool coherent = (P.coherent >= 1.0f) || (U(L.r) < P.coherent);
    if (cond) {
      A.trials++;
      wave_count(cc, CLR_C_COND_ACC_COHERENT, coherent);
      wave_count(cc, CLR_C_COND_ACC_SCAT, coherent);
    }
    if (coherent) {
      L.wcoh++;
      mie_scatter(L.r, P, pidx, L.vx, L.vy, L.vz);
      if constexpr (kLight) light_event(*D, gid, sx0, sy0, sz0, 4);
      if constexpr (kDiag)
        prov_commit(V, L, P, vix, viy, viz, sx0, sy0, sz0, 4, pidx, PCount{0, 0, 0, 0, 0u, 0.0f, -1.0f, -1.0f, 0, 0});
      return;
    }
    // morphology-wave event about the local azimuthal flow direction,
    // CLARITAS.py:4862-4881, 2783-2799 (configurations with a coherent
    // fraction below 1 run on the ROUGH variant)
    if constexpr (!kRough) {
      if (cond) wave_count(cc, CLR_C_COND_NUMERICAL, true);
      terminate(L, 4, PF_SPHEROID);
      return;
    }
    dir_only = true;
    const float fr = sqrtf(L.x * L.x + L.y * L.y);
    if (fr > 1e-12f) { fx = -L.y / fr; fy = L.x / fr; }
    rad = P.radius[pidx];
    if (rad <= 0.0f) {  // a point particle: accepted, direction kept
      if (cond) { wave_count(cc, CLR_C_COND_ACC_MORPH, true); wave_count(cc, CLR_C_COND_ACC_SCAT, true); }
      return;
    }
    if (!(P.aspect > 0.0f) || !isfinite(P.aspect)) {
      if (cond) wave_count(cc, CLR_C_COND_NUMERICAL, true);
      terminate(L, 4, PF_SPHEROID, SUB_SPHEROID_ASPECT);
      return;
    }
    if (fabsf(P.aspect - 1.0f) < 1e-7f) {
      sphere_sample(L.r, L.vx, L.vy, L.vz, gx, gy, gz);
      dcx = -(double)rad * (double)gx; dcy = -(double)rad * (double)gy; dcz = -(double)rad * (double)gz;
    } else {
      sph = true;
    }
  } else {
    L.gic++;
    if (cond) A.trials++;
    pidx = sel >= 0 ? sel : select_bin(L.r, P, false, L.x, L.y, L.z, LG);
    rad = P.radius[pidx];
    sphere_sample(L.r, L.vx, L.vy, L.vz, gx, gy, gz);
    dcx = (double)L.x - (double)rad * (double)gx;
    dcy = (double)L.y - (double)rad * (double)gy;
    dcz = (double)L.z - (double)rad * (double)gz;
    const bool accessible = sphere_accessible(P, dcx, dcy, dcz, rad);
    if (!accessible && cond) {
      A.inacc++;
      cbin = pidx; ctrials = P.cond_trials - 1; fallback = true;
    } else if (!accessible) {
      L.ic--; L.gic--; A.n_null++;
      if constexpr (kLight) light_mark(*D, gid, 8);
      return;
    }
  }
  if (cbin >= 0) {
    CondOut h;
    CondState st;
    if (cond_begin<kRough>(P, L.r, cbin, L.x, L.y, L.z, L.vx, L.vy, L.vz, ctrials, st, h)) {
      closure_end<kKind>(L, P, A, D, V, cc, h, fallback, cbin);
      return;
    }
    // the trials run in CLOSE passes, one per pass
    const int tid = threadIdx.x;
    s_clo[0][tid] = (uint32_t)cbin | (fallback ? 0x10000u : 0u) | (st.target_abs ? 0x20000u : 0u);
    s_clo[1][tid] = (uint32_t)ctrials;
    s_clo[2][tid] = 0u; s_clo[3][tid] = 0u; s_clo[4][tid] = 0u; s_clo[5][tid] = 0u;
    L.region = R_WCLOSE;
    return;
  }
  PCount pc{0, 0, 0, 0, 0u};
  int ppf = 0, ppfs = 0;
  float px = dir_only ? 0.0f : L.x, py = dir_only ? 0.0f : L.y, pz = dir_only ? 0.0f : L.z;
  float pvx = L.vx, pvy = L.vy, pvz = L.vz;
  const int ps = particle_event<kRough>(P, L.r, sph, P.n_water, P.n_real[pidx], P.k_imag[pidx], rad, gx, gy, gz, dcx, dcy,
                                dcz, fx, fy, 0.0f, px, py, pz, pvx, pvy, pvz, pc, ppf, ppfs);
  if (!dir_only) { L.x = px; L.y = py; L.z = pz; }
  // a direction-only sphere keeps the incoming direction unless it escaped
  if (!dir_only || sph || ps == 1) { L.vx = pvx; L.vy = pvy; L.vz = pvz; }
  L.pr += pc.pr;
  if constexpr (kDiag) { V.per += pc.per; V.pir += pc.pir; }
  if (pc.per) L.flags |= F_PER;
  if (pc.pir) L.flags |= F_PIR;
  if (pc.pabs) L.flags |= F_PABS;
  A.n_abs += pc.n_abs;
  if (cond) {
    wave_count(cc, CLR_C_COND_ACC_GEOM, ps == 1 && !dir_only);
    wave_count(cc, CLR_C_COND_ACC_MORPH, ps == 1 && dir_only);
    wave_count(cc, CLR_C_COND_ACC_SCAT, ps == 1);
    wave_count(cc, CLR_C_COND_NUMERICAL, ps != 1 && ps != 2);
  }
  if (ps == 2) {
    if constexpr (kLight) light_mark(*D, gid, 9);
    terminate(L, 5);
    return;
  }
  if (ps != 1) {
    if constexpr (kLight) light_mark(*D, gid, 10);
    terminate(L, 4, ppf ? ppf : PF_OTHER, ppfs);
    return;
  }
  {
    const int fres = pc.per ? 1 : (pc.pir ? 3 : 2);
    const int type = dir_only ? fres + 4 : fres;
    if constexpr (kLight) light_event(*D, gid, sx0, sy0, sz0, type);
    if constexpr (kDiag) prov_commit(V, L, P, vix, viy, viz, sx0, sy0, sz0, type, pidx, pc);
  }
  }
  return;
}

// the boundary interaction at the end of a flight (CLARITAS.py:4982-5066)
__device__ __forceinline__ void water_iface(Lane& L, const DevParams& P, Acc& A) {
  L.region = R_WATER;  // reflections stay in the water; transmissions set the region
  const bool axial_top = (L.ev == EV_TOP), axial_bot = (L.ev == EV_BOTTOM);
  bool tir = false;
  A.n_iface++;
  if (axial_top) {
    float snx = 0.0f, sny = 0.0f, snz = -1.0f;
    free_surface_normal(P, L.x, L.y, snx, sny, snz);
    if (!fresnel_transmit(L.r, L.vx, L.vy, L.vz, snx, sny, snz, P.n_water, P.n_air, tir)) {
      L.tr++;
      if (tir) L.flags |= F_TTIR;
      if (!water_surface_offset(P, L.x, L.y, L.z, snx, sny, snz)) terminate(L, 4, PF_INVALID_MEDIUM, SUB_WATER_OUTSIDE);
      return;
    }
    L.x -= kEpsF * snx; L.y -= kEpsF * sny; L.z -= kEpsF * snz;
    L.region = R_HEADSPACE; L.hctx = H_TOP; L.hit = 0;
    return;
  }
  if (axial_bot) {
    // water -> bottom disc, CLARITAS.py:3211-3228
    const double r2 = (double)L.x * (double)L.x + (double)L.y * (double)L.y;
    if (!(r2 < (double)P.rin * (double)P.rin)) { terminate(L, 4, PF_BOTTOM_GEOMETRY, SUB_BOTTOM_TOPOLOGY); return; }
    if (!fresnel_transmit(L.r, L.vx, L.vy, L.vz, 0.0f, 0.0f, 1.0f, P.n_water, P.n_acrylic, tir)) {
      L.bir++;
      if (tir) L.flags |= F_BTIR;
      L.z = P.zmin + kEpsF;
      return;
    }
    L.z = P.zmin - kEpsF;
    enter_acrylic(L, A_BOTTOM);
    return;
  }
  // water -> acrylic at the side wall, CLARITAS.py:5038-5065
  float nx, ny;
  radial_normal(L.x, L.y, nx, ny);
  if (!fresnel_transmit(L.r, L.vx, L.vy, L.vz, -nx, -ny, 0.0f, P.n_water, P.n_acrylic, tir)) {
    L.cir++;
    if (tir) L.flags |= F_CTIR;
    L.x -= kEpsF * nx; L.y -= kEpsF * ny;
    return;
  }
  L.x += kEpsF * nx; L.y += kEpsF * ny;
  enter_acrylic(L, A_WATER);
}

// ------------------------------------------------------------------ tallies --
constexpr int kDetWords = CLR_MAX_GROUPS * 4 * CLR_N_CHANNELS;
constexpr int kAuxWords = CLR_MAX_GROUPS * 2 * CLR_MAX_AUX;

struct LdsTally {
  uint32_t det[kDetWords];
  uint32_t aux[kAuxWords];
  uint32_t auxw[kAuxWords];
  uint32_t status[16];
  uint32_t fail[32];
  uint32_t ich[CLR_N_ICHIST];
  float inc[2 * CLR_N_CHANNELS];
  unsigned long long cnt[CLR_N_COUNTERS];
};

__device__ __forceinline__ void lds_add64(unsigned long long* p, uint32_t v) {
  if (v) atomicAdd(p, (unsigned long long)v);
}

// Score a history that has ended (CLARITAS.py:4397-4435) and fold its
// per-ray reductions (CLARITAS.py:8659-8733) into the workgroup's LDS.  In
// the ROUGH and LIGHT variants the reflecting detector bores, the ring-face
// cavity and the SFH213 weights draw from the history's stream and write
// their (rare) stray counts and photodiode sums straight to `Gt`
// tallies.  Lanes of one wave that hit the same bin are serialised by the
// LDS atomic unit, which is cheaper than aggregating them in software.
template <int kKind>
__device__ __forceinline__ void finish_lane(Lane& L, const DevParams& P, LdsTally& T, Acc& A, clr_record* rec,
                                            uint64_t n_rec, uint64_t hidx, int g, clr_tallies* __restrict__ Gt,
                                            const Prov& V, const TransportLaunch& Lp) {
  constexpr bool kBores = kind_rough<kKind>();
  int nat = -1, mir = -1, aux_bits = 0;
  bool scored = false;
  float inc_n = 0.0f, inc_m = 0.0f, ang_n = -1.0f, ang_m = -1.0f;
  {
    if (L.flags & F_ENTERED) A.entered++;
    if (L.st == 1) {
      const float ex = L.x, ey = L.y, ez = L.z, evx = L.vx, evy = L.vy, evz = L.vz;
      A.n_exit++;
      if (isfinite(ex) && isfinite(evx)) A.valid_air++;
      if (isfinite(ex) && isfinite(ey) && isfinite(ez) && isfinite(evx) && isfinite(evy) && isfinite(evz)) {
        int nb = 0, mb = 0, nfb = 0, mfb = 0;
        bool bores = false;
        if constexpr (kBores) bores = (P.det_bore | P.face_model) != 0;
        if (bores) {
          // direct acceptance first (native then mirror), then the ring-face
          // cavity for a channel that failed, CLARITAS.py:4398-4413
          if (P.det_bore) {
            nat = detector_collimator(L.r, P, ex, ey, ez, evx, evy, evz, false, nb);
            mir = detector_collimator(L.r, P, ex, ey, ez, evx, evy, evz, true, mb);
          } else {
            hardware_detectors(P, ex, ey, ez, evx, evy, evz, nat, mir);
          }
          if (nat < 0 && P.face_model) nat = ring_face_path(L.r, P, ex, ey, ez, evx, evy, evz, false, nb, nfb);
          if (mir < 0 && P.face_model) mir = ring_face_path(L.r, P, ex, ey, ez, evx, evy, evz, true, mb, mfb);
        } else {
          hardware_detectors(P, ex, ey, ez, evx, evy, evz, nat, mir);
        }
        if (nat >= 0)
          inc_n = acosf(fminf(fmaxf(evx * P.det_ux[nat] + evy * P.det_uy[nat], -1.0f), 1.0f)) * (180.0f / kPiF);
        if (mir >= 0)
          inc_m = acosf(fminf(fmaxf(-evx * P.det_ux[mir] + evy * P.det_uy[mir], -1.0f), 1.0f)) * (180.0f / kPiF);
        if constexpr (kBores) {
          if (nat >= 0) {
            if (nb > 0 || nfb > 0) atomicAdd(&Gt->stray[0][nat], 1ull);
            if (nfb > 0) atomicAdd(&Gt->face_stray[0][nat], 1ull);
            if (P.pd_model) atomicAdd(&Gt->photodiode[g][0][nat], (double)sfh213_response(inc_n));
          }
          if (mir >= 0) {
            if (mb > 0 || mfb > 0) atomicAdd(&Gt->stray[1][mir], 1ull);
            if (mfb > 0) atomicAdd(&Gt->face_stray[1][mir], 1ull);
            if (P.pd_model) atomicAdd(&Gt->photodiode[g][1][mir], (double)sfh213_response(inc_m));
          }
        }
        for (int a = 0; a < P.n_aux; ++a) {
          float t;
          if (bore_accept(P, ex, ey, ez, evx, evy, evz, P.aux_ux[a], P.aux_uy[a], t)) aux_bits |= 1 << a;
          if (bore_accept(P, -ex, ey, ez, -evx, evy, evz, P.aux_ux[a], P.aux_uy[a], t)) aux_bits |= 16 << a;
        }
        // the mirrored exit point's azimuth is the negated one
        ang_n = atan2f(ex, ey) * (180.0f / kPiF);
        ang_m = -ang_n;
        if (ang_n < 0.0f) ang_n += 360.0f;
        if (ang_m < 0.0f) ang_m += 360.0f;
        scored = true;
      }
    }
    if (L.flags & F_ENTERED) {
      // rarely non-zero sums: plain LDS atomics
      lds_add64(&T.cnt[CLR_C_SUM_IC], (uint32_t)L.ic);
      lds_add64(&T.cnt[CLR_C_SUM_GIC], (uint32_t)L.gic);
      lds_add64(&T.cnt[CLR_C_SUM_WIC], (uint32_t)L.wic);
      lds_add64(&T.cnt[CLR_C_SUM_WCOH], (uint32_t)L.wcoh);
      lds_add64(&T.cnt[CLR_C_SUM_PR], (uint32_t)L.pr);
      lds_add64(&T.cnt[CLR_C_SUM_CIR], (uint32_t)L.cir);
      lds_add64(&T.cnt[CLR_C_SUM_COR], (uint32_t)L.cor);
      lds_add64(&T.cnt[CLR_C_SUM_TOPR], (uint32_t)L.tr);
      lds_add64(&T.cnt[CLR_C_SUM_BIR], (uint32_t)L.bir);
      lds_add64(&T.cnt[CLR_C_SUM_BOR], (uint32_t)L.bor);
    }
    // history flags: one LDS atomic per wave and counter
    const bool ent = (L.flags & F_ENTERED) != 0;
    wave_count(T.cnt, CLR_C_RAYS_PER, ent && (L.flags & F_PER));
    wave_count(T.cnt, CLR_C_RAYS_PIR, ent && (L.flags & F_PIR));
    wave_count(T.cnt, CLR_C_RAYS_PABS, ent && (L.flags & F_PABS));
    wave_count(T.cnt, CLR_C_RAYS_CTIR, ent && (L.flags & F_CTIR));
    wave_count(T.cnt, CLR_C_RAYS_TOPTIR, ent && (L.flags & F_TTIR));
    wave_count(T.cnt, CLR_C_RAYS_BTIR, ent && (L.flags & F_BTIR));
    wave_count(T.cnt, CLR_C_BALLISTIC, ent && L.ic == 0);
    lds_add64(&T.cnt[CLR_C_N_SCAT], (uint32_t)(L.ic > 0 ? L.ic : 0));
    if constexpr (kind_diag<kKind>()) {
      // a failed history's sample (rare: one atomic per failure)
      if (L.st == 4 && Lp.fail_count != nullptr) {
        const unsigned long long k = atomicAdd(Lp.fail_count, 1ull);
        if (k < Lp.fail_cap) {
          clr_failure_sample f;
          f.history = rng_counter(L.r); f.code = L.pf & 255; f.subcode = L.pf >> 8;
          f.trace_iteration = L.wstep; f.pad_ = 0;
          f.x = L.x; f.y = L.y; f.z = L.z; f.vx = L.vx; f.vy = L.vy; f.vz = L.vz;
          Lp.fail_rec[k] = f;
        }
      }
      // causal record of a detected history: one append per wave
      const bool want = Lp.causal != nullptr && (nat >= 0 || mir >= 0);
      const uint64_t m = __ballot(want);
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned long long base = 0ull;
        if ((int)__lane_id() == leader) base = atomicAdd(Lp.causal_count, (unsigned long long)__popcll(m));
        base = __shfl(base, leader, 64);
        const uint64_t k = base + (uint64_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        if (want && k < Lp.causal_cap) prov_record(Lp.causal + k, L, V, rng_counter(L.r), nat, mir);
      }
    }
    A.draws += rng_draws(L.r);
    if (rec && hidx < n_rec) {
      clr_record q;
      q.x = L.x; q.y = L.y; q.z = L.z; q.vx = L.vx; q.vy = L.vy; q.vz = L.vz;
      q.status = L.st; q.native_id = nat; q.mirror_id = mir; q.interactions = L.ic;
      q.failure = L.pf & 255; q.draws = (int32_t)rng_draws(L.r);
      rec[hidx] = q;
    }
  }
  atomicAdd(&T.status[L.st], 1u);
  if (L.st == 4) {
    const int code = L.pf & 255;
    atomicAdd(&T.fail[code < CLR_N_FAILURE ? code : PF_OTHER], 1u);
    atomicAdd(&Gt->failure_subcodes[subcode_slot(L.pf >> 8)], 1ull);  // rare: global
  }
  if (L.flags & F_ENTERED) atomicAdd(&T.ich[L.ic < CLR_N_ICHIST - 1 ? L.ic : CLR_N_ICHIST - 1], 1u);
  if (nat >= 0) {
    atomicAdd(&T.det[(g * 4 + 0) * CLR_N_CHANNELS + nat], 1u);
    atomicAdd(&T.inc[nat], inc_n);
  }
  if (mir >= 0) {
    atomicAdd(&T.det[(g * 4 + 1) * CLR_N_CHANNELS + mir], 1u);
    atomicAdd(&T.inc[CLR_N_CHANNELS + mir], inc_m);
  }
  if (aux_bits) {
    for (int a = 0; a < CLR_MAX_AUX; ++a) {
      if (aux_bits & (1 << a)) atomicAdd(&T.aux[(g * 2 + 0) * CLR_MAX_AUX + a], 1u);
      if (aux_bits & (16 << a)) atomicAdd(&T.aux[(g * 2 + 1) * CLR_MAX_AUX + a], 1u);
    }
  }
  // legacy exit-position windows |angle - 10 j| <= acceptance; with the
  // acceptance capped at 15 deg only j within +-2 of angle/10 can match.
  // The same rule at the auxiliary angles fills aux_window.
  if (scored) {
    for (int m = 0; m < 2; ++m) {
      const float ang = m ? ang_m : ang_n;
      if (ang < 0.0f || ang > 180.0f) continue;
      const int j0 = (int)rintf(ang * 0.1f);
      for (int k = -2; k <= 2; ++k) {
        const int j = j0 + k;
        if (j >= 0 && j < CLR_N_CHANNELS && !(L.ic <= 0 && j >= 9) && fabsf(ang - 10.0f * (float)j) <= P.legacy_acc)
          atomicAdd(&T.det[(g * 4 + 2 + m) * CLR_N_CHANNELS + j], 1u);
      }
      for (int a = 0; a < P.n_aux; ++a) {
        const float deg = P.aux_deg[a];
        if (!(L.ic <= 0 && deg >= 90.0f) && fabsf(ang - deg) <= P.legacy_acc)
          atomicAdd(&T.auxw[(g * 2 + m) * CLR_MAX_AUX + a], 1u);
      }
    }
  }
}

__device__ void zero_lds(LdsTally& T) {
  uint32_t* w = reinterpret_cast<uint32_t*>(&T);
  const int nw = (int)(sizeof(LdsTally) / 4);
  for (int i = threadIdx.x; i < nw; i += blockDim.x) w[i] = 0u;
}

// flush LDS tallies of the block into the configuration's global tallies
__device__ void flush_lds(LdsTally& T, clr_tallies* __restrict__ G) {
  __syncthreads();
  unsigned long long* gdet = reinterpret_cast<unsigned long long*>(&G->det[0][0][0]);
  for (int i = threadIdx.x; i < kDetWords; i += blockDim.x)
    if (T.det[i]) { atomicAdd(gdet + i, (unsigned long long)T.det[i]); T.det[i] = 0u; }
  unsigned long long* gaux = reinterpret_cast<unsigned long long*>(&G->aux[0][0][0]);
  for (int i = threadIdx.x; i < kAuxWords; i += blockDim.x)
    if (T.aux[i]) { atomicAdd(gaux + i, (unsigned long long)T.aux[i]); T.aux[i] = 0u; }
  unsigned long long* gauxw = reinterpret_cast<unsigned long long*>(&G->aux_window[0][0][0]);
  for (int i = threadIdx.x; i < kAuxWords; i += blockDim.x)
    if (T.auxw[i]) { atomicAdd(gauxw + i, (unsigned long long)T.auxw[i]); T.auxw[i] = 0u; }
  for (int i = threadIdx.x; i < CLR_N_STATUS; i += blockDim.x)
    if (T.status[i]) { atomicAdd((unsigned long long*)&G->status[i], (unsigned long long)T.status[i]); T.status[i] = 0u; }
  for (int i = threadIdx.x; i < CLR_N_FAILURE; i += blockDim.x)
    if (T.fail[i]) { atomicAdd((unsigned long long*)&G->failures[i], (unsigned long long)T.fail[i]); T.fail[i] = 0u; }
  for (int i = threadIdx.x; i < CLR_N_ICHIST; i += blockDim.x)
    if (T.ich[i]) { atomicAdd((unsigned long long*)&G->ic_hist[i], (unsigned long long)T.ich[i]); T.ich[i] = 0u; }
  for (int i = threadIdx.x; i < 2 * CLR_N_CHANNELS; i += blockDim.x)
    if (T.inc[i] != 0.0f) { atomicAdd(&G->incidence_sum[0][0] + i, (double)T.inc[i]); T.inc[i] = 0.0f; }
  for (int i = threadIdx.x; i < CLR_N_COUNTERS; i += blockDim.x)
    if (T.cnt[i]) { atomicAdd((unsigned long long*)&G->counters[i], T.cnt[i]); T.cnt[i] = 0ull; }
  __syncthreads();
}

// fold the per-lane work counters into LDS (wave shuffles, then one LDS
// atomic per wave and counter)
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ void fold_acc(Acc& A, LdsTally& T) {
  const uint32_t vals[12] = {A.n_flight, A.n_null, A.n_iface, A.n_abs, A.steps, A.trials,
                             A.rejects, A.inacc, A.draws, A.entered, A.n_exit, A.valid_air};
  const int idx[12] = {CLR_C_N_FLIGHT, CLR_C_N_NULL, CLR_C_N_IFACE, CLR_C_N_ABS, CLR_C_N_STEPS,
                       CLR_C_COND_TRIALS, CLR_C_COND_REJECT, CLR_C_COND_INACCESSIBLE, CLR_C_RNG_DRAWS,
                       CLR_C_ENTERED, CLR_C_N_EXIT, CLR_C_VALID_AIR};
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const uint32_t s = wave_sum(vals[k]);
    if (__lane_id() == 0 && s) atomicAdd(&T.cnt[idx[k]], (unsigned long long)s);
  }
  A = Acc{};
}

constexpr int kMaxBlock = 1024;

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Occupancy targets (waves per SIMD).  Long dependent FP64/transcendental
// chains and table loads leave issue slots that only other resident waves
// fill, so the register budget is capped to keep more waves resident: the
// plain kernel at 128 VGPRs (4 waves) spills a little and runs faster than at
// 168 (3 waves), while 96 VGPRs (5 waves) spills too much.
#ifndef CLR_PLAIN_WAVES_PER_SIMD
#define CLR_PLAIN_WAVES_PER_SIMD 4
#endif
#ifndef CLR_GENERIC_WAVES_PER_SIMD
#define CLR_GENERIC_WAVES_PER_SIMD 2
#endif
template <int kKind>
__global__ __launch_bounds__(256, kKind == CLR_KERNEL_PLAIN ? CLR_PLAIN_WAVES_PER_SIMD : CLR_GENERIC_WAVES_PER_SIMD) void transport_kernel(TransportLaunch Lp,
                                                        const DevParams* __restrict__ params,
                                                        const WorkItem* __restrict__ items) {
  __shared__ LdsTally T;
  __shared__ int s_item;
  __shared__ unsigned int s_next;
  zero_lds(T);
  // per-wave pass-cost estimates in cycles (initial values are typical
  // costs; they adapt to the configuration within a few passes).  Costs stay
  // in [64, 2^24] and a group has at most 64 lanes, so the 32-bit products
  // of the group choice below stay under 2^31: the whole choice runs on the
  // scalar unit (gfx9 has no scalar 64-bit magnitude compare)
  uint32_t cost[K_N] = {12000, 12000, 3000, 6000, 12000, 12000};
  int cur_cfg = -1;
  Acc A{};
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const int it = (int)atomicAdd(Lp.item_counter, 1u);
      s_item = it;
      s_next = 0u;
      if (it < Lp.n_items) {
        s_key[0] = (uint32_t)items[it].seed;
        s_key[1] = (uint32_t)(items[it].seed >> 32);
      }
    }
    __syncthreads();
    // item fields are block-uniform: readfirstlane makes them SGPR values so
    // every DevParams access below is a scalar (SMEM) load
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= Lp.n_items) break;
    const WorkItem* wp = items + item;
    const int cfg = __builtin_amdgcn_readfirstlane(wp->config);
    const uint64_t w_start = uniform64(wp->start), w_count = uniform64(wp->count);
    const uint64_t w_total = uniform64(wp->n_total), w_offset = uniform64(wp->offset);
    if (cfg != cur_cfg) {
      if (cur_cfg >= 0) {
        fold_acc(A, T);
        flush_lds(T, Lp.tallies + cur_cfg);
      }
      cur_cfg = cfg;
      for (int i = threadIdx.x; i < CLR_MAJ_SLABS; i += blockDim.x) s_slab_mu[i] = params[cfg].slab_mu[i];
      __syncthreads();
    }
    const DevParams& P = params[cfg];
    const uint64_t G = (uint64_t)P.n_groups;
    clr_record* rec = (cfg == 0) ? Lp.records : nullptr;
    Lane L;
    Prov V;
    bool alive = false, exhausted = false;
    for (;;) {
      // Region-coherent scheduling: each pass runs one handler for one group
      // of lanes in the wave.  A pass costs about the same however many
      // lanes are active, so handlers that a lane runs back to back share a
      // pass (RECYCLE scores finished histories and starts new ones; WATER
      // runs a pending particle event, then the next flight; RECYCLE and
      // IFACE passes end with one step of every lane in the acrylic, which
      // new histories and most water-boundary transmissions enter), and the
      // group chosen is the one with the most lanes per cycle of its
      // handler's measured pass cost.
      const bool fin = alive && L.region == R_DONE;
      const uint64_t m_rc = __ballot(fin || (!alive && !exhausted));
      const uint64_t m_w = __ballot(alive && (L.region == R_WATER || L.region == R_WPART));
      const uint64_t m_i = __ballot(alive && L.region == R_WIFACE);
      const uint64_t m_a = __ballot(alive && L.region == R_ACRYLIC);
      const uint64_t m_h = __ballot(alive && L.region == R_HEADSPACE);
      const uint64_t m_c = (kKind == CLR_KERNEL_PLAIN) ? 0ull : __ballot(alive && L.region == R_WCLOSE);
      if (!(m_rc | m_w | m_i | m_a | m_h | m_c)) break;
      int op = K_ACRYLIC;
      uint32_t bn = (uint32_t)__popcll(m_a), bc = cost[K_ACRYLIC];
      const uint32_t cnt[K_N] = {(uint32_t)__popcll(m_rc), (uint32_t)__popcll(m_w), (uint32_t)__popcll(m_i),
                                 (uint32_t)__popcll(m_a), (uint32_t)__popcll(m_h), (uint32_t)__popcll(m_c)};
#ifndef CLR_WATER_BIAS
#define CLR_WATER_BIAS 8
#endif
      // the water pass's cost is weighted up by CLR_WATER_BIAS/4 so lanes parked
      // at cheap handlers return to the water group sooner
#pragma unroll
      for (int k = 0; k < K_N; ++k) {
        const uint32_t ck = (k == K_WATER) ? (cost[k] * CLR_WATER_BIAS) >> 2 : cost[k];
        const uint32_t ek = (cnt[k] && (k == K_RECYCLE || k == K_IFACE)) ? cnt[k] + cnt[K_ACRYLIC] : cnt[k];
        if (ek * bc > bn * ck) { op = k; bn = ek; bc = ck; }
      }
      const uint64_t t_pass = __builtin_amdgcn_s_memtime();
      if (op == K_RECYCLE) {
        if (fin) {
          // history number and jackknife group from the stream counter
          const uint64_t hidx = rng_counter(L.r) - w_offset;
          finish_lane<kKind>(L, P, T, A, rec, Lp.n_records, hidx, (int)((hidx * G) / w_total), Lp.tallies + cfg, V, Lp);
          alive = false;
        }
        // refill: one LDS atomic per wave hands consecutive history numbers
        // to the lanes that need one, in lane order
        const bool need = !alive && !exhausted;
        const uint64_t m_need = __ballot(need);
        if (m_need) {
          const uint32_t n_need = (uint32_t)__popcll(m_need);
          const int leader = __ffsll((unsigned long long)m_need) - 1;
          uint32_t base = 0u;
          if ((int)__lane_id() == leader) base = atomicAdd(&s_next, n_need);
          base = __shfl(base, leader, 64);
          const uint32_t rank = (uint32_t)__popcll(m_need & ((1ull << __lane_id()) - 1ull));
          if (need) {
            const unsigned int k = base + rank;
            if ((uint64_t)k < w_count) {
              rng_seed(L.r, w_offset + w_start + k);
              L.region = R_WATER; L.st = 0; L.pf = PF_OK; L.wstep = 0;
              L.ic = L.gic = L.wic = L.wcoh = L.pr = L.cir = L.cor = L.tr = L.bir = L.bor = 0;
              L.flags = 0u;
              if constexpr (kind_diag<kKind>()) prov_init(V);
              L.actx = A_LAUNCH; L.hctx = H_LAUNCH; L.it = 0; L.hit = 0; L.stuck = 0; L.last_surf = -1;
              int sbi = 0, sbr = 0;
              launch_history<kKind>(L, P, A, sbi, sbr);  // may end it at once: scored by a later RECYCLE pass
              if constexpr (kind_rough<kKind>()) {
                if (P.src_bore) {
                  lds_add64(&T.cnt[CLR_C_SRC_BORE_INTER], (uint32_t)sbi);
                  lds_add64(&T.cnt[CLR_C_SRC_BORE_REFL], (uint32_t)sbr);
                  if (L.st != 3 && sbr > 0) atomicAdd(&T.cnt[CLR_C_SRC_REFL_ACCEPT], 1ull);
                }
              }
              alive = true;
            } else {
              exhausted = true;
            }
          }
        }
      } else if (op == K_WATER) {
        if (alive && (L.region == R_WPART || L.region == R_WATER)) rng_topup(L.r);
        if (alive && L.region == R_WPART) water_particle<kKind>(L, P, A, Lp.light, V, T.cnt);
        if (alive && L.region == R_WATER) water_flight<kKind>(L, P, A, Lp.light, V);
      } else if (op == K_IFACE) {
        if (alive && L.region == R_WIFACE) water_iface(L, P, A);
      } else if (op == K_CLOSE) {
        if constexpr (kKind != CLR_KERNEL_PLAIN) {
          if (alive && L.region == R_WCLOSE) rng_topup(L.r);
          if (alive && L.region == R_WCLOSE) water_closure<kKind>(L, P, A, Lp.light, V, T.cnt);
        }
      } else if (op == K_HEAD) {
        const bool was_head = alive && L.region == R_HEADSPACE;
        if (alive && L.region == R_HEADSPACE) headspace_step(L, P, A);
        if constexpr (kind_diag<kKind>())
          if (was_head && L.region == R_WATER && L.ic > 0) V.ret++;
      }
      if (op == K_ACRYLIC || op == K_RECYCLE || op == K_IFACE) {
        const bool was_acrylic = alive && L.region == R_ACRYLIC;
        if (alive && L.region == R_ACRYLIC) acrylic_step(L, P, A);
        // A chain of kLongChain bounces in one acrylic visit is almost always
        // trapped by total internal reflection and runs on to the emergency
        // cap; stepping it to the end now (the rest of the wave waits) keeps
        // such a chain from trailing past the end of the launch.  The same
        // applies when acrylic lanes are all this wave has left.
        const bool only = op == K_ACRYLIC && !(m_rc | m_w | m_i | m_h | m_c);
        while (__any(alive && L.region == R_ACRYLIC && (only || L.it >= kLongChain))) {
          if (alive && L.region == R_ACRYLIC && (only || L.it >= kLongChain)) acrylic_step(L, P, A);
        }
        // a return to the water after a particle event (CLARITAS.py:5009-5058)
        if constexpr (kind_diag<kKind>())
          if (was_acrylic && L.region == R_WATER && L.ic > 0) V.ret++;
      }
      // running estimate of the pass cost (cycles, 1/8 weight to the new
      // pass).  Constant indices keep cost[] in scalar registers: an index
      // by op would move the array through vector registers every pass.
      const int32_t dt = (int32_t)min((int64_t)(__builtin_amdgcn_s_memtime() - t_pass), (int64_t)(1 << 24));
#pragma unroll
      for (int k = 0; k < K_N; ++k)
        if (k == op) cost[k] = (uint32_t)max((int32_t)cost[k] + ((dt - (int32_t)cost[k]) >> 3), 64);
    }
  }
  if (cur_cfg >= 0) {
    fold_acc(A, T);
    flush_lds(T, Lp.tallies + cur_cfg);
  }
}

}  // namespace

namespace {
// the reference's diagnostic test kernels (CLARITAS.py:3470-3485): mode 0
// deposits segment t of data[n][7] (x0, y0, z0, vx, vy, vz, dist) with
// particle order t & 1; mode 1 records event site t of data[n][3]
__global__ void light_probe_kernel(const LightDiag* __restrict__ D, const float* __restrict__ data, int n, int mode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (mode == 2) {  // heat-map raster of segment t, z axis to D->zmax
    const float* q = data + 7 * (size_t)t;
    heat_raster(*D, D->zmax, q[0], q[1], q[2], q[3], q[4], q[5], q[6]);
  } else if (mode == 0) {
    const float* q = data + 7 * (size_t)t;
    light_segment(*D, (uint64_t)t, q[0], q[1], q[2], q[3], q[4], q[5], q[6], t & 1);
  } else {
    const float* q = data + 3 * (size_t)t;
    light_site(*D, (uint64_t)t, q[0], q[1], q[2], D->event_weight, D->event_count, D->event_total);
  }
}

// The transport's building blocks on explicit inputs, one item per thread
// (the counterpart of the reference's test kernels free_surface_geometry_test_kernel,
// sediment_field_test_kernel, bore_wall_test_kernel and
// water_surface_offset_test_kernel, CLARITAS.py:4438-4455, 4228-4235,
// 2952-2956, 2980-2985).  Item i reads in[i * in_stride ..] and writes
// out[i * out_stride ..]; random streams are (seed, history i).  Modes and
// layouts are listed with clr_probe in include/claritas.h.
__global__ void probe_kernel(const DevParams* __restrict__ PP, int mode, int n, const float* __restrict__ in,
                             int in_stride, double* __restrict__ out, int out_stride, uint64_t seed) {
  if (threadIdx.x == 0) { s_key[0] = (uint32_t)seed; s_key[1] = (uint32_t)(seed >> 32); }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevParams& P = *PP;
  const float* q = in + (size_t)i * in_stride;
  double* o = out + (size_t)i * out_stride;
  Rng r;
  rng_seed(r, (uint64_t)i);
  switch (mode) {
    case CLR_PROBE_SURFACE: {
      o[0] = free_surface_z(P, q[0], q[1]);
      const float t = free_surface_hit(P, q[0], q[1], q[2], q[3], q[4], q[5]);
      o[1] = t;
      float nx = 0.0f, ny = 0.0f, nz = 0.0f, zh = kInfF;
      if (t < 0.5f * kInfF) {
        const float hx = q[0] + t * q[3], hy = q[1] + t * q[4];
        zh = free_surface_z(P, hx, hy);
        free_surface_normal(P, hx, hy, nx, ny, nz);
      }
      o[2] = zh; o[3] = nx; o[4] = ny; o[5] = nz;
      break;
    }
    case CLR_PROBE_WALL: {
      o[0] = bore_wall_hit_d(P, q[0], q[1], q[2], q[3], q[4], q[5], P.rin, P.zmin, P.zwall_top);
      o[1] = finite_cylinder_forward_d(P, q[0], q[1], q[2], q[3], q[4], q[5], P.rout, P.zlow, P.zwall_top);
      double ta = 0.0;
      const int surf = acrylic_next_surface(P, q[0], q[1], q[2], q[3], q[4], q[5], ta);
      o[2] = ta; o[3] = surf;
      break;
    }
    case CLR_PROBE_OFFSET: {
      float x = q[0], y = q[1], z = q[2];
      const bool ok = water_surface_offset(P, x, y, z, q[3], q[4], q[5]);
      o[0] = ok ? 1.0 : 0.0; o[1] = x; o[2] = y; o[3] = z;
      break;
    }
    case CLR_PROBE_SEDIMENT: {
      const int b = (int)q[0];
      o[0] = (b >= 0 && b < P.n_bins) ? sediment_scale(P, b, q[1], q[2], q[3]) : -1.0;
      float f = 1.0f;
      if (P.sed_model == CLR_SED_RZ_FIELD)
        f = rz_field_f32(P.field, P.fn0, P.fn1, q[1], q[2], q[3], P.rin, P.zmin, P.field_inv0, P.field_inv1);
      else if (P.sed_model == CLR_SED_XYZ_FIELD)
        f = xyz_field_f32(P.field, P.fn0, P.fn1, P.fn2, q[1], q[2], q[3], P.field_half, P.zmin, P.field_inv0,
                          P.field_inv1, P.field_inv2);
      o[1] = f;
      break;
    }
    case CLR_PROBE_DETECTOR:
      o[0] = hardware_detector(P, q[0], q[1], q[2], q[3], q[4], q[5], false);
      o[1] = hardware_detector(P, q[0], q[1], q[2], q[3], q[4], q[5], true);
      break;
    case CLR_PROBE_SAMPLER: {
      const int k = (int)q[0];
      o[0] = k == 0 ? gamma_draw(r, q[1]) : k == 1 ? beta_draw(r, q[1], q[2])
                                         : k == 2 ? beta_radiance_polar(r, q[1], q[2]) : normal_bm(r);
      o[1] = rng_draws(r);
      break;
    }
    case CLR_PROBE_SOURCE: {
      float x, y, z, vx, vy, vz;
      source_sample(r, P, x, y, z, vx, vy, vz);
      o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
      break;
    }
    case CLR_PROBE_MIE_THETA: {
      const int b = (int)q[0];
      o[0] = (b >= 0 && b < P.n_bins) ? sample_mie_theta(r, P, b) : -1.0;
      break;
    }
    case CLR_PROBE_CONDITIONAL: {
      const int b = (int)q[0];
      if (b < 0 || b >= P.n_bins) { o[0] = -1.0; break; }
      const bool rough = P.rough > 1e-7f || P.coherent < 1.0f;
      const CondOut c = rough ? conditional_event<true>(P, r, b, q[1], q[2], q[3], q[4], q[5], q[6], P.cond_trials)
                              : conditional_event<false>(P, r, b, q[1], q[2], q[3], q[4], q[5], q[6], P.cond_trials);
      o[0] = c.status; o[1] = c.channel; o[2] = c.trials; o[3] = c.rejects; o[4] = c.inacc; o[5] = c.failure;
      o[6] = c.vx; o[7] = c.vy; o[8] = c.vz;
      break;
    }
    default:
      break;
  }
}
}  // namespace

hipError_t clr_launch_light_probe(const LightDiag* D, const float* data, int n, int mode, hipStream_t s) {
  hipLaunchKernelGGL(light_probe_kernel, dim3((n + 63) / 64), dim3(64), 0, s, D, data, n, mode);
  return hipGetLastError();
}

hipError_t clr_launch_probe(const DevParams* P, int mode, int n, const float* in, int in_stride, double* out,
                            int out_stride, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(probe_kernel, dim3((n + 63) / 64), dim3(64), 0, s, P, mode, n, in, in_stride, out, out_stride,
                     seed);
  return hipGetLastError();
}

hipError_t clr_launch_transport(const TransportLaunch& L, hipStream_t s) {
  if (L.kind == CLR_KERNEL_PLAIN)
    hipLaunchKernelGGL(transport_kernel<CLR_KERNEL_PLAIN>, dim3(L.grid), dim3(L.block), 0, s, L, L.params, L.items);
  else if (L.kind == CLR_KERNEL_ROUGH)
    hipLaunchKernelGGL(transport_kernel<CLR_KERNEL_ROUGH>, dim3(L.grid), dim3(L.block), 0, s, L, L.params, L.items);
  else if (L.kind == CLR_KERNEL_LIGHT)
    hipLaunchKernelGGL(transport_kernel<CLR_KERNEL_LIGHT>, dim3(L.grid), dim3(L.block), 0, s, L, L.params, L.items);
  else if (L.kind == CLR_KERNEL_CAUSAL)
    hipLaunchKernelGGL(transport_kernel<CLR_KERNEL_CAUSAL>, dim3(L.grid), dim3(L.block), 0, s, L, L.params, L.items);
  else
    hipLaunchKernelGGL(transport_kernel<CLR_KERNEL_GENERIC>, dim3(L.grid), dim3(L.block), 0, s, L, L.params,
                       L.items);
  return hipGetLastError();
}

int clr_transport_occupancy(int block, int kind) {
  int n = 0;
  hipError_t e;
  if (kind == CLR_KERNEL_PLAIN)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, transport_kernel<CLR_KERNEL_PLAIN>, block, 0);
  else if (kind == CLR_KERNEL_ROUGH)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, transport_kernel<CLR_KERNEL_ROUGH>, block, 0);
  else if (kind == CLR_KERNEL_LIGHT)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, transport_kernel<CLR_KERNEL_LIGHT>, block, 0);
  else if (kind == CLR_KERNEL_CAUSAL)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, transport_kernel<CLR_KERNEL_CAUSAL>, block, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, transport_kernel<CLR_KERNEL_GENERIC>, block, 0);
  if (e != hipSuccess) return 1;
  return n > 0 ? n : 1;
}
