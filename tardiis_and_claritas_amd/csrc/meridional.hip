// meridional.hip -- stationary meridional advection-diffusion sediment
// field on gfx950: one r-z plane per particle bin, nr x nz cells, z fastest.
//
// The reference assembles an upwind finite-volume system per bin and relaxes
// it with one CUDA launch per red/black half-sweep, a single-thread
// normalisation launch every 20 iterations and a host synchronisation every
// check interval (CLARITAS.py:4091-4215 kernels, 6451-6725 driver).  Here one
// workgroup owns one bin for a whole check interval: the plane sits in LDS,
// the five coefficients of each thread's cells sit in registers, a
// half-sweep is one pass over the thread's cells of that colour between two
// workgroup barriers, and the mass and residual are workgroup reductions.
// The host sees the device once per check interval (residuals), as the
// reference's driver does.
//
// Planes with more than 3 cells of each colour per 1024-thread workgroup
// (above 6144 cells; the default plane is 48 x 96 = 4608) run the same sweep
// on the grid in global memory (L2-resident) with the coefficients read per
// update.
#include <hip/hip_runtime.h>

#include "meridional.h"

namespace {

constexpr unsigned kE = 1u, kW = 2u, kN = 4u, kS = 8u, kActive = 16u, kRelax = 32u;
constexpr int kIShift = 8;       // row index i stored above the flag bits
constexpr int kMaxK = 3;         // colour cells per thread in the LDS sweep (4 would spill at 128 VGPRs)
constexpr int kMaxLdsBytes = 65536;

// free-surface height and slope at radius r, CLARITAS.py:4046-4057
// (fsm 1 parabolic, 2 Scully)
__device__ __forceinline__ void surface_geometry(int fsm, double r, double zmean, double R, double D, double a,
                                                 double& zs, double& hp) {
  const double rr = fmin(fmax(r, 0.0), R), rr2 = rr * rr, R2 = R * R;
  zs = zmean;
  hp = 0.0;
  if (!(D > 0.0) || fsm == 0) return;
  if (fsm == 1) {
    zs = zmean + D * (rr2 / R2 - 0.5);
    hp = 2.0 * D * rr / R2;
    return;
  }
  const double a2 = a * a, Hinf = D * (a2 + R2) / R2, hbar = Hinf * (1.0 - (a2 / R2) * log1p(R2 / a2));
  zs = zmean - hbar + Hinf * rr2 / (a2 + rr2);
  hp = 2.0 * Hinf * a2 * rr / ((a2 + rr2) * (a2 + rr2));
}

// streamfunction velocity psi = -2U r^2 (1-r^2/R^2)^2 eta(1-eta),
// CLARITAS.py:4059-4073
__device__ __forceinline__ void velocity(const MeridionalLaunch& L, double r, double z, double& ur, double& uz) {
  ur = 0.0;
  uz = 0.0;
  if (!(L.upeak > 0.0) || !(L.R > 0.0) || !(L.dh > 0.0) || (L.fs_model == 2 && !(L.core > 0.0))) return;
  const double R = L.R, rr = fmin(fmax(r, 0.0), R), R2 = R * R, rr2 = rr * rr;
  double zs, hp;
  surface_geometry(L.fs_model, rr, L.zmean, R, L.dh, L.core, zs, hp);
  const double H = zs - L.zmin;
  if (!(H > 0.0)) return;
  double eta = (z - L.zmin) / H;
  eta = fmin(fmax(eta, 0.0), 1.0);
  const double s = rr2 / R2, f = eta * (1.0 - eta), fp = 1.0 - 2.0 * eta, A = 2.0 * L.upeak, om = 1.0 - s;
  ur = A * rr * om * om * fp / H;
  uz = -A * (2.0 * om * (1.0 - 3.0 * s) * f - rr * om * om * fp * eta * hp / H);
}

// linear lookup of the particle-relative radial slip, CLARITAS.py:4075-4081
__device__ __forceinline__ double radial_slip(const double* WR, int nrad, double r, double R) {
  if (nrad < 2 || !(R > 0.0)) return 0.0;
  const double q = fmin(fmax(r / R, 0.0), 1.0) * (double)(nrad - 1);
  int j = (int)floor(q);
  double t = q - (double)j;
  if (j >= nrad - 1) { j = nrad - 2; t = 1.0; }
  return (1.0 - t) * WR[j] + t * WR[j + 1];
}

// CLARITAS.py:4083-4088
__device__ __forceinline__ bool physical_fluid(const MeridionalLaunch& L, double r, double z) {
  if (z < L.zmin) return false;
  double zs, hp;
  surface_geometry(L.fs_model, r, L.zmean, L.R, L.dh, L.core, zs, hp);
  return z < zs;
}

// finite-Re drift-diffusion scale at a float point on the x axis (the
// reference's sediment_local_scale with model 2, CLARITAS.py:4228-4262):
// the solver's initial guess, CLARITAS.py:4100-4110
__device__ double analytic_scale(const double* radlog, int nrad, double B, double norm, float x, float z, float R,
                                 float zmin) {
  double rr2 = (double)x * (double)x;
  const double R2 = (double)R * (double)R;
  if (rr2 > R2) rr2 = R2;
  const double rr = sqrt(fmax(rr2, 0.0));
  const double u = (R > 0.0f) ? rr / (double)R * (double)(nrad - 1) : 0.0;
  int j = (int)floor(u);
  j = min(max(j, 0), nrad - 2);
  const double t = u - (double)j;
  double rl = (1.0 - t) * radlog[j] + t * radlog[j + 1];
  if (rl > 0.0) rl = 0.0;
  const double h = fmax((double)z - (double)zmin, 0.0);
  if (!(norm > 0.0)) return 0.0;
  return exp(rl - B * h) / norm;
}

// bilinear sample of one bin's double grid, CLARITAS.py:4205-4226 (y = 0)
__device__ double grid_sample(const double* G, int NR, int NZ, float x, float z, float R, float zmin, float zmax) {
  if (NR < 2 || NZ < 2 || !(zmax > zmin)) return 1.0;
  double rr = sqrt(fmax((double)x * (double)x, 0.0));
  if (rr > (double)R) rr = (double)R;
  const double dr = (double)R / (double)NR, dz = ((double)zmax - (double)zmin) / (double)NZ;
  double ur = rr / dr - 0.5, uz = ((double)z - (double)zmin) / dz - 0.5;
  ur = fmin(fmax(ur, 0.0), (double)(NR - 1));
  uz = fmin(fmax(uz, 0.0), (double)(NZ - 1));
  int i = (int)floor(ur), k = (int)floor(uz);
  double tr = ur - (double)i, tz = uz - (double)k;
  if (i >= NR - 1) { i = NR - 2; tr = 1.0; }
  if (k >= NZ - 1) { k = NZ - 2; tz = 1.0; }
  const double q00 = G[(size_t)i * NZ + k], q10 = G[(size_t)(i + 1) * NZ + k];
  const double q01 = G[(size_t)i * NZ + k + 1], q11 = G[(size_t)(i + 1) * NZ + k + 1];
  const double q0 = (1.0 - tr) * q00 + tr * q10, q1 = (1.0 - tr) * q01 + tr * q11;
  return fmax((1.0 - tz) * q0 + tz * q1, 0.0);
}

// meridional_active_mask_kernel, CLARITAS.py:4091-4098
__global__ void mask_kernel(MeridionalLaunch L) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, plane = L.nr * L.nz;
  if (t >= plane) return;
  const int i = t / L.nz, k = t - i * L.nz;
  const double dr = L.R / (double)L.nr, dz = (L.zgrid - L.zmin) / (double)L.nz;
  L.active[t] = physical_fluid(L, ((double)i + 0.5) * dr, L.zmin + ((double)k + 0.5) * dz) ? 1 : 0;
}

// Initial guess and coefficients of one (bin, cell): the reference's
// meridional_grid_initialize_kernel and meridional_build_coefficients_kernel
// (CLARITAS.py:4100-4148).  Each face is upwinded on its outward volume flux
// F = V . n A with the eddy diffusion D A / h; a face counts only when both
// cells are fluid and the face point lies in the fluid (no-flux walls).
__global__ void setup_kernel(MeridionalLaunch L) {
  const int NR = L.nr, NZ = L.nz;
  const size_t plane = (size_t)NR * NZ, total = (size_t)L.n * plane;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int p = (int)(t / plane), c = (int)(t - (size_t)p * plane), i = c / NZ, k = c - i * NZ;
  double* AP = L.coeff + t;
  double* AE = AP + total;
  double* AW = AE + total;
  double* AN = AW + total;
  double* AS = AN + total;
  *AP = 0.0; *AE = 0.0; *AW = 0.0; *AN = 0.0; *AS = 0.0;
  if (!L.active[c]) { L.grid[t] = 0.0; return; }
  const double R = L.R, dr = R / (double)NR, dz = (L.zgrid - L.zmin) / (double)NZ;
  const int nrad = L.n_radial;
  const float xf = (float)(((double)i + 0.5) * dr), zf = (float)(L.zmin + ((double)k + 0.5) * dz);
  L.grid[t] = analytic_scale(L.radlog + (size_t)p * nrad, nrad, L.sedB[p], L.sednorm[p], xf, zf, (float)R,
                             (float)L.zmin);
  const double* WR = L.radwr + (size_t)p * nrad;
  const double rc = ((double)i + 0.5) * dr, zc = L.zmin + ((double)k + 0.5) * dz;
  const double rw = (double)i * dr, re = (double)(i + 1) * dr;
  const double Az = 0.5 * (re * re - rw * rw), eddy = L.eddy, ws = L.ws[p];
  double ap = 0.0, ur, uz, V, F, diff;
  if (i + 1 < NR && L.active[(i + 1) * NZ + k] && physical_fluid(L, re, zc)) {
    velocity(L, re, zc, ur, uz);
    V = ur + radial_slip(WR, nrad, re, R); F = V * re * dz; diff = eddy * re * dz / dr;
    *AE = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
  }
  if (i > 0 && L.active[(i - 1) * NZ + k] && physical_fluid(L, rw, zc)) {
    velocity(L, rw, zc, ur, uz);
    V = ur + radial_slip(WR, nrad, rw, R); F = -V * rw * dz; diff = eddy * rw * dz / dr;
    *AW = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
  }
  double zface = zc + 0.5 * dz;
  if (k + 1 < NZ && L.active[i * NZ + k + 1] && physical_fluid(L, rc, zface)) {
    velocity(L, rc, zface, ur, uz);
    V = uz - ws; F = V * Az; diff = eddy * Az / dz;
    *AN = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
  }
  zface = zc - 0.5 * dz;
  if (k > 0 && L.active[i * NZ + k - 1] && physical_fluid(L, rc, zface)) {
    velocity(L, rc, zface, ur, uz);
    V = uz - ws; F = -V * Az; diff = eddy * Az / dz;
    *AS = diff + fmax(-F, 0.0); ap += diff + fmax(F, 0.0);
  }
  *AP = ap;
}

// Cell h of colour q ((i + k) & 1 == q) in row-pair order: rows 2m and
// 2m+1 together hold NZ cells of each colour, row 2m's first (k = q, q+2,
// ...), then row 2m+1's (k = 1-q, 3-q, ...).  -1 past the last row.
__device__ __forceinline__ int colour_cell(int h, int q, int NR, int NZ) {
  const int m = h / NZ, r = h - m * NZ, c0 = (NZ - q + 1) >> 1;
  int i, k;
  if (r < c0) { i = 2 * m; k = q + 2 * r; }
  else { i = 2 * m + 1; k = (1 - q) + 2 * (r - c0); }
  return i < NR ? i * NZ + k : -1;
}

__device__ __forceinline__ unsigned cell_flags(const MeridionalLaunch& L, int c, double ap) {
  const int NR = L.nr, NZ = L.nz, i = c / NZ, k = c - i * NZ;
  unsigned f = (i + 1 < NR ? kE : 0u) | (i > 0 ? kW : 0u) | (k + 1 < NZ ? kN : 0u) | (k > 0 ? kS : 0u);
  if (L.active[c]) f |= kActive | (ap > 0.0 ? kRelax : 0u);
  return f | ((unsigned)i << kIShift);
}

// five-point neighbour sum in the reference's order (E, W, N, S),
// CLARITAS.py:4156-4162
__device__ __forceinline__ double stencil(const double* G, int c, int NZ, unsigned f, double ae, double aw, double an,
                                          double as) {
  double num = 0.0;
  if (f & kE) num += ae * G[c + NZ];
  if (f & kW) num += aw * G[c - NZ];
  if (f & kN) num += an * G[c + 1];
  if (f & kS) num += as * G[c - 1];
  return num;
}

// SOR update of one cell, CLARITAS.py:4163-4164: (1 - w) G + w num / AP,
// with w / AP precomputed (one multiply instead of an FP64 divide per
// update; the result moves by at most an ulp of the relaxation term)
__device__ __forceinline__ void relax_cell(double* G, int c, int NZ, unsigned f, double w_ap, double ae, double aw,
                                           double an, double as, double omega) {
  const double num = stencil(G, c, NZ, f, ae, aw, an, as);
  const double nv = (1.0 - omega) * G[c] + num * w_ap;
  G[c] = (!isfinite(nv) || nv < 1e-14) ? 1e-14 : nv;
}

// relative residual of one cell, CLARITAS.py:4185-4187
__device__ __forceinline__ double cell_residual(const double* G, int c, int NZ, unsigned f, double ap, double ae,
                                                double aw, double an, double as) {
  const double num = stencil(G, c, NZ, f, ae, aw, an, as);
  const double den = fmax(fabs(ap * G[c]), 1e-30);
  return fabs(ap * G[c] - num) / den;
}

__device__ __forceinline__ double cell_volume(unsigned f, double dr, double dz) {
  const int i = (int)(f >> kIShift);
  const double rw = (double)i * dr, re = (double)(i + 1) * dr;
  return 0.5 * (re * re - rw * rw) * dz;
}

// Workgroup sum / max; every thread returns the same value (the partials
// are combined in wave order by each thread).  The caller places a barrier
// between two reductions.
__device__ __forceinline__ double block_sum(double v, double* red, int nw) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}
__device__ __forceinline__ double block_max(double v, double* red, int nw) {
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int w = 1; w < nw; ++w) s = fmax(s, red[w]);
  return s;
}

// K > 0: plane in LDS, K cells of each colour per thread with their
// coefficients in registers.  K == 0: plane in global memory, cells looped
// with a stride of the workgroup size, coefficients loaded per update.
template <int K>
__global__ __launch_bounds__(1024) void sweep_kernel(MeridionalLaunch L, int it0, int it1, int initial) {
  extern __shared__ double smem[];
  const int p = blockIdx.x, tid = threadIdx.x, bs = blockDim.x, nw = bs >> 6;
  const int NR = L.nr, NZ = L.nz, plane = NR * NZ, H = ((NR + 1) >> 1) * NZ;
  const size_t total = (size_t)L.n * plane, base = (size_t)p * plane;
  const double* A0 = L.coeff + base;
  const double dr = L.R / (double)NR, dz = (L.zgrid - L.zmin) / (double)NZ, omega = L.omega;
  double* G = (K > 0) ? smem : L.grid + base;
  double* red = (K > 0) ? smem + plane : smem;

  int cell[2][K > 0 ? K : 1];
  unsigned fl[2][K > 0 ? K : 1];
  double cap[2][K > 0 ? K : 1], cae[2][K > 0 ? K : 1], caw[2][K > 0 ? K : 1], can[2][K > 0 ? K : 1],
      cas[2][K > 0 ? K : 1];
  if constexpr (K > 0) {
    for (int c = tid; c < plane; c += bs) G[c] = L.grid[base + c];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int h = tid + j * bs;
        const int c = (h < H) ? colour_cell(h, q, NR, NZ) : -1;
        cell[q][j] = c < 0 ? 0 : c;
        fl[q][j] = 0u;
        cap[q][j] = cae[q][j] = caw[q][j] = can[q][j] = cas[q][j] = 0.0;
        if (c >= 0) {
          const double ap = A0[c];
          cap[q][j] = (ap > 0.0) ? omega / ap : 0.0;  // w / AP for the update
          cae[q][j] = A0[total + c];
          caw[q][j] = A0[2 * total + c];
          can[q][j] = A0[3 * total + c];
          cas[q][j] = A0[4 * total + c];
          fl[q][j] = cell_flags(L, c, ap);
        }
      }
    __syncthreads();
  }

  // mass renormalisation, CLARITAS.py:4166-4177; false (grid unchanged)
  // when the mass is not positive and finite -- the same verdict on every
  // thread, so the loop below stays uniform
  auto normalize = [&]() -> bool {
    double part = 0.0;
    if constexpr (K > 0) {
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j)
          if (fl[q][j] & kActive) part += G[cell[q][j]] * cell_volume(fl[q][j], dr, dz);
    } else {
      for (int c = tid; c < plane; c += bs)
        if (L.active[c]) part += G[c] * cell_volume(cell_flags(L, c, 0.0), dr, dz);
    }
    const double mass = block_sum(part, red, nw);
    if (!(mass > 0.0) || !(L.target > 0.0) || !isfinite(mass)) return false;
    const double scale = L.target / mass;
    if constexpr (K > 0) {
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j)
          if (fl[q][j] & kActive) G[cell[q][j]] *= scale;
    } else {
      for (int c = tid; c < plane; c += bs)
        if (L.active[c]) G[c] *= scale;
    }
    __syncthreads();
    return true;
  };
  auto half_sweep = [&](int q) {
    if constexpr (K > 0) {
#pragma unroll
      for (int j = 0; j < K; ++j)
        if (fl[q][j] & kRelax)
          relax_cell(G, cell[q][j], NZ, fl[q][j], cap[q][j], cae[q][j], caw[q][j], can[q][j], cas[q][j], omega);
    } else {
      for (int h = tid; h < H; h += bs) {
        const int c = colour_cell(h, q, NR, NZ);
        if (c < 0) continue;
        const double ap = A0[c];
        const unsigned f = cell_flags(L, c, ap);
        if (f & kRelax) relax_cell(G, c, NZ, f, omega / ap, A0[total + c], A0[2 * total + c], A0[3 * total + c],
                                   A0[4 * total + c], omega);
      }
    }
    __syncthreads();
  };

  bool ok = initial ? normalize() : true;
  int failed_at = ok ? -1 : 0;
  for (int it = it0 + 1; ok && it <= it1; ++it) {
    half_sweep(0);
    half_sweep(1);
    if (it % 20 == 0) {
      ok = normalize();
      if (!ok) failed_at = it;
    }
  }
  if (ok) {
    ok = normalize();
    if (!ok) failed_at = it1;
  }
  double mr = 0.0;
  if (ok) {
    if constexpr (K > 0) {
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j)
          if (fl[q][j] & kRelax)
            mr = fmax(mr, cell_residual(G, cell[q][j], NZ, fl[q][j], A0[cell[q][j]], cae[q][j], caw[q][j],
                                        can[q][j], cas[q][j]));
    } else {
      for (int c = tid; c < plane; c += bs) {
        const double ap = A0[c];
        const unsigned f = cell_flags(L, c, ap);
        if (f & kRelax)
          mr = fmax(mr, cell_residual(G, c, NZ, f, ap, A0[total + c], A0[2 * total + c], A0[3 * total + c],
                                      A0[4 * total + c]));
      }
    }
    mr = block_max(mr, red, nw);
  }
  if constexpr (K > 0) {
    for (int c = tid; c < plane; c += bs) L.grid[base + c] = G[c];
  }
  if (tid == 0) {
    L.residual[p] = ok ? mr : INFINITY;
    if (!ok) {
      atomicExch(L.error_flag, 41);
      atomicExch(L.error_flag + 1, failed_at);
    }
  }
}

// meridional_extend_grid_kernel, CLARITAS.py:4190-4196: cells above the
// free surface take the column's top fluid value (interpolation only)
__global__ void extend_kernel(MeridionalLaunch L) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, NR = L.nr, NZ = L.nz;
  if (t >= L.n * NR) return;
  const int p = t / NR, i = t - p * NR;
  double* col = L.grid + (size_t)p * NR * NZ + (size_t)i * NZ;
  int top = -1;
  for (int k = NZ - 1; k >= 0; --k)
    if (L.active[i * NZ + k]) { top = k; break; }
  if (top < 0) {
    for (int k = 0; k < NZ; ++k) col[k] = 1.0;
    return;
  }
  const double q = col[top];
  for (int k = top + 1; k < NZ; ++k) col[k] = q;
}

// meridional_finalize_kernel, CLARITAS.py:4198-4209: extrema over fluid
// cells and the four point scales; one workgroup of 256 per bin
__global__ __launch_bounds__(256) void finalize_kernel(MeridionalLaunch L) {
  __shared__ double red[8];
  const int p = blockIdx.x, plane = L.nr * L.nz;
  const double* G = L.grid + (size_t)p * plane;
  double mx = 0.0, mn = 1e300;
  for (int c = threadIdx.x; c < plane; c += blockDim.x)
    if (L.active[c]) { mx = fmax(mx, G[c]); mn = fmin(mn, G[c]); }
  mx = block_max(mx, red, 4);
  __syncthreads();
  mn = -block_max(-mn, red + 4, 4);
  if (threadIdx.x == 0) {
    if (!(mx > 0.0)) { mx = 1.0; mn = 1.0; }
    double* st = L.stats + 6 * (size_t)p;
    st[0] = mx;
    st[1] = mn;
    const float R = (float)L.R, z0 = (float)L.zmin, zt = (float)L.zgrid, xw = (float)(0.95 * L.R);
    st[2] = grid_sample(G, L.nr, L.nz, 0.0f, 0.0f, R, z0, zt);
    st[3] = grid_sample(G, L.nr, L.nz, xw, 0.0f, R, z0, zt);
    st[4] = grid_sample(G, L.nr, L.nz, 0.0f, z0, R, z0, zt);
    st[5] = grid_sample(G, L.nr, L.nz, xw, z0, R, z0, zt);
  }
}

}  // namespace

void clr_meridional_shape(MeridionalLaunch& L) {
  const int plane = L.nr * L.nz, H = ((L.nr + 1) / 2) * L.nz;
  const int k = (H + 1023) / 1024;
  if (k <= kMaxK) {
    const int per = (H + k - 1) / k, block = ((per + 63) / 64) * 64;
    if ((size_t)(plane + block / 64) * sizeof(double) <= (size_t)kMaxLdsBytes) {
      L.use_lds = 1;
      L.k_cells = k;
      L.block = block;
      return;
    }
  }
  L.use_lds = 0;
  L.k_cells = 0;
  L.block = 1024;
}

hipError_t clr_meridional_setup(const MeridionalLaunch& L, hipStream_t s) {
  const int plane = L.nr * L.nz;
  hipLaunchKernelGGL(mask_kernel, dim3((plane + 255) / 256), dim3(256), 0, s, L);
  const size_t total = (size_t)plane * L.n;
  hipLaunchKernelGGL(setup_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, L);
  return hipGetLastError();
}

hipError_t clr_meridional_sweep(const MeridionalLaunch& L, int it0, int it1, bool initial, hipStream_t s) {
  const int nw = L.block / 64, plane = L.nr * L.nz;
  const size_t lds = (size_t)((L.use_lds ? plane : 0) + nw) * sizeof(double);
  const dim3 grid((unsigned)L.n), block((unsigned)L.block);
  const int init = initial ? 1 : 0;
  if (!L.use_lds) {
    hipLaunchKernelGGL(sweep_kernel<0>, grid, block, lds, s, L, it0, it1, init);
  } else {
    switch (L.k_cells) {
      case 1: hipLaunchKernelGGL(sweep_kernel<1>, grid, block, lds, s, L, it0, it1, init); break;
      case 2: hipLaunchKernelGGL(sweep_kernel<2>, grid, block, lds, s, L, it0, it1, init); break;
      case 3: hipLaunchKernelGGL(sweep_kernel<3>, grid, block, lds, s, L, it0, it1, init); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t clr_meridional_finish(const MeridionalLaunch& L, hipStream_t s) {
  const int cols = L.n * L.nr;
  hipLaunchKernelGGL(extend_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, L);
  hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)L.n), dim3(256), 0, s, L);
  return hipGetLastError();
}
