// precompute.hip -- per-bin transport split and sediment coefficients.
//
// Device side of transport_precompute_kernel (CLARITAS/CLARITAS.py:3845-3921)
// with the trace-channel choice of CLARITAS.py:6368-6384, and of
// sediment_transport_precompute_kernel (CLARITAS.py:3940-4038).  The reference
// runs both as a single CUDA thread.  Here the sediment coefficients run one
// thread per size bin (each bin's Newton solves, radial potential and 4096-
// point volume quadrature are independent) and the order-sensitive sums
// (normalisation, channel CDFs, majorants) are taken in bin order by one lane
// so they reproduce the reference's summation order exactly.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "precompute.h"

namespace {
constexpr double kPi = 3.141592653589793238462643383279502884;

// One workgroup; lane 0 keeps the reference's sequential sums.
__global__ void transport_split_kernel(PrepLaunch L) {
  if (blockIdx.x != 0) return;
  const int n = L.n, em = L.event_model;
  const int t = threadIdx.x;
  __shared__ double s_sw, s_avg;
  __shared__ int s_err;
  if (t == 0) {
    double sw = 0.0;
    for (int i = 0; i < n; ++i) sw += L.weight[i];
    s_sw = sw;
    s_err = (sw > 0.0) ? 0 : 4;
    double avg = 0.0;
    if (sw > 0.0)
      for (int i = 0; i < n; ++i) {
        const double r = 0.5 * L.diameter[i];
        avg += (L.weight[i] / sw) * (L.density * (4.0 / 3.0) * kPi * r * r * r);
      }
    s_avg = avg;
  }
  __syncthreads();
  if (s_err) {
    if (t == 0) atomicExch(L.error_flag, 4);
    return;
  }
  for (int i = t; i < n; i += blockDim.x) {
    const double r = 0.5 * L.diameter[i];
    const double srcw = L.weight[i] / s_sw;
    const double mass = L.density * (4.0 / 3.0) * kPi * r * r * r;
    const double nd = (L.weight_mode == 0) ? L.concentration * srcw / mass
                                           : (s_avg > 0.0 ? L.concentration / s_avg : 0.0) * srcw;
    const double qe = L.qext[i], qs = L.qsca[i], k = L.k_imag[i];
    const double g0 = kPi * r * r, s0 = fmax(qs, 0.0) * g0, e0 = fmax(qe, 0.0) * g0;
    const double tol = 1e-10 * fmax(1.0, fmax(fabs(qe), fabs(qs)));
    if (!isfinite(k) || k < 0.0 || !isfinite(qe) || !isfinite(qs) || qe < -tol || qs < -tol || qs > qe + tol ||
        (k == 0.0 && fabs(qe - qs) > tol))
      atomicExch(L.error_flag, 5);
    double w = 0.0;
    if (em == 3) w = L.hybrid_wave_scale * fmax(qs - 1.0, 0.0) * g0;
    else if (em == 4) {
      const double rate = (k > 0.0) ? e0 : s0;
      w = (k > 0.0) ? ((s0 > 0.0) ? rate * fmax(s0 - g0, 0.0) / s0 : 0.0) : fmax(s0 - g0, 0.0);
    }
    double mb;
    if (em == 0) mb = nd * g0;
    else if (em == 1) mb = nd * s0;
    else if (em == 2 || em == 6) mb = nd * e0;
    else if (em == 3) mb = nd * g0 + nd * w;
    else mb = (k > 0.0) ? nd * e0 : nd * s0;
    L.radius[i] = r;
    L.number_density[i] = nd;
    L.mu_geom[i] = nd * g0;
    L.mu_sca[i] = nd * s0;
    L.mu_ext[i] = nd * e0;
    L.mu_wave[i] = nd * w;
    L.mu_event[i] = mb;
    double tg, tw;
    if (em == 3) { tg = nd * g0; tw = nd * w; }
    else if (em == 4) { tg = fmax(mb - nd * w, 0.0); tw = nd * w; }
    else { tg = mb; tw = 0.0; }
    L.trace_g[i] = tg;
    L.trace_w[i] = tw;
    L.radius_f[i] = (float)r;
    L.n_real_f[i] = (float)L.n_real[i];
    L.k_imag_f[i] = (float)k;
    L.albedo_f[i] = (float)(qe > 0.0 ? qs / qe : 1.0);
  }
  __syncthreads();
  if (t != 0) return;
  double mg = 0, ms = 0, me = 0, mw = 0, mt = 0;
  for (int i = 0; i < n; ++i) {
    mg += L.mu_geom[i]; ms += L.mu_sca[i]; me += L.mu_ext[i]; mw += L.mu_wave[i]; mt += L.mu_event[i];
  }
  const bool hyb = (em == 3 || em == 4);
  const double mtg = (em == 4) ? fmax(mt - mw, 0.0) : ((em == 3) ? mg : mt);
  const double mtw = hyb ? mw : 0.0;
  double cg = 0.0, cw = 0.0, sg = 0.0, swv = 0.0;
  for (int i = 0; i < n; ++i) {
    const double mgi = (em == 4) ? fmax(L.mu_event[i] - L.mu_wave[i], 0.0) : L.mu_geom[i];
    const double pg = hyb ? (mtg > 0.0 ? mgi / mtg : 0.0) : (mt > 0.0 ? L.mu_event[i] / mt : 0.0);
    const double pw = (hyb && mtw > 0.0) ? L.mu_wave[i] / mtw : 0.0;
    cg += pg; cw += pw;
    L.geom_cdf[i] = cg;
    L.wave_cdf[i] = cw;
    sg += L.trace_g[i];
    swv += L.trace_w[i];
  }
  if (n > 0) {
    if (cg > 0.0) L.geom_cdf[n - 1] = 1.0;
    if (cw > 0.0) L.wave_cdf[n - 1] = 1.0;
  }
  for (int i = 0; i < n; ++i) {
    L.geom_cdf_f[i] = (float)L.geom_cdf[i];
    L.wave_cdf_f[i] = (float)L.wave_cdf[i];
  }
  L.mu_scalars[0] = mtg; L.mu_scalars[1] = mtw; L.mu_scalars[2] = mg; L.mu_scalars[3] = ms;
  L.mu_scalars[4] = me; L.mu_scalars[5] = mtw; L.mu_scalars[6] = mt;
  double cwave = 1.0;
  if (hyb && L.wave_rms_height > 0.0 && L.wavelength > 0.0 && L.n_medium > 0.0) {
    const double a = 4.0 * kPi * L.n_medium * L.wave_rms_height / L.wavelength;
    cwave = fmin(fmax(exp(-a * a), 0.0), 1.0);
  }
  L.mu_scalars[7] = cwave;
  L.trace_sums[0] = sg;
  L.trace_sums[1] = swv;
}

__device__ double schiller_naumann(double v0, double d, double rho_w, double mu) {
  if (!(v0 > 0.0)) return 0.0;
  double v = v0;
  const double K = rho_w * d / mu;
  for (int it = 0; it < 24; ++it) {
    const double Re = fmax(K * v, 0.0), corr = 0.15 * pow(Re, 0.687);
    const double f = v * (1.0 + corr) - v0, df = 1.0 + 1.687 * corr;
    const double vn = fmax(v - f / df, 0.0);
    if (fabs(vn - v) <= 2e-14 * fmax(1.0, fabs(v))) { v = vn; break; }
    v = vn;
  }
  return v;
}

// One workgroup per bin: settling speed, radial slip and log-potential
// tables, vertical decay B = w_s / D_t and the volume normalisation
// (CLARITAS.py:3940-4031).  The radial points' Newton solves run across the
// workgroup; the trapezoid integral of the slip is a chunked scan (each
// thread sums a contiguous chunk, one thread scans the chunk sums, each
// thread then writes its chunk); the normalisation quadrature is a
// workgroup reduction.
constexpr int kSedThreads = 256;
__device__ __forceinline__ double sed_block_sum(double v, double* red) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kSedThreads / 64; ++w) s += red[w];
  return s;
}
__global__ __launch_bounds__(kSedThreads) void sediment_bins_kernel(PrepLaunch L) {
  __shared__ double part[kSedThreads];
  __shared__ double red[kSedThreads / 64];
  __shared__ double s_total;
  const int i = blockIdx.x, tid = threadIdx.x;
  if (i >= L.n) return;
  const int tm = L.sed_model, fsm = L.fs_model, nrad = L.n_radial, NQ = L.n_quad;
  const double R = L.R, R2 = R * R, a2 = L.core * L.core, DH = L.dh;
  double* rl = L.radlog + (size_t)i * nrad;
  double* wrt = L.radwr ? L.radwr + (size_t)i * nrad : nullptr;
  if (tm != CLR_SED_STOKES && tm != CLR_SED_SCHILLER && tm != CLR_SED_GRID_PER_BIN) {
    if (tid == 0) { L.sedA[i] = 0.0; L.sedB[i] = 0.0; L.sednorm[i] = 1.0; L.sedws[i] = 0.0; }
    for (int j = tid; j < nrad; j += kSedThreads) {
      rl[j] = 0.0;
      if (wrt) wrt[j] = 0.0;
    }
    return;
  }
  double Hinf = 0.0, hbar = 0.0, om2 = 0.0;
  if (DH > 0.0 && fsm == 1) { Hinf = DH; hbar = 0.5 * DH; om2 = 2.0 * L.gravity * DH / R2; }
  else if (DH > 0.0 && fsm == 2) {
    Hinf = DH * (a2 + R2) / R2; hbar = Hinf * (1.0 - (a2 / R2) * log1p(R2 / a2)); om2 = 2.0 * L.gravity * Hinf / a2;
  }
  const double fR = (fsm == 2) ? R2 / (a2 + R2) : R2;
  const double Hmean = L.zmean - L.zmin, dr = R / (double)NQ, drad = R / (double)(nrad - 1);
  const double d = L.diameter[i], rho_p = L.particle_density, rho_w = L.water_density, mu = L.viscosity;
  const double tau = rho_p * d * d / (18.0 * mu);
  const double ws0 = (rho_p - rho_w) * L.gravity * d * d / (18.0 * mu);
  const bool finite_re = (tm != CLR_SED_STOKES);
  const double ws = finite_re ? schiller_naumann(ws0, d, rho_w, mu) : ws0;
  const double B = ws / L.eddy_diffusivity;
  double A;
  auto radial_acc = [&](double r) -> double {
    if (!(r > 0.0)) return 0.0;
    return (fsm == 1) ? om2 * r : ((fsm == 2) ? om2 * r / pow(1.0 + r * r / a2, 2.0) : 0.0);
  };
  if (!finite_re) {
    A = tau * om2 * ((fsm == 2) ? a2 : 1.0) / (2.0 * L.eddy_diffusivity);
    for (int j = tid; j < nrad; j += kSedThreads) {
      const double r = (double)j * drad, rr2 = r * r, f = (fsm == 2) ? rr2 / (a2 + rr2) : rr2;
      rl[j] = A * (f - fR);
      if (wrt) wrt[j] = tau * radial_acc(r);
    }
  } else {
    // finite-Re slip w_r at every radial point (rl holds w_r until the scan)
    for (int j = tid; j < nrad; j += kSedThreads) {
      const double wr = schiller_naumann(tau * radial_acc((double)j * drad), d, rho_w, mu);
      rl[j] = wr;
      if (wrt) wrt[j] = wr;
    }
    __syncthreads();
    // psi(r_j) = [int_0^r_j w - int_0^R w] / D_t by the trapezoid rule
    const int chunk = (nrad + kSedThreads - 1) / kSedThreads;
    const int j0 = min(tid * chunk, nrad), j1 = min(j0 + chunk, nrad);
    double loc = 0.0;
    for (int j = max(j0, 1); j < j1; ++j) loc += 0.5 * (rl[j - 1] + rl[j]) * drad;
    part[tid] = loc;
    const double prev0 = (j0 > 0 && j0 < nrad) ? rl[j0 - 1] : 0.0;
    __syncthreads();
    if (tid == 0) {
      double run = 0.0;
      for (int t = 0; t < kSedThreads; ++t) { const double v = part[t]; part[t] = run; run += v; }
      s_total = run;
    }
    __syncthreads();
    const double total = s_total;
    double cum = part[tid], prev = prev0;
    for (int j = j0; j < j1; ++j) {
      const double wr = rl[j];
      if (j > 0) cum += 0.5 * (prev + wr) * drad;
      prev = wr;
      rl[j] = fmin((cum - total) / L.eddy_diffusivity, 0.0);
    }
    __syncthreads();
    if (tid == 0) rl[nrad - 1] = 0.0;
    A = (fR > 0.0) ? -fmin(-total / L.eddy_diffusivity, 0.0) / fR : 0.0;
    __syncthreads();
  }
  double integ = 0.0;
  for (int j = tid; j < NQ; j += kSedThreads) {
    const double r = ((double)j + 0.5) * dr, rr2 = r * r;
    double zs = L.zmean;
    if (fsm == 1 && DH > 0.0) zs = L.zmean + DH * (rr2 / R2 - 0.5);
    else if (fsm == 2 && DH > 0.0) zs = L.zmean - hbar + Hinf * rr2 / (a2 + rr2);
    const double H = fmax(zs - L.zmin, 0.0);
    double lr;
    if (!finite_re) { const double f = (fsm == 2) ? rr2 / (a2 + rr2) : rr2; lr = A * (f - fR); }
    else {
      const double u = r / R * (double)(nrad - 1);
      int k = (int)floor(u);
      k = min(max(k, 0), nrad - 2);
      const double tt = u - (double)k;
      lr = fmin((1.0 - tt) * rl[k] + tt * rl[k + 1], 0.0);
    }
    const double rad = exp(lr), iz = (fabs(B) < 1e-14) ? rad * H : rad * (-expm1(-B * H)) / B;
    integ += r * iz;
  }
  integ = sed_block_sum(integ, red);
  if (tid == 0) {
    const double norm = (2.0 / (R2 * Hmean)) * integ * dr;
    if (!(norm > 0.0) || !isfinite(norm)) atomicExch(L.error_flag, 34);
    L.sedA[i] = A; L.sedB[i] = B; L.sedws[i] = ws; L.sednorm[i] = norm;
  }
}

// maximum of a non-negative float field (IEEE order == integer order)
__global__ void field_max_kernel(const float* __restrict__ f, size_t n, size_t plane,
                                 unsigned int* __restrict__ out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float v = fmaxf(f[i], 0.0f);
    const unsigned int bits = __float_as_uint(v);
    const size_t b = plane ? i / plane : 0;
    atomicMax(out + b, bits);
  }
}

// Woodcock majorants in bin order, CLARITAS.py:4033-4037
__global__ void majorant_kernel(PrepLaunch L, const unsigned int* __restrict__ fmax_bits) {
  if (blockIdx.x || threadIdx.x) return;
  const int tm = L.sed_model;
  double mg = 0.0, mw = 0.0, maxA = 0.0;
  if (tm == CLR_SED_UNIFORM) { mg = L.trace_sums[0]; mw = L.trace_sums[1]; }
  else if (tm == CLR_SED_RZ_FIELD || tm == CLR_SED_XYZ_FIELD) {
    const double mx = (double)__uint_as_float(fmax_bits[0]);
    mg = L.trace_sums[0] * mx;
    mw = L.trace_sums[1] * mx;
  } else if (tm == CLR_SED_GRID_PER_BIN) {
    for (int i = 0; i < L.n; ++i) {
      const double mx = (double)__uint_as_float(fmax_bits[i]);
      mg += L.trace_g[i] * mx;
      mw += L.trace_w[i] * mx;
    }
  } else {
    for (int i = 0; i < L.n; ++i) {
      mg += L.trace_g[i] / L.sednorm[i];
      mw += L.trace_w[i] / L.sednorm[i];
      if (L.sedA[i] > maxA) maxA = L.sedA[i];
    }
  }
  L.majorants[0] = mg; L.majorants[1] = mw; L.majorants[2] = mg + mw; L.majorants[3] = maxA;
  if (!(L.majorants[2] >= 0.0) || !isfinite(L.majorants[2])) atomicExch(L.error_flag, 35);
}
}  // namespace

hipError_t clr_launch_sediment_bins(const PrepLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL(sediment_bins_kernel, dim3(L.n), dim3(kSedThreads), 0, s, L);
  return hipGetLastError();
}

hipError_t clr_launch_precompute(const PrepLaunch& L, unsigned int* fmax_bits, hipStream_t s) {
  hipLaunchKernelGGL(transport_split_kernel, dim3(1), dim3(256), 0, s, L);
  hipLaunchKernelGGL(sediment_bins_kernel, dim3(L.n), dim3(kSedThreads), 0, s, L);
  if (L.field && L.field_len) {
    const size_t plane = (L.sed_model == CLR_SED_GRID_PER_BIN) ? L.field_len / (size_t)L.n : 0;
    size_t blocks = (L.field_len + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(field_max_kernel, dim3((unsigned)blocks), dim3(256), 0, s, L.field, L.field_len, plane,
                       fmax_bits);
  }
  hipLaunchKernelGGL(majorant_kernel, dim3(1), dim3(1), 0, s, L, fmax_bits);
  return hipGetLastError();
}

namespace {
// lower bound of u in a nondecreasing float row: the smallest i with
// u <= cdf[i], or n-1 when there is none (the sampler's own search)
__device__ int row_lower_bound(const float* __restrict__ cdf, int n, float u) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (u <= cdf[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}
__global__ void guide_kernel(const float* __restrict__ cdf, int n_rows, int n, unsigned int* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n_rows * kGuideN) return;
  const int row = (int)(t / kGuideN), j = (int)(t % kGuideN);
  const float* c = cdf + (size_t)row * n;
  const unsigned int shift = 23u - (unsigned int)kGuideBits;
  const unsigned int m_lo = (unsigned int)j << shift, m_hi = (((unsigned int)j + 1u) << shift) - 1u;
  // same float expression as the kernel's uniform draw
  const float u_lo = ((float)m_lo + 0.5f) * 1.1920928955078125e-7f;
  const float u_hi = ((float)m_hi + 0.5f) * 1.1920928955078125e-7f;
  const int lo = row_lower_bound(c, n, u_lo), hi = row_lower_bound(c, n, u_hi);
  out[t] = (unsigned int)lo | ((unsigned int)hi << 16);
}
}  // namespace

hipError_t clr_launch_guides(const float* cdf, int n_rows, int n, unsigned int* out, hipStream_t s) {
  if (n_rows <= 0 || n <= 0 || n > 65535) return hipErrorInvalidValue;
  const size_t total = (size_t)n_rows * kGuideN;
  const int block = 256;
  hipLaunchKernelGGL(guide_kernel, dim3((unsigned int)((total + block - 1) / block)), dim3(block), 0, s, cdf, n_rows,
                     n, out);
  return hipGetLastError();
}
