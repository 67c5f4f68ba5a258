// mie.hip -- FP64 Lorenz-Mie tables on gfx950.
//
// Device side of the reference's phase-table precompute
// (mie_coefficients_efficiencies_kernel, build_theta_grid_kernel,
// mie_phase_pdf_kernel, mie_phase_cdf_kernel; CLARITAS/CLARITAS.py:3488-3592).
// Differences by design:
//  - the expansion order is sized per call from the largest size parameter
//    instead of a fixed MIE_MAX_ORDER=4096 workspace (CLARITAS.py:1137,3515);
//  - the phase PDF kernel stages each bin's a_n, b_n through LDS in tiles so
//    the 256 angles of a workgroup share one read of the coefficients;
//  - the phase CDF is one workgroup per bin: per-thread trapezoid partial
//    sums, a block scan, then normalisation with the monotone clamp.
// FP64 throughout: the recurrences need it for x up to ~5e3.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mie.h"

namespace {

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx mk(double a, double b) { return cplx{a, b}; }
__device__ __forceinline__ cplx add(cplx a, cplx b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cplx sub(cplx a, cplx b) { return mk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cplx mul(cplx a, cplx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cplx scale(cplx a, double s) { return mk(a.re * s, a.im * s); }
// complex quotient with the reference's 1e-300 denominator floor
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  double d = b.re * b.re + b.im * b.im;
  if (d < 1e-300) d = 1e-300;
  return mk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

constexpr double kPi = 3.141592653589793238462643383279502884;

// One thread per size bin: downward D_n, upward psi/chi, a_n and b_n, then
// Q_ext, Q_sca, Q_abs, Q_back, g and albedo.  Work arrays are strided by
// `stride` (>= largest n_mx + 1).
__global__ void mie_coefficients_kernel(const double* __restrict__ diameter,
                                        const double* __restrict__ n_real,
                                        const double* __restrict__ k_imag, double n_medium,
                                        double wavelength, int n_bins, int stride,
                                        double2* __restrict__ D, double2* __restrict__ a,
                                        double2* __restrict__ b, int* __restrict__ nstop_out,
                                        double* __restrict__ q /* [7][n_bins] */,
                                        int* __restrict__ error_flag) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_bins) return;
  double x = kPi * n_medium * diameter[p] / wavelength;
  q[6 * n_bins + p] = x;
  if (!(x > 0.0) || !(n_real[p] > 0.0) || !(n_medium > 0.0) || !(wavelength > 0.0)) {
    nstop_out[p] = 0;
    for (int k = 0; k < 5; ++k) q[k * n_bins + p] = 0.0;
    q[5 * n_bins + p] = 1.0;
    atomicExch(error_flag, 2);
    return;
  }
  const double mr = n_real[p] / n_medium, mi = k_imag[p] / n_medium;
  int nstop = (int)(x + 4.0 * cbrt(x) + 2.0);
  if (nstop < 1) nstop = 1;
  int nmx = (int)fmax((double)nstop, hypot(mr * x, mi * x)) + 128;
  if (nmx < nstop + 1) nmx = nstop + 1;
  if (nmx >= stride) {  // host sized the workspace too small: fail loudly
    nstop_out[p] = 0;
    atomicExch(error_flag, 1);
    return;
  }
  nstop_out[p] = nstop;
  double2* Dp = D + (size_t)p * stride;
  double2* ap = a + (size_t)p * stride;
  double2* bp = b + (size_t)p * stride;
  const cplx z = mk(mr * x, mi * x), one = mk(1.0, 0.0);
  cplx dn = mk(0.0, 0.0);
  Dp[nmx] = make_double2(0.0, 0.0);
  for (int n = nmx; n >= 1; --n) {
    cplx nz = cdiv(mk((double)n, 0.0), z);
    dn = sub(nz, cdiv(one, add(dn, nz)));
    Dp[n - 1] = make_double2(dn.re, dn.im);
  }
  double psi_m1 = sin(x), psi = sin(x) / x - cos(x);
  double chi_m1 = cos(x), chi = cos(x) / x + sin(x);
  const cplx m = mk(mr, mi);
  for (int j = 0; j < nstop; ++j) {
    const int n = j + 1;
    double2 dd = Dp[n];
    cplx d = mk(dd.x, dd.y);
    cplx xi_m1 = mk(psi_m1, -chi_m1), xi = mk(psi, -chi);
    cplx ca = add(cdiv(d, m), mk((double)n / x, 0.0));
    cplx cb = add(mul(m, d), mk((double)n / x, 0.0));
    cplx an = cdiv(sub(scale(ca, psi), mk(psi_m1, 0.0)), sub(mul(ca, xi), xi_m1));
    cplx bn = cdiv(sub(scale(cb, psi), mk(psi_m1, 0.0)), sub(mul(cb, xi), xi_m1));
    ap[j] = make_double2(an.re, an.im);
    bp[j] = make_double2(bn.re, bn.im);
    double psi_p1 = ((2.0 * n + 1.0) / x) * psi - psi_m1;
    double chi_p1 = ((2.0 * n + 1.0) / x) * chi - chi_m1;
    psi_m1 = psi;
    psi = psi_p1;
    chi_m1 = chi;
    chi = chi_p1;
  }
  double se = 0.0, ss = 0.0, sbr = 0.0, sbi = 0.0, t1 = 0.0, t2 = 0.0;
  double2 cur_a = ap[0], cur_b = bp[0];
  for (int j = 0; j < nstop; ++j) {
    const int n = j + 1;
    const double cn = 2.0 * n + 1.0;
    se += cn * (cur_a.x + cur_b.x);
    ss += cn * (cur_a.x * cur_a.x + cur_a.y * cur_a.y + cur_b.x * cur_b.x + cur_b.y * cur_b.y);
    const double sg = (n & 1) ? -1.0 : 1.0;
    sbr += cn * sg * (cur_a.x - cur_b.x);
    sbi += cn * sg * (cur_a.y - cur_b.y);
    t2 += (cn / ((double)n * (n + 1.0))) * (cur_a.x * cur_b.x + cur_a.y * cur_b.y);
    if (j + 1 < nstop) {
      double2 na = ap[j + 1], nb = bp[j + 1];
      t1 += ((double)n * (n + 2.0) / (n + 1.0)) *
            ((cur_a.x * na.x + cur_a.y * na.y) + (cur_b.x * nb.x + cur_b.y * nb.y));
      cur_a = na;
      cur_b = nb;
    }
  }
  const double x2 = x * x;
  double qe = 2.0 * se / x2, qs = 2.0 * ss / x2, qb = (sbr * sbr + sbi * sbi) / x2;
  if (qe < 0.0 && qe > -1e-12) qe = 0.0;
  if (qs < 0.0 && qs > -1e-12) qs = 0.0;
  double qa = qe - qs;
  if (qa < 0.0 && qa > -1e-10) qa = 0.0;
  double gg = qs > 0.0 ? 4.0 * (t1 + t2) / (x2 * qs) : 0.0;
  gg = fmin(fmax(gg, -1.0), 1.0);
  q[0 * n_bins + p] = qe;
  q[1 * n_bins + p] = qs;
  q[2 * n_bins + p] = qa;
  q[3 * n_bins + p] = qb;
  q[4 * n_bins + p] = gg;
  q[5 * n_bins + p] = qe > 0.0 ? qs / qe : 1.0;
}

// theta_j = (pi/2)(1 - cos(pi j/(N-1))): clustered at 0 and pi
__global__ void theta_grid_kernel(double* __restrict__ theta, int nt) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  if (j == 0) { theta[j] = 0.0; return; }
  if (j == nt - 1) { theta[j] = kPi; return; }
  theta[j] = 0.5 * kPi * (1.0 - cos(kPi * ((double)j / (double)(nt - 1))));
}

constexpr int kPdfBlock = 256;
constexpr int kCoefTile = 256;

// grid = (ceil(nt/256), n_bins).  Each workgroup owns one bin and 256
// angles; a_n, b_n are staged through LDS in tiles of 256 orders.
__global__ __launch_bounds__(kPdfBlock) void phase_pdf_kernel(
    const double* __restrict__ theta, int nt, int stride, const int* __restrict__ nstop,
    const double2* __restrict__ a, const double2* __restrict__ b, double* __restrict__ pdf) {
  __shared__ double2 sa[kCoefTile], sb[kCoefTile];
  const int p = blockIdx.y;
  const int j = blockIdx.x * kPdfBlock + threadIdx.x;
  const int ns = nstop[p];
  const double th = (j < nt) ? theta[j] : 0.0;
  const double mu = cos(th);
  double s1r = 0.0, s1i = 0.0, s2r = 0.0, s2i = 0.0, pm1 = 0.0, pn = 1.0;
  const double2* ap = a + (size_t)p * stride;
  const double2* bp = b + (size_t)p * stride;
  for (int base = 0; base < ns; base += kCoefTile) {
    const int cnt = min(kCoefTile, ns - base);
    __syncthreads();
    if (threadIdx.x < cnt) {
      sa[threadIdx.x] = ap[base + threadIdx.x];
      sb[threadIdx.x] = bp[base + threadIdx.x];
    }
    __syncthreads();
    for (int k = 0; k < cnt; ++k) {
      const double nf = (double)(base + k + 1);
      const double tau = nf * mu * pn - (nf + 1.0) * pm1;
      const double c = (2.0 * nf + 1.0) / (nf * (nf + 1.0));
      const double2 ak = sa[k], bk = sb[k];
      s1r += c * (ak.x * pn + bk.x * tau);
      s1i += c * (ak.y * pn + bk.y * tau);
      s2r += c * (ak.x * tau + bk.x * pn);
      s2i += c * (ak.y * tau + bk.y * pn);
      const double pp1 = ((2.0 * nf + 1.0) / nf) * mu * pn - ((nf + 1.0) / nf) * pm1;
      pm1 = pn;
      pn = pp1;
    }
  }
  if (j < nt) pdf[(size_t)p * nt + j] = 0.5 * (s1r * s1r + s1i * s1i + s2r * s2r + s2i * s2i) * sin(th);
}

constexpr int kCdfBlock = 256;

// One workgroup per bin: trapezoid areas, block-wide prefix sum, normalise,
// monotone clamp, last entry exactly 1.  Writes the FP64 CDF and the float
// copy the transport kernel samples from.
__global__ __launch_bounds__(kCdfBlock) void phase_cdf_kernel(const double* __restrict__ theta,
                                                              const double* __restrict__ pdf, int nt,
                                                              double* __restrict__ cdf,
                                                              float* __restrict__ cdf_f,
                                                              int* __restrict__ error_flag) {
  __shared__ double part[kCdfBlock];
  __shared__ int bad;
  const int p = blockIdx.x, t = threadIdx.x;
  const double* pp = pdf + (size_t)p * nt;
  double* cp = cdf + (size_t)p * nt;
  float* cf = cdf_f + (size_t)p * nt;
  // segment [lo, hi) of interval indices 1..nt-1 for this thread
  const int nint = nt - 1;
  const int per = (nint + kCdfBlock - 1) / kCdfBlock;
  const int lo = 1 + t * per, hi = min(nt, lo + per);
  double s = 0.0;
  for (int j = lo; j < hi; ++j) {
    double area = 0.5 * (pp[j] + pp[j - 1]) * (theta[j] - theta[j - 1]);
    if (area < 0.0 && area > -1e-18) area = 0.0;
    s += area;
  }
  part[t] = s;
  if (t == 0) bad = 0;
  __syncthreads();
  // inclusive Hillis-Steele scan of the 256 partial sums
  for (int off = 1; off < kCdfBlock; off <<= 1) {
    double v = (t >= off) ? part[t - off] : 0.0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  const double total = part[kCdfBlock - 1];
  double run = (t > 0) ? part[t - 1] : 0.0;
  if (!(total > 0.0) || !isfinite(total)) {
    if (t == 0) atomicExch(error_flag, 3);
    for (int j = t; j < nt; j += kCdfBlock) {
      double v = (double)j / (double)(nt - 1);
      cp[j] = v;
      cf[j] = (float)v;
    }
    return;
  }
  if (t == 0) cp[0] = 0.0;
  for (int j = lo; j < hi; ++j) {
    double area = 0.5 * (pp[j] + pp[j - 1]) * (theta[j] - theta[j - 1]);
    if (area < 0.0 && area > -1e-18) area = 0.0;
    run += area;
    cp[j] = run;
  }
  __syncthreads();
  // normalise + monotone clamp: a running max over the array.  The raw
  // cumulative sums are non-decreasing up to rounding, so the clamp only
  // removes rounding dips; it is applied serially by one lane per bin.
  if (t == 0) {
    double prev = 0.0;
    for (int j = 0; j < nt; ++j) {
      double v = cp[j] / total;
      if (v < prev) v = prev;
      if (v > 1.0) v = 1.0;
      cp[j] = v;
      prev = v;
    }
    cp[nt - 1] = 1.0;
    (void)bad;
  }
  __syncthreads();
  for (int j = t; j < nt; j += kCdfBlock) cf[j] = (float)cp[j];
}

__global__ void to_float_kernel(const double* __restrict__ src, float* __restrict__ dst, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

}  // namespace

extern "C++" hipError_t clr_launch_mie(const MieLaunch& L, hipStream_t s) {
  const int nb = L.n_bins;
  hipLaunchKernelGGL(mie_coefficients_kernel, dim3((nb + 63) / 64), dim3(64), 0, s, L.diameter,
                     L.n_real, L.k_imag, L.n_medium, L.wavelength, nb, L.stride, L.D, L.a, L.b,
                     L.nstop, L.q, L.error_flag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || L.n_theta < 2) return e;
  hipLaunchKernelGGL(theta_grid_kernel, dim3((L.n_theta + 255) / 256), dim3(256), 0, s, L.theta,
                     L.n_theta);
  hipLaunchKernelGGL(phase_pdf_kernel, dim3((L.n_theta + kPdfBlock - 1) / kPdfBlock, nb),
                     dim3(kPdfBlock), 0, s, L.theta, L.n_theta, L.stride, L.nstop, L.a, L.b,
                     L.pdf);
  hipLaunchKernelGGL(phase_cdf_kernel, dim3(nb), dim3(kCdfBlock), 0, s, L.theta, L.pdf, L.n_theta,
                     L.cdf, L.cdf_f, L.error_flag);
  hipLaunchKernelGGL(to_float_kernel, dim3((L.n_theta + 255) / 256), dim3(256), 0, s, L.theta,
                     L.theta_f, L.n_theta);
  return hipGetLastError();
}
