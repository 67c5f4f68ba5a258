// precompute.h -- launch description of the transport-split and sediment
// coefficient kernels (precompute.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/claritas.h"

struct PrepLaunch {
  int n, weight_mode, event_model, sed_model, fs_model, n_quad, n_radial, pad0;
  double concentration, density, n_medium, wavelength, hybrid_wave_scale, wave_rms_height;
  double particle_density, water_density, viscosity, gravity, eddy_diffusivity;
  double R, zmin, zmean, dh, core;
  // inputs (device)
  const double *diameter, *weight, *n_real, *k_imag, *qext, *qsca;
  const float* field;
  size_t field_len;
  // outputs (device)
  double *radius, *number_density, *mu_geom, *mu_sca, *mu_ext, *mu_wave, *mu_event;
  double *trace_g, *trace_w, *geom_cdf, *wave_cdf, *mu_scalars, *trace_sums;
  double *sedA, *sedB, *sednorm, *sedws, *radlog, *majorants;
  double* radwr;  // optional [n][n_radial] radial slip w_r (the meridional solve's face slips)
  float *radius_f, *n_real_f, *k_imag_f, *albedo_f, *geom_cdf_f, *wave_cdf_f;
  int* error_flag;
};

// fmax_bits: device scratch of max(1, n) words, zeroed by the caller.
hipError_t clr_launch_precompute(const PrepLaunch& L, unsigned int* fmax_bits, hipStream_t s);
// Only the per-bin sediment coefficients (settling speed, A, B, norm, radial
// tables); reads n, the sediment/geometry scalars and `diameter`.
hipError_t clr_launch_sediment_bins(const PrepLaunch& L, hipStream_t s);

// Guide tables for inverse-CDF sampling on the transport kernel's uniform
// u = (m + 1/2) 2^-23 (m a 23-bit integer).  Bucket j of a row covers
// m in [j 2^(23-kGuideBits), (j+1) 2^(23-kGuideBits)); its entry packs the
// lower-bound index of the bucket's smallest u (low 16 bits) and of its
// largest u (high 16 bits), so a search restricted to that index range
// returns exactly the full binary search's index.  Rows need n <= 65535.
constexpr int kGuideBits = 11;
constexpr int kGuideN = 1 << kGuideBits;
// cdf: [n_rows][n] floats, out: [n_rows][kGuideN]
hipError_t clr_launch_guides(const float* cdf, int n_rows, int n, unsigned int* out, hipStream_t s);
