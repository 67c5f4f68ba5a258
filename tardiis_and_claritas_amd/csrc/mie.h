// mie.h -- launch description for the FP64 Mie table kernels (mie.hip).
#pragma once
#include <hip/hip_runtime.h>

struct MieLaunch {
  const double* diameter;  // [n_bins]
  const double* n_real;    // [n_bins]
  const double* k_imag;    // [n_bins]
  double n_medium, wavelength;
  int n_bins, n_theta, stride;  // stride >= largest expansion order + 1
  double2 *D, *a, *b;           // [n_bins * stride] workspaces / coefficients
  int* nstop;                   // [n_bins]
  double* q;                    // [7 * n_bins]: qext qsca qabs qback g albedo x
  double* theta;                // [n_theta]
  float* theta_f;               // [n_theta]
  double* pdf;                  // [n_bins * n_theta]
  double* cdf;                  // [n_bins * n_theta]
  float* cdf_f;                 // [n_bins * n_theta]
  int* error_flag;
};

hipError_t clr_launch_mie(const MieLaunch& L, hipStream_t s);
