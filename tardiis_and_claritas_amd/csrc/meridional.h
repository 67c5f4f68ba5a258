// meridional.h -- device solve of the stationary meridional
// advection-diffusion sediment field (one r-z grid per particle bin).
#pragma once
#include <hip/hip_runtime.h>

struct MeridionalLaunch {
  int n, nr, nz, fs_model, n_radial, block, k_cells, use_lds;
  double R, zmin, zgrid, zmean, dh, core, upeak, eddy, omega, target;
  const double *sedB, *sednorm, *ws;  // [n]
  const double *radlog, *radwr;       // [n][n_radial]
  unsigned char* active;              // [nr][nz]
  double* grid;                       // [n][nr][nz], z fastest
  double* coeff;                      // [5][n][nr][nz]: AP, AE, AW, AN, AS
  double* residual;                   // [n]
  double* stats;                      // [n][6]: max, min, 4 point scales
  int* error_flag;                    // [2]: code, iteration of the failure
};

// Sweep shape for an nr x nz plane: cells of one colour per thread
// (k_cells), workgroup size and whether the plane lives in LDS.
void clr_meridional_shape(MeridionalLaunch& L);
// Fluid mask, initial guess (analytic finite-Re drift-diffusion field) and
// the finite-volume coefficients.
hipError_t clr_meridional_setup(const MeridionalLaunch& L, hipStream_t s);
// Iterations it0+1 .. it1 of red-black SOR on every bin (mass renormalised
// every 20 iterations), then the check: renormalise and write the max
// relative residual per bin.  `initial` renormalises before the first sweep.
hipError_t clr_meridional_sweep(const MeridionalLaunch& L, int it0, int it1, bool initial, hipStream_t s);
// Extension above the free surface, per-bin extrema and point scales.
hipError_t clr_meridional_finish(const MeridionalLaunch& L, hipStream_t s);
