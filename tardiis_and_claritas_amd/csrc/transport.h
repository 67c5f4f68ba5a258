// transport.h -- device parameter block and launch interface of the
// persistent photon-transport kernel (transport.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/claritas.h"

// Everything the transport kernel reads for one configuration.  Scalars are
// stored as the float32 values the reference passes to trace_kernel
// (CLARITAS.py:7642-7760); tables live in device memory.
constexpr int CLR_MAJ_SLABS = 32;

struct DevParams {
  float n_water, n_acrylic, n_air, n_bottom_ext;
  float rin, rout, zmin, zmean, zwall_top, zlow;
  float ring_in, throat_out, ring_out, source_r, throat_r, cb_r;
  float dh, core_r, alpha1, alpha2, sigma, beam_r, legacy_acc, wavelength;
  float mu_geom, mu_wave, mu_track, coherent, zs_min, field_zmax, field_half;
  float det_ux[CLR_N_CHANNELS], det_uy[CLR_N_CHANNELS];
  float aux_ux[CLR_MAX_AUX], aux_uy[CLR_MAX_AUX], aux_deg[CLR_MAX_AUX];
  double rk_K, rk_C0, sum_trace_g, sum_trace_w;
  // shared FP32 fields (sediment models 4, 5): inverse cell sizes and the
  // bin-summed channel rates in float
  float field_inv0, field_inv1, field_inv2, sum_g_f, sum_w_f;
  // collimator bores, ring face and photodiode (CLARITAS.py:5558-5562):
  // 0 absorbing / ideal, 1 smooth Fresnel nylon / SFH213 fit
  int src_bore, det_bore, face_model, pd_model, max_bore_bounces, max_face_bounces;
  float n_bore, n_face;
  int fs_model, src_model, event_model, sed_model;
  int max_pb, max_iter, cond_trials, n_aux;
  float rough, aspect, kappa;     // particle Beckmann alpha, spheroid a/b, axis-law kappa
  // shared FP32 fields: piecewise-constant majorant over CLR_MAJ_SLABS z slabs
  // [slab_z0 + j w, slab_z0 + (j+1) w) (the end slabs are open), n_slabs = 0
  // when the single majorant mu_track is used
  float slab_mu[CLR_MAJ_SLABS];
  float slab_z0, slab_w, slab_inv_w, slab_mu_max;
  int n_slabs;
  int orientation;                // spheroid axis law (0 isotropic, 1 local-tangential vMF)
  int n_groups, n_bins, n_theta, n_radial;
  int fn0, fn1, fn2;
  // bores tested either side of the exit azimuth's nearest bore (hardware_detector)
  int det_window;
  // least alignment u.v of a ray that can pass both throat apertures of a bore
  float det_cos_min;
  // free-surface height as a function of q = x^2 + y^2 in float (surface_floor):
  // Rankine z = fs_c0 + fs_k q / fs_a2 (q < fs_a2), fs_c0 + fs_k (2 - fs_a2 / q);
  // Scully z = fs_c0 + fs_k q / (fs_a2 + q); parabolic z = fs_c0 + fs_k q
  float fs_c0, fs_k, fs_a2, fs_qpad;
  // per-bin tables
  const float* radius;
  const float* n_real;
  const float* k_imag;
  const float* albedo;
  const float* gcdf;
  const float* wcdf;
  const double* qsca;
  const double* qext;
  const double* trace_g;
  const double* trace_w;
  const double* sedA;
  const double* sedB;
  const double* sednorm;
  const double* radlog;
  // Mie phase table (float, sampled by inverse CDF)
  const float* cdf;    // [n_bins * n_theta]
  const float* theta;  // [n_theta]
  // guide tables of the inverse-CDF searches (layout in precompute.h)
  const unsigned int* theta_guide;  // [n_bins * kGuideN] over cdf rows
  const unsigned int* gguide;       // [kGuideN] over gcdf
  const unsigned int* wguide;       // [kGuideN] over wcdf
  const float* field;  // sediment scale field or nullptr
};

// One unit of the device work queue: histories [start, start+count) of
// configuration `config`, whose whole run has n_total histories.
struct WorkItem {
  int config;
  int pad;
  uint64_t start, count, n_total, seed, offset;
};

// Light-transport diagnostic buffers (device memory), CLARITAS.py:7232-7300.
// Voxel grids are [nz][ny][nx] over [-rin, rin]^2 x [zmin, zmax]; the
// trajectory arrays are [rep_rays][max_segments].  Histories with id <
// max_rays are recorded; the first rep_rays also store their segments.
struct LightDiag {
  float *total, *post, *event_weight;
  unsigned int *event_count, *absorption_count;
  double *segment_total, *deposited_total;
  unsigned long long *segment_count, *event_total, *absorption_total;
  unsigned int *error_count, *boundary_adjust, *outside;  // outside[8]
  float *tx0, *ty0, *tz0, *tx1, *ty1, *tz1;
  int *torder, *tevent, *tcount, *toverflow;
  uint64_t max_rays;
  int nx, ny, nz, rep_rays, max_segments, pad_;
  float rin, zmin, zmax, pad2_;
  float *heat_xy, *heat_zy;  // water-path heat maps [vis][vis] (vis 0: none)
  int vis, pad3_;
};

struct TransportLaunch {
  const DevParams* params;   // [n_configs]
  const WorkItem* items;     // [n_items]
  int n_items;
  unsigned int* item_counter;
  clr_tallies* tallies;      // [n_configs], device layout == clr_tallies
  clr_record* records;       // optional, config 0 only
  uint64_t n_records;
  int grid, block;
  int kind;   // kernel variant, CLR_KERNEL_*
  const LightDiag* light;    // CLR_KERNEL_LIGHT only
  // causal provenance records of detected histories (CAUSAL, and LIGHT when
  // non-null): appended at *causal_count, written while below causal_cap
  clr_causal_record* causal;
  unsigned long long* causal_count;
  uint64_t causal_cap;
  // failed-history samples (CAUSAL/LIGHT with a causal target)
  clr_failure_sample* fail_rec;
  unsigned long long* fail_count;
  uint64_t fail_cap;
};

// Kernel variants: PLAIN when every configuration uses the plain-Mie law and a
// uniform or shared FP32 field; ROUGH when one needs the Smith microsurface
// walk (rough particles) or the spheroid-morphology wave channel; GENERIC
// otherwise.  The rare, FP64-heavy rough-surface code lives only in ROUGH, so
// it costs the other variants no registers.  LIGHT is ROUGH plus the
// light-transport voxel diagnostics; it runs only when they are requested.
// LIGHT and CAUSAL also keep each history's causal path provenance (event
// types, deflections, interface angles, sites, acrylic returns) and record
// it for detected histories; CAUSAL is ROUGH plus that alone.
enum : int { CLR_KERNEL_GENERIC = 0, CLR_KERNEL_PLAIN = 1, CLR_KERNEL_ROUGH = 2, CLR_KERNEL_LIGHT = 3,
             CLR_KERNEL_CAUSAL = 4 };

hipError_t clr_launch_transport(const TransportLaunch& L, hipStream_t s);
// Diagnostic deposition of explicit inputs (mode 0 segments [n][7], mode 1
// event sites [n][3]) into the buffers of D (device pointers).
hipError_t clr_launch_light_probe(const LightDiag* D, const float* data, int n, int mode, hipStream_t s);
// Device probes of the transport's building blocks on explicit inputs
// (clr_probe, include/claritas.h): mode, n items of `in_stride` floats in,
// `out_stride` doubles out per item, random streams keyed by seed.
hipError_t clr_launch_probe(const DevParams* P, int mode, int n, const float* in, int in_stride, double* out,
                            int out_stride, uint64_t seed, hipStream_t s);
// Resident workgroups per CU for `block` threads (occupancy query).
int clr_transport_occupancy(int block, int kind);
