/*
 * claritas.h -- C ABI of libclaritas_hip.so, the gfx950 photon-transport
 * library for the CLARITAS sample cell.
 *
 * The library replaces the device side of CLARITAS's forward model: the
 * CuPy RawModule kernels that `ClaritasForwardModel` compiles from CUDA_SRC
 * (CLARITAS/CLARITAS.py:6000-6034) and launches from
 * `_prepare_gpu_particle_physics` (CLARITAS/CLARITAS.py:6036-6965) and from the
 * chunk loop of `simulate` (CLARITAS/CLARITAS.py:7614-7848).  The Python
 * package `tardiis_and_claritas_amd` binds these entry points with ctypes and
 * keeps the reference's `simulate(...) -> ForwardResult` surface on top.
 *
 * Conventions
 *  - Plain C types only: fixed-width integers, doubles, floats, pointers.
 *  - Every entry point returns 0 on success.  Positive codes reuse the
 *    reference's precompute error numbering (1 Mie order cap, 2 invalid Mie
 *    input, 3 phase CDF failure, 4/5 transport precompute, 31-35 sediment,
 *    41 meridional mass normalisation, see CLARITAS/CLARITAS.py:3505,3515,
 *    3590,3868,3877,3958-4037,4180);
 *    negative codes are API/HIP errors (CLR_E_*).  clr_last_error() gives text.
 *  - The library owns every device buffer inside a clr_ctx; the caller owns
 *    the host structs and output arrays it passes in.
 *  - Calls on one context are synchronous and must not overlap.
 *  - Per-history failures never raise: they are counted by production
 *    failure code (CLARITAS/CLARITAS.py:1846-1866) in clr_tallies.failures.
 */
#ifndef CLARITAS_H
#define CLARITAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLR_ABI_VERSION 11

#define CLR_N_CHANNELS 18      /* detectors at 0..170 deg, CLARITAS.py:420 */
#define CLR_MAX_GROUPS 64      /* jackknife batches per configuration */
#define CLR_MAX_AUX 4          /* extra hardware bores at arbitrary angles */
#define CLR_N_COUNTERS 64
#define CLR_N_ICHIST 64        /* interaction-order histogram, last bin = overflow */
#define CLR_N_STATUS 9         /* status 0..8, CLARITAS.py:4678-5068 */
#define CLR_N_FAILURE 19       /* PROD_* codes 0..18, CLARITAS.py:1846-1866 */

/* error codes (negative) */
#define CLR_E_ARG      (-1)
#define CLR_E_HIP      (-2)
#define CLR_E_NOMEM    (-3)
#define CLR_E_STATE    (-4)
#define CLR_E_RANGE    (-5)
#define CLR_E_UNSUPPORTED (-6)
#define CLR_E_NOT_CONVERGED (-7) /* meridional solve ended above its tolerance */

/* ---- model codes (values identical to CLARITAS.py:5543-5575) ---- */
/* event law, CLARITAS.py:5543-5550; CLR_EVENT_PLAIN_MIE has no reference code */
#define CLR_EVENT_GEOMETRIC 0
#define CLR_EVENT_MIE_QSCA 1
#define CLR_EVENT_MIE_QEXT 2
#define CLR_EVENT_HYBRID_EXCESS 3
#define CLR_EVENT_HYBRID_EXACT_RATE 4
#define CLR_EVENT_PLAIN_MIE 6
/* source angular law, CLARITAS.py:5553-5557; 3 = collimated uniform disk */
#define CLR_SOURCE_LEGACY_POLAR 0
#define CLR_SOURCE_BETA_RADIANCE 1
#define CLR_SOURCE_THESIS_PLANAR 2
#define CLR_SOURCE_COLLIMATED_DISK 3
/* free surface, CLARITAS.py:5565-5569; 3 = volume-preserving Rankine vortex */
#define CLR_SURFACE_FLAT 0
#define CLR_SURFACE_PARABOLIC 1
#define CLR_SURFACE_SCULLY 2
#define CLR_SURFACE_RANKINE 3
/* sediment field, CLARITAS.py:5570-5575.  3 = per-bin r-z grid supplied by
 * the caller (the reference fills it with its meridional solver);
 * 4 = one FP32 r-z scale field shared by every bin (e.g. a Rouse profile);
 * 5 = one FP32 Cartesian x-y-z scale field shared by every bin. */
#define CLR_SED_UNIFORM 0
#define CLR_SED_STOKES 1
#define CLR_SED_SCHILLER 2
#define CLR_SED_GRID_PER_BIN 3
#define CLR_SED_RZ_FIELD 4
#define CLR_SED_XYZ_FIELD 5

/* ---- counters reported per configuration (clr_tallies.counters) ----
 * Sums over histories that entered the water unless stated.  These are the
 * per-ray integer arrays the reference reduces on the host
 * (CLARITAS.py:8665-8733), reduced on the device instead. */
enum {
  CLR_C_HISTORIES = 0,     /* histories launched (all) */
  CLR_C_ENTERED,           /* entered the water */
  CLR_C_VALID_AIR,         /* status 1 with finite exit state */
  CLR_C_SUM_IC,            /* particle interactions */
  CLR_C_SUM_GIC,           /* geometric-channel interactions */
  CLR_C_SUM_WIC,           /* wave-channel interactions */
  CLR_C_SUM_WCOH,          /* coherent Mie wave interactions */
  CLR_C_BALLISTIC,         /* entered with zero interactions */
  CLR_C_SUM_PR,            /* particle Fresnel reflections */
  CLR_C_RAYS_PER,          /* histories with >=1 entry reflection */
  CLR_C_RAYS_PIR,          /* histories with >=1 internal reflection */
  CLR_C_RAYS_PABS,         /* histories absorbed inside a particle */
  CLR_C_SUM_CIR,           /* cell inner-wall reflections */
  CLR_C_SUM_COR,           /* cell outer-wall reflections */
  CLR_C_RAYS_CTIR,         /* histories with >=1 cell TIR */
  CLR_C_SUM_TOPR,          /* free-surface reflections */
  CLR_C_RAYS_TOPTIR,       /* histories with >=1 free-surface TIR */
  CLR_C_SUM_BIR,           /* bottom inner reflections */
  CLR_C_SUM_BOR,           /* bottom outer reflections */
  CLR_C_RAYS_BTIR,         /* histories with >=1 bottom TIR */
  /* algorithmic work counters over all histories (FLOP model) */
  CLR_C_N_FLIGHT,          /* free-flight segments in water */
  CLR_C_N_NULL,            /* Woodcock null collisions */
  CLR_C_N_SCAT,            /* real particle events */
  CLR_C_N_IFACE,           /* Fresnel interface decisions */
  CLR_C_N_ABS,             /* absorption roulette samples */
  CLR_C_N_EXIT,            /* scored air exits */
  CLR_C_COND_TRIALS,       /* conditional-law candidate trials: closure trials plus one per
                              unconditional encounter (CLARITAS.py conditional_stats[0]) */
  CLR_C_COND_REJECT,       /* conditional outcome rejections */
  CLR_C_COND_INACCESSIBLE, /* inaccessible particle placements */
  CLR_C_RNG_DRAWS,         /* 32-bit Philox words consumed */
  CLR_C_N_STEPS,           /* state-machine steps (device) or loop passes */
  /* reflecting source collimator (source_bore_model 1), CLARITAS.py:8677-8688 */
  CLR_C_SRC_BORE_INTER,    /* source bore wall interactions */
  CLR_C_SRC_BORE_REFL,     /* source bore wall reflections */
  CLR_C_SRC_REFL_ACCEPT,   /* histories through the collimator after >= 1 reflection */
  /* the rest of the conditional law's sampler counts, CLARITAS.py:9119-9125
     (conditional_stats[3..9]; counted only under hybrid_mie_exact_rate) */
  CLR_C_COND_ACC_GEOM,     /* accepted geometric-channel encounters */
  CLR_C_COND_ACC_COHERENT, /* accepted coherent-wave encounters */
  CLR_C_COND_ACC_MORPH,    /* accepted morphology-wave encounters */
  CLR_C_COND_ACC_ABS,      /* accepted encounters that absorb */
  CLR_C_COND_ACC_SCAT,     /* accepted encounters that scatter */
  CLR_C_COND_NUMERICAL,    /* encounters ended by a numerical failure */
  CLR_C_COND_EXHAUSTED,    /* closure calls that ran out of trials */
  CLR_C_COUNT_
};

/* ---- configuration: the scalar parameter set of trace_kernel
 * (CLARITAS.py:4551-4605) in sensor-plane coordinates (metres).
 * Derivation from user-facing kwargs follows CLARITAS.py:5660-5680,10024-10052. */
typedef struct clr_config {
  double n_water, n_acrylic, n_air, bottom_external_n;
  double rin, rout, zmin, zmean, zwall_top, bottom_disc_thickness;
  double ring_in, throat_out, ring_out, source_r, throat_r, counterbore_r;
  double vortex_delta_h, vortex_core_radius;
  double alpha1, alpha2, beam_sigma, beam_radius;
  double legacy_accept_deg;
  double wavelength;            /* vacuum wavelength (m) */
  double rough_alpha;           /* particle Beckmann alpha = tan(rms slope), CLARITAS.py:5701-5704 */
  double aux_angles_deg[CLR_MAX_AUX];
  double field_zmax;            /* top of the sediment grid (m) */
  double field_xy_half;         /* half-width of the Cartesian field (m) */
  double spheroid_aspect;       /* wave_spheroid_aspect_ratio a/b, CLARITAS.py:5706 */
  double spheroid_kappa;        /* wave_spheroid_orientation_kappa, CLARITAS.py:5710 */
  int32_t free_surface_model;
  int32_t source_angular_model;
  int32_t event_model;
  int32_t sediment_model;
  int32_t max_particle_bounces; /* max_internal_bounces */
  int32_t max_cell_bounces;
  int32_t max_iterations;
  int32_t conditional_max_trials;
  int32_t n_aux;
  int32_t n_groups;             /* jackknife groups, 1..CLR_MAX_GROUPS */
  int32_t field_n0, field_n1, field_n2; /* r-z: (nr,nz,-); xyz: (nx,ny,nz) */
  int32_t field_nbins;          /* 1 or n_bins (CLR_SED_GRID_PER_BIN) */
  int32_t spheroid_orientation; /* 0 isotropic, 1 local_tangential_axial_vmf, CLARITAS.py:5564 */
  int32_t pad_;
  /* collimator bores, ring face, photodiode (CLARITAS.py:5558-5562, 5596-5603) */
  double bore_n;                /* bore_refractive_index (> 1) */
  double ring_face_n;           /* ring_inner_face_refractive_index (> 0) */
  int32_t source_bore_model;    /* 0 absorbing, 1 smooth_fresnel_nylon */
  int32_t detector_bore_model;  /* 0 absorbing, 1 smooth_fresnel_nylon */
  int32_t ring_face_model;      /* 0 absorbing, 1 smooth_fresnel_nylon */
  int32_t photodiode_model;     /* 0 ideal_hard_aperture, 1 sfh213_datasheet_fit */
  int32_t max_bore_bounces;     /* >= 0 */
  int32_t max_ring_face_bounces;/* >= 1 */
} clr_config;

/* ---- particle population (one PSD, CLARITAS.py:6036-6060) ---- */
typedef struct clr_particles {
  const double* diameter;       /* [n] m */
  const double* weight;         /* [n] raw PSD weights (normalised inside) */
  const double* n_real;         /* [n] */
  const double* k_imag;         /* [n] */
  double concentration;         /* kg/m^3 (= g/L) */
  double density;               /* particle density kg/m^3 */
  double n_medium;
  double wavelength;            /* vacuum wavelength, m */
  double hybrid_wave_scale;
  double wave_rms_height_m;
  int32_t n;
  int32_t n_theta;              /* Mie phase grid size */
  int32_t weight_mode;          /* 0 mass fraction, 1 number fraction */
  int32_t event_model;
} clr_particles;

/* ---- sediment transport inputs (CLARITAS.py:3940-3948) ---- */
typedef struct clr_sediment {
  double particle_density, water_density, water_viscosity, gravity;
  double eddy_diffusivity;
  int32_t n_quad, n_radial;
} clr_sediment;

/* ---- Mie outputs (CLARITAS.py:3495-3592); every array is caller-owned ---- */
typedef struct clr_mie_out {
  double* size_parameter;       /* [n] */
  double* qext;                 /* [n] */
  double* qsca;                 /* [n] */
  double* qabs;                 /* [n] */
  double* qback;                /* [n] */
  double* asym_g;               /* [n] */
  double* albedo;               /* [n] */
  int32_t* nstop;               /* [n] */
  double* theta;                /* [n_theta] or NULL */
  double* phase_pdf;            /* [n*n_theta] or NULL */
  double* phase_cdf;            /* [n*n_theta] or NULL */
} clr_mie_out;

/* ---- host copy of the prepared transport tables ---- */
typedef struct clr_tables_out {
  double* radius;               /* [n] */
  double* number_density;       /* [n] */
  double* mu_geom_bin;          /* [n] physical projected-area rate */
  double* mu_sca_bin;           /* [n] */
  double* mu_ext_bin;           /* [n] */
  double* trace_geom_bin;       /* [n] traced geometric-channel rate */
  double* trace_wave_bin;       /* [n] traced wave-channel rate */
  double* geom_cdf;             /* [n] */
  double* wave_cdf;             /* [n] */
  double* sed_A;                /* [n] */
  double* sed_B;                /* [n] */
  double* sed_norm;             /* [n] */
  double* sed_settling;         /* [n] settling velocity */
  double* sed_radial_log;       /* [n*n_radial] or NULL */
  double mu_scalars[8];         /* CLARITAS.py:3911-3913 layout */
  double majorants[4];          /* CLARITAS.py:4035 layout */
} clr_tables_out;

/* Production failure subcodes (CLARITAS.py:464-521) by tally slot: slot k
 * counts the failed histories whose subcode is CLR_SUBCODE_VALUES[k]; the
 * last slot counts any other subcode. */
#define CLR_N_SUBCODE 64
#define CLR_SUBCODE_VALUES                                                                            \
  {0,    10,   11,   12,   13,   14,   20,   21,   22,   23,   24,   25,   26,   27,   28,   29,   30, \
   31,   32,   33,   34,   35,   36,   37,   38,   39,   40,   41,   1001, 1002, 1003, 1004, 1005,   \
   1101, 1102, 1103, 1104, 1105, 1106, 1107, 2001, 2002, 2101, 2102, 2103, 2104, 3001, 3002, 3003, \
   3004, 3005, 3006, 4001, 9001, 9002, 9003}
#define CLR_N_SUBCODE_KNOWN 56

/* ---- tallies for one configuration ----
 * det[g][k][c]: k = 0 hardware native, 1 hardware mirror, 2 legacy native,
 * 3 legacy mirror; channel c = detector at 10*c degrees (CLARITAS.py:4378-4436).
 * aux[g][k][a]: k = 0 native, 1 mirror, extra bore at aux_angles_deg[a];
 * aux_window[g][k][a]: legacy exit-position window at aux_angles_deg[a]. */
typedef struct clr_tallies {
  uint64_t det[CLR_MAX_GROUPS][4][CLR_N_CHANNELS];
  uint64_t aux[CLR_MAX_GROUPS][2][CLR_MAX_AUX];
  /* exit-position windows |angle - aux_angles_deg[a]| <= legacy_accept_deg
   * (native, mirror), the legacy rule of CLARITAS.py:4426-4435 applied at the
   * extra angles (ballistic histories skip windows at >= 90 deg) */
  uint64_t aux_window[CLR_MAX_GROUPS][2][CLR_MAX_AUX];
  uint64_t group_histories[CLR_MAX_GROUPS];
  /* hardware hits after >= 1 bore or ring-face reflection, and those via
   * the ring-face cavity (native, mirror; CLARITAS.py:4414-4423) */
  uint64_t stray[2][CLR_N_CHANNELS];
  uint64_t face_stray[2][CLR_N_CHANNELS];
  double incidence_sum[2][CLR_N_CHANNELS];   /* degrees, native / mirror */
  /* photodiode-weighted hits per group (native, mirror): SFH213 relative
   * response at the incidence angle, or 1 per hit for the ideal aperture */
  double photodiode[CLR_MAX_GROUPS][2][CLR_N_CHANNELS];
  uint64_t status[CLR_N_STATUS];
  uint64_t failures[CLR_N_FAILURE];
  uint64_t failure_subcodes[CLR_N_SUBCODE];  /* failed histories by subcode slot (CLR_SUBCODE_VALUES) */
  uint64_t counters[CLR_N_COUNTERS];
  uint64_t ic_hist[CLR_N_ICHIST];
  int32_t n_groups;
  int32_t pad_;
} clr_tallies;

/* ---- per-history record, written for the first n_records histories
 * when requested (history-by-history parity tests) ---- */
typedef struct clr_record {
  float x, y, z, vx, vy, vz;    /* exit state (status 1) or last state */
  int32_t status;
  int32_t native_id, mirror_id;
  int32_t interactions;
  int32_t failure;
  int32_t draws;                /* Philox words consumed */
} clr_record;

/* ---- one item of a batched launch ---- */
typedef struct clr_batch_item {
  int32_t config;               /* index into the prepared configurations */
  int32_t pad_;
  uint64_t n_histories;
  uint64_t seed;
  uint64_t history_offset;
} clr_batch_item;

/* ---- launch shape (autotuned or fixed) ---- */
typedef struct clr_launch {
  int32_t block;                /* threads per workgroup: 64..1024 */
  int32_t blocks_per_cu;        /* resident workgroups per CU (grid sizing) */
  int32_t chunk;                /* histories per work-queue item (0: sized by the queued work) */
  int32_t queue_order;          /* batches: 0 as given, 1 most expensive configuration first */
} clr_launch;

/* ---- meridional advection-diffusion solve: the producer of the per-bin
 * r-z grid of CLR_SED_GRID_PER_BIN (the reference's sediment model
 * 'schiller_naumann_meridional_advection_diffusion', code 3).  Parameters
 * and limits follow CLARITAS.py:5625-5631,5873-5887. ---- */
typedef struct clr_meridional {
  double circulation_ratio;     /* chi in [0, 0.5]: peak meridional speed = chi * max v_theta */
  double tolerance;             /* max relative residual, in (0, 1e-2) */
  double sor;                   /* over-relaxation factor, in [1, 1.95) */
  int32_t nr, nz;               /* cells in r and z, >= 16 and >= 32 */
  int32_t max_iterations;       /* >= 250 */
  int32_t check_interval;       /* iterations between residual checks, >= 10 */
} clr_meridional;

/* Outputs of clr_meridional_solve; every array is caller-owned and may be
 * NULL (skipped).  Cell (i, k) has centre r = (i + 1/2) R/nr,
 * z = zmin + (k + 1/2) (grid_zmax - zmin)/nz; arrays are [bin][i][k],
 * k fastest. */
typedef struct clr_meridional_out {
  double* grid;                 /* [n][nr][nz] density scale; cells above the surface hold the column's top value */
  double* residual;             /* [n] max relative residual at the last check */
  double* scale_max;            /* [n] max over fluid cells (sediment_max_density_scale) */
  double* scale_min;            /* [n] min over fluid cells */
  double* point_scales;         /* [n][4] at the sensor centre, sensor wall (0.95 R), bottom centre, bottom wall */
  uint8_t* active;              /* [nr][nz] 1 for fluid cells (centre below the free surface) */
  double* coeff;                /* [5][n][nr][nz] finite-volume coefficients AP, AE, AW, AN, AS */
  double grid_zmax;             /* top of the grid = free-surface height at the wall (m) */
  double vtheta_peak;           /* maximum azimuthal speed of the vortex (m/s) */
  double peak_speed;            /* chi * vtheta_peak (m/s) */
  int32_t iterations;           /* SOR iterations run (a multiple of check_interval, or max_iterations) */
  int32_t pad_;
} clr_meridional_out;

/* ---- light-transport voxel diagnostics (the reference's
 * light_transport_config, CLARITAS.py:7232-7300) ---- */
typedef struct clr_light_config {
  uint64_t max_rays;            /* histories with id < max_rays are recorded (0: no voxel
                                   diagnostics, only the heat maps) */
  int32_t nx, ny, nz;           /* voxels over [-rin, rin]^2 x [zmin, zmax], each >= 4 */
  int32_t representative_rays;  /* the first histories whose segments are stored (<= max_rays) */
  int32_t max_segments;         /* segments stored per representative history (>= 1) */
  int32_t heatmap_size;         /* side of the water-path heat maps (0: none), CLARITAS.py:3296-3309 */
  double zmax;                  /* top of the box: max(centre, wall) height of the free surface */
} clr_light_config;

/* Outputs; arrays are caller-owned, grids [nz][ny][nx] (x fastest),
 * trajectory arrays [representative_rays][max_segments].  Any array may be
 * NULL (skipped). */
typedef struct clr_light_out {
  float* track_length;               /* water-path track length per voxel (m) */
  float* post_scatter_track_length;  /* the part after the first particle event */
  float* scatter_event_weight;       /* surviving particle events (weight 1 each) */
  uint32_t* scatter_event_count;
  uint32_t* absorption_event_count;  /* accepted extinction encounters ending in absorption */
  float* traj_x0; float* traj_y0; float* traj_z0;   /* clipped segment ends (NaN unused) */
  float* traj_x1; float* traj_y1; float* traj_z1;
  int32_t* traj_order;               /* particle events before the segment (-1 unused) */
  int32_t* traj_event;               /* terminal event code (-1 unused; 0 propagation/boundary,
                                        1-3 geometric external reflection / transmission /
                                        internal reflection, 4 coherent Mie, 5-7 morphology,
                                        8 inaccessible placement, 9 absorption, 10 failure) */
  int32_t* traj_count;               /* [representative_rays] segments stored */
  int32_t* traj_overflow;            /* [representative_rays] 1 when segments were dropped */
  double segment_total;              /* sum of clipped segment lengths (m) */
  double deposited_total;            /* sum of deposited lengths (m) */
  uint64_t segment_count, scatter_event_total, absorption_event_total;
  uint32_t error_count;              /* traversal-integrity failures and out-of-box sites */
  uint32_t boundary_adjustment_count;/* benign endpoint snaps */
  float* heatmap_xy;                 /* [heatmap_size][heatmap_size] water-path samples every 10 um
                                        of every history, x-y projection (row: y from the top) */
  float* heatmap_zy;                 /* the z-y projection (column: z from zmin to the surface
                                        height at the wall) */
  uint32_t outside[8];               /* out-of-box sites: count, first x, y, z (FP32 bits),
                                        history id, tolerance (FP32 bits), recorded flag */
} clr_light_out;

/* ---- causal path provenance of one hardware-detected history: the
 * per-ray fields trace_kernel writes for the reference's detector-path,
 * causal and H170 diagnostics (CLARITAS.py:4629-4647 state, 5076-5103
 * write-out, with its sentinels: -1 for an undefined angle, path or
 * radius, 0 for an undefined position or class).  Event type codes
 * (CLARITAS.py:5325-5334): 1 geometric external (entry) reflection,
 * 2 geometric transmission/refraction, 3 geometric internal-reflection
 * path, 4 coherent Mie wave, 5-7 the morphology-wave counterparts of 1-3.
 * Bins are particle-size (PSD) bins, -1 when none. ---- */
typedef struct clr_causal_record {
  uint64_t history;             /* history id (Philox counter) */
  int32_t native_id, mirror_id; /* hardware detector channels, -1 none */
  int32_t ic, gic, wic, wcoh;   /* particle events: all, geometric, wave, coherent wave */
  int32_t per, pir, pabs;       /* entry reflections, internal reflections, absorptions */
  int32_t acrylic_reflections;  /* cell inner + outer + bottom inner + outer reflections */
  int32_t acrylic_returns;      /* returns to the water from the acrylic or headspace after a particle event */
  int32_t events_after_return;  /* particle events after such a return */
  int32_t first_type, last_type, last_order, first_backscatter, first_source_hemisphere;
  int32_t first_bin, last_bin, first_geometric_bin, last_geometric_bin, first_wave_bin, last_wave_bin;
  int32_t branch_transitions, type_transitions;
  int32_t small_deflections, medium_deflections, large_deflections;  /* < 30, < 90, >= 90 deg */
  int32_t max_type, max_order, max_bin, max_macro_bounces, max_micro_order;
  float first_deflection, last_deflection;       /* deg */
  float internal_path, internal_opl;             /* summed over the history's particle events (m) */
  float water_path_after_last;                   /* water path after the last particle event (m) */
  float cumulative_deflection, cumulative_deflection_sq, max_deflection;
  float turn_x, turn_y, turn_z;  /* site of the largest turn (m) */
  float first_x, first_y, first_z, last_x, last_y, last_z;  /* first / last event sites (m) */
  float mean_r, mean_z, min_r, max_r, min_z, max_z;
  float water_path_first, water_path_last;       /* water path at the first / last event (m) */
  float max_internal_path, max_path_over_diameter, max_incidence, max_exit;  /* largest turn (m, -, deg, deg) */
} clr_causal_record;

/* One failed history (status 4) of a causal run: the reference's failure
 * sample fields (CLARITAS.py:8505-8640) without the particle drill-down. */
typedef struct clr_failure_sample {
  uint64_t history;
  int32_t code, subcode;        /* production failure code and subcode */
  int32_t trace_iteration;      /* water-loop iterations completed */
  int32_t pad_;
  float x, y, z, vx, vy, vz;    /* photon state at the failure */
} clr_failure_sample;

/* Caller-owned record buffer of clr_run_causal.  The call traces the
 * histories [range_start, range_start + range_count) of an n-history run
 * (range_count 0: all n), with the run's jackknife groups, so that the
 * tallies of consecutive ranges add up to those of the whole run.  count
 * receives the number of detected histories in the range; min(count,
 * capacity) records are written, in no particular order. */
typedef struct clr_causal_out {
  clr_causal_record* records;
  uint64_t capacity;
  uint64_t count;
  uint64_t range_start, range_count;
  clr_failure_sample* failures;  /* optional: the first failed histories appended (any order) */
  uint64_t failure_capacity;
  uint64_t failure_count;        /* failed histories in the range */
} clr_causal_out;

struct clr_ctx;
typedef struct clr_ctx clr_ctx;

/* ABI identification and struct sizes (arity test, cf. the reference's
 * launch-arity check validate_claritas.py:972-1001). id: 0 config,
 * 1 particles, 2 sediment, 3 mie_out, 4 tables_out, 5 tallies, 6 record,
 * 7 batch_item, 8 launch, 9 meridional, 10 meridional_out, 11 light_config,
 * 12 light_out, 13 causal_record, 14 failure_sample. Returns 0 for an unknown id. */
int clr_abi_version(void);
size_t clr_sizeof(int struct_id);

/* Context lifetime.  Replaces the CuPy device selection and NVRTC compile
 * of ClaritasForwardModel.__init__ (CLARITAS.py:5653-5656, 6000-6034). */
int clr_create(clr_ctx** out, int device);
void clr_destroy(clr_ctx* ctx);
const char* clr_last_error(const clr_ctx* ctx);
/* Fill *info with: [0] CU count, [1] wavefront size, [2] LDS bytes per
 * workgroup, [3] clock kHz, [4] gfx major*100+minor. */
int clr_device_info(clr_ctx* ctx, int64_t info[8]);

/* Mie coefficients, efficiencies, theta grid, phase PDF and CDF in FP64 on
 * the device.  Replaces mie_coefficients_efficiencies_kernel,
 * build_theta_grid_kernel, mie_phase_pdf_kernel and mie_phase_cdf_kernel
 * (CLARITAS.py:3488-3592) as launched at CLARITAS.py:6186-6245, without the
 * reference's MIE_MAX_ORDER=4096 cap (order is sized per call). */
int clr_mie_tables(clr_ctx* ctx, const clr_particles* p, clr_mie_out* out);

/* Full precompute for configuration slot `slot` (0 for single runs):
 * Mie tables, transport split (transport_precompute_kernel,
 * CLARITAS.py:3845-3921), sediment coefficients
 * (sediment_transport_precompute_kernel, CLARITAS.py:3940-4038) and the
 * device tables the transport kernel reads.  `field` is the optional FP32
 * sediment scale grid for CLR_SED_GRID_PER_BIN / RZ / XYZ models, laid out
 * [bin][n0][n1] (r-z) or [n0][n1][n2] (x-y-z), z fastest.
 * Replaces _prepare_gpu_particle_physics (CLARITAS.py:6036-6965). */
int clr_prepare(clr_ctx* ctx, int slot, const clr_config* cfg,
                const clr_particles* p, const clr_sediment* sed,
                const float* field, size_t field_len);
/* Number of prepared configuration slots the context holds. */
int clr_n_slots(clr_ctx* ctx);
/* Copy the prepared tables of `slot` to host arrays (NULL fields skipped). */
int clr_get_tables(clr_ctx* ctx, int slot, clr_tables_out* out);

/* Trace n histories of slot 0 with Philox key `seed`, histories numbered
 * history_offset .. history_offset+n-1 (64-bit; the reference's uint32 ray
 * ids end at 2^32, CLARITAS.py:7191), and add detector tallies into *out
 * (which the call zeroes first).  Replaces the trace_kernel +
 * score_detectors_kernel chunk loop and its tally reduction
 * (CLARITAS.py:7614-7864, 8773-8817).  `records` (optional) receives the
 * first n_records per-history records. `launch` may be NULL (defaults). */
int clr_run(clr_ctx* ctx, uint64_t n, uint64_t seed, uint64_t history_offset,
            const clr_launch* launch, clr_tallies* out,
            clr_record* records, uint64_t n_records);

/* Batched launch: every item traced in one persistent kernel fed from a
 * device work queue, captured once in a hipGraph and replayed.  out[i]
 * receives the tallies of items[i].  No reference counterpart: the
 * reference runs one configuration per simulate() call. */
int clr_run_batch(clr_ctx* ctx, const clr_batch_item* items, int n_items,
                  const clr_launch* launch, clr_tallies* out);

/* clr_run plus the light-transport voxel diagnostics of the histories with
 * id < light->max_rays: exact track-length deposition of every water
 * segment by 3-D voxel traversal (all path, and path after the first
 * particle event), scatter-event and absorption site grids, closure totals
 * and the segments of the first representative histories.  The tallies
 * equal clr_run's (same histories, same random numbers); the diagnostics
 * run on a separate kernel variant.  Replaces trace_kernel's
 * LIGHT_DIAG_ENABLED path (CLARITAS.py:3296-3485 at 4611, 4766, 4850,
 * 4926-4946) and its readback (CLARITAS.py:7232-7300, 9455-9560).
 * Diagnostic memory is capped at 2 GB. */
int clr_run_light(clr_ctx* ctx, uint64_t n, uint64_t seed, uint64_t history_offset,
                  const clr_launch* launch, const clr_light_config* light,
                  clr_tallies* out, clr_light_out* light_out);

/* The diagnostics' deposition on explicit inputs, as the reference's test
 * kernels light_diag_segment_test_kernel / light_diag_event_test_kernel
 * (CLARITAS.py:3470-3485): mode 0 deposits segment i of data[n][7]
 * (x0, y0, z0, vx, vy, vz, dist; history id i, particle order i & 1), mode 1
 * records event site i of data[n][3] into the scatter-event grids, mode 2
 * rasters segment i of data[n][7] into the heat maps (z axis up to
 * light->zmax), in the box [-rin, rin]^2 x [zmin, light->zmax].  Outputs as
 * clr_run_light. */
/* clr_run with the causal path provenance of every history that reaches a
 * hardware detector (native or mirror), for the reference's detector-path,
 * causal and H170 diagnostics (CLARITAS.py:7352-7613 state, 8017-8480
 * reductions).  The provenance is observational: the histories, their
 * random numbers and the tallies are those of clr_run.  With light non-NULL
 * the light-transport diagnostics are recorded in the same pass (as
 * clr_run_light).  Returns CLR_E_ARG when causal is NULL. */
int clr_run_causal(clr_ctx* ctx, uint64_t n, uint64_t seed, uint64_t history_offset,
                   const clr_launch* launch, const clr_light_config* light, clr_tallies* out,
                   clr_light_out* light_out, clr_causal_out* causal);

/* Host binning of causal records into the per-channel score tables of the
 * detector-path / causal / H170 diagnostics (replaces the reductions of
 * CLARITAS.py:8017-8480).  Table t of clr_causal_layout has K[t] keys x 18
 * channels of FP64, laid out one after another; clr_causal_bin adds the n
 * records into `tables` (the caller zeroes it once), each at its native and
 * mirror channel with weight 1/2 (symmetric != 0) or at the native channel
 * with weight 1.  r_edges and z_edges are the 6 radial and axial edges of
 * the H170 collision maps.  clr_causal_layout returns the number of tables
 * and, for 0 <= index < that number, the table's name and K. */
int clr_causal_layout(int n_bins, int index, char* name, int name_cap, int* K);
int clr_causal_bin(const clr_causal_record* records, uint64_t n, int n_bins, int symmetric, const double* r_edges,
                   const double* z_edges, double* tables);

int clr_light_probe(clr_ctx* ctx, int mode, int n, const float* data, const clr_light_config* light, double rin,
                    double zmin, clr_light_out* light_out);

/* ---- device probes of the transport's building blocks on explicit
 * inputs.  Replaces the reference's test kernels
 * free_surface_geometry_test_kernel (CLARITAS.py:4438-4455),
 * sediment_field_test_kernel (CLARITAS.py:4228-4235), bore_wall_test_kernel
 * (CLARITAS.py:2952-2956) and water_surface_offset_test_kernel
 * (CLARITAS.py:2980-2985), and exposes the samplers and the detector
 * acceptance of trace_kernel for statistical and closed-form tests.  Item i
 * reads in[i*IN ..] and writes out[i*OUT ..] (IN floats, OUT doubles, see
 * clr_probe_shape); random draws come from the Philox stream (seed,
 * history i), exactly as a history's first draws.  Every mode uses the
 * prepared configuration `slot`.
 *   SURFACE     in x y z vx vy vz -> surface z at (x,y), hit distance (>= 3e38: none),
 *                                     surface z at the hit, outward normal (3)
 *   WALL        in x y z vx vy vz -> water-side exit through r = rin on [zmin, wall top],
 *                                     forward hit of r = rout on [bottom of the disc, wall top],
 *                                     next acrylic surface distance and its code (1 inner wall,
 *                                     2 outer wall, 3 disc bottom, 4 wall top, 5 water base);
 *                                     1e300: no hit, -1: start outside the water
 *   OFFSET      in x y z nx ny nz    -> accepted (1/0), offset position (3)
 *   SEDIMENT    in bin x y z         -> local density scale of the bin (FP64 path of
 *                                     CLARITAS.py:1537-1609), FP32 shared-field scale
 *   DETECTOR    in x y z vx vy vz    -> hardware bore of the exit ray, native and mirror (-1 none)
 *   SAMPLER     in kind a b          -> draw, words used; kind 0 gamma(a), 1 beta(a, b),
 *                                     2 Beta-radiance polar angle (a, b), 3 standard normal
 *                                     (CLARITAS.py:1149-1212)
 *   SOURCE      in -                 -> launch position and direction before the collimator
 *   MIE_THETA   in bin               -> polar scattering angle sampled from the bin's table
 *                                     (CLARITAS.py:2632-2641)
 *   CONDITIONAL in bin x y z vx vy vz -> status (1 scattered, 2 absorbed, other: failed),
 *                                     channel, trials, outcome rejections, inaccessible
 *                                     placements, failure code, direction (3) of the
 *                                     conditional closure (CLARITAS.py:4479-4536) ---- */
#define CLR_PROBE_SURFACE 0
#define CLR_PROBE_WALL 1
#define CLR_PROBE_OFFSET 2
#define CLR_PROBE_SEDIMENT 3
#define CLR_PROBE_DETECTOR 4
#define CLR_PROBE_SAMPLER 5
#define CLR_PROBE_SOURCE 6
#define CLR_PROBE_MIE_THETA 7
#define CLR_PROBE_CONDITIONAL 8
#define CLR_PROBE_N_MODES 9
/* Input floats and output doubles per item of a probe mode; returns 0, or
 * CLR_E_ARG for an unknown mode. */
int clr_probe_shape(int mode, int* in_stride, int* out_stride);
/* Run n items of probe `mode` on slot `slot` (host arrays in[n*IN],
 * out[n*OUT]); at most 2^24 items per call. */
int clr_probe(clr_ctx* ctx, int slot, int mode, int n, uint64_t seed, const float* in, double* out);

/* Time one launch shape: trace n histories of slot 0 and return device
 * milliseconds in *ms (HIP events on the library stream). */
int clr_time_run(clr_ctx* ctx, uint64_t n, uint64_t seed,
                 const clr_launch* launch, int repeats, double* ms);

/* Average duration (ms) of the last transport-kernel launch, measured with
 * HIP events on the stream the kernel runs on. */
double clr_last_kernel_ms(clr_ctx* ctx);

/* Stationary meridional advection-diffusion sediment field, one r-z grid
 * per bin of `p` (only p->n and p->diameter are read), on the free surface
 * of `cfg` (parabolic or Scully vortex with vortex_delta_h > 0).  Upwind
 * finite volumes with no-flux walls, red-black SOR with every bin's mass
 * sum G (re^2 - rw^2)/2 dz over fluid cells renormalised to
 * R^2 (zmean - zmin)/2 (mean scale 1) every 20 iterations and at each
 * residual check; the solve stops at the first check at or after
 * min(1000, max_iterations) iterations whose max residual over bins is
 * <= tolerance.  Replaces meridional_active_mask_kernel ..
 * meridional_finalize_kernel (CLARITAS.py:4091-4215) and their driver loop
 * (CLARITAS.py:6451-6725).  Returns CLR_E_NOT_CONVERGED (outputs filled)
 * when the last check is above the tolerance, 41 when a bin's mass is not
 * positive and finite.  The grid feeds clr_prepare as the CLR_SED_GRID_PER_BIN
 * field (converted to FP32) with field_zmax = grid_zmax. */
int clr_meridional_solve(clr_ctx* ctx, const clr_config* cfg, const clr_particles* p,
                         const clr_sediment* sed, const clr_meridional* mp,
                         clr_meridional_out* out);

/* ---- surrogate: ensemble of MLPs, BF16 weights, FP32 accumulate on MFMA.
 * No reference counterpart.  Layout: n_members networks of
 * in_dim -> hidden -> hidden -> hidden -> out_dim with GELU (tanh form). */
typedef struct clr_mlp_shape {
  int32_t n_members, in_dim, hidden, out_dim, n_hidden_layers, pad_;
} clr_mlp_shape;

/* Create/replace the ensemble, weights initialised from `seed`. */
int clr_mlp_init(clr_ctx* ctx, const clr_mlp_shape* shape, uint64_t seed);
/* Copy master FP32 parameters out / in (flat, member-major, per layer W then b). */
size_t clr_mlp_param_count(clr_ctx* ctx);
int clr_mlp_get_params(clr_ctx* ctx, float* out, size_t n);
int clr_mlp_set_params(clr_ctx* ctx, const float* in, size_t n);
/* Train: `epochs` passes over (x[n][in_dim], y[n][out_dim]) in minibatches
 * of `batch`, Adam(lr, b1, b2, eps), weighted MSE loss
 * sum w (out - y)^2 / (rows * out_dim) with per-element weights w[n][out_dim]
 * (w = NULL: all ones); each member sees its own shuffle (seed). Returns the
 * final mean loss per member in loss[n_members]. */
int clr_mlp_train(clr_ctx* ctx, const float* x, const float* y, const float* w,
                  int64_t n, int32_t batch, int32_t epochs, float lr,
                  uint64_t seed, float* loss);
/* One forward+backward on a minibatch without an optimizer step; gradients
 * of the weighted mean MSE loss (w may be NULL) copied to grads (same layout
 * as params). */
int clr_mlp_grad(clr_ctx* ctx, const float* x, const float* y, const float* w,
                 int32_t n, float* loss, float* grads);
/* Inference: mean[n][out_dim] and std[n][out_dim] over the members. */
int clr_mlp_infer(clr_ctx* ctx, const float* x, int64_t n, float* mean,
                  float* stdev);
/* Inference that allocates no device memory.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same results as
 * clr_mlp_infer (the same fused kernel); the queries and outputs are staged
 * in the library's static 4 MiB device array, one per GPU, so one call takes
 * at most n rows with n * (in_dim + 2 out_dim) * 4 bytes <= 4 MiB - 768
 * (74,898 queries of the config-4 ensemble) and CLR_E_RANGE beyond that;
 * ensembles outside the fused shape (hidden <= 256, in_dim <= 32,
 * members * out_dim <= 80) get CLR_E_UNSUPPORTED.  Calls that overlap on
 * one GPU share the array and race. */
int clr_mlp_infer_static(clr_ctx* ctx, const float* x, int64_t n, float* mean, float* stdev);
/* Device-resident inference timing: n queries, `repeats` launches; *ms is
 * the average per launch (HIP events). */
int clr_mlp_time_infer(clr_ctx* ctx, int64_t n, int repeats, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* CLARITAS_H */
