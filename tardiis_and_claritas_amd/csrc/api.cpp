// api.cpp -- C ABI of libclaritas_hip.so (declared in include/claritas.h).
//
// Host orchestration only: argument checks, device buffers, work-queue
// construction and launches of the HIP kernels in mie.hip, precompute.hip,
// meridional.hip and transport.hip.  Physics runs on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/claritas.h"
#include "meridional.h"
#include "mie.h"
#include "mlp.h"
#include "graph_key.h"
#include "precompute.h"
#include "transport.h"

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr int kMaxBins = 4096;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Device tables of one prepared configuration.
struct Slot {
  bool ready = false;
  clr_config cfg{};
  int n = 0, nt = 0, nrad = 0;
  DevBuf in_d, in_w, in_nr, in_k;                  // inputs (double)
  DevBuf q;                                        // [7][n] Mie efficiencies
  DevBuf nstop;                                    // [n] expansion orders
  DevBuf theta, theta_f, pdf, cdf, cdf_f;          // phase tables
  DevBuf radius, nd, mug, mus, mue, muw, muev, trace_g, trace_w, gcdf, wcdf;
  DevBuf mu_scalars, trace_sums, sedA, sedB, sednorm, sedws, radlog, majorants;
  DevBuf radius_f, n_f, k_f, albedo_f, gcdf_f, wcdf_f;
  DevBuf field;
  DevBuf theta_guide, gguide, wguide;  // inverse-CDF guide tables
  size_t field_len = 0;
  DevParams params{};
  std::vector<float> slab_max;  // field maximum per majorant slab (shared fields)
  std::vector<double> mie_key;  // optical inputs of the Mie tables held by this slot
  double mu_scalars_h[8]{};
  double majorants_h[4]{};
  void release() {
    DevBuf* all[] = {&in_d, &in_w, &in_nr, &in_k, &q, &nstop, &theta, &theta_f, &pdf, &cdf, &cdf_f, &radius, &nd,
                     &mug, &mus, &mue, &muw, &muev, &trace_g, &trace_w, &gcdf, &wcdf, &mu_scalars,
                     &trace_sums, &sedA, &sedB, &sednorm, &sedws, &radlog, &majorants, &radius_f, &n_f,
                     &k_f, &albedo_f, &gcdf_f, &wcdf_f, &field, &theta_guide, &gguide, &wguide};
    for (DevBuf* b : all) b->release();
    mie_key.clear();
    ready = false;
  }
};

// The captured clear + transport graph of clr_run_batch, replayed while
// every argument it captured is unchanged (graph_key.h).
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  GraphKey key{};
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key = GraphKey{};
  }
};

}  // namespace

// Device buffers of the light-transport diagnostics (grown on demand).
struct LightBufs {
  DevBuf grids, scalars, traj, desc, probe, heat;
};

struct clr_ctx {
  int device = 0;
  int n_cu = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  std::vector<Slot*> slots;
  DevBuf params, tallies, items, counter, records, scratch_u32, mie_work;
  double last_ms = 0.0;
  GraphCache graph;
  MlpState* mlp = nullptr;
  void* static_scratch = nullptr;  // this device's copy of mlp_static.hip's staging array
  LightBufs light;
  // causal provenance records of the run in progress (clr_run_causal)
  DevBuf causal;
  clr_causal_record* causal_rec = nullptr;
  unsigned long long* causal_count = nullptr;
  uint64_t causal_cap = 0;
  clr_failure_sample* fail_rec = nullptr;
  unsigned long long* fail_count = nullptr;
  uint64_t fail_cap = 0;
};

namespace {

int fail(clr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
#define HIPCHK(ctx, expr)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail((ctx), CLR_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

// Expansion order the Mie kernel will need for one bin (same rule as the
// kernel: n_mx = max(n_stop, |m x|) + 128).
int mie_order(double d, double nr, double k, double n_med, double wl) {
  const double x = kPi * n_med * d / wl;
  if (!(x > 0.0)) return 2;
  int nstop = (int)(x + 4.0 * std::cbrt(x) + 2.0);
  if (nstop < 1) nstop = 1;
  int nmx = (int)std::fmax((double)nstop, std::hypot(nr / n_med * x, k / n_med * x)) + 128;
  return std::max(nmx, nstop + 1);
}

int check_particles(clr_ctx* c, const clr_particles* p) {
  if (!p || p->n <= 0 || p->n > kMaxBins) return fail(c, CLR_E_ARG, "particles: n must be 1..4096");
  if (!p->diameter || !p->weight || !p->n_real || !p->k_imag) return fail(c, CLR_E_ARG, "particles: null array");
  if (p->n_theta != 0 && p->n_theta < 2) return fail(c, CLR_E_ARG, "particles: n_theta must be >= 2");
  if (!(p->n_medium > 0.0) || !(p->wavelength > 0.0)) return fail(c, CLR_E_ARG, "particles: n_medium, wavelength");
  for (int i = 0; i < p->n; ++i) {
    if (!(p->diameter[i] > 0.0) || !std::isfinite(p->diameter[i])) return fail(c, CLR_E_ARG, "particles: diameter");
    if (!(p->n_real[i] > 0.0) || !(p->k_imag[i] >= 0.0)) return fail(c, CLR_E_ARG, "particles: complex index");
    if (!(p->weight[i] >= 0.0) || !std::isfinite(p->weight[i])) return fail(c, CLR_E_ARG, "particles: weight");
  }
  return 0;
}

// Run the Mie kernels for particles p into slot-owned buffers.
int run_mie(clr_ctx* c, Slot& s, const clr_particles* p) {
  const int n = p->n, nt = std::max(p->n_theta, 2);
  // The Mie tables depend on the sizes, indices, medium, wavelength and grid
  // only (not on PSD weights or concentration): a slot prepared again with
  // the same optics keeps its tables and only takes the new weights.
  std::vector<double> key = {(double)n, (double)nt, p->n_medium, p->wavelength};
  key.insert(key.end(), p->diameter, p->diameter + n);
  key.insert(key.end(), p->n_real, p->n_real + n);
  key.insert(key.end(), p->k_imag, p->k_imag + n);
  if (!s.mie_key.empty() && key == s.mie_key && s.n == n && s.nt == nt) {
    HIPCHK(c, hipMemcpyAsync(s.in_w.p, p->weight, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    return 0;
  }
  s.mie_key.clear();
  int stride = 2;
  for (int i = 0; i < n; ++i)
    stride = std::max(stride, mie_order(p->diameter[i], p->n_real[i], p->k_imag[i], p->n_medium, p->wavelength) + 2);
  if (stride > (1 << 22)) return fail(c, 1, "Mie expansion order exceeds 2^22");
  HIPCHK(c, s.in_d.ensure(sizeof(double) * n));
  HIPCHK(c, s.in_w.ensure(sizeof(double) * n));
  HIPCHK(c, s.in_nr.ensure(sizeof(double) * n));
  HIPCHK(c, s.in_k.ensure(sizeof(double) * n));
  HIPCHK(c, hipMemcpyAsync(s.in_d.p, p->diameter, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.in_w.p, p->weight, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.in_nr.p, p->n_real, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.in_k.p, p->k_imag, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, s.q.ensure(sizeof(double) * 7 * n));
  HIPCHK(c, s.theta.ensure(sizeof(double) * nt));
  HIPCHK(c, s.theta_f.ensure(sizeof(float) * nt));
  HIPCHK(c, s.pdf.ensure(sizeof(double) * (size_t)n * nt));
  HIPCHK(c, s.cdf.ensure(sizeof(double) * (size_t)n * nt));
  HIPCHK(c, s.cdf_f.ensure(sizeof(float) * (size_t)n * nt));
  // workspace: D, a, b (double2) + nstop + error flag
  const size_t coef = sizeof(double2) * (size_t)n * stride;
  HIPCHK(c, c->mie_work.ensure(3 * coef + sizeof(int)));
  HIPCHK(c, s.nstop.ensure(sizeof(int) * n));
  char* w = c->mie_work.as<char>();
  int* nstop = s.nstop.as<int>();
  int* eflag = reinterpret_cast<int*>(w + 3 * coef);
  HIPCHK(c, hipMemsetAsync(eflag, 0, sizeof(int), c->stream));
  MieLaunch L{};
  L.diameter = s.in_d.as<double>();
  L.n_real = s.in_nr.as<double>();
  L.k_imag = s.in_k.as<double>();
  L.n_medium = p->n_medium;
  L.wavelength = p->wavelength;
  L.n_bins = n;
  L.n_theta = nt;
  L.stride = stride;
  L.D = reinterpret_cast<double2*>(w);
  L.a = reinterpret_cast<double2*>(w + coef);
  L.b = reinterpret_cast<double2*>(w + 2 * coef);
  L.nstop = nstop;
  L.q = s.q.as<double>();
  L.theta = s.theta.as<double>();
  L.theta_f = s.theta_f.as<float>();
  L.pdf = s.pdf.as<double>();
  L.cdf = s.cdf.as<double>();
  L.cdf_f = s.cdf_f.as<float>();
  L.error_flag = eflag;
  HIPCHK(c, clr_launch_mie(L, c->stream));
  int ecode = 0;
  HIPCHK(c, hipMemcpyAsync(&ecode, eflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ecode) return fail(c, ecode, "Mie precompute failed with error code " + std::to_string(ecode));
  s.n = n;
  s.nt = nt;
  s.mie_key = std::move(key);
  return 0;
}

float det_u(double deg, bool cosine) {
  const float th = (float)deg * (3.14159265358979323846f / 180.0f);
  return cosine ? (float)std::cos((double)th) : (float)std::sin((double)th);
}

// Bores hardware_detector (transport.hip) must test either side of the bore
// nearest to the azimuth phi where an exit ray crosses r = ring_in.  In the
// frame of bore j (axis y, lateral x) an accepted ray meets the ring_in plane
// at lateral offset |l| <= throat_r and the throat_out plane within throat_r
// of the axis, so its x-y slope against the axis is at most
// s = 2 throat_r / (throat_out - ring_in).  Between the circle crossing C and
// the plane the ray advances tau = ring_in - sqrt(ring_in^2 - x_C^2) along the
// axis, so |x_C| <= X with X the fixed point of X = throat_r + s tau(X), and
// |phi - 10 j| <= delta = asin(X / ring_in).  The nearest bore jc = rint(phi/10)
// then satisfies |jc - j| <= floor(1/2 + delta/10); one bore and half a degree
// are added against FP32 rounding of phi.  The window covers the full scan of
// CLARITAS.py:4241-4253 (18 bores) when the bound fails to converge.
int window_for_cell(double ring_in, double throat_out, double throat_r) {
  const double s = 2.0 * throat_r / (throat_out - ring_in);
  double X = throat_r;
  for (int it = 0; it < 200; ++it) {
    if (!(X < ring_in)) return CLR_N_CHANNELS;
    const double tau = ring_in - std::sqrt(ring_in * ring_in - X * X);
    const double nx = throat_r + s * tau;
    if (std::fabs(nx - X) <= 1e-15 * ring_in) { X = nx; break; }
    X = nx;
  }
  if (!(X < ring_in)) return CLR_N_CHANNELS;
  const double delta = std::asin(X / ring_in) * (180.0 / kPi) + 0.5;
  const int w = (int)std::floor(0.5 + delta / 10.0) + 1;
  return std::min(w, CLR_N_CHANNELS);
}

int validate_config(clr_ctx* c, const clr_config* g, const clr_particles* p) {
  if (!g) return fail(c, CLR_E_ARG, "config is NULL");
  // CLARITAS.py:5797-5826
  if (!(g->rough_alpha >= 0.0) || !std::isfinite(g->rough_alpha))
    return fail(c, CLR_E_ARG, "rough_alpha must be finite and >= 0");
  if (!(g->spheroid_aspect >= 0.2 && g->spheroid_aspect <= 5.0))
    return fail(c, CLR_E_ARG, "spheroid_aspect must be in [0.2, 5]");
  if (!(g->spheroid_kappa >= 0.0 && g->spheroid_kappa <= 100.0))
    return fail(c, CLR_E_ARG, "spheroid_kappa must be in [0, 100]");
  if (g->spheroid_orientation != 0 && g->spheroid_orientation != 1)
    return fail(c, CLR_E_ARG, "spheroid_orientation must be 0 (isotropic) or 1 (local tangential vMF)");
  // CLARITAS.py:5966-5991
  if ((g->source_bore_model != 0 && g->source_bore_model != 1) ||
      (g->detector_bore_model != 0 && g->detector_bore_model != 1))
    return fail(c, CLR_E_ARG, "bore surface model must be absorbing or smooth_fresnel_nylon");
  if (g->ring_face_model != 0 && g->ring_face_model != 1)
    return fail(c, CLR_E_ARG, "ring inner face surface model must be absorbing or smooth_fresnel_nylon");
  if (g->photodiode_model != 0 && g->photodiode_model != 1)
    return fail(c, CLR_E_ARG, "photodiode response model must be ideal_hard_aperture or sfh213_datasheet_fit");
  if (!(g->ring_face_n > 0.0) || !std::isfinite(g->ring_face_n))
    return fail(c, CLR_E_ARG, "ring_inner_face_refractive_index must be finite and positive");
  if (g->max_ring_face_bounces < 1) return fail(c, CLR_E_ARG, "max_ring_face_bounces must be >=1");
  if (!(g->bore_n > 1.0)) return fail(c, CLR_E_ARG, "bore_refractive_index must be >1");
  if (g->max_bore_bounces < 0) return fail(c, CLR_E_ARG, "max_bore_bounces must be >=0");
  const int em = g->event_model;
  if (!(em == 0 || em == 1 || em == 2 || em == 3 || em == 4 || em == 6))
    return fail(c, CLR_E_UNSUPPORTED, "unsupported particle event model");
  if (p && p->event_model != em) return fail(c, CLR_E_ARG, "config and particles disagree on event model");
  if (g->free_surface_model < 0 || g->free_surface_model > 3) return fail(c, CLR_E_ARG, "free surface model");
  if (g->source_angular_model < 0 || g->source_angular_model > 3) return fail(c, CLR_E_ARG, "source model");
  if (g->sediment_model < 0 || g->sediment_model > 5) return fail(c, CLR_E_ARG, "sediment model");
  if (g->n_aux < 0 || g->n_aux > CLR_MAX_AUX) return fail(c, CLR_E_ARG, "n_aux");
  if (g->n_groups < 1 || g->n_groups > CLR_MAX_GROUPS) return fail(c, CLR_E_ARG, "n_groups must be 1..64");
  if (!(g->rin > 0.0 && g->rin < g->rout && g->rout < g->ring_in && g->ring_in < g->throat_out &&
        g->throat_out <= g->source_r))
    return fail(c, CLR_E_ARG, "geometry must satisfy 0 < rin < rout < ring_in < throat_out <= source_r");
  if (!(g->zwall_top > g->zmean && g->zmean > g->zmin)) return fail(c, CLR_E_ARG, "heights: zwall_top > zmean > zmin");
  if (!(g->bottom_disc_thickness > 0.0)) return fail(c, CLR_E_ARG, "bottom disc thickness");
  if (!(0.0 < g->throat_r && g->throat_r < g->counterbore_r)) return fail(c, CLR_E_ARG, "bore radii");
  if (g->free_surface_model == CLR_SURFACE_RANKINE && !(g->vortex_core_radius > 0.0 && g->vortex_core_radius < g->rin))
    return fail(c, CLR_E_ARG, "Rankine vortex needs 0 < core radius < inner radius");
  if (g->free_surface_model == CLR_SURFACE_SCULLY && g->vortex_delta_h > 0.0 && !(g->vortex_core_radius > 0.0))
    return fail(c, CLR_E_ARG, "Scully vortex needs a positive core radius");
  if (g->max_iterations < 1 || g->max_particle_bounces < 0) return fail(c, CLR_E_ARG, "iteration caps");
  return 0;
}

int validate_sediment(clr_ctx* c, const clr_config* g, const clr_sediment* sed, size_t field_len, int n) {
  const int tm = g->sediment_model, fsm = g->free_surface_model;
  if (tm == CLR_SED_STOKES || tm == CLR_SED_SCHILLER || tm == CLR_SED_GRID_PER_BIN) {
    if (!sed) return fail(c, CLR_E_ARG, "sediment parameters required");
    if (!(sed->particle_density > sed->water_density) || !(sed->water_density > 0.0) || !(sed->water_viscosity > 0.0) ||
        !(sed->gravity > 0.0) || !(sed->eddy_diffusivity > 0.0) || sed->n_quad < 64 || sed->n_radial < 65)
      return fail(c, 31, "sediment precompute failed with error code 31");
    if (fsm < 0 || fsm > 2 || (g->vortex_delta_h > 0.0 && fsm == 0) || (fsm == 2 && !(g->vortex_core_radius > 0.0)))
      return fail(c, 32, "sediment precompute failed with error code 32");
    if (!(g->zmean - g->zmin > 0.0)) return fail(c, 33, "sediment precompute failed with error code 33");
  }
  const size_t need_rz = (size_t)std::max(g->field_n0, 0) * (size_t)std::max(g->field_n1, 0);
  if (tm == CLR_SED_GRID_PER_BIN && (g->field_n0 < 2 || g->field_n1 < 2 || field_len < need_rz * (size_t)n))
    return fail(c, CLR_E_ARG, "per-bin grid needs n_bins*n0*n1 values");
  if (tm == CLR_SED_RZ_FIELD && (g->field_n0 < 2 || g->field_n1 < 2 || field_len < need_rz))
    return fail(c, CLR_E_ARG, "r-z field needs n0*n1 values");
  if (tm == CLR_SED_XYZ_FIELD &&
      (g->field_n0 < 2 || g->field_n1 < 2 || g->field_n2 < 2 || field_len < need_rz * (size_t)g->field_n2 ||
       !(g->field_xy_half > 0.0)))
    return fail(c, CLR_E_ARG, "x-y-z field needs n0*n1*n2 values and a half-width");
  if ((tm == CLR_SED_GRID_PER_BIN || tm == CLR_SED_RZ_FIELD || tm == CLR_SED_XYZ_FIELD) && !(g->field_zmax > g->zmin))
    return fail(c, CLR_E_ARG, "field_zmax must exceed zmin");
  return 0;
}

void fill_params(Slot& s, const clr_config& g) {
  DevParams& P = s.params;
  P.n_water = (float)g.n_water; P.n_acrylic = (float)g.n_acrylic; P.n_air = (float)g.n_air;
  P.n_bottom_ext = (float)g.bottom_external_n;
  P.rin = (float)g.rin; P.rout = (float)g.rout; P.zmin = (float)g.zmin; P.zmean = (float)g.zmean;
  P.zwall_top = (float)g.zwall_top; P.zlow = P.zmin - (float)g.bottom_disc_thickness;
  P.ring_in = (float)g.ring_in; P.throat_out = (float)g.throat_out; P.ring_out = (float)g.ring_out;
  P.source_r = (float)g.source_r; P.throat_r = (float)g.throat_r; P.cb_r = (float)g.counterbore_r;
  P.dh = (float)g.vortex_delta_h; P.core_r = (float)g.vortex_core_radius;
  P.alpha1 = (float)g.alpha1; P.alpha2 = (float)g.alpha2; P.sigma = (float)g.beam_sigma;
  P.beam_r = (float)g.beam_radius; P.legacy_acc = (float)g.legacy_accept_deg; P.wavelength = (float)g.wavelength;
  P.mu_geom = (float)s.mu_scalars_h[0];
  P.mu_wave = (float)s.mu_scalars_h[1];
  P.coherent = std::fmin(std::fmax((float)s.mu_scalars_h[7], 0.0f), 1.0f);
  P.mu_track = (g.sediment_model == CLR_SED_UNIFORM) ? (P.mu_geom + P.mu_wave)
                                                     : ((float)s.majorants_h[0] + (float)s.majorants_h[1]);
  P.field_zmax = (float)g.field_zmax;
  P.field_half = (float)g.field_xy_half;
  for (int j = 0; j < CLR_N_CHANNELS; ++j) {
    P.det_ux[j] = det_u(10.0 * j, false);
    P.det_uy[j] = det_u(10.0 * j, true);
  }
  P.det_window = window_for_cell(g.ring_in, g.throat_out, g.throat_r);
  {
    // cos(atan(2 r_t / D)) with D the distance between the throat planes,
    // lowered by 1e-4 so that rounding never rejects a ray the apertures pass
    const double D = g.throat_out - g.ring_in, w = 2.0 * g.throat_r;
    P.det_cos_min = (float)(D / std::sqrt(D * D + w * w) - 1e-4);
  }
  for (int a = 0; a < CLR_MAX_AUX; ++a) {
    P.aux_ux[a] = det_u(g.aux_angles_deg[a], false);
    P.aux_uy[a] = det_u(g.aux_angles_deg[a], true);
    P.aux_deg[a] = (float)g.aux_angles_deg[a];
  }
  if (g.free_surface_model == CLR_SURFACE_RANKINE) {
    const double R = g.rin, a = g.vortex_core_radius, R2 = R * R, a2 = a * a;
    const double K = g.vortex_delta_h / (2.0 - a2 / R2);
    P.rk_K = K;
    P.rk_C0 = g.zmean - (2.0 * K / R2) * (R2 - 0.75 * a2 - a2 * std::log(R / a));
  } else {
    P.rk_K = 0.0;
    P.rk_C0 = g.zmean;
  }
  P.fs_model = g.free_surface_model; P.src_model = g.source_angular_model;
  P.event_model = g.event_model; P.sed_model = g.sediment_model;
  P.max_pb = g.max_particle_bounces; P.max_iter = g.max_iterations; P.cond_trials = g.conditional_max_trials;
  P.rough = (float)g.rough_alpha; P.aspect = (float)g.spheroid_aspect; P.kappa = (float)g.spheroid_kappa;
  P.orientation = g.spheroid_orientation;
  P.src_bore = g.source_bore_model; P.det_bore = g.detector_bore_model; P.face_model = g.ring_face_model;
  P.pd_model = g.photodiode_model; P.max_bore_bounces = g.max_bore_bounces;
  P.max_face_bounces = g.max_ring_face_bounces;
  P.n_bore = (float)g.bore_n; P.n_face = (float)g.ring_face_n;
  P.n_aux = g.n_aux; P.n_groups = g.n_groups; P.n_bins = s.n; P.n_theta = s.nt; P.n_radial = s.nrad;
  P.fn0 = g.field_n0; P.fn1 = g.field_n1; P.fn2 = g.field_n2;
  // lowest free-surface height (host float arithmetic of the device formula)
  float zs = P.zmean;
  if (P.fs_model == CLR_SURFACE_PARABOLIC && P.dh > 0.0f) zs = P.zmean - 0.5f * P.dh;
  else if (P.fs_model == CLR_SURFACE_SCULLY && P.dh > 0.0f) {
    const float a2 = P.core_r * P.core_r, R2 = P.rin * P.rin, H = P.dh * (a2 + R2) / R2;
    zs = P.zmean - H * (1.0f - (a2 / R2) * std::log1p(R2 / a2));
  } else if (P.fs_model == CLR_SURFACE_RANKINE && P.dh > 0.0f) zs = (float)P.rk_C0;
  P.zs_min = zs - 1e-6f * std::fmax(std::fabs(zs), 1e-3f);
  P.fs_c0 = P.zmean; P.fs_k = 0.0f; P.fs_a2 = 1.0f;
  P.fs_qpad = 1e-6f * P.rin * P.rin;
  if (P.fs_model == CLR_SURFACE_PARABOLIC && P.dh > 0.0f) {
    P.fs_c0 = P.zmean - 0.5f * P.dh; P.fs_k = P.dh / (P.rin * P.rin);
  } else if (P.fs_model == CLR_SURFACE_SCULLY && P.dh > 0.0f) {
    const float a2 = P.core_r * P.core_r;
    P.fs_c0 = zs; P.fs_k = P.dh * (a2 + P.rin * P.rin) / (P.rin * P.rin); P.fs_a2 = a2;
  } else if (P.fs_model == CLR_SURFACE_RANKINE && P.dh > 0.0f) {
    P.fs_c0 = (float)P.rk_C0; P.fs_k = (float)P.rk_K; P.fs_a2 = P.core_r * P.core_r;
  }
  P.radius = s.radius_f.as<float>();
  P.n_real = s.n_f.as<float>();
  P.k_imag = s.k_f.as<float>();
  P.albedo = s.albedo_f.as<float>();
  P.gcdf = s.gcdf_f.as<float>();
  P.wcdf = s.wcdf_f.as<float>();
  P.qsca = s.q.as<double>() + s.n;
  P.qext = s.q.as<double>();
  P.trace_g = s.trace_g.as<double>();
  P.trace_w = s.trace_w.as<double>();
  P.sedA = s.sedA.as<double>();
  P.sedB = s.sedB.as<double>();
  P.sednorm = s.sednorm.as<double>();
  P.radlog = s.radlog.as<double>();
  P.cdf = s.cdf_f.as<float>();
  P.theta = s.theta_f.as<float>();
  P.theta_guide = s.theta_guide.as<unsigned int>();
  P.gguide = s.gguide.as<unsigned int>();
  P.wguide = s.wguide.as<unsigned int>();
  P.field = s.field.as<float>();
  // the reference traces the Rouse/xyz scale with the bin-sum rates
  double sums[2] = {0.0, 0.0};
  (void)hipMemcpy(sums, s.trace_sums.p, sizeof(sums), hipMemcpyDeviceToHost);
  P.sum_trace_g = sums[0];
  P.sum_trace_w = sums[1];
  P.sum_g_f = (float)sums[0];
  P.sum_w_f = (float)sums[1];
  // piecewise majorant over z slabs: mu_track's formula with each slab's maximum
  P.n_slabs = 0;
  P.slab_z0 = P.zmin;
  P.slab_w = P.slab_inv_w = 0.0f;
  for (int j = 0; j < CLR_MAJ_SLABS; ++j) P.slab_mu[j] = P.mu_track;
  if (!s.slab_max.empty()) {
    const int S = (int)s.slab_max.size();
    P.n_slabs = S;
    P.slab_w = (float)((g.field_zmax - g.zmin) / S);
    P.slab_inv_w = (float)(S / (g.field_zmax - g.zmin));
    for (int j = 0; j < S; ++j)
      P.slab_mu[j] = (float)(sums[0] * (double)s.slab_max[j]) + (float)(sums[1] * (double)s.slab_max[j]);
  }
  P.slab_mu_max = 0.0f;
  for (int j = 0; j < P.n_slabs; ++j) P.slab_mu_max = std::max(P.slab_mu_max, P.slab_mu[j]);
  // inverse cell sizes of the shared fields, in float as the oracle computes them
  P.field_inv0 = P.field_inv1 = P.field_inv2 = 0.0f;
  if (g.sediment_model == CLR_SED_RZ_FIELD && g.field_n0 > 0 && g.field_n1 > 0) {
    P.field_inv0 = (float)g.field_n0 / P.rin;
    P.field_inv1 = (float)g.field_n1 / (P.field_zmax - P.zmin);
  } else if (g.sediment_model == CLR_SED_XYZ_FIELD && g.field_n0 > 0) {
    P.field_inv0 = (float)g.field_n0 / (2.0f * P.field_half);
    P.field_inv1 = (float)g.field_n1 / (2.0f * P.field_half);
    P.field_inv2 = (float)g.field_n2 / (P.field_zmax - P.zmin);
  }
}

int ensure_slots(clr_ctx* c, int need) {
  const size_t before = c->slots.size();
  while ((int)c->slots.size() < need) c->slots.push_back(new Slot());
  if (c->slots.size() == before && c->params.p) return 0;
  // DevBuf::ensure reallocates without keeping the contents, so after growing
  // the parameter array every prepared slot's parameters are uploaded again
  // from their host copies (the tallies are cleared at every launch).
  HIPCHK(c, c->params.ensure(sizeof(DevParams) * c->slots.size()));
  HIPCHK(c, c->tallies.ensure(sizeof(clr_tallies) * c->slots.size()));
  for (size_t k = 0; k < before; ++k)
    if (c->slots[k]->ready)
      HIPCHK(c, hipMemcpy(c->params.as<DevParams>() + k, &c->slots[k]->params, sizeof(DevParams),
                          hipMemcpyHostToDevice));
  return 0;
}

// The ideal hard-aperture photodiode weighs every hit 1 (CLARITAS.py:4418):
// its sums are the hardware counts, filled here rather than on the device.
void ideal_photodiode(const clr_config& g, clr_tallies& t) {
  if (g.photodiode_model != 0) return;
  for (int q = 0; q < CLR_MAX_GROUPS; ++q)
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < CLR_N_CHANNELS; ++j) t.photodiode[q][k][j] = (double)t.det[q][k][j];
}

void group_sizes(uint64_t n, int G, uint64_t* out) {
  for (int g = 0; g < CLR_MAX_GROUPS; ++g) out[g] = 0;
  auto lo = [&](int g) -> uint64_t {
    // ceil(g*n/G) without overflow for n < 2^57
    const unsigned __int128 num = (unsigned __int128)g * n;
    return (uint64_t)((num + (unsigned)G - 1) / (unsigned)G);
  };
  for (int g = 0; g < G; ++g) out[g] = lo(g + 1) - lo(g);
}

__global__ void clear_kernel(unsigned long long* __restrict__ p, size_t n_words, unsigned int* __restrict__ counter) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n_words; k += stride) p[k] = 0ull;
  if (i == 0) *counter = 0u;
}

struct RunShape {
  int block, grid, chunk, min_chunk, kind;
  bool fixed;  // chunk given by the caller: every item has that size
};

// Guided item sizes: an item holds the histories queued behind it divided by
// kGuidedSplit x grid, clamped to [kMinChunk, chunk] and rounded down to whole
// blocks, so the items a launch hands out last are small and its blocks run out
// of work together instead of waiting on one block's last large item.
constexpr uint64_t kGuidedSplit = 2;
constexpr int kMinChunk = 2048;

RunShape pick_shape(clr_ctx* c, const clr_launch* launch, uint64_t total, int kind) {
  RunShape r{};
  r.kind = kind;
  r.block = (launch && launch->block > 0) ? launch->block : 256;
  r.block = std::min(256, std::max(64, (r.block / 64) * 64));
  const int occ = clr_transport_occupancy(r.block, kind);
  int per_cu = (launch && launch->blocks_per_cu > 0) ? std::min(launch->blocks_per_cu, occ) : occ;
  r.grid = std::max(1, c->n_cu * per_cu);
  r.fixed = launch && launch->chunk > 0;
  if (r.fixed) r.chunk = launch->chunk;
  else {
    const uint64_t want = total / ((uint64_t)r.grid * 8ull);
    r.chunk = (int)std::min<uint64_t>(65536, std::max<uint64_t>(kMinChunk, want));
  }
  r.chunk = std::max(r.block, (r.chunk / r.block) * r.block);
  r.min_chunk = std::min(r.chunk, std::max(r.block, (kMinChunk / r.block) * r.block));
  return r;
}

// Builds the device work queue in the order items are taken.  Item sizes are
// guided by the estimated work queued behind them (histories times a
// per-history cost), so the items handed out last are small in work, not just
// in histories.
struct QueueBuilder {
  const RunShape& sh;
  double remaining;  // estimated work queued from the next item on
  std::vector<WorkItem> items;
  QueueBuilder(const RunShape& s, double total_work) : sh(s), remaining(total_work) {}
  // histories [start, start + n) of a configuration whose run numbers n_total;
  // `cost` is its estimated work per history
  void add(int cfg, uint64_t start, uint64_t n, uint64_t n_total, uint64_t seed, uint64_t offset, double cost = 1.0) {
    for (uint64_t s = 0; s < n;) {
      uint64_t c = (uint64_t)sh.chunk;
      if (!sh.fixed) {
        const double g = remaining / ((double)kGuidedSplit * (double)sh.grid * cost);
        c = std::min<uint64_t>(c, std::max<uint64_t>((uint64_t)sh.min_chunk, (uint64_t)std::min(g, 1e18)));
        c = std::max<uint64_t>((uint64_t)sh.block, (c / sh.block) * sh.block);
      }
      c = std::min<uint64_t>(c, n - s);
      WorkItem w{};
      w.config = cfg;
      w.start = start + s;
      w.count = c;
      w.n_total = n_total;
      w.seed = seed;
      w.offset = offset;
      items.push_back(w);
      s += c;
      remaining = std::max(0.0, remaining - cost * (double)c);
    }
  }
};

// Estimated work per history of a prepared configuration, in flights: a few
// for launch, walls and exit, plus the majorant collisions (real and null)
// along a path of about one cell diameter.
double history_cost(const Slot& s) {
  const double tau = (double)s.params.mu_track * 2.0 * (double)s.params.rin;
  return 4.0 + (std::isfinite(tau) && tau > 0.0 ? tau : 0.0);
}

// Field maximum over each z slab of the piecewise-constant Woodcock majorant
// (shared r-z and x-y-z fields).  Slab j covers [zmin + j w, zmin + (j+1) w),
// w = (zmax - zmin) / S, and the end slabs extend to the clamped field edges.
// The interpolation at z uses the z nodes floor(u) and floor(u) + 1 with
// u = (z - zmin) nz / (zmax - zmin) - 1/2; one extra node on each side keeps
// the bound safe against rounding of u.  Each slab's maximum is over every
// r (x, y) node, so it bounds the interpolated field anywhere in the slab.
void slab_maxima(const clr_config& g, const float* field, size_t len, std::vector<float>& out) {
  out.clear();
  const bool rz = g.sediment_model == CLR_SED_RZ_FIELD, xyz = g.sediment_model == CLR_SED_XYZ_FIELD;
  if (!(rz || xyz) || !field) return;
  const int nz = rz ? g.field_n1 : g.field_n2;
  const size_t planes = rz ? (size_t)g.field_n0 : (size_t)g.field_n0 * g.field_n1;
  if (nz < 2 || planes == 0 || len < planes * nz || !(g.field_zmax > g.zmin)) return;
  const int S = std::min(CLR_MAJ_SLABS, nz);
  out.assign(S, 0.0f);
  for (int j = 0; j < S; ++j) {
    int k0 = (int)std::floor((double)j * nz / S - 0.5) - 1, k1 = (int)std::floor((double)(j + 1) * nz / S - 0.5) + 2;
    if (j == 0) k0 = 0;
    if (j == S - 1) k1 = nz - 1;
    k0 = std::max(k0, 0);
    k1 = std::min(k1, nz - 1);
    float m = 0.0f;
    for (size_t pl = 0; pl < planes; ++pl)
      for (int k = k0; k <= k1; ++k) m = std::max(m, field[pl * nz + k]);
    out[j] = m;
  }
}

// kernel variant that serves every configuration of a launch (CLR_KERNEL_*)
int kernel_kind(clr_ctx* c, const std::vector<int>& used) {
  bool plain = true, rough = false;
  for (int k : used) {
    const Slot& s = *c->slots[k];
    const clr_config& g = s.cfg;
    if (g.event_model != CLR_EVENT_PLAIN_MIE) plain = false;
    if (!(g.sediment_model == CLR_SED_UNIFORM || g.sediment_model == CLR_SED_RZ_FIELD ||
          g.sediment_model == CLR_SED_XYZ_FIELD))
      plain = false;
    if (g.rough_alpha > 1e-7 || (g.event_model != CLR_EVENT_PLAIN_MIE && s.params.coherent < 1.0f)) rough = true;
    // reflecting collimators, the ring-face cavity and the SFH213 weights
    // are compiled into the ROUGH variant only
    if (g.source_bore_model || g.detector_bore_model || g.ring_face_model || g.photodiode_model) {
      plain = false;
      rough = true;
    }
  }
  return plain ? CLR_KERNEL_PLAIN : rough ? CLR_KERNEL_ROUGH : CLR_KERNEL_GENERIC;
}

// Upload items, clear tallies, launch the persistent kernel (optionally
// through a cached hipGraph), time it with events on the library stream.
int launch_items(clr_ctx* c, const std::vector<WorkItem>& items, const RunShape& sh, bool use_graph,
                 clr_record* d_records, uint64_t n_records, const LightDiag* light = nullptr) {
  const int n_slots = (int)c->slots.size();
  HIPCHK(c, c->items.ensure(sizeof(WorkItem) * std::max<size_t>(items.size(), 1)));
  HIPCHK(c, c->counter.ensure(sizeof(unsigned int)));
  HIPCHK(c, hipMemcpyAsync(c->items.p, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice, c->stream));
  TransportLaunch L{};
  L.params = c->params.as<DevParams>();
  L.items = c->items.as<WorkItem>();
  L.n_items = (int)items.size();
  L.item_counter = c->counter.as<unsigned int>();
  L.tallies = c->tallies.as<clr_tallies>();
  L.records = d_records;
  L.n_records = n_records;
  L.grid = sh.grid;
  L.block = sh.block;
  L.kind = sh.kind;
  L.light = light;
  L.causal = c->causal_rec;
  L.causal_count = c->causal_count;
  L.causal_cap = c->causal_cap;
  L.fail_rec = c->fail_rec;
  L.fail_count = c->fail_count;
  L.fail_cap = c->fail_cap;
  const size_t words = sizeof(clr_tallies) * n_slots / 8;
  if (use_graph) {
    GraphCache& G = c->graph;
    const GraphKey key = graph_key(L, n_slots);
    if (!(G.exec && same_graph_key(G.key, key))) {
      G.reset();
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      hipLaunchKernelGGL(clear_kernel, dim3(256), dim3(256), 0, c->stream, c->tallies.as<unsigned long long>(), words,
                         L.item_counter);
      hipError_t le = clr_launch_transport(L, c->stream);
      hipError_t ce = hipStreamEndCapture(c->stream, &G.graph);
      if (le != hipSuccess) return fail(c, CLR_E_HIP, std::string("transport launch: ") + hipGetErrorString(le));
      if (ce != hipSuccess) return fail(c, CLR_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
      HIPCHK(c, hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
      G.key = key;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(G.exec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  } else {
    hipLaunchKernelGGL(clear_kernel, dim3(256), dim3(256), 0, c->stream, c->tallies.as<unsigned long long>(), words,
                       L.item_counter);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, clr_launch_transport(L, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  }
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0.0f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  return 0;
}

// Validate a light-diagnostic configuration, size and clear its device
// buffers and upload the descriptor (CLARITAS.py:7232-7300).
int light_setup(clr_ctx* c, const clr_light_config* lc, float rin, float zmin, LightDiag& D) {
  if (lc->heatmap_size < 0 || lc->heatmap_size > 4096) return fail(c, CLR_E_ARG, "heatmap_size must be in [0, 4096]");
  if (lc->max_rays < 1 && lc->heatmap_size < 1)
    return fail(c, CLR_E_ARG, "light diagnostic max_rays must be >=1 (or heat maps requested)");
  if (std::min(lc->nx, std::min(lc->ny, lc->nz)) < 4)
    return fail(c, CLR_E_ARG, "light diagnostic grid dimensions must each be >=4");
  if (lc->representative_rays < 0 || (uint64_t)lc->representative_rays > lc->max_rays || lc->max_segments < 1)
    return fail(c, CLR_E_ARG, "invalid representative-ray diagnostic configuration");
  if (!(lc->zmax > (double)zmin) || !(rin > 0.0f)) return fail(c, CLR_E_ARG, "light diagnostic box is empty");
  const size_t vox = (size_t)lc->nx * lc->ny * lc->nz, rep = (size_t)lc->representative_rays;
  const size_t traj = std::max<size_t>(rep * lc->max_segments, 1), rep1 = std::max<size_t>(rep, 1);
  const double bytes = (double)vox * 20 + (double)traj * 32 + (double)rep1 * 8 + 8;
  if (bytes > 1e9)
    return fail(c, CLR_E_RANGE, "light diagnostic allocation is too large; reduce grid dimensions, "
                                "representative rays, or segments per ray");
  HIPCHK(c, hipSetDevice(c->device));
  LightBufs& B = c->light;
  HIPCHK(c, B.grids.ensure(vox * 20));
  HIPCHK(c, B.scalars.ensure(128));
  HIPCHK(c, B.traj.ensure(traj * 32 + rep1 * 8));
  HIPCHK(c, B.desc.ensure(sizeof(LightDiag)));
  const size_t vis2 = (size_t)lc->heatmap_size * lc->heatmap_size;
  HIPCHK(c, B.heat.ensure(std::max<size_t>(vis2, 1) * 8));
  char* gp = B.grids.as<char>();
  char* sp = B.scalars.as<char>();
  D = LightDiag{};
  D.total = reinterpret_cast<float*>(gp);
  D.post = D.total + vox;
  D.event_weight = D.post + vox;
  D.event_count = reinterpret_cast<unsigned int*>(D.event_weight + vox);
  D.absorption_count = D.event_count + vox;
  // scalars: 2 doubles, 3 u64, error and snap counts, outside[8] (80 bytes)
  D.segment_total = reinterpret_cast<double*>(sp);
  D.deposited_total = D.segment_total + 1;
  D.segment_count = reinterpret_cast<unsigned long long*>(sp + 16);
  D.event_total = D.segment_count + 1;
  D.absorption_total = D.segment_count + 2;
  D.error_count = reinterpret_cast<unsigned int*>(sp + 40);
  D.boundary_adjust = D.error_count + 1;
  D.outside = D.error_count + 2;
  float* tf = B.traj.as<float>();
  D.tx0 = tf; D.ty0 = tf + traj; D.tz0 = tf + 2 * traj;
  D.tx1 = tf + 3 * traj; D.ty1 = tf + 4 * traj; D.tz1 = tf + 5 * traj;
  D.torder = reinterpret_cast<int*>(tf + 6 * traj);
  D.tevent = D.torder + traj;
  D.tcount = D.tevent + traj;
  D.toverflow = D.tcount + rep1;
  D.max_rays = lc->max_rays;
  D.nx = lc->nx; D.ny = lc->ny; D.nz = lc->nz;
  D.rep_rays = lc->representative_rays; D.max_segments = lc->max_segments;
  D.rin = rin; D.zmin = zmin; D.zmax = (float)lc->zmax;
  D.vis = lc->heatmap_size;
  D.heat_xy = B.heat.as<float>();
  D.heat_zy = D.heat_xy + vis2;
  if (vis2) HIPCHK(c, hipMemsetAsync(D.heat_xy, 0, vis2 * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(gp, 0, vox * 20, c->stream));
  HIPCHK(c, hipMemsetAsync(sp, 0, 128, c->stream));
  HIPCHK(c, hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(tf), 0x7fc00000, traj * 6, c->stream));  // NaN
  HIPCHK(c, hipMemsetAsync(D.torder, 0xff, traj * 8, c->stream));                                         // -1
  HIPCHK(c, hipMemsetAsync(D.tcount, 0, rep1 * 8, c->stream));
  HIPCHK(c, hipMemcpyAsync(B.desc.p, &D, sizeof(LightDiag), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// Copy the diagnostics of D into the caller's arrays (NULL ones skipped).
int light_readback(clr_ctx* c, const clr_light_config* lc, const LightDiag& D, clr_light_out* lo) {
  const size_t vox = (size_t)lc->nx * lc->ny * lc->nz, rep = (size_t)lc->representative_rays;
  auto get = [&](void* dst, const void* src, size_t nb) -> hipError_t {
    return dst ? hipMemcpyAsync(dst, src, nb, hipMemcpyDeviceToHost, c->stream) : hipSuccess;
  };
  HIPCHK(c, get(lo->track_length, D.total, vox * 4));
  HIPCHK(c, get(lo->post_scatter_track_length, D.post, vox * 4));
  HIPCHK(c, get(lo->scatter_event_weight, D.event_weight, vox * 4));
  HIPCHK(c, get(lo->scatter_event_count, D.event_count, vox * 4));
  HIPCHK(c, get(lo->absorption_event_count, D.absorption_count, vox * 4));
  if (rep) {
    const size_t tb = rep * lc->max_segments;
    HIPCHK(c, get(lo->traj_x0, D.tx0, tb * 4)); HIPCHK(c, get(lo->traj_y0, D.ty0, tb * 4));
    HIPCHK(c, get(lo->traj_z0, D.tz0, tb * 4)); HIPCHK(c, get(lo->traj_x1, D.tx1, tb * 4));
    HIPCHK(c, get(lo->traj_y1, D.ty1, tb * 4)); HIPCHK(c, get(lo->traj_z1, D.tz1, tb * 4));
    HIPCHK(c, get(lo->traj_order, D.torder, tb * 4)); HIPCHK(c, get(lo->traj_event, D.tevent, tb * 4));
    HIPCHK(c, get(lo->traj_count, D.tcount, rep * 4)); HIPCHK(c, get(lo->traj_overflow, D.toverflow, rep * 4));
  }
  const size_t vis2 = (size_t)lc->heatmap_size * lc->heatmap_size;
  if (vis2) {
    HIPCHK(c, get(lo->heatmap_xy, D.heat_xy, vis2 * 4));
    HIPCHK(c, get(lo->heatmap_zy, D.heat_zy, vis2 * 4));
  }
  unsigned char sc[80];
  HIPCHK(c, hipMemcpyAsync(sc, D.segment_total, 80, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::memcpy(&lo->segment_total, sc, 8);
  std::memcpy(&lo->deposited_total, sc + 8, 8);
  std::memcpy(&lo->segment_count, sc + 16, 8);
  std::memcpy(&lo->scatter_event_total, sc + 24, 8);
  std::memcpy(&lo->absorption_event_total, sc + 32, 8);
  std::memcpy(&lo->error_count, sc + 40, 4);
  std::memcpy(&lo->boundary_adjustment_count, sc + 44, 4);
  std::memcpy(lo->outside, sc + 48, 32);
  return 0;
}

// Trace n histories of slot 0 on kernel variant `kind` (with the light
// diagnostics when `light` is given) and reduce the tallies into *out.
// histories [r0, r0 + rn) of an n-history run (rn 0: all), jackknife
// groups numbered over the whole run
int run_single(clr_ctx* c, uint64_t n, uint64_t seed, uint64_t offset, const clr_launch* launch, clr_tallies* out,
               clr_record* records, uint64_t n_records, int kind, const LightDiag* light, uint64_t r0 = 0,
               uint64_t rn = 0) {
  if (n == 0) return fail(c, CLR_E_ARG, "n must be positive");
  if (n > (1ull << 56) || offset > (1ull << 62)) return fail(c, CLR_E_RANGE, "history range too large");
  if (rn == 0) { r0 = 0; rn = n; }
  if (r0 >= n || rn > n - r0) return fail(c, CLR_E_RANGE, "history sub-range outside the run");
  HIPCHK(c, hipSetDevice(c->device));
  const RunShape sh = pick_shape(c, launch, rn, kind);
  // one launch covers at most 2^31 histories so LDS u32 tallies cannot wrap
  const uint64_t per_launch = 1ull << 31;
  const int n_slots = (int)c->slots.size();
  std::vector<clr_tallies> acc(1);
  memset(&acc[0], 0, sizeof(clr_tallies));
  std::vector<clr_tallies> host(n_slots);
  clr_record* d_rec = nullptr;
  const uint64_t nrec = std::min<uint64_t>(n_records, n);
  if (records && nrec) {
    HIPCHK(c, c->records.ensure(sizeof(clr_record) * nrec));
    d_rec = c->records.as<clr_record>();
  }
  double total_ms = 0.0;
  const int G = c->slots[0]->cfg.n_groups;
  for (uint64_t done = r0; done < r0 + rn; done += per_launch) {
    const uint64_t m = std::min(per_launch, r0 + rn - done);
    // histories of this launch keep their global group by numbering them
    // within the whole run: items carry start offsets into [0, n)
    QueueBuilder q(sh, (double)m);
    q.add(0, done, m, n, seed, offset);
    if (int e = launch_items(c, q.items, sh, false, done == r0 ? d_rec : nullptr, done == r0 ? nrec : 0, light))
      return e;
    total_ms += c->last_ms;
    HIPCHK(c, hipMemcpy(host.data(), c->tallies.p, sizeof(clr_tallies), hipMemcpyDeviceToHost));
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&host[0]);
    uint64_t* dst = reinterpret_cast<uint64_t*>(&acc[0]);
    // u64 fields up to incidence_sum, FP64 sums up to status, u64 after
    for (size_t k = 0; k < offsetof(clr_tallies, incidence_sum) / 8; ++k) dst[k] += src[k];
    const double* fs = reinterpret_cast<const double*>(&host[0]);
    double* fd = reinterpret_cast<double*>(&acc[0]);
    for (size_t k = offsetof(clr_tallies, incidence_sum) / 8; k < offsetof(clr_tallies, status) / 8; ++k) fd[k] += fs[k];
    const size_t a0 = offsetof(clr_tallies, status) / 8, a1 = offsetof(clr_tallies, n_groups) / 8;
    for (size_t k = a0; k < a1; ++k) dst[k] += src[k];
  }
  c->last_ms = total_ms;
  *out = acc[0];
  out->n_groups = G;
  group_sizes(n, G, out->group_histories);
  if (rn != n) {  // the range's share of each group: i in [r0, r0 + rn) with floor(i G / n) == g
    uint64_t lo[CLR_MAX_GROUPS + 1] = {0};
    for (int g = 0; g <= G; ++g) {  // first history of group g: ceil(g n / G)
      const unsigned __int128 v = (unsigned __int128)g * n;
      lo[g] = (uint64_t)((v + (unsigned)G - 1) / (unsigned)G);
    }
    for (int g = 0; g < G; ++g) {
      const uint64_t a = std::max(lo[g], r0), b = std::min(lo[g + 1], r0 + rn);
      out->group_histories[g] = b > a ? b - a : 0;
    }
  }
  ideal_photodiode(c->slots[0]->cfg, *out);
  out->counters[CLR_C_HISTORIES] = rn;  // launched histories (the device counts per-event work only)
  if (d_rec) HIPCHK(c, hipMemcpy(records, d_rec, sizeof(clr_record) * nrec, hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace

extern "C" {

int clr_abi_version(void) { return CLR_ABI_VERSION; }

size_t clr_sizeof(int id) {
  switch (id) {
    case 0: return sizeof(clr_config);
    case 1: return sizeof(clr_particles);
    case 2: return sizeof(clr_sediment);
    case 3: return sizeof(clr_mie_out);
    case 4: return sizeof(clr_tables_out);
    case 5: return sizeof(clr_tallies);
    case 6: return sizeof(clr_record);
    case 7: return sizeof(clr_batch_item);
    case 8: return sizeof(clr_launch);
    case 9: return sizeof(clr_meridional);
    case 10: return sizeof(clr_meridional_out);
    case 11: return sizeof(clr_light_config);
    case 12: return sizeof(clr_light_out);
    case 13: return sizeof(clr_causal_record);
    case 14: return sizeof(clr_failure_sample);
    default: return 0;
  }
}

int clr_create(clr_ctx** out, int device) {
  if (!out) return CLR_E_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CLR_E_HIP;
  if (device < 0 || device >= count) return CLR_E_ARG;
  clr_ctx* c = new clr_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return CLR_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
  if (mlp_static_scratch(device, &c->static_scratch, c->err) != 0) {
    clr_destroy(c);
    return CLR_E_HIP;
  }
  *out = c;
  return 0;
}

void clr_destroy(clr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->graph.reset();
  for (Slot* s : c->slots) { s->release(); delete s; }
  DevBuf* bufs[] = {&c->params,      &c->tallies,      &c->items,         &c->counter,         &c->records,
                    &c->scratch_u32, &c->mie_work,     &c->light.grids,   &c->light.scalars,   &c->light.traj,
                    &c->light.desc,  &c->light.probe};
  for (DevBuf* b : bufs) b->release();
  clr_mlp_free(c->mlp);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* clr_last_error(const clr_ctx* c) { return c ? c->err.c_str() : "null context"; }

int clr_device_info(clr_ctx* c, int64_t info[8]) {
  if (!c || !info) return CLR_E_ARG;
  hipDeviceProp_t p;
  HIPCHK(c, hipGetDeviceProperties(&p, c->device));
  info[0] = p.multiProcessorCount;
  info[1] = p.warpSize;
  info[2] = (int64_t)p.sharedMemPerBlock;
  info[3] = p.clockRate;
  info[4] = (int64_t)p.major * 100 + p.minor;
  info[5] = (int64_t)(p.totalGlobalMem >> 20);
  info[6] = clr_transport_occupancy(256, CLR_KERNEL_PLAIN);
  info[7] = clr_transport_occupancy(256, CLR_KERNEL_GENERIC);
  return 0;
}

int clr_mie_tables(clr_ctx* c, const clr_particles* p, clr_mie_out* o) {
  if (!c || !o) return fail(c, CLR_E_ARG, "null argument");
  if (int e = check_particles(c, p)) return e;
  HIPCHK(c, hipSetDevice(c->device));
  Slot s;
  int e = run_mie(c, s, p);
  if (!e) {
    const int n = p->n, nt = s.nt;
    std::vector<double> q((size_t)7 * n);
    std::vector<int> nstop(n);
    hipError_t he = hipMemcpy(q.data(), s.q.p, sizeof(double) * 7 * n, hipMemcpyDeviceToHost);
    if (he == hipSuccess) he = hipMemcpy(nstop.data(), s.nstop.p, sizeof(int) * n, hipMemcpyDeviceToHost);
    double* dst[7] = {o->qext, o->qsca, o->qabs, o->qback, o->asym_g, o->albedo, o->size_parameter};
    for (int k = 0; k < 7 && he == hipSuccess; ++k)
      if (dst[k]) memcpy(dst[k], q.data() + (size_t)k * n, sizeof(double) * n);
    if (he == hipSuccess && o->nstop) memcpy(o->nstop, nstop.data(), sizeof(int) * n);
    if (he == hipSuccess && o->theta && p->n_theta >= 2) he = hipMemcpy(o->theta, s.theta.p, sizeof(double) * nt, hipMemcpyDeviceToHost);
    if (he == hipSuccess && o->phase_pdf && p->n_theta >= 2)
      he = hipMemcpy(o->phase_pdf, s.pdf.p, sizeof(double) * (size_t)n * nt, hipMemcpyDeviceToHost);
    if (he == hipSuccess && o->phase_cdf && p->n_theta >= 2)
      he = hipMemcpy(o->phase_cdf, s.cdf.p, sizeof(double) * (size_t)n * nt, hipMemcpyDeviceToHost);
    if (he != hipSuccess) e = fail(c, CLR_E_HIP, std::string("copy Mie tables: ") + hipGetErrorString(he));
  }
  s.release();
  return e;
}

int clr_prepare(clr_ctx* c, int slot, const clr_config* g, const clr_particles* p, const clr_sediment* sed,
                const float* field, size_t field_len) {
  if (!c) return CLR_E_ARG;
  if (slot < 0 || slot > 4095) return fail(c, CLR_E_ARG, "slot out of range");
  if (int e = check_particles(c, p)) return e;
  if (p->n_theta < 2) return fail(c, CLR_E_ARG, "particles: n_theta must be >= 2 for transport");
  if (int e = validate_config(c, g, p)) return e;
  if (int e = validate_sediment(c, g, sed, field_len, p->n)) return e;
  if ((g->sediment_model >= CLR_SED_GRID_PER_BIN) && !field) return fail(c, CLR_E_ARG, "field is NULL");
  HIPCHK(c, hipSetDevice(c->device));
  c->graph.reset();
  if (int e = ensure_slots(c, slot + 1)) return e;
  Slot& s = *c->slots[slot];
  s.ready = false;
  if (int e = run_mie(c, s, p)) return e;
  const int n = p->n;
  s.nrad = (sed && (g->sediment_model == CLR_SED_STOKES || g->sediment_model == CLR_SED_SCHILLER ||
                    g->sediment_model == CLR_SED_GRID_PER_BIN))
               ? sed->n_radial
               : 2;
  DevBuf* dn[] = {&s.radius, &s.nd, &s.mug, &s.mus, &s.mue, &s.muw, &s.muev, &s.trace_g, &s.trace_w, &s.gcdf,
                  &s.wcdf, &s.sedA, &s.sedB, &s.sednorm, &s.sedws};
  for (DevBuf* b : dn) HIPCHK(c, b->ensure(sizeof(double) * n));
  DevBuf* fn[] = {&s.radius_f, &s.n_f, &s.k_f, &s.albedo_f, &s.gcdf_f, &s.wcdf_f};
  for (DevBuf* b : fn) HIPCHK(c, b->ensure(sizeof(float) * n));
  HIPCHK(c, s.mu_scalars.ensure(sizeof(double) * 8));
  HIPCHK(c, s.trace_sums.ensure(sizeof(double) * 2));
  HIPCHK(c, s.majorants.ensure(sizeof(double) * 4));
  HIPCHK(c, s.radlog.ensure(sizeof(double) * (size_t)n * s.nrad));
  s.field_len = 0;
  if (field && field_len) {
    HIPCHK(c, s.field.ensure(sizeof(float) * field_len));
    HIPCHK(c, hipMemcpyAsync(s.field.p, field, sizeof(float) * field_len, hipMemcpyHostToDevice, c->stream));
    s.field_len = field_len;
  }
  slab_maxima(*g, field, field_len, s.slab_max);
  HIPCHK(c, c->scratch_u32.ensure(sizeof(unsigned int) * (size_t)(n + 2)));
  HIPCHK(c, hipMemsetAsync(c->scratch_u32.p, 0, sizeof(unsigned int) * (size_t)(n + 2), c->stream));
  int* eflag = c->scratch_u32.as<int>() + n + 1;
  PrepLaunch L{};
  L.n = n; L.weight_mode = p->weight_mode; L.event_model = p->event_model; L.sed_model = g->sediment_model;
  L.fs_model = g->free_surface_model; L.n_quad = sed ? sed->n_quad : 64; L.n_radial = s.nrad;
  L.concentration = p->concentration; L.density = p->density; L.n_medium = p->n_medium; L.wavelength = p->wavelength;
  L.hybrid_wave_scale = p->hybrid_wave_scale; L.wave_rms_height = p->wave_rms_height_m;
  if (sed) {
    L.particle_density = sed->particle_density; L.water_density = sed->water_density;
    L.viscosity = sed->water_viscosity; L.gravity = sed->gravity; L.eddy_diffusivity = sed->eddy_diffusivity;
  }
  L.R = g->rin; L.zmin = g->zmin; L.zmean = g->zmean; L.dh = g->vortex_delta_h; L.core = g->vortex_core_radius;
  L.diameter = s.in_d.as<double>(); L.weight = s.in_w.as<double>(); L.n_real = s.in_nr.as<double>();
  L.k_imag = s.in_k.as<double>(); L.qext = s.q.as<double>(); L.qsca = s.q.as<double>() + n;
  L.field = s.field_len ? s.field.as<float>() : nullptr;
  L.field_len = s.field_len;
  L.radius = s.radius.as<double>(); L.number_density = s.nd.as<double>(); L.mu_geom = s.mug.as<double>();
  L.mu_sca = s.mus.as<double>(); L.mu_ext = s.mue.as<double>(); L.mu_wave = s.muw.as<double>();
  L.mu_event = s.muev.as<double>(); L.trace_g = s.trace_g.as<double>(); L.trace_w = s.trace_w.as<double>();
  L.geom_cdf = s.gcdf.as<double>(); L.wave_cdf = s.wcdf.as<double>(); L.mu_scalars = s.mu_scalars.as<double>();
  L.trace_sums = s.trace_sums.as<double>(); L.sedA = s.sedA.as<double>(); L.sedB = s.sedB.as<double>();
  L.sednorm = s.sednorm.as<double>(); L.sedws = s.sedws.as<double>(); L.radlog = s.radlog.as<double>();
  L.majorants = s.majorants.as<double>(); L.radius_f = s.radius_f.as<float>(); L.n_real_f = s.n_f.as<float>();
  L.k_imag_f = s.k_f.as<float>(); L.albedo_f = s.albedo_f.as<float>(); L.geom_cdf_f = s.gcdf_f.as<float>();
  L.wave_cdf_f = s.wcdf_f.as<float>(); L.error_flag = eflag;
  HIPCHK(c, clr_launch_precompute(L, c->scratch_u32.as<unsigned int>(), c->stream));
  if (s.nt > 65535 || n > 65535) return fail(c, CLR_E_ARG, "transport needs n_theta and the bin count <= 65535");
  HIPCHK(c, s.theta_guide.ensure(sizeof(unsigned int) * (size_t)n * kGuideN));
  HIPCHK(c, s.gguide.ensure(sizeof(unsigned int) * kGuideN));
  HIPCHK(c, s.wguide.ensure(sizeof(unsigned int) * kGuideN));
  HIPCHK(c, clr_launch_guides(s.cdf_f.as<float>(), n, s.nt, s.theta_guide.as<unsigned int>(), c->stream));
  HIPCHK(c, clr_launch_guides(s.gcdf_f.as<float>(), 1, n, s.gguide.as<unsigned int>(), c->stream));
  HIPCHK(c, clr_launch_guides(s.wcdf_f.as<float>(), 1, n, s.wguide.as<unsigned int>(), c->stream));
  int ecode = 0;
  HIPCHK(c, hipMemcpyAsync(&ecode, eflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.mu_scalars_h, s.mu_scalars.p, sizeof(double) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.majorants_h, s.majorants.p, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ecode) return fail(c, ecode, "transport/sediment precompute failed with error code " + std::to_string(ecode));
  s.cfg = *g;
  fill_params(s, *g);
  HIPCHK(c, hipMemcpy(c->params.as<DevParams>() + slot, &s.params, sizeof(DevParams), hipMemcpyHostToDevice));
  s.ready = true;
  return 0;
}

int clr_n_slots(clr_ctx* c) { return c ? (int)c->slots.size() : 0; }

int clr_get_tables(clr_ctx* c, int slot, clr_tables_out* o) {
  if (!c || !o || slot < 0 || slot >= (int)c->slots.size() || !c->slots[slot]->ready)
    return fail(c, CLR_E_STATE, "slot not prepared");
  HIPCHK(c, hipSetDevice(c->device));
  Slot& s = *c->slots[slot];
  const size_t nb = sizeof(double) * s.n;
  struct { double* dst; DevBuf* src; } m[] = {
      {o->radius, &s.radius}, {o->number_density, &s.nd}, {o->mu_geom_bin, &s.mug}, {o->mu_sca_bin, &s.mus},
      {o->mu_ext_bin, &s.mue}, {o->trace_geom_bin, &s.trace_g}, {o->trace_wave_bin, &s.trace_w},
      {o->geom_cdf, &s.gcdf}, {o->wave_cdf, &s.wcdf}, {o->sed_A, &s.sedA}, {o->sed_B, &s.sedB},
      {o->sed_norm, &s.sednorm}, {o->sed_settling, &s.sedws}};
  for (auto& k : m)
    if (k.dst) HIPCHK(c, hipMemcpy(k.dst, k.src->p, nb, hipMemcpyDeviceToHost));
  if (o->sed_radial_log) HIPCHK(c, hipMemcpy(o->sed_radial_log, s.radlog.p, nb * s.nrad, hipMemcpyDeviceToHost));
  memcpy(o->mu_scalars, s.mu_scalars_h, sizeof(s.mu_scalars_h));
  memcpy(o->majorants, s.majorants_h, sizeof(s.majorants_h));
  return 0;
}

int clr_probe_shape(int mode, int* in_stride, int* out_stride) {
  static const int kIn[CLR_PROBE_N_MODES] = {6, 6, 6, 4, 6, 3, 1, 1, 7};
  static const int kOut[CLR_PROBE_N_MODES] = {6, 4, 4, 2, 2, 2, 6, 1, 9};
  if (mode < 0 || mode >= CLR_PROBE_N_MODES) return CLR_E_ARG;
  if (in_stride) *in_stride = kIn[mode];
  if (out_stride) *out_stride = kOut[mode];
  return 0;
}

int clr_probe(clr_ctx* c, int slot, int mode, int n, uint64_t seed, const float* in, double* out) {
  if (!c || !out || (n > 0 && !in)) return fail(c, CLR_E_ARG, "null argument");
  int si = 0, so = 0;
  if (clr_probe_shape(mode, &si, &so) != 0) return fail(c, CLR_E_ARG, "unknown probe mode");
  if (n < 0 || n > (1 << 24)) return fail(c, CLR_E_RANGE, "probe item count must be in [0, 2^24]");
  if (slot < 0 || slot >= (int)c->slots.size() || !c->slots[slot]->ready) return fail(c, CLR_E_STATE, "slot not prepared");
  if (n == 0) return 0;
  HIPCHK(c, hipSetDevice(c->device));
  void* din = nullptr;
  void* dout = nullptr;
  const size_t bin = sizeof(float) * (size_t)n * si, bout = sizeof(double) * (size_t)n * so;
  HIPCHK(c, hipMalloc(&din, bin));
  hipError_t e = hipMalloc(&dout, bout);
  if (e == hipSuccess) e = hipMemcpyAsync(din, in, bin, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = clr_launch_probe(c->params.as<DevParams>() + slot, mode, n, static_cast<const float*>(din), si,
                         static_cast<double*>(dout), so, seed, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, bout, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) return fail(c, CLR_E_HIP, std::string("probe: ") + hipGetErrorString(e));
  return 0;
}

int clr_run(clr_ctx* c, uint64_t n, uint64_t seed, uint64_t offset, const clr_launch* launch, clr_tallies* out,
            clr_record* records, uint64_t n_records) {
  if (!c || !out) return fail(c, CLR_E_ARG, "null argument");
  if (c->slots.empty() || !c->slots[0]->ready) return fail(c, CLR_E_STATE, "clr_prepare(slot 0) has not run");
  return run_single(c, n, seed, offset, launch, out, records, n_records, kernel_kind(c, {0}), nullptr);
}

int clr_run_light(clr_ctx* c, uint64_t n, uint64_t seed, uint64_t offset, const clr_launch* launch,
                  const clr_light_config* lc, clr_tallies* out, clr_light_out* lo) {
  if (!c || !out || !lc || !lo) return fail(c, CLR_E_ARG, "null argument");
  if (c->slots.empty() || !c->slots[0]->ready) return fail(c, CLR_E_STATE, "clr_prepare(slot 0) has not run");
  const clr_config& g = c->slots[0]->cfg;
  LightDiag D{};
  if (int e = light_setup(c, lc, (float)g.rin, (float)g.zmin, D)) return e;
  if (int e = run_single(c, n, seed, offset, launch, out, nullptr, 0, CLR_KERNEL_LIGHT, c->light.desc.as<LightDiag>()))
    return e;
  return light_readback(c, lc, D, lo);
}

int clr_run_causal(clr_ctx* c, uint64_t n, uint64_t seed, uint64_t offset, const clr_launch* launch,
                   const clr_light_config* lc, clr_tallies* out, clr_light_out* lo, clr_causal_out* co) {
  if (!c || !out || !co || (lc && !lo)) return fail(c, CLR_E_ARG, "null argument");
  if (co->capacity && !co->records) return fail(c, CLR_E_ARG, "causal records buffer is NULL");
  if (co->capacity > (1ull << 40)) return fail(c, CLR_E_RANGE, "causal capacity too large");
  if (c->slots.empty() || !c->slots[0]->ready) return fail(c, CLR_E_STATE, "clr_prepare(slot 0) has not run");
  HIPCHK(c, hipSetDevice(c->device));
  if (co->failure_capacity && !co->failures) return fail(c, CLR_E_ARG, "failure samples buffer is NULL");
  if (co->failure_capacity > (1ull << 24)) return fail(c, CLR_E_RANGE, "failure sample capacity too large");
  // record area, failure samples, then the two 8-byte append counters
  const size_t rec_bytes = sizeof(clr_causal_record) * (size_t)co->capacity;
  const size_t fs_bytes = sizeof(clr_failure_sample) * (size_t)co->failure_capacity;
  HIPCHK(c, c->causal.ensure(rec_bytes + fs_bytes + 256));
  char* base = static_cast<char*>(c->causal.p);
  c->causal_rec = reinterpret_cast<clr_causal_record*>(base);
  c->fail_rec = reinterpret_cast<clr_failure_sample*>(base + rec_bytes);
  c->causal_count = reinterpret_cast<unsigned long long*>(base + rec_bytes + fs_bytes);
  c->fail_count = c->causal_count + 1;
  c->causal_cap = co->capacity;
  c->fail_cap = co->failure_capacity;
  HIPCHK(c, hipMemsetAsync(c->causal_count, 0, 2 * sizeof(unsigned long long), c->stream));
  LightDiag D{};
  int e = 0;
  if (lc) {
    const clr_config& g = c->slots[0]->cfg;
    e = light_setup(c, lc, (float)g.rin, (float)g.zmin, D);
    if (!e)
      e = run_single(c, n, seed, offset, launch, out, nullptr, 0, CLR_KERNEL_LIGHT, c->light.desc.as<LightDiag>(),
                     co->range_start, co->range_count);
    if (!e) e = light_readback(c, lc, D, lo);
  } else {
    e = run_single(c, n, seed, offset, launch, out, nullptr, 0, CLR_KERNEL_CAUSAL, nullptr, co->range_start,
                   co->range_count);
  }
  c->causal_rec = nullptr; c->causal_count = nullptr; c->causal_cap = 0;
  c->fail_rec = nullptr; c->fail_count = nullptr; c->fail_cap = 0;
  if (e) return e;
  unsigned long long cnts[2] = {0, 0};
  HIPCHK(c, hipMemcpy(cnts, base + rec_bytes + fs_bytes, sizeof(cnts), hipMemcpyDeviceToHost));
  const unsigned long long cnt = cnts[0];
  co->count = cnt;
  co->failure_count = cnts[1];
  const uint64_t fw = std::min<uint64_t>(cnts[1], co->failure_capacity);
  if (fw) HIPCHK(c, hipMemcpy(co->failures, base + rec_bytes, sizeof(clr_failure_sample) * fw, hipMemcpyDeviceToHost));
  const uint64_t w = std::min<uint64_t>(cnt, co->capacity);
  if (w) HIPCHK(c, hipMemcpy(co->records, base, sizeof(clr_causal_record) * w, hipMemcpyDeviceToHost));
  return 0;
}

int clr_light_probe(clr_ctx* c, int mode, int n, const float* data, const clr_light_config* lc, double rin,
                    double zmin, clr_light_out* lo) {
  if (!c || !data || !lc || !lo || n <= 0 || mode < 0 || mode > 2) return fail(c, CLR_E_ARG, "null argument");
  if (mode == 2 && lc->heatmap_size < 1) return fail(c, CLR_E_ARG, "heat-map probe needs heatmap_size >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  LightDiag D{};
  if (int e = light_setup(c, lc, (float)rin, (float)zmin, D)) return e;
  const size_t nb = sizeof(float) * (size_t)n * (mode == 1 ? 3 : 7);
  HIPCHK(c, c->light.probe.ensure(nb));
  HIPCHK(c, hipMemcpyAsync(c->light.probe.p, data, nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, clr_launch_light_probe(c->light.desc.as<LightDiag>(), c->light.probe.as<float>(), n, mode, c->stream));
  return light_readback(c, lc, D, lo);
}

int clr_run_batch(clr_ctx* c, const clr_batch_item* bi, int n_items, const clr_launch* launch, clr_tallies* out) {
  if (!c || !bi || !out || n_items <= 0) return fail(c, CLR_E_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<int> used;
  uint64_t total = 0;
  for (int i = 0; i < n_items; ++i) {
    const int k = bi[i].config;
    if (k < 0 || k >= (int)c->slots.size() || !c->slots[k]->ready) return fail(c, CLR_E_STATE, "batch item refers to an unprepared slot");
    if (bi[i].n_histories == 0 || bi[i].n_histories > (1ull << 31)) return fail(c, CLR_E_RANGE, "batch item size must be 1..2^31");
    if (std::find(used.begin(), used.end(), k) != used.end()) return fail(c, CLR_E_ARG, "each slot may appear once per batch");
    used.push_back(k);
    total += bi[i].n_histories;
  }
  const RunShape sh = pick_shape(c, launch, total, kernel_kind(c, used));
  // caller's order, or most expensive configurations first (longest
  // processing time first) on request
  std::vector<int> order(n_items);
  std::vector<double> cost(n_items);
  double work = 0.0;
  for (int i = 0; i < n_items; ++i) {
    order[i] = i;
    cost[i] = history_cost(*c->slots[bi[i].config]);
    work += cost[i] * (double)bi[i].n_histories;
  }
  if (launch && launch->queue_order == 1)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  QueueBuilder q(sh, work);
  for (int i : order)
    q.add(bi[i].config, 0, bi[i].n_histories, bi[i].n_histories, bi[i].seed, bi[i].history_offset, cost[i]);
  if (int e = launch_items(c, q.items, sh, true, nullptr, 0)) return e;
  const int n_slots = (int)c->slots.size();
  std::vector<clr_tallies> host(n_slots);
  HIPCHK(c, hipMemcpy(host.data(), c->tallies.p, sizeof(clr_tallies) * n_slots, hipMemcpyDeviceToHost));
  for (int i = 0; i < n_items; ++i) {
    const int k = bi[i].config;
    out[i] = host[k];
    out[i].n_groups = c->slots[k]->cfg.n_groups;
    group_sizes(bi[i].n_histories, out[i].n_groups, out[i].group_histories);
    ideal_photodiode(c->slots[k]->cfg, out[i]);
    out[i].counters[CLR_C_HISTORIES] = bi[i].n_histories;
  }
  return 0;
}

int clr_time_run(clr_ctx* c, uint64_t n, uint64_t seed, const clr_launch* launch, int repeats, double* ms) {
  if (!c || !ms || repeats < 1) return fail(c, CLR_E_ARG, "null argument");
  clr_tallies t;
  double sum = 0.0;
  for (int r = 0; r < repeats; ++r) {
    if (int e = clr_run(c, n, seed + (uint64_t)r, 0, launch, &t, nullptr, 0)) return e;
    sum += c->last_ms;
  }
  *ms = sum / repeats;
  return 0;
}

double clr_last_kernel_ms(clr_ctx* c) { return c ? c->last_ms : 0.0; }

int clr_meridional_solve(clr_ctx* c, const clr_config* g, const clr_particles* p, const clr_sediment* sed,
                         const clr_meridional* mp, clr_meridional_out* out) {
  if (!c) return CLR_E_ARG;
  if (!g || !p || !sed || !mp || !out) return fail(c, CLR_E_ARG, "meridional: null argument");
  if (p->n <= 0 || p->n > kMaxBins || !p->diameter) return fail(c, CLR_E_ARG, "particles: n must be 1..4096");
  // CLARITAS.py:5873-5887
  if (!(mp->circulation_ratio >= 0.0 && mp->circulation_ratio <= 0.5))
    return fail(c, CLR_E_ARG, "meridional_circulation_ratio must be finite and in [0,0.5]");
  if (mp->nr < 16 || mp->nz < 32) return fail(c, CLR_E_ARG, "meridional solver grid must be at least 16 x 32");
  if (mp->nr > 4096 || mp->nz > 4096 || (int64_t)mp->nr * mp->nz * p->n > (int64_t)1 << 28)
    return fail(c, CLR_E_RANGE, "meridional grid too large");
  if (mp->max_iterations < 250 || mp->check_interval < 10)
    return fail(c, CLR_E_ARG, "invalid meridional solver iteration controls");
  if (!(mp->tolerance > 0.0 && mp->tolerance < 1e-2))
    return fail(c, CLR_E_ARG, "meridional_solver_tolerance must be in (0,1e-2)");
  if (!(mp->sor >= 1.0 && mp->sor < 1.95)) return fail(c, CLR_E_ARG, "meridional_solver_sor must be in [1,1.95)");
  const int fsm = g->free_surface_model;
  const double R = g->rin, D = g->vortex_delta_h, a = g->vortex_core_radius;
  // CLARITAS.py:5911-5921
  if ((fsm != CLR_SURFACE_PARABOLIC && fsm != CLR_SURFACE_SCULLY) || !(D > 0.0) ||
      (fsm == CLR_SURFACE_SCULLY && !(a > 0.0)))
    return fail(c, CLR_E_ARG, "meridional transport requires a non-zero supported vortex");
  if (!(R > 0.0) || !(g->zmean > g->zmin)) return fail(c, CLR_E_ARG, "geometry: rin > 0 and zmean > zmin");
  if (!(sed->particle_density > sed->water_density) || !(sed->water_density > 0.0) || !(sed->water_viscosity > 0.0) ||
      !(sed->gravity > 0.0) || !(sed->eddy_diffusivity > 0.0) || sed->n_quad < 64 || sed->n_radial < 65)
    return fail(c, 31, "sediment precompute failed with error code 31");
  // vortex_surface_geometry (CLARITAS.py:708-792): wall and centre heights,
  // peak azimuthal speed
  double wall_z, center_z, vtheta;
  const double R2 = R * R, grav = sed->gravity;
  if (fsm == CLR_SURFACE_PARABOLIC) {
    wall_z = g->zmean + 0.5 * D;
    center_z = g->zmean - 0.5 * D;
    vtheta = std::sqrt(2.0 * grav * D / R2) * R;
  } else {
    const double a2 = a * a, h_inf = D * (a2 + R2) / R2, h_bar = h_inf * (1.0 - (a2 / R2) * std::log1p(R2 / a2));
    wall_z = g->zmean - h_bar + D;
    center_z = g->zmean - h_bar;
    vtheta = 0.5 * std::sqrt(2.0 * grav * h_inf / a2) * a;
  }
  if (!(center_z > g->zmin))
    return fail(c, CLR_E_ARG, "vortex depression reaches or passes the acrylic bottom");
  if (!(wall_z < g->zwall_top)) return fail(c, CLR_E_ARG, "free surface reaches or exceeds the physical acrylic wall top");
  out->grid_zmax = wall_z;
  out->vtheta_peak = vtheta;
  out->peak_speed = mp->circulation_ratio * vtheta;
  out->iterations = 0;
  HIPCHK(c, hipSetDevice(c->device));

  const int n = p->n, NR = mp->nr, NZ = mp->nz, nrad = sed->n_radial;
  const size_t plane = (size_t)NR * NZ, total = plane * (size_t)n;
  struct Bufs {
    DevBuf d, A, B, norm, ws, radlog, radwr, active, grid, coeff, res, stats, err;
    ~Bufs() {
      DevBuf* all[] = {&d, &A, &B, &norm, &ws, &radlog, &radwr, &active, &grid, &coeff, &res, &stats, &err};
      for (DevBuf* b : all) b->release();
    }
  } b;
  HIPCHK(c, b.d.ensure(sizeof(double) * n));
  DevBuf* per_bin[] = {&b.A, &b.B, &b.norm, &b.ws, &b.res};
  for (DevBuf* x : per_bin) HIPCHK(c, x->ensure(sizeof(double) * n));
  HIPCHK(c, b.radlog.ensure(sizeof(double) * (size_t)n * nrad));
  HIPCHK(c, b.radwr.ensure(sizeof(double) * (size_t)n * nrad));
  HIPCHK(c, b.active.ensure(plane));
  HIPCHK(c, b.grid.ensure(sizeof(double) * total));
  HIPCHK(c, b.coeff.ensure(sizeof(double) * 5 * total));
  HIPCHK(c, b.stats.ensure(sizeof(double) * 6 * n));
  HIPCHK(c, b.err.ensure(sizeof(int) * 2));
  HIPCHK(c, hipMemcpyAsync(b.d.p, p->diameter, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(b.err.p, 0, sizeof(int) * 2, c->stream));

  // per-bin finite-Re coefficients: the reference runs
  // sediment_transport_precompute_kernel with model 3 (CLARITAS.py:6399-6443)
  PrepLaunch P{};
  P.n = n; P.sed_model = CLR_SED_GRID_PER_BIN; P.fs_model = fsm; P.n_quad = sed->n_quad; P.n_radial = nrad;
  P.particle_density = sed->particle_density; P.water_density = sed->water_density;
  P.viscosity = sed->water_viscosity; P.gravity = grav; P.eddy_diffusivity = sed->eddy_diffusivity;
  P.R = R; P.zmin = g->zmin; P.zmean = g->zmean; P.dh = D; P.core = a;
  P.diameter = b.d.as<double>();
  P.sedA = b.A.as<double>(); P.sedB = b.B.as<double>(); P.sednorm = b.norm.as<double>(); P.sedws = b.ws.as<double>();
  P.radlog = b.radlog.as<double>(); P.radwr = b.radwr.as<double>(); P.error_flag = b.err.as<int>();
  HIPCHK(c, clr_launch_sediment_bins(P, c->stream));
  int ecode[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(ecode, b.err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ecode[0]) return fail(c, ecode[0], "sediment precompute failed with error code " + std::to_string(ecode[0]));

  MeridionalLaunch L{};
  L.n = n; L.nr = NR; L.nz = NZ; L.fs_model = fsm; L.n_radial = nrad;
  L.R = R; L.zmin = g->zmin; L.zgrid = wall_z; L.zmean = g->zmean; L.dh = D; L.core = a;
  L.upeak = out->peak_speed; L.eddy = sed->eddy_diffusivity; L.omega = mp->sor;
  L.target = 0.5 * R2 * (g->zmean - g->zmin);
  L.sedB = b.B.as<double>(); L.sednorm = b.norm.as<double>(); L.ws = b.ws.as<double>();
  L.radlog = b.radlog.as<double>(); L.radwr = b.radwr.as<double>();
  L.active = b.active.as<unsigned char>(); L.grid = b.grid.as<double>(); L.coeff = b.coeff.as<double>();
  L.residual = b.res.as<double>(); L.stats = b.stats.as<double>(); L.error_flag = b.err.as<int>();
  clr_meridional_shape(L);
  HIPCHK(c, clr_meridional_setup(L, c->stream));

  // driver loop, CLARITAS.py:6568-6700: one launch per check interval
  std::vector<double> res((size_t)n, 0.0);
  const int maxit = mp->max_iterations, check = mp->check_interval, min_it = std::min(1000, maxit);
  int it = 0, iterations = 0;
  double maxres = 0.0;
  int rc = 0;
  while (it < maxit) {
    const int next = std::min((it / check + 1) * check, maxit);
    HIPCHK(c, clr_meridional_sweep(L, it, next, it == 0, c->stream));
    HIPCHK(c, hipMemcpyAsync(res.data(), b.res.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ecode, b.err.p, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    it = next;
    if (ecode[0]) {
      rc = fail(c, ecode[0], std::string(ecode[1] == 0 ? "meridional initial normalization" : "meridional normalization") +
                                 " failed with error code " + std::to_string(ecode[0]));
      break;
    }
    maxres = *std::max_element(res.begin(), res.end());
    if (it >= min_it && maxres <= mp->tolerance) { iterations = it; break; }
  }
  if (iterations == 0) iterations = it;
  out->iterations = iterations;
  if (!rc) {
    HIPCHK(c, clr_meridional_finish(L, c->stream));
    if (!(maxres <= mp->tolerance)) {
      char msg[160];
      std::snprintf(msg, sizeof msg, "meridional transport solver did not converge: max residual=%.6g after %d iterations",
                    maxres, iterations);
      rc = fail(c, CLR_E_NOT_CONVERGED, msg);
    }
  }
  if (out->grid) HIPCHK(c, hipMemcpyAsync(out->grid, b.grid.p, sizeof(double) * total, hipMemcpyDeviceToHost, c->stream));
  if (out->coeff)
    HIPCHK(c, hipMemcpyAsync(out->coeff, b.coeff.p, sizeof(double) * 5 * total, hipMemcpyDeviceToHost, c->stream));
  if (out->active) HIPCHK(c, hipMemcpyAsync(out->active, b.active.p, plane, hipMemcpyDeviceToHost, c->stream));
  std::vector<double> st(6 * (size_t)n, 0.0);
  HIPCHK(c, hipMemcpyAsync(st.data(), b.stats.p, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out->residual) std::memcpy(out->residual, res.data(), sizeof(double) * n);
  for (int q = 0; q < n; ++q) {
    if (out->scale_max) out->scale_max[q] = st[6 * q];
    if (out->scale_min) out->scale_min[q] = st[6 * q + 1];
    if (out->point_scales)
      for (int j = 0; j < 4; ++j) out->point_scales[4 * q + j] = st[6 * q + 2 + j];
  }
  return rc;
}

}  // extern "C"

// ---- surrogate entry points (ensemble kernels in mlp.hip) ----
extern "C" {
int clr_mlp_init(clr_ctx* c, const clr_mlp_shape* shape, uint64_t seed) {
  if (!c || !shape) return fail(c, CLR_E_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  clr_mlp_free(c->mlp);
  c->mlp = nullptr;
  return mlp_init(&c->mlp, shape, seed, c->stream, c->err);
}
size_t clr_mlp_param_count(clr_ctx* c) { return (c && c->mlp) ? mlp_param_count(c->mlp) : 0; }
int clr_mlp_get_params(clr_ctx* c, float* out, size_t n) {
  if (!c || !c->mlp || !out) return fail(c, CLR_E_STATE, "no ensemble");
  HIPCHK(c, hipSetDevice(c->device));
  return mlp_get_params(c->mlp, out, n, c->stream, c->err);
}
int clr_mlp_set_params(clr_ctx* c, const float* in, size_t n) {
  if (!c || !c->mlp || !in) return fail(c, CLR_E_STATE, "no ensemble");
  HIPCHK(c, hipSetDevice(c->device));
  return mlp_set_params(c->mlp, in, n, c->stream, c->err);
}
int clr_mlp_train(clr_ctx* c, const float* x, const float* y, const float* w, int64_t n, int32_t batch,
                  int32_t epochs, float lr, uint64_t seed, float* loss) {
  if (!c || !c->mlp || !x || !y) return fail(c, CLR_E_STATE, "no ensemble");
  HIPCHK(c, hipSetDevice(c->device));
  return mlp_train(c->mlp, x, y, w, n, batch, epochs, lr, seed, loss, c->stream, c->err);
}
int clr_mlp_grad(clr_ctx* c, const float* x, const float* y, const float* w, int32_t n, float* loss, float* grads) {
  if (!c || !c->mlp || !x || !y) return fail(c, CLR_E_STATE, "no ensemble");
  HIPCHK(c, hipSetDevice(c->device));
  return mlp_grad(c->mlp, x, y, w, n, loss, grads, c->stream, c->err);
}
int clr_mlp_infer(clr_ctx* c, const float* x, int64_t n, float* mean, float* stdev) {
  if (!c || !c->mlp || !x || !mean) return fail(c, CLR_E_STATE, "no ensemble");
  HIPCHK(c, hipSetDevice(c->device));
  return mlp_infer(c->mlp, x, n, mean, stdev, c->stream, c->err);
}
int clr_mlp_infer_static(clr_ctx* c, const float* x, int64_t n, float* mean, float* stdev) {
  if (!c || !c->mlp || !x || !mean) return fail(c, CLR_E_STATE, "no ensemble");
  if (!c->static_scratch) return fail(c, CLR_E_STATE, "the static staging array is not set up on this device");
  HIPCHK(c, hipSetDevice(c->device));
  return mlp_infer_static(c->mlp, c->static_scratch, x, n, mean, stdev, c->stream, c->err);
}
int clr_mlp_time_infer(clr_ctx* c, int64_t n, int repeats, double* ms) {
  if (!c || !c->mlp || !ms) return fail(c, CLR_E_STATE, "no ensemble");
  HIPCHK(c, hipSetDevice(c->device));
  return mlp_time_infer(c->mlp, n, repeats, ms, c->stream, c->err);
}
}  // extern "C"
