// causal_bin.cpp -- single-pass host binning of causal provenance records into
// the per-channel score tables of the detector-path / causal / H170
// diagnostics (CLARITAS.py:8017-8480).  Each table is K keys x 18 channels of
// FP64; a record adds its weight at (key, native channel) and (key, mirror
// channel), 1/2 each with the symmetry average, or at the native channel
// alone.  The table names, their order and K are those of clr_causal_layout.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/claritas.h"

#pragma clang fp contract(off)

namespace {

constexpr int kCh = CLR_N_CHANNELS;

const char* const kPathKeys[25] = {
    "any_particle", "any_geometric", "any_wave", "any_coherent_wave", "any_morphology_wave",
    "any_entry_reflection", "any_internal_reflection", "ballistic", "single_interaction",
    "single_entry_reflection", "multiple_interaction", "interaction_order_2", "interaction_order_3",
    "interaction_order_4plus", "exactly_one_geometric", "exactly_one_wave", "single_coherent_wave_escape",
    "single_external_entry_reflection_escape", "single_internal_reflection_escape",
    "single_surface_transmission_escape", "mixed_geometric_wave", "geometric_only_multiple",
    "wave_only_multiple", "multiple_with_entry_reflection", "multiple_with_internal_reflection"};
const char* const kSums[7] = {"interactions", "geometric_interactions", "wave_interactions",
                              "coherent_wave_interactions", "morphology_wave_interactions", "entry_reflections",
                              "internal_reflections"};
const char* const kBins[6] = {"first_bin", "last_bin", "first_geometric_bin", "last_geometric_bin",
                              "first_wave_bin", "last_wave_bin"};
const char* const kMetrics[23] = {
    "cumulative_particle_deflection_deg", "cumulative_particle_deflection_sq_deg2", "max_particle_deflection_deg",
    "first_particle_r_m", "first_particle_z_m", "last_particle_r_m", "last_particle_z_m",
    "particle_event_mean_r_m", "particle_event_mean_z_m", "particle_event_min_r_m", "particle_event_max_r_m",
    "particle_event_min_z_m", "particle_event_max_z_m", "water_path_at_first_particle_event_m",
    "water_path_at_last_particle_event_m", "particle_water_path_span_first_to_last_m", "particle_radial_span_m",
    "particle_axial_span_m", "particle_branch_transition_count", "particle_event_type_transition_count",
    "particle_small_deflection_count", "particle_medium_deflection_count", "particle_large_deflection_count"};

const double kCumEdges[14] = {0.0, 30.0, 60.0, 90.0, 120.0, 180.0, 240.0, 360.0, 540.0, 720.0, 1080.0,
                              1800.0, 3600.0, 1.0e9};
const double kLdEdges[13] = {0.0, 1.0e-12, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0, 10.0, 1.0e9};
const double kIncEdges[7] = {0.0, 15.0, 30.0, 45.0, 60.0, 75.0, 90.0001};
const double kAngEdges[7] = {0.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0001};

// b with edges[b] <= v < edges[b + 1] (v finite), else -1
int bin_index(double v, const double* e, int ne) {
  if (!std::isfinite(v)) return -1;
  int b = 0;
  while (b < ne && e[b] <= v) ++b;
  b -= 1;
  return (b >= 0 && b < ne - 1) ? b : -1;
}

struct Layout {
  std::vector<std::string> names;
  std::vector<int> K, off;
  int total = 0;
  int add(const std::string& n, int k) {
    names.push_back(n); K.push_back(k); off.push_back(total);
    total += k * kCh;
    return (int)names.size() - 1;
  }
};

// table ids, in layout order
struct Ids {
  int path[25], hardware, sum[7], epo, ego, bin[6], wall, after_ret, acr, last_ev, first_ev, last_ang_sum,
      last_ang_cnt, fback, fhemi, fas, fac, las, lac, ips, ios, pls, los, pcnt, msum[23], mcnt[23], defl, branch,
      btrans, etrans, h_cum, h_first, h_last, h_max, mtype, morder, mbin, mmacro, h_tmacro, mmicro, ld_sum,
      macro_sum, micro_sum, h_ld, inc_sum, inc_def, exit_sum, exit_def, h_inc, h_exit, h_tinc, h_texit, h_pinc,
      h_pexit, h_tpsd, h_last_all, h_last_type, h_first_ang;
};

Layout make_layout(int nb, Ids& I) {
  Layout L;
  for (int k = 0; k < 25; ++k) I.path[k] = L.add(std::string("path:") + kPathKeys[k], 1);
  I.hardware = L.add("hardware", 1);
  for (int k = 0; k < 7; ++k) I.sum[k] = L.add(std::string("sum:") + kSums[k], 1);
  I.epo = L.add("exact_particle_order", 6);
  I.ego = L.add("exact_geometric_order", 6);
  for (int k = 0; k < 6; ++k) I.bin[k] = L.add(std::string("bin:") + kBins[k], nb);
  I.wall = L.add("wall_path", 3);
  I.after_ret = L.add("particle_after_acrylic_return", 1);
  I.acr = L.add("acrylic_reflection_order", 5);
  I.last_ev = L.add("last_event", 8);
  I.first_ev = L.add("first_event", 8);
  I.last_ang_sum = L.add("last_event_angle_sum", 8);
  I.last_ang_cnt = L.add("last_event_angle_count", 8);
  I.fback = L.add("first_backscatter", 1);
  I.fhemi = L.add("first_source_hemisphere", 1);
  I.fas = L.add("first_angle_sum", 1);
  I.fac = L.add("first_angle_count", 1);
  I.las = L.add("last_angle_sum", 1);
  I.lac = L.add("last_angle_count", 1);
  I.ips = L.add("internal_path_sum", 1);
  I.ios = L.add("internal_opl_sum", 1);
  I.pls = L.add("post_last_water_path_sum", 1);
  I.los = L.add("last_order_sum", 1);
  I.pcnt = L.add("path_count", 1);
  for (int k = 0; k < 23; ++k) {
    I.msum[k] = L.add(std::string("metric_sum:") + kMetrics[k], 1);
    I.mcnt[k] = L.add(std::string("metric_count:") + kMetrics[k], 1);
  }
  I.defl = L.add("deflection_history", 4);
  I.branch = L.add("branch_sequence", 7);
  I.btrans = L.add("branch_transition", 6);
  I.etrans = L.add("event_transition", 6);
  I.h_cum = L.add("h170_cumulative", 13);
  I.h_first = L.add("h170_first_map", 25);
  I.h_last = L.add("h170_last_map", 25);
  I.h_max = L.add("h170_max_map", 25);
  I.mtype = L.add("max_type", 8);
  I.morder = L.add("max_order", 5);
  I.mbin = L.add("max_bin", nb);
  I.mmacro = L.add("max_macro", 5);
  I.h_tmacro = L.add("h170_type_by_macro", 40);
  I.mmicro = L.add("max_micro", 6);
  I.ld_sum = L.add("max_ld_sum", 1);
  I.macro_sum = L.add("max_macro_sum", 1);
  I.micro_sum = L.add("max_micro_sum", 1);
  I.h_ld = L.add("h170_ld", 12);
  I.inc_sum = L.add("incidence_sum", 1);
  I.inc_def = L.add("incidence_defined", 1);
  I.exit_sum = L.add("exit_sum", 1);
  I.exit_def = L.add("exit_defined", 1);
  I.h_inc = L.add("h170_incidence", 7);
  I.h_exit = L.add("h170_exit", 7);
  I.h_tinc = L.add("h170_type_by_incidence", 56);
  I.h_texit = L.add("h170_type_by_exit", 56);
  I.h_pinc = L.add("h170_psd_by_incidence", nb * 7);
  I.h_pexit = L.add("h170_psd_by_exit", nb * 7);
  I.h_tpsd = L.add("h170_type_by_psd", 8 * nb);
  I.h_last_all = L.add("h170_last_angle_all", 6);
  I.h_last_type = L.add("h170_last_angle_type", 48);
  I.h_first_ang = L.add("h170_first_angle", 6);
  return L;
}

}  // namespace

extern "C" {

int clr_causal_layout(int n_bins, int index, char* name, int name_cap, int* K) {
  if (n_bins < 1) return CLR_E_ARG;
  Ids I;
  const Layout L = make_layout(n_bins, I);
  const int n = (int)L.names.size();
  if (index >= 0 && index < n) {
    if (name && name_cap > 0) {
      std::strncpy(name, L.names[index].c_str(), (size_t)name_cap - 1);
      name[name_cap - 1] = '\0';
    }
    if (K) *K = L.K[index];
  }
  return n;
}

int clr_causal_bin(const clr_causal_record* rec, uint64_t n, int n_bins, int symmetric, const double* r_edges,
                   const double* z_edges, double* tables) {
  if ((n && !rec) || !tables || !r_edges || !z_edges || n_bins < 1) return CLR_E_ARG;
  Ids I;
  const Layout L = make_layout(n_bins, I);
  const int nb = n_bins;
  for (uint64_t i = 0; i < n; ++i) {
    const clr_causal_record& q = rec[i];
    const int chs[2] = {q.native_id, q.mirror_id};
    const double scales[2] = {symmetric ? 0.5 : 1.0, 0.5};
    const int n_sides = symmetric ? 2 : 1;
    // keys and weights of this record, added at each valid channel
    struct Add { int t; long key; double w; };
    Add adds[256];
    int na = 0;
    auto put = [&](int t, long key, double w) {
      if (key >= 0 && key < L.K[t]) adds[na++] = Add{t, key, w};
    };
    const long ic = q.ic, gic = q.gic, wic = q.wic, wcoh = q.wcoh, wmorph = wic - wcoh;
    const long per = q.per, pir = q.pir, pabs = q.pabs;
    const bool single = ic == 1, multiple = ic >= 2, surface = gic > 0 || wmorph > 0;
    const bool m[25] = {ic > 0, gic > 0, wic > 0, wcoh > 0, wmorph > 0, per > 0, pir > 0, ic == 0, single,
                        single && per > 0, multiple, ic == 2, ic == 3, ic >= 4,
                        single && gic == 1 && wic == 0, single && gic == 0 && wic == 1, single && wcoh == 1,
                        single && per > 0, single && pir > 0,
                        single && surface && per == 0 && pir == 0 && pabs == 0,
                        multiple && gic > 0 && wic > 0, multiple && gic > 0 && wic == 0,
                        multiple && gic == 0 && wic > 0, multiple && per > 0, multiple && pir > 0};
    for (int k = 0; k < 25; ++k) if (m[k]) put(I.path[k], 0, 1.0);
    put(I.hardware, 0, 1.0);
    const double sv[7] = {(double)ic, (double)gic, (double)wic, (double)wcoh, (double)wmorph, (double)per,
                          (double)pir};
    for (int k = 0; k < 7; ++k) put(I.sum[k], 0, sv[k]);
    put(I.epo, ic < 5 ? ic : 5, 1.0);
    put(I.ego, gic < 5 ? gic : 5, 1.0);
    const int bins[6] = {q.first_bin, q.last_bin, q.first_geometric_bin, q.last_geometric_bin, q.first_wave_bin,
                         q.last_wave_bin};
    for (int k = 0; k < 6; ++k) put(I.bin[k], bins[k], 1.0);
    put(I.wall, ic == 0 ? 0 : (q.acrylic_returns > 0 ? 2 : 1), 1.0);
    if (q.events_after_return > 0) put(I.after_ret, 0, 1.0);
    put(I.acr, q.acrylic_reflections < 4 ? q.acrylic_reflections : 4, 1.0);
    const int ft = q.first_type, lt = q.last_type;
    const double fa = q.first_deflection, la = q.last_deflection;
    const bool valid = lt > 0;
    put(I.last_ev, lt, 1.0);
    put(I.first_ev, ft, 1.0);
    const bool lav_any = std::isfinite(la) && la >= 0.0;
    if (lav_any) { put(I.last_ang_sum, lt, la); put(I.last_ang_cnt, lt, 1.0); }
    if (valid && q.first_backscatter > 0) put(I.fback, 0, 1.0);
    if (valid && q.first_source_hemisphere > 0) put(I.fhemi, 0, 1.0);
    const bool fav = valid && std::isfinite(fa) && fa >= 0.0, lav = valid && lav_any;
    if (fav) { put(I.fas, 0, fa); put(I.fac, 0, 1.0); }
    if (lav) { put(I.las, 0, la); put(I.lac, 0, 1.0); }
    const double ip = q.internal_path, io = q.internal_opl, wpa = q.water_path_after_last;
    if (valid && std::isfinite(ip) && std::isfinite(io) && std::isfinite(wpa) && wpa >= 0.0) {
      put(I.ips, 0, ip); put(I.ios, 0, io); put(I.pls, 0, wpa); put(I.los, 0, (double)q.last_order);
      put(I.pcnt, 0, 1.0);
    }
    // FP32 radii as the reference computes them from the FP32 sites
    const float fr = std::sqrt(std::fmax(q.first_x * q.first_x + q.first_y * q.first_y, 0.0f));
    const float lr = std::sqrt(std::fmax(q.last_x * q.last_x + q.last_y * q.last_y, 0.0f));
    const float mr = std::sqrt(std::fmax(q.turn_x * q.turn_x + q.turn_y * q.turn_y, 0.0f));
    const double mv[23] = {q.cumulative_deflection, q.cumulative_deflection_sq, q.max_deflection, fr, q.first_z, lr,
                           q.last_z, q.mean_r, q.mean_z, q.min_r, q.max_r, q.min_z, q.max_z, q.water_path_first,
                           q.water_path_last, (double)(q.water_path_last - q.water_path_first),
                           (double)q.max_r - (double)q.min_r, (double)q.max_z - (double)q.min_z,
                           (double)q.branch_transitions, (double)q.type_transitions, (double)q.small_deflections,
                           (double)q.medium_deflections, (double)q.large_deflections};
    for (int k = 0; k < 23; ++k)
      if (valid && std::isfinite(mv[k])) { put(I.msum[k], 0, mv[k]); put(I.mcnt[k], 0, 1.0); }
    int dcls = -1;
    if (ic == 0) dcls = 0;
    if (valid && q.medium_deflections == 0 && q.large_deflections == 0) dcls = 1;
    if (valid && q.medium_deflections > 0 && q.large_deflections == 0) dcls = 2;
    if (valid && q.large_deflections > 0) dcls = 3;
    put(I.defl, dcls, 1.0);
    const int fb = (ft >= 1 && ft <= 3) ? 1 : (ft >= 4 ? 2 : 0), lb = (lt >= 1 && lt <= 3) ? 1 : (lt >= 4 ? 2 : 0);
    int bcls = -1;
    if (ic == 0) bcls = 0;
    if (valid && gic > 0 && wic == 0) bcls = 1;
    if (valid && wic > 0 && gic == 0) bcls = 2;
    if (valid && gic > 0 && wic > 0) {
      if (fb == 1 && lb == 1) bcls = 3;
      if (fb == 1 && lb == 2) bcls = 4;
      if (fb == 2 && lb == 1) bcls = 5;
      if (fb == 2 && lb == 2) bcls = 6;
    }
    put(I.branch, bcls, 1.0);
    if (valid) {
      put(I.btrans, q.branch_transitions < 5 ? q.branch_transitions : 5, 1.0);
      put(I.etrans, q.type_transitions < 5 ? q.type_transitions : 5, 1.0);
      put(I.h_cum, bin_index(q.cumulative_deflection, kCumEdges, 14), 1.0);
      auto cell = [&](double r, double z) {
        const int ri = bin_index(r, r_edges, 6), zi = bin_index(z, z_edges, 6);
        return (ri >= 0 && zi >= 0) ? ri * 5 + zi : -1;
      };
      put(I.h_first, cell(fr, q.first_z), 1.0);
      put(I.h_last, cell(lr, q.last_z), 1.0);
      put(I.h_max, cell(mr, q.turn_z), 1.0);
      const int mt = q.max_type, mo = q.max_order, mbn = q.max_bin;
      const int mac = q.max_macro_bounces, mic = q.max_micro_order;
      put(I.mtype, mt, 1.0);
      put(I.morder, mo >= 1 ? (mo < 5 ? mo : 5) - 1 : -1, 1.0);
      put(I.mbin, mbn, 1.0);
      const int mcls = mac < 4 ? mac : 4;
      put(I.mmacro, mcls, 1.0);
      put(I.h_tmacro, mt >= 0 ? mt * 5 + mcls : -1, 1.0);
      put(I.mmicro, mic < 5 ? mic : 5, 1.0);
      const double ld = q.max_path_over_diameter;
      put(I.ld_sum, 0, ld); put(I.macro_sum, 0, (double)mac); put(I.micro_sum, 0, (double)mic);
      put(I.h_ld, bin_index(ld, kLdEdges, 13), 1.0);
      const double inc = q.max_incidence, ext = q.max_exit;
      const bool inc_ok = std::isfinite(inc) && inc >= 0.0 && inc <= 90.0001;
      const bool ext_ok = std::isfinite(ext) && ext >= 0.0 && ext <= 90.0001;
      if (inc_ok) { put(I.inc_sum, 0, inc); put(I.inc_def, 0, 1.0); }
      if (ext_ok) { put(I.exit_sum, 0, ext); put(I.exit_def, 0, 1.0); }
      int icls = inc_ok ? bin_index(inc, kIncEdges, 7) + 1 : 0, ecls = ext_ok ? bin_index(ext, kIncEdges, 7) + 1 : 0;
      if (icls < 0) icls = 0;
      if (ecls < 0) ecls = 0;
      put(I.h_inc, icls, 1.0);
      put(I.h_exit, ecls, 1.0);
      if (mt >= 1) { put(I.h_tinc, mt * 7 + icls, 1.0); put(I.h_texit, mt * 7 + ecls, 1.0); }
      if (mbn >= 0) { put(I.h_pinc, (long)mbn * 7 + icls, 1.0); put(I.h_pexit, (long)mbn * 7 + ecls, 1.0); }
      if (mt >= 1 && mbn >= 0) put(I.h_tpsd, (long)mt * nb + mbn, 1.0);
    }
    const int lbin = bin_index(la, kAngEdges, 7);
    if (lav) {
      put(I.h_last_all, lbin, 1.0);
      if (lbin >= 0 && lt >= 1) put(I.h_last_type, lt * 6 + lbin, 1.0);
    }
    if (fav) put(I.h_first_ang, bin_index(fa, kAngEdges, 7), 1.0);
    for (int s = 0; s < n_sides; ++s) {
      const int ch = chs[s];
      if (ch < 0 || ch >= kCh) continue;
      for (int a = 0; a < na; ++a)
        tables[L.off[adds[a].t] + adds[a].key * kCh + ch] += scales[s] * adds[a].w;
    }
  }
  return 0;
}

}  // extern "C"
