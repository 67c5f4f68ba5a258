// graph_key.h -- what a captured transport graph depends on.
//
// A hipGraph captured from clear_kernel + the persistent transport kernel
// holds every kernel argument by value: the device addresses of the
// parameter array, work items, item counter, tallies, per-history records,
// diagnostics and causal buffers, and the launch shape.  A replay is valid
// only while all of them are unchanged; any buffer that was reallocated since
// the capture (a non-graph run that needed more work items, a grown slot
// array) invalidates it.  GraphKey collects those values so the cache can
// compare the whole set.  Header-only and free of HIP so that the
// comparison can be checked on the host.
#pragma once
#include <stdint.h>

struct GraphKey {
  uintptr_t ptr[10];
  uint64_t len[3];
  int32_t dims[6];
  bool valid;
};

// Launch is any struct with the fields of TransportLaunch (transport.h).
template <class Launch>
inline GraphKey graph_key(const Launch& L, int n_slots) {
  GraphKey k{};
  k.ptr[0] = (uintptr_t)L.params;
  k.ptr[1] = (uintptr_t)L.items;
  k.ptr[2] = (uintptr_t)L.item_counter;
  k.ptr[3] = (uintptr_t)L.tallies;
  k.ptr[4] = (uintptr_t)L.records;
  k.ptr[5] = (uintptr_t)L.light;
  k.ptr[6] = (uintptr_t)L.causal;
  k.ptr[7] = (uintptr_t)L.causal_count;
  k.ptr[8] = (uintptr_t)L.fail_rec;
  k.ptr[9] = (uintptr_t)L.fail_count;
  k.len[0] = (uint64_t)L.n_records;
  k.len[1] = (uint64_t)L.causal_cap;
  k.len[2] = (uint64_t)L.fail_cap;
  k.dims[0] = L.n_items;
  k.dims[1] = L.grid;
  k.dims[2] = L.block;
  k.dims[3] = L.kind;
  k.dims[4] = n_slots;
  k.dims[5] = 0;
  k.valid = true;
  return k;
}

inline bool same_graph_key(const GraphKey& a, const GraphKey& b) {
  if (!a.valid || !b.valid) return false;
  for (int i = 0; i < 10; ++i)
    if (a.ptr[i] != b.ptr[i]) return false;
  for (int i = 0; i < 3; ++i)
    if (a.len[i] != b.len[i]) return false;
  for (int i = 0; i < 6; ++i)
    if (a.dims[i] != b.dims[i]) return false;
  return true;
}
