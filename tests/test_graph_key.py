"""The captured-graph cache key of clr_run_batch (csrc/graph_key.h), checked
on the host: a replay is reused only when every captured kernel argument
is unchanged, so a work-item buffer reallocated by an intervening run (or
any other buffer, count or launch dimension) forces a new capture."""
import shutil
import subprocess
from pathlib import Path

import pytest

HEADER_DIR = Path(__file__).resolve().parents[1] / 'tardiis_and_claritas_amd' / 'csrc'

SRC = r'''
#include <cstdio>
#include <cstdint>
#include "graph_key.h"
struct Launch {  // the fields of TransportLaunch, host-only stand-ins
  const void* params; const void* items; int n_items; unsigned* item_counter; void* tallies;
  void* records; uint64_t n_records; int grid, block, kind; const void* light;
  void* causal; unsigned long long* causal_count; uint64_t causal_cap;
  void* fail_rec; unsigned long long* fail_count; uint64_t fail_cap;
};
static char a[64];
int main() {
  Launch L{a + 0, a + 1, 7, (unsigned*)(a + 8), a + 3, nullptr, 0, 1024, 256, 0, nullptr,
           nullptr, nullptr, 0, nullptr, nullptr, 0};
  const GraphKey k0 = graph_key(L, 3);
  int bad = 0;
  if (!same_graph_key(k0, graph_key(L, 3))) bad |= 1;        // identical arguments reuse
  if (same_graph_key(k0, GraphKey{})) bad |= 2;               // an empty cache never matches
  if (same_graph_key(k0, graph_key(L, 4))) bad |= 4;          // slot count
  int fails = 0, cases = 0;
#define CHANGE(stmt) do { Launch M = L; stmt; ++cases; if (same_graph_key(k0, graph_key(M, 3))) ++fails; } while (0)
  CHANGE(M.params = a + 20);
  CHANGE(M.items = a + 21);            // work items reallocated by another run
  CHANGE(M.n_items = 8);
  CHANGE(M.item_counter = (unsigned*)(a + 24));
  CHANGE(M.tallies = a + 22);
  CHANGE(M.records = a + 23);
  CHANGE(M.n_records = 5);
  CHANGE(M.grid = 512);
  CHANGE(M.block = 128);
  CHANGE(M.kind = 1);
  CHANGE(M.light = a + 30);
  CHANGE(M.causal = a + 31);
  CHANGE(M.causal_count = (unsigned long long*)(a + 32));
  CHANGE(M.causal_cap = 9);
  CHANGE(M.fail_rec = a + 33);
  CHANGE(M.fail_count = (unsigned long long*)(a + 40));
  CHANGE(M.fail_cap = 2);
  std::printf("%d %d %d\n", bad, fails, cases);
  return 0;
}
'''


def test_graph_key_invalidates_on_every_captured_argument(tmp_path):
    gxx = shutil.which('g++')
    if gxx is None:
        pytest.skip('g++ not available')
    src = tmp_path / 'key.cpp'
    src.write_text(SRC)
    exe = tmp_path / 'key'
    subprocess.run([gxx, '-std=c++17', '-O1', '-I', str(HEADER_DIR), str(src), '-o', str(exe)], check=True)
    bad, fails, cases = map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                                check=True).stdout.split())
    assert bad == 0 and fails == 0 and cases == 17
