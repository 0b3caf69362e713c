// Host-side helpers shared by the instance types of the library: device
// allocation bookkeeping and the preparation of the grid tables from the
// caller's descriptor.
#pragma once
#include <math.h>
#include <stdio.h>

#include <map>
#include <string>
#include <vector>

#include "fimhip.h"
#include "fimhip_internal.h"

#define HIPCHK(x)                                                    \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) {                                          \
      fprintf(stderr, "fimhip: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return FIMHIP_EHIP;                                            \
    }                                                                \
  } while (0)

namespace fimhip {

// Device allocations of one instance, zeroed on its stream and freed together.
// Each array starts at a different offset (a multiple of 4352 bytes below
// 512 KiB) from its 2 MiB-aligned allocation, so that kernels streaming many
// arrays at the same index do not send every stream to the same memory
// channel at the same time.
struct Arena {
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  int raw(void **p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    const size_t pad = (allocs.size() * 4352) % (512 * 1024);
    void *base = nullptr;
    if (hipMalloc(&base, bytes + pad) != hipSuccess) return FIMHIP_ENOMEM;
    allocs.push_back(base);
    *p = (char *)base + pad;
    if (hipMemsetAsync(*p, 0, bytes, stream) != hipSuccess) return FIMHIP_EHIP;
    return 0;
  }
  template <class T>
  int alloc(T **p, size_t n) { return raw((void **)p, n * sizeof(T)); }
  template <class T>
  int upload(T **p, const std::vector<T> &h) {
    int r = alloc(p, h.size());
    if (r) return r;
    if (hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream) != hipSuccess)
      return FIMHIP_EHIP;
    return 0;
  }
  void release() {
    for (void *p : allocs) (void)hipFree(p);
    allocs.clear();
  }
};

// Per-kernel launch timing with HIP events on the launching stream: each
// bracketed launch records a pair of events, and flush() adds the elapsed
// times to per-name totals.
struct KTimer {
  struct Rec { const char *name; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long>> acc;
  hipEvent_t get() {
    hipEvent_t e = nullptr;
    if (!pool.empty()) { e = pool.back(); pool.pop_back(); }
    else if (hipEventCreate(&e) != hipSuccess) e = nullptr;
    return e;
  }
  int flush(hipStream_t st) {
    if (hipStreamSynchronize(st) != hipSuccess) return FIMHIP_EHIP;
    for (auto &r : recs) {
      float ms = 0.f;
      if (r.a && r.b && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        auto &v = acc[r.name];
        v.first += ms;
        v.second += 1;
      }
      if (r.a) pool.push_back(r.a);
      if (r.b) pool.push_back(r.b);
    }
    recs.clear();
    return 0;
  }
  ~KTimer() {
    for (auto &r : recs) { if (r.a) (void)hipEventDestroy(r.a); if (r.b) (void)hipEventDestroy(r.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// Grid tables derived on the host from the caller's descriptor:
// 0-based neighbour tables with slot 6 of a pentagon equal to slot 5
// (dyn_init.F90:546-548), corio/rarea/grndspd (dyn_init.F90:753-757),
// dfflen and smoo_coeff (dyn_init.F90:309-321).
struct HostGrid {
  std::vector<int> prox, proxs, nprox;
  std::vector<double> rarea, corio, grndspd, dfflen, smoo;
};

int host_grid_prepare(const fimhip_grid_desc *gd, HostGrid &h);
// upload the descriptor's metrics and the derived tables into g
int dev_grid_upload(Arena &a, const fimhip_grid_desc *gd, const HostGrid &h, DevGrid &g);

}  // namespace fimhip
