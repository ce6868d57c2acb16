// Event-pair kernel timer behind AFX_LAUNCH (see afx_profile.h).
#include "afx_common.h"
#include "afx_profile.h"
#include "aquaflux_amd.h"

#include <atomic>
#include <mutex>
#include <vector>

namespace afx {

const char* const kKernelNames[kKNumKernels] = {
    "geom_face",     "geom_approx",   "geom_orient",     "geom_volume",  "flow_grad",
    "flow_ap",       "flow_res",      "flow_res_adj",    "flow_ap_adj",  "flow_grad_adj",
    "scalar_grad",   "scalar_res",    "scalar_res_adj",  "scalar_grad_adj", "geom_adj_cell",
    "rans_closure",  "rans_flow_res", "rans_scalar_res_adj", "rans_closure_adj", "bcsr_assemble",
    "bcsr_spmv",     "pbj_setup",     "pbj_apply",       "blas",
    "rans_grad",     "rans_ap",       "rans_ap_adj",     "rans_grad_adj",   "scalar_lim_adj",
    "geom_adj_face", "geom_adj_node",   "rans_scalar_res", "rans_flow_res_adj",
    "limiter",       "cgg_sweep",     "cgg_sweep_adj",   "probe_scatter",
    "amg",
    "krylov",
    "face_flux",
};

namespace {
struct Rec {
  int id;
  hipEvent_t start, stop;
};
std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<hipEvent_t> g_pool;  // free events
std::vector<Rec> g_recs;         // recorded pairs, in launch order
size_t g_cap = 0;
}  // namespace

bool profile_begin(int id, hipStream_t s, hipEvent_t* stop) {
  if (!g_on.load(std::memory_order_relaxed)) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_pool.size() < 2) return false;
  hipEvent_t a = g_pool.back();
  g_pool.pop_back();
  hipEvent_t b = g_pool.back();
  g_pool.pop_back();
  if (hipEventRecord(a, s) != hipSuccess) {
    g_pool.push_back(a);
    g_pool.push_back(b);
    return false;
  }
  g_recs.push_back(Rec{id, a, b});
  *stop = b;
  return true;
}

void profile_end(hipEvent_t stop, hipStream_t s) { (void)hipEventRecord(stop, s); }

}  // namespace afx

using namespace afx;

// Enables timing with a pool of `max_launches` event pairs (0 disables and frees the pool).
extern "C" int afx_profile_enable(int64_t max_launches) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on.store(false);
  for (auto& r : g_recs) {
    g_pool.push_back(r.start);
    g_pool.push_back(r.stop);
  }
  g_recs.clear();
  if (max_launches <= 0) {
    for (auto e : g_pool) (void)hipEventDestroy(e);
    g_pool.clear();
    g_cap = 0;
    return 0;
  }
  while (g_cap < (size_t)max_launches * 2) {
    hipEvent_t e;
    AFX_CHECK_HIP(hipEventCreate(&e));
    g_pool.push_back(e);
    ++g_cap;
  }
  g_on.store(true);
  return 0;
}

// After the recorded work has completed: total milliseconds and launch count per kernel id
// (arrays of length afx_profile_kernel_count()); clears the recorded pairs.
extern "C" int afx_profile_collect(double* total_ms, int64_t* count) {
  AFX_REQUIRE(total_ms && count, "afx_profile_collect: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  for (int i = 0; i < kKNumKernels; ++i) {
    total_ms[i] = 0.0;
    count[i] = 0;
  }
  for (auto& r : g_recs) {
    float ms = 0.0f;
    AFX_CHECK_HIP(hipEventSynchronize(r.stop));
    AFX_CHECK_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    total_ms[r.id] += ms;
    count[r.id] += 1;
    g_pool.push_back(r.start);
    g_pool.push_back(r.stop);
  }
  g_recs.clear();
  return 0;
}

extern "C" int afx_profile_kernel_count(void) { return kKNumKernels; }

extern "C" const char* afx_profile_kernel_name(int id) {
  return (id >= 0 && id < kKNumKernels) ? kKernelNames[id] : "";
}

namespace afx {
dim3 cell_grid(int64_t n, int resident_per_cu) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      cus = v;
    if (cus <= 0) cus = 256;
  }
  const int64_t nb = (n + kBlock - 1) / kBlock;
  if (resident_per_cu <= 0) return dim3((unsigned)nb);
  int64_t g = (int64_t)cus * resident_per_cu;
  g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
  return dim3((unsigned)(nb < g ? nb : g));
}
}  // namespace afx
