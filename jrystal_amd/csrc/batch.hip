// One optimiser step over a batch of plans of different shapes (grids,
// k-meshes, band and plane-wave counts): every member's fused energy +
// gradient evaluation, then one grouped Adam launch over all members'
// parameters, with the step count kept on the device so that the step can
// be captured once into a hipGraph and replayed.  Members are spread over
// auxiliary streams (default priority) between a fork and a join event, so
// the captured graph holds independent per-member chains that the device
// runs concurrently.
#include <math.h>

#include <vector>

#include "common.h"

namespace jrx {
namespace {

constexpr int kAdamBlock = 256;
constexpr int kAdamPerThread = 4;

struct AdamItem {
  double* p;
  const double* g;
  double* m;
  double* v;
  int64_t n;
};

// Grouped optax.adam update of every item; off[i] is the start of item i in
// the concatenated index space (off[nitems] = total).  The bias
// corrections use the device step count.
__global__ void __launch_bounds__(kAdamBlock)
grouped_adam_kernel(const AdamItem* __restrict__ items,
                    const int64_t* __restrict__ off, int nitems,
                    const int64_t* __restrict__ count, double lr, double b1,
                    double b2, double eps, double eps_root) {
  const double t = (double)count[0];
  const double c1 = 1.0 - pow(b1, t);
  const double c2 = 1.0 - pow(b2, t);
  const int64_t total = off[nitems];
  const int64_t base =
      (int64_t)blockIdx.x * kAdamBlock * kAdamPerThread + threadIdx.x;
  for (int r = 0; r < kAdamPerThread; ++r) {
    const int64_t i = base + (int64_t)r * kAdamBlock;
    if (i >= total) return;
    int lo = 0, hi = nitems - 1;  // last item with off <= i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const AdamItem it = items[lo];
    const int64_t j = i - off[lo];
    const double gi = it.g[j];
    const double mi = b1 * it.m[j] + (1.0 - b1) * gi;
    const double vi = b2 * it.v[j] + (1.0 - b2) * gi * gi;
    it.m[j] = mi;
    it.v[j] = vi;
    it.p[j] -= lr * (mi / c1) / (sqrt(vi / c2 + eps_root) + eps);
  }
}

__global__ void count_kernel(int64_t* count, int64_t add, int64_t set) {
  if (threadIdx.x == 0 && blockIdx.x == 0) count[0] = set >= 0 ? set : count[0] + add;
}

}  // namespace
}  // namespace jrx

using namespace jrx;

struct jrx_batch {
  std::vector<jrx_batch_member> members;
  jrx_adam_config adam;
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  int64_t* d_count = nullptr;
  AdamItem* d_items = nullptr;
  int64_t* d_off = nullptr;
  int nitems = 0;
  int64_t total = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void batch_free(jrx_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (auto s : b->aux) (void)hipStreamDestroy(s);
  if (b->d_count) (void)hipFree(b->d_count);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_off) (void)hipFree(b->d_off);
  delete b;
}

// Record one optimiser step on `s` (eager or under capture).
static int batch_record(jrx_batch* b, hipStream_t s) {
  int rc;
  hipLaunchKernelGGL(count_kernel, dim3(1), dim3(64), 0, s, b->d_count,
                     (int64_t)1, (int64_t)-1);
  JRX_LAUNCH_CHECK();
  const int na = (int)b->aux.size();
  if (na > 0) {
    JRX_CHECK_HIP(hipEventRecord(b->fork, s));
    for (int a = 0; a < na; ++a)
      JRX_CHECK_HIP(hipStreamWaitEvent(b->aux[a], b->fork, 0));
  }
  for (size_t i = 0; i < b->members.size(); ++i) {
    const jrx_batch_member& m = b->members[i];
    hipStream_t ms = na > 0 ? b->aux[i % na] : s;
    if ((rc = jrx_energy_grad(m.plan, m.w_re, m.w_im, m.energies, m.dw_re,
                              m.dw_im, nullptr, nullptr, m.workspace,
                              m.workspace_bytes, ms)))
      return rc;
  }
  if (na > 0) {
    for (int a = 0; a < na; ++a) {
      JRX_CHECK_HIP(hipEventRecord(b->join[a], b->aux[a]));
      JRX_CHECK_HIP(hipStreamWaitEvent(s, b->join[a], 0));
    }
  }
  if (b->total > 0) {
    const int64_t per_block = (int64_t)kAdamBlock * kAdamPerThread;
    const int64_t nblk = (b->total + per_block - 1) / per_block;
    JRX_REQUIRE(nblk <= 2147483647, "batch too large for one Adam launch");
    hipLaunchKernelGGL(grouped_adam_kernel, dim3((unsigned)nblk),
                       dim3(kAdamBlock), 0, s, b->d_items, b->d_off,
                       b->nitems, b->d_count, b->adam.lr, b->adam.b1,
                       b->adam.b2, b->adam.eps, b->adam.eps_root);
    JRX_LAUNCH_CHECK();
  }
  return JRX_OK;
}

extern "C" {

int jrx_batch_create(jrx_batch** out, const jrx_batch_member* members, int n,
                     const jrx_adam_config* adam, int streams) {
  JRX_REQUIRE(out && members && adam && n >= 1, "invalid batch arguments");
  JRX_REQUIRE(streams >= 0 && streams <= 16, "streams must be in [0, 16]");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const jrx_batch_member& m = members[i];
    JRX_REQUIRE(m.plan && m.w_re && m.w_im && m.dw_re && m.dw_im && m.m_re &&
                    m.m_im && m.v_re && m.v_im && m.energies && m.workspace,
                "batch member with a NULL buffer");
    size_t need = 0;
    int rc = jrx_plan_workspace_bytes(m.plan, &need);
    if (rc) return rc;
    JRX_REQUIRE(m.workspace_bytes >= need, "batch member workspace too small");
  }
  jrx_batch* b = new jrx_batch();
  b->members.assign(members, members + n);
  b->adam = *adam;
  auto fail = [&](int rc) {
    batch_free(b);
    return rc;
  };
  std::vector<AdamItem> items;
  std::vector<int64_t> off;
  for (int i = 0; i < n; ++i) {
    const jrx_batch_member& m = members[i];
    jrx_plan_info info;
    int rc = jrx_plan_get_info(m.plan, &info);
    if (rc) return fail(rc);
    (void)info;
    int64_t np = 0;
    rc = jrx_plan_param_count(m.plan, &np);
    if (rc) return fail(rc);
    for (int part = 0; part < 2; ++part) {
      AdamItem it;
      it.p = part ? m.w_im : m.w_re;
      it.g = part ? m.dw_im : m.dw_re;
      it.m = part ? m.m_im : m.m_re;
      it.v = part ? m.v_im : m.v_re;
      it.n = np;
      off.push_back(b->total);
      items.push_back(it);
      b->total += np;
    }
  }
  off.push_back(b->total);
  b->nitems = (int)items.size();
  hipError_t e;
  if ((e = hipMalloc(&b->d_count, sizeof(int64_t))) != hipSuccess ||
      (e = hipMalloc(&b->d_items, items.size() * sizeof(AdamItem))) !=
          hipSuccess ||
      (e = hipMalloc(&b->d_off, off.size() * sizeof(int64_t))) != hipSuccess ||
      (e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(AdamItem),
                     hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(b->d_off, off.data(), off.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice)) != hipSuccess) {
    set_error(std::string("batch allocation failed: ") + hipGetErrorString(e));
    return fail(JRX_ERR_HIP);
  }
  const int64_t zero = 0;
  if ((e = hipMemcpy(b->d_count, &zero, sizeof(int64_t),
                     hipMemcpyHostToDevice)) != hipSuccess) {
    set_error(std::string("batch set-up failed: ") + hipGetErrorString(e));
    return fail(JRX_ERR_HIP);
  }
  for (int a = 0; a < streams; ++a) {
    hipStream_t st;
    hipEvent_t ev;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) !=
        hipSuccess) {
      set_error(std::string("stream creation failed: ") + hipGetErrorString(e));
      return fail(JRX_ERR_HIP);
    }
    b->aux.push_back(st);
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) !=
        hipSuccess) {
      set_error(std::string("event creation failed: ") + hipGetErrorString(e));
      return fail(JRX_ERR_HIP);
    }
    b->join.push_back(ev);
  }
  if (streams > 0 &&
      (e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) !=
          hipSuccess) {
    set_error(std::string("event creation failed: ") + hipGetErrorString(e));
    return fail(JRX_ERR_HIP);
  }
  *out = b;
  return JRX_OK;
}

int jrx_batch_destroy(jrx_batch* b) {
  if (b) batch_free(b);
  return JRX_OK;
}

int jrx_batch_set_step_count(jrx_batch* b, int64_t count, void* stream) {
  JRX_REQUIRE(b && count >= 0, "invalid arguments");
  hipLaunchKernelGGL(count_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     b->d_count, (int64_t)0, count);
  JRX_LAUNCH_CHECK();
  return JRX_OK;
}

int jrx_batch_get_step_count(const jrx_batch* b, int64_t* count,
                             void* stream) {
  JRX_REQUIRE(b && count, "invalid arguments");
  JRX_CHECK_HIP(hipMemcpyAsync(count, b->d_count, sizeof(int64_t),
                               hipMemcpyDeviceToHost, (hipStream_t)stream));
  JRX_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return JRX_OK;
}

int jrx_batch_step(jrx_batch* b, void* stream) {
  JRX_REQUIRE(b, "null batch");
  return batch_record(b, (hipStream_t)stream);
}

int jrx_batch_capture(jrx_batch* b, void* stream) {
  JRX_REQUIRE(b && stream, "capture needs a batch and a non-null stream");
  hipStream_t s = (hipStream_t)stream;
  if (b->exec) {
    JRX_CHECK_HIP(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  if (b->graph) {
    JRX_CHECK_HIP(hipGraphDestroy(b->graph));
    b->graph = nullptr;
  }
  JRX_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = batch_record(b, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
    return JRX_ERR_HIP;
  }
  b->graph = g;
  JRX_CHECK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return JRX_OK;
}

int jrx_batch_replay(jrx_batch* b, void* stream) {
  JRX_REQUIRE(b && b->exec, "no captured step (call jrx_batch_capture)");
  JRX_CHECK_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return JRX_OK;
}

}  // extern "C"
