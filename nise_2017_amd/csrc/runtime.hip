// Host runtime of libnise_hip: error state, device buffers, the layered
// rotation schedule, device and stream setup.
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "nise_internal.h"

namespace nise {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char *get_error() { return g_err; }

int DevBuf::alloc(size_t n) {
  if (n <= bytes && p) return NISE_OK;
  release();
  if (n == 0) return NISE_OK;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    return NISE_ERR_NOMEM;
  }
  bytes = n;
  return NISE_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// ASAP layering of the lexicographic rotation sequence: rotation k goes to
// layer 1 + max(last layer touching a_k, last layer touching b_k).  Every
// site therefore sees its rotations in the original order, and rotations of
// one layer share no site.
void build_schedule(int N, const std::vector<uint8_t> &mask, Schedule &s) {
  s.N = N;
  s.ab.clear();
  std::vector<int> la, lb, lay;
  std::vector<int> last(N, -1);
  int nl = 0;
  int64_t p = 0;
  for (int a = 0; a < N; ++a)
    for (int b = a + 1; b < N; ++b, ++p) {
      if (!mask[p]) continue;
      int l = 1 + (last[a] > last[b] ? last[a] : last[b]);
      last[a] = last[b] = l;
      la.push_back(a), lb.push_back(b), lay.push_back(l);
      if (l + 1 > nl) nl = l + 1;
    }
  s.K = s.Kreal = (int)la.size();
  s.layer_off.assign(nl + 1, 0);
  for (int l : lay) s.layer_off[l + 1]++;
  for (int l = 0; l < nl; ++l) s.layer_off[l + 1] += s.layer_off[l];
  std::vector<int> fill(s.layer_off.begin(), s.layer_off.end() - 1);
  s.ab.assign(s.K, 0);
  for (int k = 0; k < s.K; ++k)  // stable within a layer
    s.ab[fill[lay[k]]++] = (uint32_t)la[k] | ((uint32_t)lb[k] << 16);
}

// Layers of at most R rotations by list scheduling on the dependency graph
// of the lexicographic sweep: a rotation follows the previous rotation on
// each of its two sites, so every site keeps its order.  Layer by layer, the
// ready rotations with the longest remaining chain (ties: lexicographic
// order) fill the R rows; this keeps the layer count near
// max(critical path, rotations / R).  Each layer is padded to exactly R
// entries; the padding entry of row q is the dummy pair (N + q, N + q),
// which the table kernel turns into an identity and the kernels apply to a
// scratch site.  The layer count is padded to a multiple of kMinLayers.
void build_schedule_rows(int N, const std::vector<uint8_t> &mask, int R, Schedule &s) {
  constexpr int kMinLayers = 8;
  s.N = N;
  std::vector<uint32_t> rot;  // lexicographic order
  int64_t p = 0;
  for (int a = 0; a < N; ++a)
    for (int b = a + 1; b < N; ++b, ++p)
      if (mask[p]) rot.push_back((uint32_t)a | ((uint32_t)b << 16));
  const int K = (int)rot.size();
  // successors: the next rotation on each site of rotation k
  std::vector<int> succ(2 * (size_t)K, -1), npred(K, 0), height(K, 1), lastk(N, -1);
  for (int k = 0; k < K; ++k) {
    const int site[2] = {(int)(rot[k] & 0xffffu), (int)(rot[k] >> 16)};
    for (int e = 0; e < 2; ++e) {
      const int j = lastk[site[e]];
      if (j >= 0) {
        succ[2 * (size_t)j + ((int)(rot[j] & 0xffffu) == site[e] ? 0 : 1)] = k;
        ++npred[k];
      }
      lastk[site[e]] = k;
    }
  }
  for (int k = K - 1; k >= 0; --k)
    for (int e = 0; e < 2; ++e) {
      const int j = succ[2 * (size_t)k + e];
      if (j >= 0) height[k] = std::max(height[k], height[j] + 1);
    }
  auto later = [&](int x, int y) { return height[x] != height[y] ? height[x] < height[y] : x > y; };
  std::vector<int> ready;
  for (int k = 0; k < K; ++k)
    if (!npred[k]) ready.push_back(k);
  std::make_heap(ready.begin(), ready.end(), later);
  std::vector<std::vector<uint32_t>> layers;
  std::vector<int> fresh;
  while (!ready.empty()) {
    std::vector<uint32_t> L;
    fresh.clear();
    while ((int)L.size() < R && !ready.empty()) {
      std::pop_heap(ready.begin(), ready.end(), later);
      const int k = ready.back();
      ready.pop_back();
      L.push_back(rot[k]);
      for (int e = 0; e < 2; ++e) {
        const int j = succ[2 * (size_t)k + e];
        if (j >= 0 && --npred[j] == 0) fresh.push_back(j);  // ready from the next layer on
      }
    }
    layers.push_back(std::move(L));
    for (int j : fresh) {
      ready.push_back(j);
      std::push_heap(ready.begin(), ready.end(), later);
    }
  }
  // at least kMinLayers layers, and a multiple of them (the kernel's
  // operand ring runs across sweeps)
  const size_t nl = std::max<size_t>(kMinLayers, (layers.size() + kMinLayers - 1) / kMinLayers * kMinLayers);
  layers.resize(nl);
  s.Kreal = K;
  s.K = (int)layers.size() * R;
  s.ab.clear();
  s.layer_off.assign(1, 0);
  for (auto &L : layers) {
    for (int q = 0; q < R; ++q)
      s.ab.push_back(q < (int)L.size() ? L[q] : (uint32_t)(N + q) | ((uint32_t)(N + q) << 16));
    s.layer_off.push_back((int)s.ab.size());
  }
}

static std::once_flag g_dev_once;
static int g_dev_status = NISE_OK;

int ensure_device() {
  std::call_once(g_dev_once, [] {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
      g_dev_status = NISE_ERR_HIP;
      return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
      g_dev_status = NISE_ERR_HIP;
      return;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      fprintf(stderr, "libnise_hip: built for gfx950, device is %s\n", prop.gcnArchName);
    }
  });
  if (g_dev_status != NISE_OK) set_error("no usable HIP device");
  return g_dev_status;
}

// One non-blocking stream per host thread: the drop-in symbols may be called
// concurrently from OpenMP threads (src/propagate.c:1541-1548).
hipStream_t default_stream() {
  static thread_local hipStream_t st = nullptr;
  if (!st) {
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
  }
  return st;
}

// A second stream of the calling thread (default priority), for work that
// is independent of the default stream's and may overlap it.
hipStream_t aux_stream() {
  static thread_local hipStream_t st = nullptr;
  if (!st) {
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
  }
  return st;
}

}  // namespace nise

extern "C" const char *nise_last_error(void) { return nise::get_error(); }
extern "C" const char *nise_version(void) { return "nise-mi355x 0.1 (gfx950)"; }
