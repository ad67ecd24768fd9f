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

// LDS bank cost of a layer's row order e[0..R) (padding rows ~0u stand for
// the row's scratch site N + q): a row's G vectors occupy 2G consecutive
// dwords, so in B banks its run falls in block slot mod (B / 2G); per group
// of lanes the busiest block bounds the passes of an access.  The grouping
// the hardware uses for 8-byte accesses is not pinned here, so the cost is
// summed over 32-lane groups on 64 banks and 16-lane groups on 32 banks
// (a grouping of all 64 lanes does not depend on the row order).
struct BankModel {
  int nblk, rpg;  // blocks, rows per lane group
};
static int group_cost(const uint32_t *e, int g0, int rpg, int nblk, int N, int R) {
  int cost = 0;
  for (int side = 0; side < 2; ++side) {
    int cnt[64], mx = 0;
    for (int k = 0; k < nblk; ++k) cnt[k] = 0;
    for (int q = g0; q < std::min(R, g0 + rpg); ++q) {
      const int slot = e[q] == ~0u ? N + q : side ? (int)(e[q] >> 16) : (int)(e[q] & 0xffffu);
      mx = std::max(mx, ++cnt[slot & (nblk - 1)]);
    }
    cost += mx;
  }
  return cost;
}

// Rotations of one layer touch disjoint sites, so their order across the R
// rows is free: passes of pairwise swaps between lane groups lower the
// layer's bank cost (a swap changes only the two groups it touches).
static void assign_rows(std::vector<uint32_t> &e, int R, int N) {
  const int G = 64 / R;
  BankModel bm[2];
  int nm = 0;
  for (const auto &m : {std::make_pair(64, 32), std::make_pair(32, 16)}) {
    const int span = 2 * G, rpg = std::max(1, m.second / G);
    if (span <= m.first && rpg < R) bm[nm++] = BankModel{m.first / span, rpg};
  }
  if (!nm) return;
  for (int pass = 0; pass < 4; ++pass) {
    bool improved = false;
    for (int i = 0; i < R; ++i)
      for (int j = i + 1; j < R; ++j) {
        if (e[i] == ~0u && e[j] == ~0u) continue;
        int before = 0, after = 0;
        for (int k = 0; k < nm; ++k) {
          const int gi = i / bm[k].rpg * bm[k].rpg, gj = j / bm[k].rpg * bm[k].rpg;
          if (gi != gj)
            before += group_cost(e.data(), gi, bm[k].rpg, bm[k].nblk, N, R) +
                      group_cost(e.data(), gj, bm[k].rpg, bm[k].nblk, N, R);
        }
        if (!before) continue;
        std::swap(e[i], e[j]);
        for (int k = 0; k < nm; ++k) {
          const int gi = i / bm[k].rpg * bm[k].rpg, gj = j / bm[k].rpg * bm[k].rpg;
          if (gi != gj)
            after += group_cost(e.data(), gi, bm[k].rpg, bm[k].nblk, N, R) +
                     group_cost(e.data(), gj, bm[k].rpg, bm[k].nblk, N, R);
        }
        if (after < before) improved = true;
        else std::swap(e[i], e[j]);
      }
    if (!improved) break;
  }
}

// Layers of at most R rotations by list scheduling on the dependency graph
// of the lexicographic sweep: a rotation follows the previous rotation on
// each of its two sites, so every site keeps its order.  Layer by layer, the
// ready rotations with the longest remaining chain (ties: lexicographic
// order) fill the R rows; this keeps the layer count near
// max(critical path, rotations / R).  The rows of a layer are then ordered
// to spread its LDS accesses over the banks (assign_rows).  Each layer is
// padded to exactly R entries; the padding entry of row q is the dummy pair
// (N + q, N + q),
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
  std::vector<uint32_t> e(R);
  for (auto &L : layers) {
    for (int q = 0; q < R; ++q) e[q] = q < (int)L.size() ? L[q] : ~0u;
    if (R >= 4 && R <= 64) assign_rows(e, R, N);
    for (int q = 0; q < R; ++q)
      s.ab.push_back(e[q] != ~0u ? e[q] : (uint32_t)(N + q) | ((uint32_t)(N + q) << 16));
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
