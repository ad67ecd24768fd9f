// Batched absorption that allocates no device memory.
//
// nise_absorption_batched_noalloc computes what nise_absorption_batched
// computes for device-resident records and the Trotter scheme, with the same
// schedule, tables and kernels (row conveyor of the default shape), so the
// two agree bit for bit.  Every device buffer the run needs beyond its inputs
// and outputs (the coupling mask, the row schedule and its step table, the
// per-frame phase and rotation tables, the system descriptor and workgroup
// map, the per-vector partial responses and the reduced response) is carved
// out of `scratch`, a file-scope device array of kScratchBytes per GPU that
// lives until the process ends.  The call makes no hipMalloc / hipFree and
// creates no stream or event: it runs on the null stream with synchronous
// copies.  Host-side bookkeeping (schedules) uses the C++ heap.
//
// Contents of the array persist between calls (the last call's state); each
// call writes every byte it reads before reading it: the mask kernel writes
// every pair, the table kernel every (frame, entry), the conveyor every
// vector's partial row and the reduction every output.  Calls on one GPU are
// serialised by a per-device lock inside this process; launches of other
// code that reuse the array concurrently would race with it.
#include <cstring>
#include <mutex>
#include <vector>

#include "nise_internal.h"

using namespace nise;

int onfly_scheme(const t_non *non);  // api_singles.hip

namespace {

constexpr size_t kScratchBytes = (size_t)32 << 20;  // per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per device: the array's device address (hipGetSymbolAddress with that
// device current, once), the status of that lookup, and the call lock.
struct DeviceScratch {
  std::once_flag once;
  int status = NISE_OK;
  unsigned char *ptr = nullptr;
  std::mutex mu;
};
DeviceScratch g_dev[kMaxDevices];

int scratch_for_current_device(DeviceScratch **out) {
  int dev = 0;
  NISE_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error("device %d beyond the scratch table", dev), NISE_ERR_ARG;
  DeviceScratch &d = g_dev[dev];
  std::call_once(d.once, [&d] {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      d.status = NISE_ERR_HIP;
      return;
    }
    d.ptr = static_cast<unsigned char *>(p);
  });
  if (d.status) return d.status;
  *out = &d;
  return NISE_OK;
}

// Bump allocation of 256-byte aligned pieces of one device's array.
struct Arena {
  unsigned char *base = nullptr;
  size_t used = 0;
  template <class T> T *take(size_t n) {
    const size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
    if (used + bytes > kScratchBytes) return nullptr;
    T *p = reinterpret_cast<T *>(base + used);
    used += bytes;
    return p;
  }
  static size_t bytes_of(size_t n, size_t elem) { return (n * elem + 255) & ~(size_t)255; }
};

int too_big(size_t need) {
  set_error("nise_absorption_batched_noalloc: the run needs %zu bytes of scratch, the limit is %zu "
            "(use nise_absorption_batched)",
            need, kScratchBytes);
  return NISE_ERR_RANGE;
}

}  // namespace

extern "C" int nise_absorption_batched_noalloc(const t_non *non, const float *H_rec, int64_t H_stride,
                                               const float *mu_rec, int64_t mu_stride, int64_t nframes,
                                               const float *coupling_rec, float *re_S, float *im_S) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!non || !H_rec || !mu_rec || !re_S || !im_S) return set_error("null argument"), NISE_ERR_ARG;
  const int N = non->singles, T = non->tmax1, sample = non->sample, m = non->ts;
  if (N < 1 || N > 65535) return set_error("singles=%d out of range", N), NISE_ERR_ARG;
  if (T < 1 || sample < 1 || m < 1) return set_error("tmax1, sample and Trotter steps must be >= 1"), NISE_ERR_ARG;
  if (non->propagation != 1)
    return set_error("nise_absorption_batched_noalloc runs propagation 1 (Coupling); got %d", non->propagation),
           NISE_ERR_ARG;
  if (onfly_scheme(non))
    return set_error("nise_absorption_batched_noalloc takes Full or Coupling records, not %s", non->hamiltonian),
           NISE_ERR_ARG;
  const int begin = non->begin;
  int end = non->end;
  if (end == 0) end = (int)((nframes - T - 1) / sample + 1);  // src/NISE_subs.c:200,210
  if (end <= begin) return set_error("no samples (begin=%d end=%d)", begin, end), NISE_ERR_RANGE;
  const int64_t f_lo = (int64_t)begin * sample, f_hi = (int64_t)(end - 1) * sample + T;
  if (f_hi > nframes)
    return set_error("trajectory has %lld frames, run needs %lld", (long long)nframes, (long long)f_hi),
           NISE_ERR_RANGE;
  const bool ctype = std::strcmp(non->hamiltonian, "Coupling") == 0;
  if (ctype && !coupling_rec) return set_error("Coupling type needs coupling_rec"), NISE_ERR_ARG;
  if (H_stride < 1 + (ctype ? N : (int64_t)N * (N + 1) / 2) || mu_stride < 1 + 3 * (int64_t)N)
    return set_error("record stride too small"), NISE_ERR_ARG;
  for (int i = 0; i < non->SingleShiftSites; ++i)
    if (non->SingleShiftSite[i] < 0 || non->SingleShiftSite[i] >= N)
      return set_error("single shift site %d out of range", non->SingleShiftSite[i]), NISE_ERR_ARG;

  DeviceScratch *ds = nullptr;
  if ((rc = scratch_for_current_device(&ds))) return rc;
  std::lock_guard<std::mutex> lock(ds->mu);
  Arena ar{ds->ptr, 0};
  const hipStream_t st = nullptr;  // the null stream; no stream is created

  FrameSource src;
  src.H_rec = H_rec, src.H_stride = H_stride, src.coupling_type = ctype, src.C_rec = coupling_rec;
  src.shifte = non->shifte, src.cut = non->couplingcut, src.apply_shift = 1;
  src.nshift = non->SingleShiftSites;
  if (src.nshift > 0) {
    int *ss = ar.take<int>(src.nshift);
    float *sab = ar.take<float>(2 * (size_t)src.nshift);
    if (!ss || !sab) return too_big(ar.used + 512);
    NISE_HIP_CHECK(hipMemcpy(ss, non->SingleShiftSite, src.nshift * sizeof(int), hipMemcpyHostToDevice));
    NISE_HIP_CHECK(hipMemcpy(sab, non->SingleShift, 2 * src.nshift * sizeof(float), hipMemcpyHostToDevice));
    src.shift_site = ss, src.shift_ab = sab;
  }

  // union of kept couplings over the frames in use: every pair is written
  const int64_t npair = (int64_t)N * (N - 1) / 2;
  std::vector<uint8_t> mask(npair);
  if (npair) {
    const size_t mark = ar.used;
    uint8_t *mk = ar.take<uint8_t>(npair);
    if (!mk) return too_big(ar.used + Arena::bytes_of(npair, 1));
    if ((rc = launch_union_mask(src, N, f_lo, f_hi - f_lo, mk, st))) return rc;
    NISE_HIP_CHECK(hipMemcpy(mask.data(), mk, npair, hipMemcpyDeviceToHost));
    ar.used = mark;  // the mask is on the host now
  }

  // the default row conveyor of nise_abs_plan_create
  Schedule asap, sch;
  build_schedule(N, mask, asap);
  const int R = rows_for(N, asap.Kreal, asap.nlayers());
  build_schedule_rows(N, mask, R, sch);
  std::vector<uint2> tab;
  rows_step_table(sch, R, tab);
  std::vector<uint32_t> ab;
  rows_device_order(sch, R, ab);
  const int64_t nfr = f_hi - f_lo;
  const int nvec = 3 * (end - begin), G = 64 / R;
  const int nwg = (nvec + G - 1) / G;
  const size_t K = std::max<size_t>(sch.K, 1);
  const size_t need = ar.used + Arena::bytes_of(tab.size(), sizeof(uint2)) + Arena::bytes_of(ab.size(), 4) +
                      Arena::bytes_of((size_t)nfr * N, sizeof(float2)) + Arena::bytes_of((size_t)nfr * K, sizeof(float2)) +
                      Arena::bytes_of((size_t)nvec * T, sizeof(double2)) + Arena::bytes_of(2 * (size_t)T, 4) +
                      Arena::bytes_of(1, sizeof(RowsArgs)) + Arena::bytes_of(nwg, sizeof(int2));
  if (need > kScratchBytes) return too_big(need);
  uint2 *steps = ar.take<uint2>(tab.size());
  uint32_t *abd = ar.take<uint32_t>(ab.size());
  float2 *diag = ar.take<float2>((size_t)nfr * N);
  float2 *cs = ar.take<float2>((size_t)nfr * K);
  double2 *partial = ar.take<double2>((size_t)nvec * T);
  float *out = ar.take<float>(2 * (size_t)T);
  RowsArgs *sysd = ar.take<RowsArgs>(1);
  int2 *wmapd = ar.take<int2>(nwg);
  if (!steps || !abd || !diag || !cs || !partial || !out || !sysd || !wmapd) return too_big(need);
  if (!tab.empty()) NISE_HIP_CHECK(hipMemcpy(steps, tab.data(), tab.size() * sizeof(uint2), hipMemcpyHostToDevice));
  if (!ab.empty()) NISE_HIP_CHECK(hipMemcpy(abd, ab.data(), ab.size() * 4, hipMemcpyHostToDevice));

  // per-frame half-step phases and rotation tables (every entry written)
  if ((rc = launch_frame_tables(src, N, sch, abd, trotter_factor(non->deltat, 1, m), f_lo, nfr, diag, cs, st, nullptr)))
    return rc;

  RowsArgs r;
  r.N = N, r.tmax = T, r.m = m, r.nlayers = sch.nlayers(), r.K = sch.K, r.rows = R;
  r.f_lo = f_lo, r.nvec = nvec, r.begin = begin, r.sample = sample;
  r.steps = steps, r.cs = cs, r.diag = diag;
  r.mu_rec = mu_rec, r.mu_stride = ctype ? 0 : mu_stride, r.partial = partial;
  std::vector<int2> wmap;
  size_t smem = 0;
  bool pf = false;
  if ((rc = rows_prepare(&r, 1, R, wmap, smem, pf))) return rc;
  if ((int)wmap.size() != nwg) return set_error("row conveyor: workgroup map mismatch"), NISE_ERR_ARG;
  NISE_HIP_CHECK(hipMemcpy(sysd, &r, sizeof(r), hipMemcpyHostToDevice));
  NISE_HIP_CHECK(hipMemcpy(wmapd, wmap.data(), wmap.size() * sizeof(int2), hipMemcpyHostToDevice));
  if ((rc = launch_rows(sysd, wmapd, nwg, R, pf, smem, st))) return rc;
  if ((rc = launch_reduce_partials(partial, nvec, T, out, out + T, st))) return rc;
  NISE_HIP_CHECK(hipMemcpy(re_S, out, T * sizeof(float), hipMemcpyDeviceToHost));
  NISE_HIP_CHECK(hipMemcpy(im_S, out + T, T * sizeof(float), hipMemcpyDeviceToHost));
  return NISE_OK;
}
