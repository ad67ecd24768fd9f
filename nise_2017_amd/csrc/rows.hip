// Row conveyor: batched absorption propagation with one wave per workgroup.
//
// A wave holds G = 64/R vectors (consecutive (sample, polarisation) pairs)
// in LDS as psi[site][G].  Lane q*G + g = (row q, vector g): in the rotation
// sweep row q applies entry q of every layer of the row schedule
// (build_schedule_rows: layers of exactly R site-disjoint entries, padding
// entries act on the scratch site N + q), so a sweep is nlayers steps with
// no masking and no barrier: the LDS operations of one wave execute in
// program order, which orders the layers.  Every site sees its rotations in
// the reference's lexicographic order (src/propagate.c:848-888), so each
// vector is propagated with the reference's FP32 operations in the
// reference's order.
//
// The G vectors of a wave run in lock step over relative time t = 0..tmax-1
// but each at its own frame fs + t; a lane reads its vector's frame tables
// (half-step phases, rotation (cos, sin) in row-schedule order, dipoles).
// Per step t: the per-site pass applies the pending half-phase of frame
// fs+t-1, accumulates mu(fs+t).psi (src/absorption.c:130, calc_S1 :205-212),
// records checkpoints, and applies the first half-phase of frame fs+t; then
// the m Trotter sweeps of frame fs+t.  The wave's response of step t (all
// its vectors) is summed over rows and vectors in a fixed order and written
// as an FP64 partial.
#include <algorithm>
#include <cstring>
#include <vector>

#include "nise_internal.h"

namespace nise {

constexpr int kRingR = 8;  // rotation operands in flight per lane

// One rotation (src/propagate.c:883-887) of the pair at LDS byte offsets
// e.x, e.y (plus the lane's vector column) with (cos, sin) = c.
__device__ __forceinline__ void rot_step(unsigned char *smem, uint32_t gofs, uint2 e, float2 c) {
  float2 *pa_ = reinterpret_cast<float2 *>(smem + (e.x + gofs));
  float2 *pb_ = reinterpret_cast<float2 *>(smem + (e.y + gofs));
  float2 pa = *pa_, pb = *pb_;
  rotate_pair(pa, pb, c.x, c.y);
  *pa_ = pa;
  *pb_ = pb;
}

// Workgroup w runs the vectors wmap[w].y .. wmap[w].y + G - 1 of system
// wmap[w].x (systems of different sizes share one launch).
template <int R>
__global__ void __launch_bounds__(64) rows_kernel(const RowsArgs *sys, const int2 *wmap,
                                                  unsigned long long *times) {
  constexpr int G = 64 / R;
  extern __shared__ __align__(16) unsigned char smem[];
  const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
  const int2 wm = wmap[blockIdx.x];
  const RowsArgs &A = sys[wm.x];
  const int N = A.N, T = A.tmax, nl = A.nlayers;
  // lane = q * G + g: the G vectors of a row are adjacent, so a row's LDS
  // accesses cover one contiguous 8*G-byte run (fewer bank conflicts)
  const int lane = threadIdx.x, q = lane / G, g = lane % G;
  const int w = wm.y / G;  // workgroup index within the system
  const int v = wm.y + g;
  const bool has = v < A.nvec;
  const int vv = has ? v : w * G;  // inactive lanes shadow the wave's first vector
  const int s = vv / 3, x = vv % 3;
  const int64_t fs = (int64_t)(A.begin + s) * A.sample;
  const uint32_t gofs = (uint32_t)g * 8u;
  const int64_t hist0 = (A.hist && has) ? (int64_t)v * A.nck : -1;
  const int64_t K = A.K;

  // ring of (LDS byte offsets of a and b, (cos, sin)) for the next kRingR steps
  uint2 re[kRingR];
  float2 rc[kRingR];
  const uint2 *stq = A.steps + q;
  {
    const float2 *c0 = A.cs + (fs - A.f_lo) * K + q;
#pragma unroll
    for (int j = 0; j < kRingR; ++j) {
      re[j] = stq[j * R];
      rc[j] = c0[j * R];
    }
  }

  for (int t = 0; t < T; ++t) {
    const int64_t f = fs + t;
    const int64_t row = f - A.f_lo;
    const float *mu = A.mu_rec + (A.mu_stride ? f * A.mu_stride : 0) + 1 + (int64_t)x * N;
    const float2 *Dc = A.diag + row * N;
    const float2 *Dp = Dc - N;
    const bool last = t == T - 1;
    int64_t hr = -1;
    if (hist0 >= 0) {
      if (t % A.hist_every == 0) hr = hist0 + t / A.hist_every;
      else if (last) hr = hist0 + t / A.hist_every + 1;
    }
    float rr = 0.f, ri = 0.f;
    for (int i = q; i < N; i += R) {
      const float m = mu[i];
      float2 p;
      if (t == 0) p = make_float2(m, 0.f);
      else p = cmul_diag(*reinterpret_cast<const float2 *>(smem + ((uint32_t)i * G * 8u + gofs)), Dp[i]);
      rr = __fadd_rn(rr, __fmul_rn(m, p.x));
      ri = __fadd_rn(ri, __fmul_rn(m, p.y));
      if (hr >= 0) A.hist[hr * N + i] = p;
      if (!last) p = cmul_diag(p, Dc[i]);
      *reinterpret_cast<float2 *>(smem + ((uint32_t)i * G * 8u + gofs)) = p;
    }
    // response of step t: rows of a vector, then the wave's vectors in order
    if (!has) rr = 0.f, ri = 0.f;
#pragma unroll
    for (int off = G; off < 64; off <<= 1) {  // over the rows of a vector
      rr += __shfl_xor(rr, off);
      ri += __shfl_xor(ri, off);
    }
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int k = 0; k < G; ++k) {  // the wave's vectors in order
      sr += (double)__shfl(rr, k);
      si += (double)__shfl(ri, k);
    }
    if (lane == 0) A.partial[(int64_t)w * T + t] = make_double2(sr, si);
    if (last) {
      if (lane == 0 && times) {  // device wall clock at start and end
        times[2 * blockIdx.x] = t_start;
        times[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
      }
      break;
    }

    for (int sub = 0; sub < A.m; ++sub) {
      if (sub > 0)
        for (int i = q; i < N; i += R) {
          float2 *pp = reinterpret_cast<float2 *>(smem + ((uint32_t)i * G * 8u + gofs));
          *pp = cmul_diag(cmul_diag(*pp, Dc[i]), Dc[i]);
        }
      // Each ring slot is consumed and refilled with the operands kRingR
      // steps ahead: from this sweep, and in the last group (nlayers is a
      // multiple of kRingR) from the first steps of the next sweep, which
      // runs this frame again or the next frame.
      const float2 *csc = A.cs + row * K + q;
      const float2 *csn = A.cs + (row + (sub + 1 < A.m ? 0 : 1)) * K + q;
      const uint2 *sa = stq + kRingR * R;
      const float2 *ca = csc + kRingR * R;
      for (int b0 = 0; b0 < nl - kRingR; b0 += kRingR, sa += kRingR * R, ca += kRingR * R) {
#pragma unroll
        for (int j = 0; j < kRingR; ++j) {
          rot_step(smem, gofs, re[j], rc[j]);
          re[j] = sa[j * R];
          rc[j] = ca[j * R];
        }
      }
#pragma unroll
      for (int j = 0; j < kRingR; ++j) {
        rot_step(smem, gofs, re[j], rc[j]);
        re[j] = stq[j * R];
        rc[j] = csn[j * R];
      }
    }
  }
}

size_t rows_smem(int N, int R) { return (size_t)(N + R) * (64 / R) * 8; }

static const void *rows_fn(int R) {
  return R == 1 ? (const void *)rows_kernel<1>
       : R == 2 ? (const void *)rows_kernel<2>
       : R == 4 ? (const void *)rows_kernel<4>
       : R == 8 ? (const void *)rows_kernel<8>
       : R == 16 ? (const void *)rows_kernel<16> : nullptr;
}

// Checks the systems against the row count and builds the workgroup map
// (G consecutive vectors of one system per workgroup) and the LDS size.
int rows_prepare(const RowsArgs *sys, int nsys, int R, std::vector<int2> &wmap, size_t &smem) {
  if (!rows_fn(R)) return set_error("row conveyor: unsupported row count %d", R), NISE_ERR_ARG;
  smem = 0;
  wmap.clear();
  const int G = 64 / R;
  for (int k = 0; k < nsys; ++k) {
    const RowsArgs &a = sys[k];
    if (a.nlayers < kRingR || a.nlayers % kRingR)
      return set_error("row conveyor: layer count %d is not a multiple of %d", a.nlayers, kRingR), NISE_ERR_ARG;
    if (a.K != (int64_t)a.nlayers * R) return set_error("row conveyor: schedule built for other rows"), NISE_ERR_ARG;
    smem = std::max(smem, rows_smem(a.N, R));
    for (int v = 0; v < a.nvec; v += G) wmap.push_back(make_int2(k, v));
  }
  if (smem > 160 * 1024) return set_error("row conveyor: %zu B of LDS needed", smem), NISE_ERR_ARG;
  return NISE_OK;
}

// sys / wmap: device copies (rows_prepare), nwg workgroups.
int launch_rows(const RowsArgs *sys, const int2 *wmap, int nwg, int R, size_t smem, hipStream_t st,
                unsigned long long *times) {
  const void *fn = rows_fn(R);
  if (!fn) return set_error("row conveyor: unsupported row count %d", R), NISE_ERR_ARG;
  if (nwg == 0) return NISE_OK;
  NISE_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  void *kargs[] = {&sys, &wmap, &times};
  NISE_HIP_CHECK(hipLaunchKernel(fn, dim3((unsigned)nwg), dim3(64), kargs, smem, st));
  return NISE_OK;
}

// Row count for a schedule: the power of two closest to the mean number of
// rotations per layer of the unconstrained layering, raised while the LDS
// image would exceed 80 KiB (so that two waves fit on a CU).
int rows_for(int N, int Kreal, int nlayers_asap) {
  const double per = nlayers_asap ? (double)Kreal / nlayers_asap : 1.0;
  int R = 1;
  while (R < 16 && R * 1.5 < per) R <<= 1;
  while (R < 16 && rows_smem(N, R) > 80 * 1024) R <<= 1;
  return R;
}

// Static step table of a row schedule: LDS byte offsets of the two sites of
// every entry for a wave of 64/R vectors.
void rows_step_table(const Schedule &s, int R, std::vector<uint2> &tab) {
  const uint32_t G = 64 / R;
  tab.resize(s.ab.size());
  for (size_t i = 0; i < s.ab.size(); ++i)
    tab[i] = make_uint2((s.ab[i] & 0xffffu) * G * 8u, (s.ab[i] >> 16) * G * 8u);
}

}  // namespace nise

// Inspection of the row schedule (host only): entries (a | b << 16) of
// nlayers layers of R, for a kept-coupling mask over the strict upper
// triangle in lexicographic order.  Returns the entry count, or -1 if cap
// is too small.
extern "C" int nise_row_schedule(int N, const uint8_t *mask, int R, uint32_t *ab_out, int cap,
                                 int *nlayers) {
  if (N < 1 || R < 1 || !mask) return -1;
  std::vector<uint8_t> m(mask, mask + (size_t)N * (N - 1) / 2);
  nise::Schedule s;
  nise::build_schedule_rows(N, m, R, s);
  if ((int)s.ab.size() > cap) return -1;
  std::copy(s.ab.begin(), s.ab.end(), ab_out);
  if (nlayers) *nlayers = s.nlayers();
  return (int)s.ab.size();
}
