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
// Each group of B = min(G, 4) consecutive vectors (g / B) runs frame-
// synchronously: its vectors see the same frame f at the same time, vector
// g being at relative time f - fs_g (its sample's start frame fs_g) and
// active while that lies in [0, tmax); groups run time-skewed against each
// other.  A frame's rotation table is therefore the same for the B lanes of
// a row within a group, so each lane loads one entry of a block of B layers
// (lane % B picks the layer) and the group broadcasts it (DPP quad_perm,
// folded into the arithmetic that uses it): one 8-byte load per lane covers
// B layers.  With R = 16 (4 vectors per wave) the table address is
// wave-uniform; R = 32 and 64 (2 and 1 vectors per wave) shorten each
// wave's chain of layers for systems with many couplings per layer.  An
// 8-layer register ring prefetches the entries.
//
// Per step: the per-site pass applies the pending half-phase of frame
// f-1, accumulates mu(f).psi (src/absorption.c:130, calc_S1 :205-212),
// records checkpoints, and applies the first half-phase of frame f; then
// the m Trotter sweeps of frame f if any vector still propagates.  Each
// vector's response at its relative time is summed over rows and written
// as an FP64 partial of that vector.
#include <algorithm>
#include <climits>
#include <cstring>
#include <type_traits>
#include <vector>

#include "nise_internal.h"

namespace nise {

constexpr int kRingL = 8;  // layers in flight (nlayers is a multiple)
// Sites per batch of HBM loads in the per-site pass (all loads of a batch
// are in flight together): a lane walks N / R sites per frame, and each
// batch waits for its loads, so up to 16 rows take batches of 16 (at the
// N = 256 lattice with 16 rows, one batch per lane and frame: conveyor 13.7
// -> 11.8 ms; 8 gained nothing).  Wider rows leave few sites per lane.
template <int R>
constexpr int site_batch() {
  return R <= 16 ? 16 : 4;
}

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

// The value of lane j of each group of B lanes (B = 4: quad_bcast; 2: lanes
// 0, 1 of the group from lane j and lanes 2, 3 from lane 2 + j; 1: the lane's
// own value).
template <int B, int j>
__device__ __forceinline__ uint32_t grp_bcast(uint32_t x) {
  if constexpr (B == 4) return quad_bcast<j>(x);
  else if constexpr (B == 2)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, j | (j << 2) | ((2 + j) << 4) | ((2 + j) << 6), 0xF, 0xF,
                                                 false);
  else return x;
}
template <int B, int j>
__device__ __forceinline__ float grp_bcast(float x) {
  return __uint_as_float(grp_bcast<B, j>(__float_as_uint(x)));
}

// Layer j of a ring slot: the entry loaded by group lane j.
template <int B, int j>
__device__ __forceinline__ void rot_slot(unsigned char *smem, uint32_t gofs, uint2 e, float2 c) {
  rot_step(smem, gofs, make_uint2(grp_bcast<B, j>(e.x), grp_bcast<B, j>(e.y)),
           make_float2(grp_bcast<B, j>(c.x), grp_bcast<B, j>(c.y)));
}
// The B layers of a ring slot.
template <int B>
__device__ __forceinline__ void rot_block(unsigned char *smem, uint32_t gofs, uint2 e, float2 c) {
  rot_slot<B, 0>(smem, gofs, e, c);
  if constexpr (B > 1) rot_slot<B, 1>(smem, gofs, e, c);
  if constexpr (B > 2) {
    rot_slot<B, 2>(smem, gofs, e, c);
    rot_slot<B, 3>(smem, gofs, e, c);
  }
}

// The workgroup's vectors wm.y .. wm.y + G - 1 of system A (wm.x).  WIDE:
// the per-site pass's batches for few rows (site_batch); they take
// registers, so the mixed-count kernel, whose register count all its row
// counts share, keeps batches of 4.  PF: the launch prefetches each frame's
// first batch (rows_prepare decides).  (Loading each batch while the
// previous one was used, in the per-site form, ran slower for the chains:
// 0.69 vs 0.60 ms at N = 16.)
template <int R, bool WIDE = false, bool PF = false>
__device__ __forceinline__ void rows_run(const RowsArgs &A, int2 wm) {
  constexpr int G = 64 / R;
  constexpr int B = G < 4 ? G : 4;  // vectors per frame group, layers per table block
  constexpr int kSiteU = WIDE ? site_batch<R>() : 4;
  // The per-site pass in one of two forms (results identical).  Branch-free
  // batches (every load and LDS read of a batch in flight at once) and the
  // prefetch of each frame's first batch suit few rows, where a lane walks
  // many sites per frame: the N = 16 chain 0.62 -> 0.38 ms, the N = 128
  // chain 7.9 -> 5.3 ms.  With 16 rows and more (one batch per lane and
  // frame) the per-site form measured 3% faster at the N = 256 lattice
  // (11.9 vs 12.3 ms), and the prefetch gained nothing there.
  constexpr bool kPerSite = WIDE && R >= 16;
  constexpr bool kPF = PF && !kPerSite;
  constexpr int kRingS = kRingL / B;  // ring slots per lane
  static_assert(G == 1 || G == 2 || G % 4 == 0, "vectors per wave: 1, 2 or a multiple of 4");
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, T = A.tmax, nl = A.nlayers, m = A.m;
  const int64_t K = A.K;
  // lane = q * G + g: the G vectors of a row are adjacent, so a row's LDS
  // accesses cover one contiguous 8*G-byte run
  const int lane = threadIdx.x, q = lane / G, g = lane % G;
  const int v = wm.y + g;
  const bool has = v < A.nvec;
  const int x = has ? v % 3 : 0;
  const int64_t fs = has ? (int64_t)(A.begin + v / 3) * A.sample : LLONG_MAX / 4;
  const uint32_t gofs = (uint32_t)g * 8u;
  const int64_t hist0 = (A.hist && has) ? (int64_t)v * A.nck : -1;
  // Vector group h = g / B (B consecutive vectors) shares a frame: at wave
  // time tau it is at frame F0(h) + tau, F0(h) being its first vector's
  // start; groups run time-skewed, so the wave runs for the longest group
  // span (its last vector's start - F0(h) + tmax).
  const int vlast = min(wm.y + G, A.nvec) - 1;
  auto start_of = [&](int vec) { return (int64_t)(A.begin + vec / 3) * A.sample; };
  int64_t span = 0;
  for (int h = 0; B * h < G && wm.y + B * h <= vlast; ++h)
    span = max(span, start_of(min(wm.y + B * h + B - 1, vlast)) - start_of(wm.y + B * h) + T);
  const int vq = min(wm.y + (g & ~(B - 1)), vlast);  // this lane's group's first vector
  const int64_t F0 = start_of(vq);
  const int64_t Fend = start_of(min(vq + B - 1, vlast)) + T - 1;  // the group's last frame

  for (int i = lane; i < (N + R) * G; i += 64) reinterpret_cast<float2 *>(smem)[i] = make_float2(0.f, 0.f);

  // ring: slot k holds, in group lane j, entry q of layer B k + j of the
  // block in flight; loff is the lane's entry within a block of B layers
  // (device order, rows_device_order)
  const uint32_t loff = (uint32_t)(B * q + (lane & (B - 1)));
  uint2 re[kRingS];
  float2 rc[kRingS];
  int64_t ring_key = -1;  // sweep (tau * m + sub) whose first block the ring holds
  // Per-site pass: batches of kSiteU sites, sites i0 + u R of the lane's
  // frame; every load of a batch is issued (from valid rows) before the
  // first is used.  load() clamps the frame to the group's frames (a
  // finished group loads valid rows); TAIL: the batch runs past N (indexes
  // past N load site 0 and are not applied).
  struct SiteBatch {
    float mv[kSiteU];
    float2 dp[kSiteU], dc[kSiteU];
  };
  auto load = [&](SiteBatch &b, int64_t fr, int i0, auto tail) {
    constexpr bool TAIL = decltype(tail)::value;
    const int64_t fc = min(fr, Fend), rw = fc - A.f_lo;
    const float *mu = A.mu_rec + (A.mu_stride ? fc * A.mu_stride : 0) + 1 + (int64_t)x * N;
    const float2 *Dc = A.diag + rw * N;
    // the previous frame's phases (rw - 1 >= 0 whenever t > 0; at t = 0 a
    // valid row is loaded and not used)
    const float2 *Dp = A.diag + (rw > 0 ? rw - 1 : 0) * N;
#pragma unroll
    for (int u = 0; u < kSiteU; ++u) {
      const int i = i0 + u * R, ii = (!TAIL || i < N) ? i : 0;
      b.mv[u] = gld(mu + ii);
      b.dp[u] = gld(Dp + ii);
      b.dc[u] = gld(Dc + ii);
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the batch of loads together
  };
  // kPF (the branch-free form of a PF launch: N is a multiple of kSiteU R,
  // so no batch runs past N, and no history rows): nb holds the first batch
  // of the wave's frame, loaded before the previous frame's sweeps (or
  // before the loop, or by a jump)
  SiteBatch nb;
  if constexpr (kPF) load(nb, F0, q, std::false_type{});

  for (int64_t tau = 0; tau < span; ++tau) {
    const int64_t f = F0 + tau;  // this lane's frame
    const int64_t t = f - fs;    // this lane's relative time
    const bool act = t >= 0 && t < T;
    if (!__ballot(act)) {  // no vector of the wave is at its frame: jump
      int64_t nt = span;
      for (int k = 0; k < G && wm.y + k <= vlast; ++k) {
        const int64_t o = start_of(wm.y + k) - start_of(wm.y + (k & ~(B - 1)));
        if (o > tau) nt = min(nt, o);
      }
      if constexpr (kPF) load(nb, F0 + nt, q, std::false_type{});
      tau = nt - 1;
      continue;
    }
    // table row of the frame, clamped to the group's frames (a finished
    // group keeps loading valid rows)
    const int64_t row = min(f, Fend) - A.f_lo;
    const bool last = t == T - 1;
    float rr = 0.f, ri = 0.f;
    if (kPerSite && act) {
      // per-site form: each site's LDS read and phases wait for its own
      // loads, with t == 0 and the last frame branching around them
      const float *mu = A.mu_rec + (A.mu_stride ? f * A.mu_stride : 0) + 1 + (int64_t)x * N;
      const float2 *Dc = A.diag + row * N;
      const float2 *Dp = Dc - N;
      int64_t hr = -1;
      if (hist0 >= 0) {
        if (t % A.hist_every == 0) hr = hist0 + t / A.hist_every;
        else if (last) hr = hist0 + t / A.hist_every + 1;
      }
      auto load_ps = [&](SiteBatch &b, int i0) {
#pragma unroll
        for (int u = 0; u < kSiteU; ++u) {
          const int i = i0 + u * R, ii = i < N ? i : 0;
          b.mv[u] = gld(mu + ii);
          b.dp[u] = t ? gld(Dp + ii) : make_float2(0.f, 0.f);
          b.dc[u] = last ? make_float2(0.f, 0.f) : gld(Dc + ii);
        }
        __builtin_amdgcn_sched_barrier(0);
      };
      auto use_ps = [&](const SiteBatch &b, int i0) {
#pragma unroll
        for (int u = 0; u < kSiteU; ++u) {
          const int i = i0 + u * R;
          const bool ok = i < N;
          float2 *pp = reinterpret_cast<float2 *>(smem + ((uint32_t)(ok ? i : q) * G * 8u + gofs));
          const float mm = b.mv[u];
          float2 p = t == 0 ? make_float2(mm, 0.f) : cmul_diag(*pp, b.dp[u]);
          if (ok) {
            rr = __fadd_rn(rr, __fmul_rn(mm, p.x));
            ri = __fadd_rn(ri, __fmul_rn(mm, p.y));
            if (hr >= 0) gst(A.hist + hr * N + i, p);
            if (!last) p = cmul_diag(p, b.dc[u]);
            *pp = p;
          }
        }
      };
      for (int i0 = q; i0 < N; i0 += kSiteU * R) {
        SiteBatch b;
        load_ps(b, i0);
        use_ps(b, i0);
      }
    } else if (act) {
      // branch-free form: t == 0 and the last frame select their values;
      // HIST: the launch writes history rows (wave-uniform)
      int64_t hr = -1;
      if (hist0 >= 0) {
        if (t % A.hist_every == 0) hr = hist0 + t / A.hist_every;
        else if (last) hr = hist0 + t / A.hist_every + 1;
      }
      auto use = [&](const SiteBatch &b, int i0, auto tail, auto hist) {
        constexpr bool TAIL = decltype(tail)::value, HIST = decltype(hist)::value;
#pragma unroll
        for (int u = 0; u < kSiteU; ++u) {
          const int i = i0 + u * R;
          if (!TAIL || i < N) {
            const float mm = b.mv[u];
            // both values first, then a select (a conditional expression
            // would branch around the LDS read)
            const float2 pv = *reinterpret_cast<const float2 *>(smem + ((uint32_t)i * G * 8u + gofs));
            const float2 pd = cmul_diag(pv, b.dp[u]);
            float2 p = t == 0 ? make_float2(mm, 0.f) : pd;
            rr = __fadd_rn(rr, __fmul_rn(mm, p.x));
            ri = __fadd_rn(ri, __fmul_rn(mm, p.y));
            if (HIST && hr >= 0) gst(A.hist + hr * N + i, p);
            if (!last) p = cmul_diag(p, b.dc[u]);
            *reinterpret_cast<float2 *>(smem + ((uint32_t)i * G * 8u + gofs)) = p;
          }
        }
      };
      auto pass = [&](auto hist) {
        int b0 = 0;
        if constexpr (kPF) {  // the first batch is in nb
          use(nb, q, std::false_type{}, hist);
          b0 = kSiteU * R;
        }
        for (; b0 + kSiteU * R <= N; b0 += kSiteU * R) {
          SiteBatch b;
          load(b, f, b0 + q, std::false_type{});
          use(b, b0 + q, std::false_type{}, hist);
        }
        if constexpr (!kPF) {
          if (b0 < N) {
            SiteBatch b;
            load(b, f, b0 + q, std::true_type{});
            use(b, b0 + q, std::true_type{}, hist);
          }
        }
      };
      if constexpr (PF) pass(std::false_type{});  // PF launches write no history
      else if (A.hist) pass(std::true_type{});
      else pass(std::false_type{});
    }
    // response of each vector at its relative time: sum over its rows
#pragma unroll
    for (int off = G; off < 64; off <<= 1) {
      rr += __shfl_xor(rr, off);
      ri += __shfl_xor(ri, off);
    }
    if (act && q == 0) gst(A.partial + (int64_t)v * T + t, make_double2((double)rr, (double)ri));
    // kPF: the next frame's first batch, in flight during the sweeps (the
    // frame clamped to the group's frames: every lane loads valid rows)
    if constexpr (kPF) load(nb, f + 1, q, std::false_type{});
    if (!__ballot(act && !last)) continue;  // nothing propagates past this frame

    for (int sub = 0; sub < m; ++sub) {
      if (sub > 0) {
        const float2 *Dc = A.diag + row * N;
        for (int i0 = q; i0 < N; i0 += kSiteU * R) {
          float2 dc[kSiteU];
#pragma unroll
          for (int u = 0; u < kSiteU; ++u) {
            const int i = i0 + u * R;
            dc[u] = gld(Dc + (i < N ? i : 0));
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < kSiteU; ++u) {
            const int i = i0 + u * R;
            if (i < N) {
              float2 *pp = reinterpret_cast<float2 *>(smem + ((uint32_t)i * G * 8u + gofs));
              *pp = cmul_diag(cmul_diag(*pp, dc[u]), dc[u]);
            }
          }
        }
      }
      const float2 *csf = A.cs + row * K + loff;
      if (ring_key != tau * m + sub) {  // the previous sweep prefetched another one
#pragma unroll
        for (int k = 0; k < kRingS; ++k) {
          re[k] = gld(A.steps + loff + B * k * R);
          rc[k] = gld(csf + B * k * R);
        }
      }
      // Each ring slot is consumed and refilled with the entries kRingL
      // layers ahead: from this sweep, and in the last block (nlayers is a
      // multiple of kRingL) from the first block of the next sweep, which
      // runs this frame again or the next frame.
      const uint2 *sa = A.steps + loff + kRingL * R;
      const float2 *ca = csf + kRingL * R;
      for (int b0 = 0; b0 < nl - kRingL; b0 += kRingL, sa += kRingL * R, ca += kRingL * R) {
#pragma unroll
        for (int k = 0; k < kRingS; ++k) {
          rot_block<B>(smem, gofs, re[k], rc[k]);
          re[k] = gld(sa + B * k * R);
          rc[k] = gld(ca + B * k * R);
          __builtin_amdgcn_sched_barrier(0);  // the refill stays here
        }
      }
      const float2 *cn = A.cs + (sub + 1 < m ? row : min(f + 1, Fend) - A.f_lo) * K + loff;
#pragma unroll
      for (int k = 0; k < kRingS; ++k) {
        rot_block<B>(smem, gofs, re[k], rc[k]);
        re[k] = gld(A.steps + loff + B * k * R);
        rc[k] = gld(cn + B * k * R);
        __builtin_amdgcn_sched_barrier(0);
      }
      ring_key = tau * m + sub + 1;
    }
  }
}

// Workgroup w runs the vectors wmap[w].y .. wmap[w].y + G - 1 of system
// wmap[w].x (systems of different sizes share one launch), all systems with
// R rows; rows_mixed_kernel takes each system's own count (sys[k].rows).
// The descriptor is wave-uniform: it is read with scalar loads and kept in
// registers (the kernel's own stores cannot alias it).  At most 168 VGPRs
// (three waves per SIMD): the prefetched batch otherwise takes the chain
// kernels to 174-183 and two waves per SIMD (at 2 the N = 256 lattice ran
// 16.3 ms); the few values this spills are reloaded once per frame.
template <int R, bool PF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
rows_kernel(const RowsArgs *__restrict__ sys, const int2 *__restrict__ wmap, unsigned long long *times) {
  const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
  const int2 wm0 = wmap[blockIdx.x];
  const int2 wm = make_int2(__builtin_amdgcn_readfirstlane(wm0.x), __builtin_amdgcn_readfirstlane(wm0.y));
  const RowsArgs A = sys[wm.x];
  rows_run<R, true, PF>(A, wm);
  if (threadIdx.x == 0 && times) {  // device wall clock at start and end
    times[2 * blockIdx.x] = t_start;
    times[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
  }
}
__global__ void __launch_bounds__(64) rows_mixed_kernel(const RowsArgs *__restrict__ sys,
                                                        const int2 *__restrict__ wmap, unsigned long long *times) {
  const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
  const int2 wm0 = wmap[blockIdx.x];
  const int2 wm = make_int2(__builtin_amdgcn_readfirstlane(wm0.x), __builtin_amdgcn_readfirstlane(wm0.y));
  const RowsArgs A = sys[wm.x];
  switch (A.rows) {
    case 1: rows_run<1>(A, wm); break;
    case 2: rows_run<2>(A, wm); break;
    case 4: rows_run<4>(A, wm); break;
    case 8: rows_run<8>(A, wm); break;
    case 16: rows_run<16>(A, wm); break;
    case 32: rows_run<32>(A, wm); break;
    default: rows_run<64>(A, wm); break;  // rows_prepare admits 1, 2, ..., 64 only
  }
  if (threadIdx.x == 0 && times) {
    times[2 * blockIdx.x] = t_start;
    times[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
  }
}

size_t rows_smem(int N, int R) { return (size_t)(N + R) * (64 / R) * 8; }

template <bool PF>
static const void *rows_fn_pf(int R) {
  return R == 1 ? (const void *)rows_kernel<1, PF>
       : R == 2 ? (const void *)rows_kernel<2, PF>
       : R == 4 ? (const void *)rows_kernel<4, PF>
       : R == 8 ? (const void *)rows_kernel<8, PF>
       : R == 16 ? (const void *)rows_kernel<16, PF>
       : R == 32 ? (const void *)rows_kernel<32, PF>
       : R == 64 ? (const void *)rows_kernel<64, PF> : nullptr;
}
static const void *rows_fn(int R, bool pf = false) { return pf ? rows_fn_pf<true>(R) : rows_fn_pf<false>(R); }
// the sites a wave's first per-site batch spans (site_batch(R) * R)
static int site_span(int R) {
  const int u = R == 1 ? site_batch<1>() : R == 2 ? site_batch<2>() : R == 4 ? site_batch<4>()
              : R == 8 ? site_batch<8>() : R == 16 ? site_batch<16>() : R == 32 ? site_batch<32>()
              : site_batch<64>();
  return u * R;
}

// Checks the systems against the row count (R = 0: each system's own,
// sys[k].rows) and builds the workgroup map (G = 64 / R consecutive vectors
// of one system per workgroup) and the LDS size.
// pf: R > 0, every system's N a multiple of the per-site batch span and
// no history rows (the launch then takes the kernel that prefetches each
// frame's first batch).
int rows_prepare(const RowsArgs *sys, int nsys, int R, std::vector<int2> &wmap, size_t &smem, bool &pf) {
  if (R && !rows_fn(R)) return set_error("row conveyor: unsupported row count %d", R), NISE_ERR_ARG;
  smem = 0;
  wmap.clear();
  pf = R > 0;
  for (int k = 0; k < nsys; ++k) pf = pf && sys[k].N % site_span(R) == 0 && !sys[k].hist;
  for (int k = 0; k < nsys; ++k) {
    const RowsArgs &a = sys[k];
    const int Rk = R ? R : a.rows;
    if (!rows_fn(Rk)) return set_error("row conveyor: unsupported row count %d", Rk), NISE_ERR_ARG;
    const int G = 64 / Rk;
    if (a.nlayers < kRingL || a.nlayers % kRingL)
      return set_error("row conveyor: layer count %d is not a multiple of %d", a.nlayers, kRingL), NISE_ERR_ARG;
    if (a.K != (int64_t)a.nlayers * Rk) return set_error("row conveyor: schedule built for other rows"), NISE_ERR_ARG;
    smem = std::max(smem, rows_smem(a.N, Rk));
    for (int v = 0; v < a.nvec; v += G) wmap.push_back(make_int2(k, v));
  }
  if (smem > 160 * 1024) return set_error("row conveyor: %zu B of LDS needed", smem), NISE_ERR_ARG;
  return NISE_OK;
}

// sys / wmap: device copies (rows_prepare), nwg workgroups; R = 0: each
// system's own row count; pf as rows_prepare set it.
int launch_rows(const RowsArgs *sys, const int2 *wmap, int nwg, int R, bool pf, size_t smem, hipStream_t st,
                unsigned long long *times) {
  const void *fn = R ? rows_fn(R, pf) : (const void *)rows_mixed_kernel;
  if (!fn) return set_error("row conveyor: unsupported row count %d", R), NISE_ERR_ARG;
  if (nwg == 0) return NISE_OK;
  NISE_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  void *kargs[] = {&sys, &wmap, &times};
  NISE_HIP_CHECK(hipLaunchKernel(fn, dim3((unsigned)nwg), dim3(64), kargs, smem, st));
  return NISE_OK;
}

// Row count for a schedule: the power of two closest to the mean number of
// rotations per layer of the unconstrained layering, at most cap, raised
// while the LDS image would exceed 80 KiB (so that two waves fit on a CU).
int rows_for(int N, int Kreal, int nlayers_asap, int cap) {
  const double per = nlayers_asap ? (double)Kreal / nlayers_asap : 1.0;
  int R = 1;
  while (R < cap && R * 1.5 < per) R <<= 1;
  while (R < cap && rows_smem(N, R) > 80 * 1024) R <<= 1;
  return R;
}

// Device order of a row schedule's entries (the order of the step table and
// of every per-frame table built from it): with B = min(64 / R, 4), entry q
// of layer l sits at (l / B) * B R + B q + l % B, so the B layers a group
// loads at once are one contiguous run and a wave's load of a B-layer block
// is contiguous.
void rows_device_order(const Schedule &s, int R, std::vector<uint32_t> &ab) {
  ab.resize(s.ab.size());
  const int nl = s.nlayers(), B = std::min(64 / R, 4);
  for (int l = 0; l < nl; ++l)
    for (int q = 0; q < R; ++q) ab[(size_t)(l / B) * B * R + B * q + l % B] = s.ab[(size_t)l * R + q];
}

// Static step table of a row schedule in device order: LDS byte offsets of
// the two sites of every entry for a wave of 64/R vectors.
void rows_step_table(const Schedule &s, int R, std::vector<uint2> &tab) {
  const uint32_t G = 64 / R;
  std::vector<uint32_t> ab;
  rows_device_order(s, R, ab);
  tab.resize(ab.size());
  for (size_t i = 0; i < ab.size(); ++i) tab[i] = make_uint2((ab[i] & 0xffffu) * G * 8u, (ab[i] >> 16) * G * 8u);
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
