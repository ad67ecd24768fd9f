// Trotter (coupling-scheme) singles kernels for gfx950.
//
// Reference algorithm: propagate_vec_coupling_S, src/propagate.c:821-899:
//   psi <- [ D . R_kmax ... R_1 . D ]^m psi
//   D   = diag(cos(f H_aa / 2) - i sin(f H_aa / 2))
//   R_k = exp(-i f J_k sigma_x) on the pair (a_k, b_k), a_k < b_k in
//         lexicographic order, applied sequentially.
//
// Work split on the GPU:
//   frame_tables  : per frame, D (float2 per site) and (cos, sin) of every
//                   scheduled rotation, in the reference's libm precision
//                   (double trig on the float angle).  One pass per frame,
//                   shared by every vector that goes through the frame.
//   conveyor      : frame-major kernel.  A workgroup owns up to VP vectors
//                   (slots) whose windows overlap in time and walks the
//                   frames in order; each frame is read once for all its
//                   slots.  Slot state lives in LDS as psi[site][slot]
//                   (float2), so the VP lanes that serve one rotation read
//                   one contiguous 8*VP byte row.  Rotations run layer by
//                   layer (disjoint pairs per layer), one barrier per layer.
//   plain         : one sample (3 polarisations) per launch, every frame
//                   re-read and re-compacted per step, rotations in one
//                   sequential sweep: the per-sample GPU baseline.
//
// Arithmetic: compiled with -ffp-contract=off so products and sums round
// exactly as the reference's x86-64 SSE code (no FMA) does.

#include "nise_internal.h"

namespace nise {

// ---------------------------------------------------------------------------
// Union of kept couplings over a frame range (one byte per a<b pair).
// ---------------------------------------------------------------------------
// Workgroup of 256 threads: kMaskPairs consecutive pairs (a row segment of
// a frame per load) x kMaskLanes frame lanes; lane fl reads frames fl, fl +
// kMaskLanes, ..., kMaskInflight of them in flight, and stops at the first
// kept coupling.  The lanes' flags are ORed in LDS.
constexpr int kMaskPairs = 32, kMaskLanes = 8, kMaskInflight = 8;
// (a, b) of the lexicographic pair index p (a < b < N): the row from the
// quadratic's root, then corrected in integers (row a starts at a (2N - a -
// 1) / 2).  A walk along the rows took up to N steps per thread.
__device__ __forceinline__ void pair_of(int64_t p, int N, int &a, int &b) {
  auto start = [&](int64_t r) { return r * (2 * (int64_t)N - r - 1) / 2; };
  const double c = 2.0 * N - 1.0;
  int64_t r = (int64_t)((c - sqrt(fmax(c * c - 8.0 * (double)p, 0.0))) * 0.5);
  r = r < 0 ? 0 : r > N - 2 ? N - 2 : r;
  while (r > 0 && start(r) > p) --r;
  while (r < N - 2 && start(r + 1) <= p) ++r;
  a = (int)r;
  b = a + 1 + (int)(p - start(r));
}

// Static couplings (one record): one thread per pair.
__global__ void __launch_bounds__(256) union_mask_static_kernel(FrameSource s, int N, int64_t f_lo, uint8_t *mask) {
  const int64_t npair = (int64_t)N * (N - 1) / 2;
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= npair) return;
  int a, b;
  pair_of(p, N, a, b);
  mask[p] = (uint8_t)(kept_coupling(s, N, f_lo, a, b) != 0.f);
}

__global__ void __launch_bounds__(256) union_mask_kernel(FrameSource s, int N, int64_t f_lo, int64_t nfr,
                                                         uint8_t *mask) {
  __shared__ uint8_t kk[kMaskLanes][kMaskPairs];
  const int tp = threadIdx.x % kMaskPairs, fl = threadIdx.x / kMaskPairs;
  const int64_t npair = (int64_t)N * (N - 1) / 2;
  const int64_t p = (int64_t)blockIdx.x * kMaskPairs + tp;
  uint8_t keep = 0;
  if (p < npair) {
    int a, b;
    pair_of(p, N, a, b);
    const int64_t nf = s.coupling_type ? 1 : nfr;
    for (int64_t f0 = fl; f0 < nf && !keep; f0 += (int64_t)kMaskLanes * kMaskInflight) {
      float v[kMaskInflight];
#pragma unroll
      for (int u = 0; u < kMaskInflight; ++u) {
        const int64_t f = f0 + (int64_t)u * kMaskLanes;
        v[u] = f < nf ? kept_coupling(s, N, f_lo + f, a, b) : 0.f;
      }
#pragma unroll
      for (int u = 0; u < kMaskInflight; ++u) keep |= (uint8_t)(v[u] != 0.f);
    }
  }
  kk[fl][tp] = keep;
  __syncthreads();
  if (fl == 0 && p < npair) {
    uint8_t r = 0;
    for (int l = 0; l < kMaskLanes; ++l) r |= kk[l][tp];
    mask[p] = r;
  }
}

int launch_union_mask(const FrameSource &src, int N, int64_t f_lo, int64_t nfr, uint8_t *mask_dev,
                      hipStream_t st) {
  int64_t npair = (int64_t)N * (N - 1) / 2;
  if (npair == 0) return NISE_OK;
  if (src.coupling_type) {
    hipLaunchKernelGGL(union_mask_static_kernel, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, st, src, N, f_lo,
                       mask_dev);
    NISE_HIP_CHECK(hipGetLastError());
    return NISE_OK;
  }
  dim3 grid((unsigned)((npair + kMaskPairs - 1) / kMaskPairs));
  hipLaunchKernelGGL(union_mask_kernel, grid, dim3(256), 0, st, src, N, f_lo, nfr, mask_dev);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

// ---------------------------------------------------------------------------
// Per-frame tables: D phases and rotation (cos, sin) in schedule order.
// ---------------------------------------------------------------------------
__global__ void frame_tables_kernel(FrameSource s, int N, int K, const uint32_t *ab, float fac,
                                    int64_t f_lo, float2 *diag, float2 *cs, float *jf) {
  int64_t row = blockIdx.y;
  int64_t f = f_lo + row;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < N) {
    // src/propagate.c:863-866: cos/sin of (0.5 * H0 * f) in double
    float h = site_energy(s, N, f, j);
    if (s.delta) h = __fadd_rn(h, s.delta[j]);
    double arg = 0.5 * (double)h * (double)fac;
    diag[row * N + j] = make_float2((float)cos(arg), (float)(-sin(arg)));
  } else if (j < N + K) {
    int r = j - N;
    uint32_t p = ab[r];
    int a = (int)(p & 0xffffu), b = (int)(p >> 16);
    float J = a < N ? kept_coupling(s, N, f, a, b) : 0.f;  // a >= N: padding
    float co = 1.f, si = 0.f;  // absent coupling: exact identity rotation
    if (J != 0.f) {
      // src/propagate.c:879-882: J = J*f; si = -sin(J); co = sqrt(1 - si*si)
      float Jf = __fmul_rn(__fmul_rn(J, s.cscale), fac);
      si = (float)(-sin((double)Jf));
      co = (float)sqrt((double)__fsub_rn(1.0f, __fmul_rn(si, si)));
    }
    cs[row * K + r] = make_float2(co, si);
    if (jf) jf[row * K + r] = __fmul_rn(J, fac);
  }
}

int launch_frame_tables(const FrameSource &src, int N, const Schedule &sch, const uint32_t *ab_dev,
                        float fac, int64_t f_lo, int64_t nfr, float2 *diag_dev, float2 *cs_dev,
                        hipStream_t st, float *jf_dev) {
  int n = N + sch.K;
  int64_t done = 0;
  while (done < nfr) {  // grid.y is limited to 65535
    int64_t chunk = nfr - done < 65535 ? nfr - done : 65535;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)chunk);
    hipLaunchKernelGGL(frame_tables_kernel, grid, dim3(256), 0, st, src, N, sch.K, ab_dev, fac,
                       f_lo + done, diag_dev + done * N, cs_dev + done * sch.K,
                       jf_dev ? jf_dev + done * sch.K : nullptr);
    NISE_HIP_CHECK(hipGetLastError());
    done += chunk;
  }
  return NISE_OK;
}


// ---------------------------------------------------------------------------
// Conveyor kernel.
//
// Rotation work of one sweep is a list of steps (build_steps).  Step s
// covers schedule rotations [x, y & 0x7fffffff) (at most BLOCK/VP of them,
// all in one layer); a set sign bit of y closes the layer with a workgroup
// barrier.  Thread (q, v) applies rotation x + q of each step to slot v.
// The list is padded with empty steps to a multiple of kRing so that a
// register ring of the next kRing steps' operands ((cos, sin) and packed
// (a, b)) can run continuously across sweeps and frames; the barrier waits
// only for LDS traffic, so the ring's global loads stay in flight.
//
// Per frame, the sites' half-step phases D(f) and dipoles mu(f) are read
// from LDS; the next frame's are fetched into registers while the current
// frame's rotations run.
// ---------------------------------------------------------------------------
constexpr int kRing = 8;

// History row recorded at response step rel of a slot, or -1: every step
// when every == 1; else the checkpoints rel = 0, every, 2*every, ... and
// the last response step.
__device__ __forceinline__ int64_t hist_row(const SlotDesc &sd, int64_t rel, bool resp, int every) {
  if (!resp || sd.hist < 0) return -1;
  if (rel % every == 0) return (int64_t)sd.hist + rel / every;
  if (rel == sd.nresp - 1) return (int64_t)sd.hist + rel / every + 1;
  return -1;
}

struct ConvSmem {
  static __host__ __device__ size_t bytes(int N, int vp, int block, int tmax_acc, int nsteps) {
    return (size_t)N * vp * 8        // psi
           + (size_t)2 * N * 8       // D, two frames
           + (size_t)3 * N * 4       // mu
           + (size_t)tmax_acc * 16   // acc
           + (size_t)(block / 64) * vp * 8  // red
           + (size_t)nsteps * 8      // steps
           + (size_t)vp * sizeof(SlotDesc) + 64;
  }
};

template <int VP, int BLOCK>
__global__ void __launch_bounds__(BLOCK) conveyor_kernel(ConveyorArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N;
  constexpr int kWaves = BLOCK / 64;
  constexpr int kRows = BLOCK / VP;      // rotations per step, sites per pass
  constexpr int kPF = 4;                 // next-frame D prefetch per thread
  // byte offsets from smem keep every pointer in the LDS address space
  const size_t o_D = (size_t)N * VP * 8;
  const size_t o_mu = o_D + (size_t)2 * N * 8;
  const size_t o_acc = (o_mu + (size_t)3 * N * 4 + 15) & ~(size_t)15;
  const size_t o_red = o_acc + (size_t)A.tmax_acc * 16;
  const size_t o_steps = o_red + (size_t)kWaves * VP * 8;
  const size_t o_sls = o_steps + (size_t)A.nsteps * 8;
  float2 *psi = reinterpret_cast<float2 *>(smem);                 // [N][VP]
  float2 *Dsh = reinterpret_cast<float2 *>(smem + o_D);           // [2][N]
  float *mush = reinterpret_cast<float *>(smem + o_mu);           // [3][N]
  double2 *acc = reinterpret_cast<double2 *>(smem + o_acc);       // [tmax_acc]
  float2 *red = reinterpret_cast<float2 *>(smem + o_red);         // [waves][VP]
  int2 *steps = reinterpret_cast<int2 *>(smem + o_steps);         // [nsteps]
  SlotDesc *sls = reinterpret_cast<SlotDesc *>(smem + o_sls);     // [VP]

  const int g = blockIdx.x;
  const int s0 = A.wg_slot_begin[g];
  const int ns = A.wg_slot_begin[g + 1] - s0;
  const int t = threadIdx.x;
  const int v = t % VP;
  const int q = t / VP;
  const int lane = t & 63, wave = t >> 6;
  const int nsteps = A.nsteps;
  const int K = A.K;

  for (int i = t; i < nsteps; i += BLOCK) steps[i] = A.steps[i];
  for (int i = t; i < ns; i += BLOCK) sls[i] = A.slots[s0 + i];
  for (int i = t; i < A.tmax_acc; i += BLOCK) acc[i] = make_double2(0.0, 0.0);
  __syncthreads();

  const bool has = v < ns;
  SlotDesc sd = has ? sls[v] : SlotDesc{0, 0, 0, 0, 0, -1, 0, -1};
  int64_t F0 = INT64_MAX, F1 = INT64_MIN;
  for (int i = 0; i < ns; ++i) {
    const SlotDesc d = sls[i];
    const int len = d.nresp > d.nprop ? d.nresp : d.nprop;
    F0 = d.fs < F0 ? d.fs : F0;
    F1 = (int64_t)d.fs + len > F1 ? (int64_t)d.fs + len : F1;
  }
  if (ns == 0) return;

  // frame F0's D and mu into LDS
  {
    const int64_t row = F0 - A.f_lo;
    const float *mu = A.mu_rec + (A.mu_stride ? F0 * A.mu_stride : 0) + 1;
    for (int i = t; i < N; i += BLOCK) Dsh[(F0 & 1) * N + i] = A.diag[row * N + i];
    for (int i = t; i < 3 * N; i += BLOCK) mush[i] = mu[i];
  }
  // operand ring for the first sweep
  int2 rst[kRing];
  float2 rc[kRing];
  uint32_t rab[kRing];
  {
    const float2 *cs0 = A.cs + (F0 - A.f_lo) * K;
#pragma unroll
    for (int j = 0; j < kRing; ++j) {
      const int jj = j < nsteps ? j : 0;
      const int2 st = nsteps ? A.steps[jj] : make_int2(0, 0);
      const int r = max(0, min(st.x + q, (st.y & 0x7fffffff) - 1));
      rst[j] = st;
      rc[j] = nsteps ? cs0[r] : make_float2(1.f, 0.f);
      rab[j] = nsteps ? A.ab[r] : 0u;
    }
  }
  __syncthreads();

  for (int64_t f = F0; f < F1; ++f) {
    const int64_t row = f - A.f_lo;
    const int64_t rel = f - sd.fs;
    const bool starts = has && rel == 0;
    const bool prev_prop = has && rel >= 1 && rel - 1 < sd.nprop;
    const bool resp = has && rel >= 0 && rel < sd.nresp;
    const bool prop = has && rel >= 0 && rel < sd.nprop;
    const float2 *Dcur = Dsh + (f & 1) * N;
    const float2 *Dprev = Dsh + ((f + 1) & 1) * N;
    const float *mu = mush + sd.x * N;

    // Per-site pass: pending second half-phase of frame f-1, response
    // mu(f).psi (src/absorption.c:130, calc_S1 :205-212), first half-phase
    // of frame f.
    float rr = 0.f, ri = 0.f;
    const int64_t hr = hist_row(sd, rel, resp, A.hist_every);
    if (starts || prev_prop || resp || prop) {
      for (int i = q; i < N; i += kRows) {
        // the LDS values and both half-phases first, then selects: branches
        // around the LDS read made each site one LDS round trip (a starting
        // slot's state load and the history store stay conditional)
        const float2 ps = psi[i * VP + v];
        const float m = mu[i];
        const float2 pd = cmul_diag(ps, Dprev[i]);
        float2 sv = make_float2(m, 0.f);
        if (starts && sd.init != 0) sv = A.state[(int64_t)sd.out * N + i];
        float2 p = prev_prop ? pd : ps;
        p = starts ? sv : p;
        if (resp && hr >= 0) A.hist[hr * N + i] = p;
        const float rx = __fadd_rn(rr, __fmul_rn(m, p.x)), ry = __fadd_rn(ri, __fmul_rn(m, p.y));
        rr = resp ? rx : rr;
        ri = resp ? ry : ri;
        const float2 pc = cmul_diag(p, Dcur[i]);
        psi[i * VP + v] = prop ? pc : p;
      }
    }
    if (A.tmax_acc > 0) {
#pragma unroll
      for (int off = VP; off < 64; off <<= 1) {
        rr += __shfl_xor(rr, off);
        ri += __shfl_xor(ri, off);
      }
      if (lane < VP) red[wave * VP + lane] = make_float2(rr, ri);
      __syncthreads();
      // A group leader (the x=0 slot of a sample) adds its group's sums in a
      // fixed order; groups of one workgroup start on distinct frames, so
      // their response indices differ and the adds do not collide.
      if (q == 0 && resp && sd.grp > 0) {
        float sr = 0.f, si = 0.f;
        for (int gg = 0; gg < sd.grp; ++gg)
          for (int w = 0; w < kWaves; ++w) sr += red[w * VP + v + gg].x, si += red[w * VP + v + gg].y;
        acc[rel].x += (double)sr;
        acc[rel].y += (double)si;
      }
    }
    __syncthreads();

    // next frame's D and mu into registers (consumed after the sweeps)
    const int64_t fn = f + 1 < F1 ? f + 1 : f;
    float2 dn[kPF];
    float mn[3 * kPF];
#pragma unroll
    for (int k = 0; k < kPF; ++k) {
      const int i = min(t + k * BLOCK, N - 1);
      dn[k] = A.diag[(fn - A.f_lo) * N + i];
      const float *mup = A.mu_rec + (A.mu_stride ? fn * A.mu_stride : 0) + 1;
#pragma unroll
      for (int x = 0; x < 3; ++x) mn[3 * k + x] = mup[x * N + i];
    }

    for (int sub = 0; sub < A.m; ++sub) {
      if (sub > 0) {
        if (prop)
          for (int i = q; i < N; i += kRows)
            psi[i * VP + v] = cmul_diag(cmul_diag(psi[i * VP + v], Dcur[i]), Dcur[i]);
        __syncthreads();
      }
      // operands of the following sweep (next substep, else next frame)
      const float2 *csn = A.cs + ((sub + 1 < A.m) ? row : (fn - A.f_lo)) * K;
      for (int b0 = 0; b0 < nsteps; b0 += kRing) {
#pragma unroll
        for (int j = 0; j < kRing; ++j) {
          const int sidx = b0 + j;
          // the step entry is the same for every lane
          const int2 st = make_int2(__builtin_amdgcn_readfirstlane(rst[j].x),
                                    __builtin_amdgcn_readfirstlane(rst[j].y));
          if (prop && st.x + q < (st.y & 0x7fffffff)) {
            const float2 c = rc[j];
            const uint32_t pk = rab[j];
            const int a = (int)(pk & 0xffffu), b = (int)(pk >> 16);
            float2 pa = psi[a * VP + v], pb = psi[b * VP + v];
            rotate_pair(pa, pb, c.x, c.y);
            psi[a * VP + v] = pa;
            psi[b * VP + v] = pb;
          }
          {
            // refill this slot with step sidx + kRing of this sweep, or of
            // the next sweep once this one runs out (nsteps % kRing == 0)
            const int sp = sidx + kRing;
            const bool nxt = sp >= nsteps;
            const int2 stp = steps[nxt ? sp - nsteps : sp];
            const int rp = max(0, min(stp.x + q, (stp.y & 0x7fffffff) - 1));
            const float2 *src = nxt ? csn : A.cs + row * K;
            rst[j] = stp;
            rc[j] = src[rp];
            rab[j] = A.ab[rp];
          }
          if (st.y < 0) lds_barrier();
        }
      }
      __syncthreads();
    }
    // publish frame f+1's D and mu
#pragma unroll
    for (int k = 0; k < kPF; ++k) {
      const int i = t + k * BLOCK;
      if (i < N) {
        Dsh[(fn & 1) * N + i] = dn[k];
#pragma unroll
        for (int x = 0; x < 3; ++x) mush[x * N + i] = mn[3 * k + x];
      }
    }
    for (int i = t + kPF * BLOCK; i < N; i += BLOCK) {  // N > kPF*BLOCK
      Dsh[(fn & 1) * N + i] = A.diag[(fn - A.f_lo) * N + i];
      const float *mup = A.mu_rec + (A.mu_stride ? fn * A.mu_stride : 0) + 1;
      for (int x = 0; x < 3; ++x) mush[x * N + i] = mup[x * N + i];
    }
    __syncthreads();
  }

  // write back; the pending second half-phase of the last propagated frame
  // is still due only when the frame loop ended right after that frame
  if (A.writeback && has && sd.out >= 0 && sd.nprop > 0) {
    const int64_t row = (int64_t)sd.fs + sd.nprop - 1 - A.f_lo;
    const bool due = (int64_t)sd.fs + sd.nprop >= F1;
    for (int i = q; i < N; i += kRows) {
      const float2 p = psi[i * VP + v];
      A.state[(int64_t)sd.out * N + i] = due ? cmul_diag(p, A.diag[row * N + i]) : p;
    }
  }
  if (A.tmax_acc > 0)
    for (int i = t; i < A.tmax_acc; i += BLOCK) A.partial[(int64_t)g * A.tmax_acc + i] = acc[i];
}

// ---------------------------------------------------------------------------
// Single-wave conveyor.  A workgroup is one wave: LDS operations of one wave
// execute in order, so no barrier separates dependent layers.  A step is a
// chunk of one layer of at most (64/VP)*kBatch rotations; each lane applies
// up to kBatch of them to its slot, issuing all of their LDS reads before
// any arithmetic so that one LDS round trip serves kBatch rotations.
// ---------------------------------------------------------------------------
constexpr int kBatch = 4;
constexpr int kRing1 = 4;

__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }

template <int VP>
__global__ void __launch_bounds__(64) conveyor_w1_kernel(ConveyorArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N;
  constexpr int kRows = 64 / VP;
  const size_t o_D = (size_t)N * VP * 8;
  const size_t o_mu = o_D + (size_t)2 * N * 8;
  const size_t o_acc = (o_mu + (size_t)3 * N * 4 + 15) & ~(size_t)15;
  const size_t o_steps = o_acc + (size_t)A.tmax_acc * 16;
  const size_t o_sls = o_steps + (size_t)A.nsteps * 8;
  float2 *psi = reinterpret_cast<float2 *>(smem);
  float2 *Dsh = reinterpret_cast<float2 *>(smem + o_D);
  float *mush = reinterpret_cast<float *>(smem + o_mu);
  double2 *acc = reinterpret_cast<double2 *>(smem + o_acc);
  int2 *steps = reinterpret_cast<int2 *>(smem + o_steps);
  SlotDesc *sls = reinterpret_cast<SlotDesc *>(smem + o_sls);

  const int g = blockIdx.x;
  const int s0 = A.wg_slot_begin[g];
  const int ns = A.wg_slot_begin[g + 1] - s0;
  const int t = threadIdx.x;
  const int v = t % VP, q = t / VP;
  const int nsteps = A.nsteps;
  const int K = A.K;
  for (int i = t; i < nsteps; i += 64) steps[i] = A.steps[i];
  for (int i = t; i < ns; i += 64) sls[i] = A.slots[s0 + i];
  for (int i = t; i < A.tmax_acc; i += 64) acc[i] = make_double2(0.0, 0.0);
  __syncthreads();
  if (ns == 0) return;
  const bool has = v < ns;
  const SlotDesc sd = has ? sls[v] : SlotDesc{0, 0, 0, 0, 0, -1, 0, -1};
  int64_t F0 = INT64_MAX, F1 = INT64_MIN;
  for (int i = 0; i < ns; ++i) {
    const SlotDesc d = sls[i];
    const int len = d.nresp > d.nprop ? d.nresp : d.nprop;
    F0 = d.fs < F0 ? d.fs : F0;
    F1 = (int64_t)d.fs + len > F1 ? (int64_t)d.fs + len : F1;
  }
  {
    const int64_t row = F0 - A.f_lo;
    const float *mu = A.mu_rec + (A.mu_stride ? F0 * A.mu_stride : 0) + 1;
    for (int i = t; i < N; i += 64) Dsh[(F0 & 1) * N + i] = A.diag[row * N + i];
    for (int i = t; i < 3 * N; i += 64) mush[i] = mu[i];
  }
  // operand ring: kRing1 steps x kBatch rounds
  int2 rst[kRing1];
  float2 rc[kRing1][kBatch];
  uint32_t rab[kRing1][kBatch];
  {
    const float2 *cs0 = A.cs + (F0 - A.f_lo) * K;
#pragma unroll
    for (int j = 0; j < kRing1; ++j) {
      const int2 st = nsteps ? A.steps[j < nsteps ? j : 0] : make_int2(0, 0);
      rst[j] = st;
      const int e = st.y & 0x7fffffff;
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int r = max(0, min(st.x + i * kRows + q, e - 1));
        rc[j][i] = nsteps ? cs0[r] : make_float2(1.f, 0.f);
        rab[j][i] = nsteps ? A.ab[r] : 0u;
      }
    }
  }
  __syncthreads();

  for (int64_t f = F0; f < F1; ++f) {
    const int64_t row = f - A.f_lo;
    const int64_t rel = f - sd.fs;
    const bool starts = has && rel == 0;
    const bool prev_prop = has && rel >= 1 && rel - 1 < sd.nprop;
    const bool resp = has && rel >= 0 && rel < sd.nresp;
    const bool prop = has && rel >= 0 && rel < sd.nprop;
    const float2 *Dcur = Dsh + (f & 1) * N;
    const float2 *Dprev = Dsh + ((f + 1) & 1) * N;
    const float *mu = mush + sd.x * N;
    float rr = 0.f, ri = 0.f;
    const int64_t hr = hist_row(sd, rel, resp, A.hist_every);
    if (starts || prev_prop || resp || prop) {
      for (int i = q; i < N; i += kRows) {
        // the LDS values and both half-phases first, then selects: branches
        // around the LDS read made each site one LDS round trip (a starting
        // slot's state load and the history store stay conditional)
        const float2 ps = psi[i * VP + v];
        const float m = mu[i];
        const float2 pd = cmul_diag(ps, Dprev[i]);
        float2 sv = make_float2(m, 0.f);
        if (starts && sd.init != 0) sv = A.state[(int64_t)sd.out * N + i];
        float2 p = prev_prop ? pd : ps;
        p = starts ? sv : p;
        if (resp && hr >= 0) A.hist[hr * N + i] = p;
        const float rx = __fadd_rn(rr, __fmul_rn(m, p.x)), ry = __fadd_rn(ri, __fmul_rn(m, p.y));
        rr = resp ? rx : rr;
        ri = resp ? ry : ri;
        const float2 pc = cmul_diag(p, Dcur[i]);
        psi[i * VP + v] = prop ? pc : p;
      }
    }
    if (A.tmax_acc > 0) {
#pragma unroll
      for (int off = VP; off < 64; off <<= 1) {
        rr += __shfl_xor(rr, off);
        ri += __shfl_xor(ri, off);
      }
      // group leader (x = 0 slot of a sample) sums its polarisations in order
      float gr = rr, gi = ri;
      for (int gg = 1; gg < 3; ++gg) {
        const float xr = __shfl(rr, (t + gg) & 63), xi = __shfl(ri, (t + gg) & 63);
        if (gg < sd.grp) gr += xr, gi += xi;
      }
      if (q == 0 && resp && sd.grp > 0) {
        acc[rel].x += (double)gr;
        acc[rel].y += (double)gi;
      }
    }
    compiler_fence();

    const int64_t fn = f + 1 < F1 ? f + 1 : f;
    float2 dn[4];
    float mn[12];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = min(t + k * 64, N - 1);
      dn[k] = A.diag[(fn - A.f_lo) * N + i];
      const float *mup = A.mu_rec + (A.mu_stride ? fn * A.mu_stride : 0) + 1;
#pragma unroll
      for (int x = 0; x < 3; ++x) mn[3 * k + x] = mup[x * N + i];
    }

    for (int sub = 0; sub < A.m; ++sub) {
      if (sub > 0) {
        if (prop)
          for (int i = q; i < N; i += kRows)
            psi[i * VP + v] = cmul_diag(cmul_diag(psi[i * VP + v], Dcur[i]), Dcur[i]);
        compiler_fence();
      }
      const float2 *csn = A.cs + ((sub + 1 < A.m) ? row : (fn - A.f_lo)) * K;
      const float2 *csc = A.cs + row * K;
      for (int b0 = 0; b0 < nsteps; b0 += kRing1) {
#pragma unroll
        for (int j = 0; j < kRing1; ++j) {
          const int sidx = b0 + j;
          const int sx = __builtin_amdgcn_readfirstlane(rst[j].x);
          const int se = __builtin_amdgcn_readfirstlane(rst[j].y) & 0x7fffffff;
          float2 pa[kBatch], pb[kBatch];
          bool on[kBatch];
#pragma unroll
          for (int i = 0; i < kBatch; ++i) {
            on[i] = prop && sx + i * kRows + q < se;
            const uint32_t pk = rab[j][i];
            const int a = (int)(pk & 0xffffu), b = (int)(pk >> 16);
            if (on[i]) pa[i] = psi[a * VP + v], pb[i] = psi[b * VP + v];
          }
#pragma unroll
          for (int i = 0; i < kBatch; ++i) {
            if (on[i]) {
              const uint32_t pk = rab[j][i];
              const int a = (int)(pk & 0xffffu), b = (int)(pk >> 16);
              rotate_pair(pa[i], pb[i], rc[j][i].x, rc[j][i].y);
              psi[a * VP + v] = pa[i];
              psi[b * VP + v] = pb[i];
            }
          }
          compiler_fence();
          {
            const int sp = sidx + kRing1;
            const bool nxt = sp >= nsteps;
            const int2 stp = steps[nxt ? sp - nsteps : sp];
            const int e = stp.y & 0x7fffffff;
            const float2 *src = nxt ? csn : csc;
            rst[j] = stp;
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
              const int rp = max(0, min(stp.x + i * kRows + q, e - 1));
              rc[j][i] = src[rp];
              rab[j][i] = A.ab[rp];
            }
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = t + k * 64;
      if (i < N) {
        Dsh[(fn & 1) * N + i] = dn[k];
#pragma unroll
        for (int x = 0; x < 3; ++x) mush[x * N + i] = mn[3 * k + x];
      }
    }
    for (int i = t + 256; i < N; i += 64) {
      Dsh[(fn & 1) * N + i] = A.diag[(fn - A.f_lo) * N + i];
      const float *mup = A.mu_rec + (A.mu_stride ? fn * A.mu_stride : 0) + 1;
      for (int x = 0; x < 3; ++x) mush[x * N + i] = mup[x * N + i];
    }
    compiler_fence();
  }
  // write back; the pending second half-phase of the last propagated frame
  // is still due only when the frame loop ended right after that frame
  if (A.writeback && has && sd.out >= 0 && sd.nprop > 0) {
    const int64_t row = (int64_t)sd.fs + sd.nprop - 1 - A.f_lo;
    const bool due = (int64_t)sd.fs + sd.nprop >= F1;
    for (int i = q; i < N; i += kRows) {
      const float2 p = psi[i * VP + v];
      A.state[(int64_t)sd.out * N + i] = due ? cmul_diag(p, A.diag[row * N + i]) : p;
    }
  }
  if (A.tmax_acc > 0)
    for (int i = t; i < A.tmax_acc; i += 64) A.partial[(int64_t)g * A.tmax_acc + i] = acc[i];
}

// ---------------------------------------------------------------------------
// Single-wave conveyor with S slots per lane.  Lane (row q, group l) applies
// rotation x + q of each step to slots l*S .. l*S+S-1, which sit next to
// each other in LDS (psi[site][slot]), so one 8*S-byte access moves all of
// them.  A step covers up to 64*S/VP rotations of one layer.  No barriers:
// the workgroup is one wave and its LDS operations execute in order.
// ---------------------------------------------------------------------------
template <int S> struct SlotVec;
template <> struct SlotVec<2> { using T = float4; };
template <> struct SlotVec<4> { using T = float4; };

template <int S>
__device__ __forceinline__ void load_slots(const float2 *p, float2 (&o)[S]) {
  const float4 *p4 = reinterpret_cast<const float4 *>(p);
#pragma unroll
  for (int k = 0; k < S / 2; ++k) {
    const float4 w = p4[k];
    o[2 * k] = make_float2(w.x, w.y);
    o[2 * k + 1] = make_float2(w.z, w.w);
  }
}

template <int S>
__device__ __forceinline__ void store_slots(float2 *p, const float2 (&o)[S]) {
  float4 *p4 = reinterpret_cast<float4 *>(p);
#pragma unroll
  for (int k = 0; k < S / 2; ++k) p4[k] = make_float4(o[2 * k].x, o[2 * k].y, o[2 * k + 1].x, o[2 * k + 1].y);
}

constexpr int kRingS = 8;

template <int VP, int S>
__global__ void __launch_bounds__(64) conveyor_s_kernel(ConveyorArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N;
  constexpr int LR = VP / S;        // lanes per rotation row
  constexpr int R = 64 / LR;        // rotation rows per step
  const size_t o_D = (size_t)N * VP * 8;
  const size_t o_mu = o_D + (size_t)2 * N * 8;
  const size_t o_acc = (o_mu + (size_t)3 * N * 4 + 15) & ~(size_t)15;
  const size_t o_red = o_acc + (size_t)A.tmax_acc * 16;
  const size_t o_steps = o_red + (size_t)VP * 8;
  const size_t o_sls = o_steps + (size_t)A.nsteps * 8;
  float2 *psi = reinterpret_cast<float2 *>(smem);
  float2 *Dsh = reinterpret_cast<float2 *>(smem + o_D);
  float *mush = reinterpret_cast<float *>(smem + o_mu);
  double2 *acc = reinterpret_cast<double2 *>(smem + o_acc);
  float2 *red = reinterpret_cast<float2 *>(smem + o_red);
  int2 *steps = reinterpret_cast<int2 *>(smem + o_steps);
  SlotDesc *sls = reinterpret_cast<SlotDesc *>(smem + o_sls);

  const int g = blockIdx.x;
  const int s0 = A.wg_slot_begin[g];
  const int ns = A.wg_slot_begin[g + 1] - s0;
  const int t = threadIdx.x;
  const int l = t % LR, q = t / LR;
  const int v0 = l * S;                 // first slot of this lane
  const int nsteps = A.nsteps;
  const int K = A.K;
  for (int i = t; i < nsteps; i += 64) steps[i] = A.steps[i];
  for (int i = t; i < ns; i += 64) sls[i] = A.slots[s0 + i];
  for (int i = t; i < A.tmax_acc; i += 64) acc[i] = make_double2(0.0, 0.0);
  __syncthreads();
  if (ns == 0) return;
  SlotDesc sd[S];
#pragma unroll
  for (int k = 0; k < S; ++k) sd[k] = (v0 + k < ns) ? sls[v0 + k] : SlotDesc{0, 0, 0, 0, 0, -1, 0, -1};
  int64_t F0 = INT64_MAX, F1 = INT64_MIN;
  for (int i = 0; i < ns; ++i) {
    const SlotDesc d = sls[i];
    const int len = d.nresp > d.nprop ? d.nresp : d.nprop;
    F0 = d.fs < F0 ? d.fs : F0;
    F1 = (int64_t)d.fs + len > F1 ? (int64_t)d.fs + len : F1;
  }
  {
    const int64_t row = F0 - A.f_lo;
    const float *mu = A.mu_rec + (A.mu_stride ? F0 * A.mu_stride : 0) + 1;
    for (int i = t; i < N; i += 64) Dsh[(F0 & 1) * N + i] = A.diag[row * N + i];
    for (int i = t; i < 3 * N; i += 64) mush[i] = mu[i];
  }
  int2 rst[kRingS];
  float2 rc[kRingS];
  uint32_t rab[kRingS];
  {
    const float2 *cs0 = A.cs + (F0 - A.f_lo) * K;
#pragma unroll
    for (int j = 0; j < kRingS; ++j) {
      const int2 st = nsteps ? A.steps[j < nsteps ? j : 0] : make_int2(0, 0);
      const int r = max(0, min(st.x + q, (st.y & 0x7fffffff) - 1));
      rst[j] = st;
      rc[j] = nsteps ? cs0[r] : make_float2(1.f, 0.f);
      rab[j] = nsteps ? A.ab[r] : 0u;
    }
  }
  __syncthreads();

  for (int64_t f = F0; f < F1; ++f) {
    const int64_t row = f - A.f_lo;
    bool starts[S], pprop[S], resp[S], prop[S];
    int64_t hrk[S];
    bool any_site = false, any_prop = false;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const bool has = v0 + k < ns;
      const int64_t rel = f - sd[k].fs;
      starts[k] = has && rel == 0;
      pprop[k] = has && rel >= 1 && rel - 1 < sd[k].nprop;
      resp[k] = has && rel >= 0 && rel < sd[k].nresp;
      prop[k] = has && rel >= 0 && rel < sd[k].nprop;
      hrk[k] = hist_row(sd[k], rel, resp[k], A.hist_every);
      any_site |= starts[k] || pprop[k] || resp[k] || prop[k];
      any_prop |= prop[k];
    }
    const float2 *Dcur = Dsh + (f & 1) * N;
    const float2 *Dprev = Dsh + ((f + 1) & 1) * N;
    float rr[S], ri[S];
#pragma unroll
    for (int k = 0; k < S; ++k) rr[k] = 0.f, ri[k] = 0.f;
    if (any_site) {
      for (int i = q; i < N; i += R) {
        float2 p[S];
        load_slots<S>(psi + i * VP + v0, p);
        const float2 dp = Dprev[i], dc = Dcur[i];
#pragma unroll
        for (int k = 0; k < S; ++k) {
          const float m = mush[sd[k].x * N + i];
          if (starts[k]) p[k] = (sd[k].init == 0) ? make_float2(m, 0.f) : A.state[(int64_t)sd[k].out * N + i];
          else if (pprop[k]) p[k] = cmul_diag(p[k], dp);
          if (resp[k]) {
            rr[k] = __fadd_rn(rr[k], __fmul_rn(m, p[k].x));
            ri[k] = __fadd_rn(ri[k], __fmul_rn(m, p[k].y));
            if (hrk[k] >= 0) A.hist[hrk[k] * N + i] = p[k];
          }
          if (prop[k]) p[k] = cmul_diag(p[k], dc);
        }
        store_slots<S>(psi + i * VP + v0, p);
      }
    }
    if (A.tmax_acc > 0) {
#pragma unroll
      for (int k = 0; k < S; ++k) {
#pragma unroll
        for (int off = LR; off < 64; off <<= 1) {
          rr[k] += __shfl_xor(rr[k], off);
          ri[k] += __shfl_xor(ri[k], off);
        }
      }
      if (q == 0) {
#pragma unroll
        for (int k = 0; k < S; ++k) red[v0 + k] = make_float2(rr[k], ri[k]);
      }
      compiler_fence();
      // one lane per slot: group leaders (x = 0 slot of a sample) sum their
      // polarisations in order; groups start on distinct frames
      if (t < ns) {
        const SlotDesc d = sls[t];
        const int64_t rel = f - d.fs;
        if (d.grp > 0 && rel >= 0 && rel < d.nresp) {
          float gr = 0.f, gi = 0.f;
          for (int gg = 0; gg < d.grp; ++gg) gr += red[t + gg].x, gi += red[t + gg].y;
          acc[rel].x += (double)gr;
          acc[rel].y += (double)gi;
        }
      }
    }
    compiler_fence();

    const int64_t fn = f + 1 < F1 ? f + 1 : f;
    float2 dn[4];
    float mn[12];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = min(t + k * 64, N - 1);
      dn[k] = A.diag[(fn - A.f_lo) * N + i];
      const float *mup = A.mu_rec + (A.mu_stride ? fn * A.mu_stride : 0) + 1;
#pragma unroll
      for (int x = 0; x < 3; ++x) mn[3 * k + x] = mup[x * N + i];
    }

    for (int sub = 0; sub < A.m; ++sub) {
      if (sub > 0) {
        if (any_prop)
          for (int i = q; i < N; i += R) {
            float2 p[S];
            load_slots<S>(psi + i * VP + v0, p);
            const float2 dc = Dcur[i];
#pragma unroll
            for (int k = 0; k < S; ++k)
              if (prop[k]) p[k] = cmul_diag(cmul_diag(p[k], dc), dc);
            store_slots<S>(psi + i * VP + v0, p);
          }
        compiler_fence();
      }
      const float2 *csn = A.cs + ((sub + 1 < A.m) ? row : (fn - A.f_lo)) * K;
      const float2 *csc = A.cs + row * K;
      for (int b0 = 0; b0 < nsteps; b0 += kRingS) {
#pragma unroll
        for (int j = 0; j < kRingS; ++j) {
          const int sidx = b0 + j;
          const int sx = __builtin_amdgcn_readfirstlane(rst[j].x);
          const int se = __builtin_amdgcn_readfirstlane(rst[j].y) & 0x7fffffff;
          if (any_prop && sx + q < se) {
            const uint32_t pk = rab[j];
            const int a = (int)(pk & 0xffffu), b = (int)(pk >> 16);
            float2 pa[S], pb[S];
            load_slots<S>(psi + a * VP + v0, pa);
            load_slots<S>(psi + b * VP + v0, pb);
            const float co = rc[j].x, si = rc[j].y;
#pragma unroll
            for (int k = 0; k < S; ++k) {
              float2 na = pa[k], nb = pb[k];
              rotate_pair(na, nb, co, si);
              // slots that do not propagate in this frame keep their values
              pa[k] = prop[k] ? na : pa[k];
              pb[k] = prop[k] ? nb : pb[k];
            }
            store_slots<S>(psi + a * VP + v0, pa);
            store_slots<S>(psi + b * VP + v0, pb);
          }
          compiler_fence();
          {
            const int sp = sidx + kRingS;
            const bool nxt = sp >= nsteps;
            const int2 stp = steps[nxt ? sp - nsteps : sp];
            const int rp = max(0, min(stp.x + q, (stp.y & 0x7fffffff) - 1));
            const float2 *src = nxt ? csn : csc;
            rst[j] = stp;
            rc[j] = src[rp];
            rab[j] = A.ab[rp];
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = t + k * 64;
      if (i < N) {
        Dsh[(fn & 1) * N + i] = dn[k];
#pragma unroll
        for (int x = 0; x < 3; ++x) mush[x * N + i] = mn[3 * k + x];
      }
    }
    for (int i = t + 256; i < N; i += 64) {
      Dsh[(fn & 1) * N + i] = A.diag[(fn - A.f_lo) * N + i];
      const float *mup = A.mu_rec + (A.mu_stride ? fn * A.mu_stride : 0) + 1;
      for (int x = 0; x < 3; ++x) mush[x * N + i] = mup[x * N + i];
    }
    compiler_fence();
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const SlotDesc d = sd[k];
    if (A.writeback && v0 + k < ns && d.out >= 0 && d.nprop > 0) {
      const int64_t row = (int64_t)d.fs + d.nprop - 1 - A.f_lo;
      const bool due = (int64_t)d.fs + d.nprop >= F1;  // half-phase not yet applied
      for (int i = q; i < N; i += R) {
        const float2 p = psi[i * VP + v0 + k];
        A.state[(int64_t)d.out * N + i] = due ? cmul_diag(p, A.diag[row * N + i]) : p;
      }
    }
  }
  if (A.tmax_acc > 0)
    for (int i = t; i < A.tmax_acc; i += 64) A.partial[(int64_t)g * A.tmax_acc + i] = acc[i];
}

// Steps for a given shape: each layer split into chunks of at most
// block/vp rotations, the layer's last chunk flagged (sign bit of y), the
// list padded with empty steps to a multiple of the operand ring length.
void build_steps(const Schedule &s, int vp, int block, std::vector<int2> &steps) {
  // block > 0: multi-wave kernel (block threads) or, for 64, the single-wave
  // kernel; block < 0: single-wave kernel with -block slots per lane
  const bool w1 = block == 64;
  const int sl = block < 0 ? -block : 0;
  const int rpp = sl ? 64 * sl / vp : w1 ? (64 / vp) * kBatch : block / vp;
  const int ring = sl ? kRingS : w1 ? kRing1 : kRing;
  steps.clear();
  for (int l = 0; l < s.nlayers(); ++l) {
    const int r0 = s.layer_off[l], r1 = s.layer_off[l + 1];
    for (int r = r0; r < r1; r += rpp) {
      const int e = r + rpp < r1 ? r + rpp : r1;
      int2 st;
      st.x = r;
      st.y = (e == r1) ? (int)(0x80000000u | (uint32_t)e) : e;
      steps.push_back(st);
    }
  }
  while (!steps.empty() && steps.size() % ring) steps.push_back(make_int2(0, 0));
}

size_t conveyor_smem(int N, int vp, int block, int tmax_acc, int nsteps) {
  return ConvSmem::bytes(N, vp, block < 0 ? 64 : block, tmax_acc, nsteps);
}

// block == 64 selects the single-wave kernel, block = -S the single-wave
// kernel with S slots per lane
static const void *conveyor_fn(int vp, int block) {
  if (block == -2) {
    if (vp == 8) return (const void *)conveyor_s_kernel<8, 2>;
    if (vp == 16) return (const void *)conveyor_s_kernel<16, 2>;
    if (vp == 32) return (const void *)conveyor_s_kernel<32, 2>;
    return nullptr;
  }
  if (block == -4) {
    if (vp == 16) return (const void *)conveyor_s_kernel<16, 4>;
    if (vp == 32) return (const void *)conveyor_s_kernel<32, 4>;
    if (vp == 64) return (const void *)conveyor_s_kernel<64, 4>;
    return nullptr;
  }
  if (block == 64) {
    if (vp == 4) return (const void *)conveyor_w1_kernel<4>;
    if (vp == 8) return (const void *)conveyor_w1_kernel<8>;
    if (vp == 16) return (const void *)conveyor_w1_kernel<16>;
    if (vp == 32) return (const void *)conveyor_w1_kernel<32>;
    return nullptr;
  }
  const void *fn = nullptr;
#define NISE_CONV(V, B) \
  if (vp == V && block == B) fn = (const void *)conveyor_kernel<V, B>;
  NISE_CONV(4, 256) NISE_CONV(8, 256) NISE_CONV(16, 256) NISE_CONV(32, 256) NISE_CONV(64, 256)
  NISE_CONV(16, 512) NISE_CONV(32, 512) NISE_CONV(64, 512)
#undef NISE_CONV
  return fn;
}

int launch_conveyor(const ConveyorArgs &a, int vp, int block, hipStream_t st) {
  const size_t smem = conveyor_smem(a.N, vp, block, a.tmax_acc, a.nsteps);
  if (smem > 160 * 1024) {
    set_error("conveyor: N=%d with %d slots and %d steps needs %zu B of LDS (max 160 KiB)", a.N,
              vp, a.nsteps, smem);
    return NISE_ERR_ARG;
  }
  if (a.nsteps % (block < 0 ? kRingS : block == 64 ? kRing1 : kRing))
    return set_error("conveyor: step list not padded"), NISE_ERR_ARG;
  if (a.nwg == 0) return NISE_OK;
  const void *fn = conveyor_fn(vp, block);
  if (!fn) return set_error("conveyor: unsupported shape vp=%d block=%d", vp, block), NISE_ERR_ARG;
  NISE_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  ConveyorArgs args = a;
  void *kargs[] = {&args};
  NISE_HIP_CHECK(hipLaunchKernel(fn, dim3(a.nwg), dim3(block < 0 ? 64 : block), kargs, smem, st));
  return NISE_OK;
}

// Deterministic sum of the per-workgroup response partials (fixed order,
// double accumulation), rounded to float at the end.
// Sum of the partial rows per time step in a fixed order: thread (x, y) of
// a block sums rows y, y + kRedY, ... of time step blockIdx.x * kRedX + x,
// then the kRedY sums are added in y order (FP64 throughout).
constexpr int kRedX = 16, kRedY = 64;

__global__ void __launch_bounds__(kRedX * kRedY) reduce_partials_kernel(const double2 *partial, int nwg, int tmax,
                                                                        float *re, float *im) {
  __shared__ double2 acc[kRedY][kRedX];
  const int x = threadIdx.x % kRedX, y = threadIdx.x / kRedX;
  const int t1 = blockIdx.x * kRedX + x;
  double sr = 0, si = 0;
  if (t1 < tmax)
    for (int g = y; g < nwg; g += kRedY) {
      const double2 p = partial[(int64_t)g * tmax + t1];
      sr += p.x, si += p.y;
    }
  acc[y][x] = make_double2(sr, si);
  __syncthreads();
  if (y == 0 && t1 < tmax) {
    sr = 0, si = 0;
    for (int k = 0; k < kRedY; ++k) sr += acc[k][x].x, si += acc[k][x].y;
    re[t1] = (float)sr;
    im[t1] = (float)si;
  }
}

int launch_reduce_partials(const double2 *partial, int nwg, int tmax, float *re, float *im,
                           hipStream_t st) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((tmax + kRedX - 1) / kRedX), dim3(kRedX * kRedY), 0, st,
                     partial, nwg, tmax, re, im);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

// ---------------------------------------------------------------------------
// Plain per-sample kernel: the straightforward GPU port of the reference's
// per-sample loop (src/absorption.c:113-145 with src/propagate.c:821-899
// inside), used as the baseline the batched kernels are measured against.
// One launch per start sample, one workgroup holding its three polarisation
// vectors in LDS.  Per step the frame is read from HBM (half-step phases of
// the diagonal, angles of the kept couplings of the union structure), then
// the whole workgroup sweeps the layered schedule, one barrier per layer
// (rotations of a layer touch disjoint sites, so every site sees the
// reference's operations in the reference's order).
// ---------------------------------------------------------------------------
constexpr int kPlainBlock = 256;

__global__ void __launch_bounds__(kPlainBlock) plain_sample_kernel(PlainArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, K = A.K;
  float2 *psi = reinterpret_cast<float2 *>(smem);        // [3][N]
  float2 *D = psi + 3 * N;                               // [N]
  float2 *red = D + N;                                   // [3][waves]
  float2 *cs = A.cs_lds ? red + 3 * (kPlainBlock / 64) : A.ws_cs;  // [K]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int kWaves = kPlainBlock / 64;

  for (int x = 0; x < 3; ++x) {
    const float *mu0 = A.mu_rec + (A.mu_stride ? A.ti * A.mu_stride : 0) + 1 + x * N;
    for (int i = t; i < N; i += kPlainBlock) psi[x * N + i] = make_float2(mu0[i], 0.f);
  }
  __syncthreads();
  for (int t1 = 0; t1 < A.tmax; ++t1) {
    const int64_t f = A.ti + t1;
    // response for the three polarisations (src/absorption.c:205-212)
    for (int x = 0; x < 3; ++x) {
      const float *mu = A.mu_rec + (A.mu_stride ? f * A.mu_stride : 0) + 1 + x * N;
      float rr = 0.f, ri = 0.f;
      for (int i = t; i < N; i += kPlainBlock) {
        rr += mu[i] * psi[x * N + i].x;
        ri += mu[i] * psi[x * N + i].y;
      }
      for (int off = 32; off > 0; off >>= 1) rr += __shfl_xor(rr, off), ri += __shfl_xor(ri, off);
      if (lane == 0) red[x * kWaves + wave] = make_float2(rr, ri);
    }
    __syncthreads();
    if (t == 0) {
      double sr = 0, si = 0;
      for (int x = 0; x < 3; ++x)
        for (int w = 0; w < kWaves; ++w) sr += red[x * kWaves + w].x, si += red[x * kWaves + w].y;
      A.resp[t1].x += sr;
      A.resp[t1].y += si;
    }
    if (t1 == A.tmax - 1) break;  // the last propagation is never observed
    // the frame: half-step phases (src/propagate.c:863-866) and rotation
    // angles (:881-886; an absent coupling is the identity rotation)
    for (int i = t; i < N; i += kPlainBlock) {
      float h = site_energy(A.src, N, f, i);
      double arg = 0.5 * (double)h * (double)A.fac;
      D[i] = make_float2((float)cos(arg), (float)(-sin(arg)));
    }
    for (int k = t; k < K; k += kPlainBlock) {
      const uint32_t p = A.ab[k];
      const int a = (int)(p & 0xffffu), b = (int)(p >> 16);
      const float J = kept_coupling(A.src, N, f, a, b);
      float co = 1.f, si = 0.f;
      if (J != 0.f) {
        const float Jf = __fmul_rn(J, A.fac);
        si = (float)(-sin((double)Jf));
        co = (float)sqrt((double)__fsub_rn(1.0f, __fmul_rn(si, si)));
      }
      cs[k] = make_float2(co, si);
    }
    __syncthreads();
    for (int sub = 0; sub < A.m; ++sub) {
      for (int i = t; i < 3 * N; i += kPlainBlock) psi[i] = cmul_diag(psi[i], D[i % N]);
      __syncthreads();
      for (int l = 0; l < A.nlayers; ++l) {
        const int k0 = A.layer_off[l], n = 3 * (A.layer_off[l + 1] - k0);
        for (int e = t; e < n; e += kPlainBlock) {
          const int k = k0 + e / 3, x = e % 3;
          const uint32_t pk = A.ab[k];
          float2 *pv = psi + x * N;
          const int ka = (int)(pk & 0xffffu), kb = (int)(pk >> 16);
          float2 pa = pv[ka], pb = pv[kb];
          const float2 c = cs[k];
          rotate_pair(pa, pb, c.x, c.y);
          pv[ka] = pa;
          pv[kb] = pb;
        }
        __syncthreads();
      }
      for (int i = t; i < 3 * N; i += kPlainBlock) psi[i] = cmul_diag(psi[i], D[i % N]);
      __syncthreads();
    }
  }
}

size_t plain_smem(int N, int K, bool cs_lds) {
  return (size_t)4 * N * sizeof(float2) + 3 * (kPlainBlock / 64) * sizeof(float2) +
         (cs_lds ? (size_t)K * sizeof(float2) : 0);
}

int launch_plain_sample(const PlainArgs &a, hipStream_t st) {
  const size_t smem = plain_smem(a.N, a.K, a.cs_lds);
  if (smem > 160 * 1024) {
    set_error("plain kernel: N=%d needs %zu B of LDS", a.N, smem);
    return NISE_ERR_ARG;
  }
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)plain_sample_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(plain_sample_kernel, dim3(1), dim3(kPlainBlock), smem, st, a);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

}  // namespace nise
