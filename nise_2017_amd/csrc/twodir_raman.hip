// 2DIR-Raman response (src/calc_2DIRraman.c:22-663) for gfx950.
//
// The reference runs 21 polarisation items per sample, each propagating its
// own vectors.  Here every distinct vector is propagated once per sample
// (tj = sample * Samplerate + T1; the t3 frames are tj + t3, as the
// reference reads them):
//   L0[x0][t1]   mu_x0(tj - t1) through frames tj - t1 .. tj - 1 (kI, +1)
//   L[x0][t1]    L0 continued through t3 (the history): rightrr(t3) =
//                conj(L(t3)), the sign -1 step of a conjugate being the
//                conjugate of the sign +1 step bit for bit
//   M[x1]        mu_x1(tj) through t3 (history): rightnr(t3) = conj(M(t3))
//   ft1[x0][t1]  |aa> = sqrt2 mu_x0,a(tj - t1) (dipole_double_ground)
//                through frames tj - t1 .. tj - 1, then through t3 (kII;
//                held in LDS for all of it by raman_persist_kernel, which
//                also forms diagram 2 and the t1n rows)
//   fr[x1]       |aa> = sqrt2 mu_x1,a(tj) through t3 (kI)
//   t1n[x0][x1][t1]  dipole_double_last(mu_x1(tj), ft1[x0][t1](tj))
//                through t3 (+1, history)
// Diagram 1 contracts fr with the Raman component alpha_c(tj + t3), c =
// find_raman_pol(x2, x3), and dots it with conj(L); diagram 2 contracts
// ft1 with alpha_c and dots it with conj(M); diagram 3 sums alpha_c times
// t1n and is subtracted (src/calc_2DIRraman.c:303-403).  The per-item
// planes are reduced in a fixed order in FP64 into the 12 planes of
// :466-468.  The technique variants keep the reference's block conditions
// (:203-218, :253-287, :304-403, :591-633): 2DIRramanI, for one, forms
// diagram 1 from fr and rightrr as they were at t3 = 0.
//
// Two departures, both where the reference's result is not defined: the
// Raman component is find_raman_pol's return value (the reference discards
// it, :319, and reads an uninitialised index), and lifetime 0 means no
// lifetime decay (the reference divides by 2 lifetime, :106-107).  Only the
// Trotter scheme with a fixed anharmonicity is provided: the reference
// exits for a per-frame anharmonicity (:222-226, :330-334), and its sparse
// scheme here (propagate_0, :537-588) applies the transposed element lists
// in the t3 steps but not in the t1 steps and propagates t1n with sign -1
// against +1 in propagate_1.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nise_internal.h"
#include "doubles_dev.h"
#include "persist_dev.h"

namespace nise {

struct RamanArgs {
  int N = 0, T1 = 0, T3 = 0, S = 0, s0 = 0, sample = 1;
  int k1 = 0, k2 = 0, d2 = 0, d3 = 0, p1 = 0;  // block conditions, see raman_blocks
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;        // 0: static dipoles (Coupling type)
  const float *al_rec = nullptr;  // Raman tensor records, 1 + 6N words, absolute frames
  float2 *L0 = nullptr;         // [S][3][T1][N]
  const float2 *hist = nullptr; // [S][3 T1 + 3][T3][N]: L rows, then M rows
  float2 *dbl = nullptr;        // [S][T1][3][N2] ft1, then [S][3][N2] fr
  float2 *t1n = nullptr;        // [S][3][3][T1][N]
  const float2 *h3 = nullptr;   // [S][3][3][T1][T3][N]
  float *E = nullptr;           // [S][21][6][T3][T1]: diagrams 1, 2, 3 (rr, ri)
};

__device__ inline int64_t r_tj(const RamanArgs &A, int s) { return (int64_t)(A.s0 + s) * A.sample + A.T1; }
__device__ inline const float *r_mu(const RamanArgs &A, int64_t f, int x) {
  return A.mu_rec + (A.mu_stride ? f * A.mu_stride : 0) + 1 + (int64_t)x * A.N;
}
__device__ inline const float *r_al(const RamanArgs &A, int64_t f, int c) {
  return A.al_rec + f * (1 + 6 * (int64_t)A.N) + 1 + (int64_t)c * A.N;
}
// src/polar.c:107-116, as a return value
__host__ __device__ inline int raman_pol(int p2, int p3) {
  if (p2 == 0 && p3 == 0) return 0;
  if ((p2 == 0 && p3 == 1) || (p2 == 1 && p3 == 0)) return 1;
  if ((p2 == 0 && p3 == 2) || (p2 == 2 && p3 == 0)) return 2;
  if (p2 == 1 && p3 == 1) return 3;
  if ((p2 == 1 && p3 == 2) || (p2 == 2 && p3 == 1)) return 4;
  return 5;
}
__device__ inline int64_t ft1_vec(const RamanArgs &A, int s, int t1, int x0) { return ((int64_t)s * A.T1 + t1) * 3 + x0; }
__device__ inline int64_t fr_vec(const RamanArgs &A, int s, int x1) { return (int64_t)A.S * A.T1 * 3 + s * 3 + x1; }

// L0[x][0] = mu_x(tj) (the t1 = 0 seeds, unpropagated)
__global__ void raman_seed0_kernel(RamanArgs A) {
  const int s = blockIdx.x, x = blockIdx.y;
  const float *m = r_mu(A, r_tj(A, s), x);
  for (int i = threadIdx.x; i < A.N; i += blockDim.x)
    A.L0[(((int64_t)s * 3 + x) * A.T1) * A.N + i] = make_float2(m[i], 0.f);
}

// dipole_double_ground (src/NISE_subs.c:512-530, fixed anharmonicity: over =
// sqrt2 dipole): the ft1 vectors of t1 (dipoles at tj - t1), or with fr the
// fr vectors (dipoles at tj); one workgroup per (sample, component)
__global__ void raman_inject_kernel(RamanArgs A, int t1, int fr) {
  const int s = blockIdx.x / 3, x = blockIdx.x % 3, N = A.N, N2 = N * (N + 1) / 2;
  const int64_t v = fr ? fr_vec(A, s, x) : ft1_vec(A, s, t1, x);
  const float *m = r_mu(A, r_tj(A, s) - (fr ? 0 : t1), x);
  float2 *d = A.dbl + v * N2;
  for (int j = threadIdx.x; j < N2; j += blockDim.x) d[j] = make_float2(0.f, 0.f);
  __syncthreads();
  for (int a = threadIdx.x; a < N; a += blockDim.x) d[sindex(a, a, N)] = make_float2(__fmul_rn(kSqrt2, m[a]), 0.f);
}

// Stage a doubles vector and 3 dipole-like rows (with their sqrt2 overtone
// rows) in LDS.
__device__ inline void raman_stage(const float2 *src, int N2, float2 *psi, const float *r0, const float *r1,
                                   const float *r2, int N, float *dipL, float *ovL) {
  for (int j = threadIdx.x; j < N2; j += blockDim.x) psi[j] = gld(src + j);
  const float *r[3] = {r0, r1, r2};
  for (int e = threadIdx.x; e < 3 * N; e += blockDim.x) {
    const float v = r[e / N][e % N];
    dipL[e] = v, ovL[e] = __fmul_rn(kSqrt2, v);
  }
}

// t1n[x0][x1][t1] = dipole_double_last(mu_x1(tj), ft1[x0][t1](tj)) for x1 =
// 0..2 (src/calc_2DIRraman.c:280-287); one workgroup per (s, x0, t1)
__global__ void raman_t1n_kernel(RamanArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, N2 = N * (N + 1) / 2, T1 = A.T1;
  const int b = blockIdx.x, s = b / (3 * T1), x0 = (b / T1) % 3, t1 = b % T1;
  float *dipL = reinterpret_cast<float *>(smem), *ovL = dipL + 3 * N;
  float2 *psi = reinterpret_cast<float2 *>(smem + (((size_t)6 * N * 4 + 15) & ~(size_t)15));
  const int64_t tj = r_tj(A, s);
  raman_stage(A.dbl + ft1_vec(A, s, t1, x0) * N2, N2, psi, r_mu(A, tj, 0), r_mu(A, tj, 1), r_mu(A, tj, 2), N, dipL,
              ovL);
  __syncthreads();
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const float ov[3] = {ovL[k], ovL[N + k], ovL[2 * N + k]};
    float2 o[3];
    ddl_multi<3, 1>(psi, N, k, dipL, dipL + N, dipL + 2 * N, ov, o);
    for (int x1 = 0; x1 < 3; ++x1) A.t1n[((((int64_t)s * 3 + x0) * 3 + x1) * T1 + t1) * N + k] = o[x1];
  }
}

// dipole_double_last(alpha_c, f) for the six Raman components c at frame tl
// into outL[c][N] (the vector f staged in psi)
__device__ inline void raman_contract6(const RamanArgs &A, int64_t tl, const float2 *psi, float *dipL, float *ovL,
                                       float2 *outL) {
  const int N = A.N;
  for (int half = 0; half < 2; ++half) {
    for (int e = threadIdx.x; e < 3 * N; e += blockDim.x) {
      const float v = r_al(A, tl, 3 * half + e / N)[e % N];
      dipL[e] = v, ovL[e] = __fmul_rn(kSqrt2, v);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      const float ov[3] = {ovL[k], ovL[N + k], ovL[2 * N + k]};
      float2 o[3];
      ddl_multi<3, 1>(psi, N, k, dipL, dipL + N, dipL + 2 * N, ov, o);
      for (int c = 0; c < 3; ++c) outL[(3 * half + c) * N + k] = o[c];
    }
    __syncthreads();
  }
}

__host__ __device__ inline size_t raman_contract_smem(int N) {
  const int N2 = N * (N + 1) / 2;
  return (((size_t)6 * N * 4 + 15) & ~(size_t)15) + (size_t)N2 * 8 + (size_t)6 * N * 8;
}

// Diagram 1 at t3 (src/calc_2DIRraman.c:306-330): per (s, x1) the contraction
// of fr[x1] with the six components, then for every item with this x1 and
// every t1 the dot with rightrr = conj(L[x0][t1]) at t3 (at t3 = 0 without
// the t3 propagation of fr and rightrr)
__global__ void raman_d1_kernel(RamanArgs A, int t3) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, N2 = N * (N + 1) / 2, T1 = A.T1, T3 = A.T3;
  const int s = blockIdx.x / 3, x1 = blockIdx.x % 3;
  float *dipL = reinterpret_cast<float *>(smem), *ovL = dipL + 3 * N;
  float2 *psi = reinterpret_cast<float2 *>(smem + (((size_t)6 * N * 4 + 15) & ~(size_t)15));
  float2 *outL = psi + N2;
  const int64_t tl = r_tj(A, s) + t3;
  for (int j = threadIdx.x; j < N2; j += blockDim.x) psi[j] = gld(A.dbl + fr_vec(A, s, x1) * N2 + j);
  __syncthreads();
  raman_contract6(A, tl, psi, dipL, ovL, outL);
  const int nrow = 3 * T1 + 3, tr = A.p1 ? t3 : 0;
  const int64_t plane = (int64_t)T3 * T1;
  for (int j = threadIdx.x; j < 21 * T1; j += blockDim.x) {
    const int m = j / T1, t1 = j % T1;
    if (c_polar[m][1] != x1) continue;
    const int x0 = c_polar[m][0], c = raman_pol(c_polar[m][2], c_polar[m][3]);
    const float2 *l = outL + c * N;
    const float2 *L = A.hist + (((int64_t)s * nrow + x0 * T1 + t1) * T3 + tr) * N;
    float rr = 0.f, ri = 0.f;
    for (int i = 0; i < N; ++i) {
      const float2 Lv = gld(L + i);
      const float Rr = Lv.x, Ri = -Lv.y;
      rr = __fadd_rn(rr, __fadd_rn(__fmul_rn(l[i].y, Rr), __fmul_rn(l[i].x, Ri)));
      ri = __fadd_rn(ri, __fsub_rn(__fmul_rn(l[i].x, Rr), __fmul_rn(Ri, l[i].y)));
    }
    float *E = A.E + (((int64_t)s * 21 + m) * 6 + 0) * plane + (int64_t)t3 * T1 + t1;
    E[0] = rr, E[plane] = ri;
  }
}

// Diagram 2 at t3 (src/calc_2DIRraman.c:333-356): per (s, x0, t1) the
// contraction of ft1[x0][t1] with the six components, then for every item
// with this x0 the dot with rightnr = conj(M[x1]) at t3
__global__ void raman_d2_kernel(RamanArgs A, int t3) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, N2 = N * (N + 1) / 2, T1 = A.T1, T3 = A.T3;
  const int b = blockIdx.x, s = b / (3 * T1), x0 = (b / T1) % 3, t1 = b % T1;
  float *dipL = reinterpret_cast<float *>(smem), *ovL = dipL + 3 * N;
  float2 *psi = reinterpret_cast<float2 *>(smem + (((size_t)6 * N * 4 + 15) & ~(size_t)15));
  float2 *outL = psi + N2;
  const int64_t tl = r_tj(A, s) + t3;
  for (int j = threadIdx.x; j < N2; j += blockDim.x) psi[j] = gld(A.dbl + ft1_vec(A, s, t1, x0) * N2 + j);
  __syncthreads();
  raman_contract6(A, tl, psi, dipL, ovL, outL);
  const int nrow = 3 * T1 + 3;
  const int64_t plane = (int64_t)T3 * T1;
  for (int m = threadIdx.x; m < 21; m += blockDim.x) {
    if (c_polar[m][0] != x0) continue;
    const int x1 = c_polar[m][1], c = raman_pol(c_polar[m][2], c_polar[m][3]);
    const float2 *l = outL + c * N;
    const float2 *M = A.hist + (((int64_t)s * nrow + 3 * T1 + x1) * T3 + t3) * N;
    float rr = 0.f, ri = 0.f;
    for (int i = 0; i < N; ++i) {
      const float2 Mv = gld(M + i);
      const float nr = Mv.x, ni = -Mv.y;
      rr = __fadd_rn(rr, __fadd_rn(__fmul_rn(nr, l[i].y), __fmul_rn(ni, l[i].x)));
      ri = __fadd_rn(ri, __fsub_rn(__fmul_rn(nr, l[i].x), __fmul_rn(ni, l[i].y)));
    }
    float *E = A.E + (((int64_t)s * 21 + m) * 6 + 2) * plane + (int64_t)t3 * T1 + t1;
    E[0] = rr, E[plane] = ri;
  }
}

// The kII doubles vectors in LDS for all their frames (src/calc_2DIRraman.c
// :203-246 and :333-356).  Workgroup (s, t1 block) holds three waves, x0 = 0,
// 1, 2, each with the vectors ft1[s][x0][t1] of t1 = t1a .. t1a + nv - 1 in
// VW slots; the three waves share every frame.  From the earliest start
// frame tj - (t1a + nv - 1) to tj, vector t1 is formed at frame tj - t1
// (dipole_double_ground: sqrt2 mu_x0,a on |aa>, every other state 0, before
// that frame's step; a slot holds zeros until then, and rotating zeros keeps
// zeros) and every formed vector takes the step of frames tj - t1 .. tj - 1.
// At tj: t1n[x0][x1][t1] = dipole_double_last(mu_x1(tj), ft1(tj)) when d3.
// Then per t3 (when d2): the contraction with the six Raman components at
// tj + t3 (dipole_double_last with alpha_c and overtone sqrt2 alpha_c), for
// every item with this x0 the k-th rounded dot term against conj(M[x1](t3))
// at once, the sums over k in ascending order (the reference's operation
// sequence, as raman_d2_kernel), then the doubles step of frame tj + t3.
struct RamanPersistArgs {
  RamanArgs A;
  int K2 = 0, m = 1;
  const uint32_t *ab2 = nullptr;  // kept couplings (lexicographic)
  const float2 *D2 = nullptr;     // [frame - f_lo][N2]
  const float4 *cs2 = nullptr;    // [frame - f_lo][K2]
  int64_t f_lo = 0;
};
constexpr int kRamanItemBatch = 2;  // items whose dot terms share the LDS rows
struct RamanLayout {
  size_t ix, al, ov, rn, lv, d2, cs, psi, total;
};
// ix: the rotation table; then the EA rows (alpha and overtone rows of the
// six components, or the dipole rows at tj; the M rows; the dot terms of
// kRamanItemBatch items per vector) sharing bytes with the doubles step's
// rows (D2, coefficient rows); then the three waves' states.
__host__ __device__ inline RamanLayout raman_layout(int N, int VW, int K2) {
  const size_t N2 = (size_t)N * (N + 1) / 2;
  RamanLayout L;
  size_t o = 0;
  L.ix = o, o += rot_table_bytes(N, 64 / VW, K2);
  const size_t u0 = o;
  L.al = o, o += ((size_t)6 * N * 4 + 15) & ~(size_t)15;
  L.ov = o, o += ((size_t)6 * N * 4 + 15) & ~(size_t)15;
  L.rn = o, o += (size_t)3 * N * 8;
  L.lv = o, o += (size_t)3 * VW * kRamanItemBatch * N * 8;
  const size_t ea_end = o;
  o = u0;
  L.d2 = o, o += N2 * 8, o = (o + 15) & ~(size_t)15;
  L.cs = o, o += (size_t)K2 * 48;
  o = (o > ea_end ? o : ea_end), o = (o + 15) & ~(size_t)15;
  L.psi = o, o += (size_t)3 * (N2 + 2) * VW * 8;
  L.total = o;
  return L;
}

template <int NP>
__global__ void __launch_bounds__(192) raman_persist_kernel(RamanPersistArgs P) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int VW = 64 / NP;
  const RamanArgs &A = P.A;
  const int N = A.N, N2 = N * (N + 1) / 2, T1 = A.T1, T3 = A.T3, K2 = P.K2, t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63, x0 = wave;
  const int nblk = (T1 + VW - 1) / VW, s = blockIdx.x / nblk, t1a = (blockIdx.x % nblk) * VW;
  const int nv = min(VW, T1 - t1a);
  const RamanLayout Ly = raman_layout(N, VW, K2);
  uint2 *ixL = reinterpret_cast<uint2 *>(smem + Ly.ix);
  float *alL = reinterpret_cast<float *>(smem + Ly.al), *ovL = reinterpret_cast<float *>(smem + Ly.ov);
  float2 *RnL = reinterpret_cast<float2 *>(smem + Ly.rn);
  float2 *LvL = reinterpret_cast<float2 *>(smem + Ly.lv) + (size_t)wave * VW * kRamanItemBatch * N;
  float2 *d2 = reinterpret_cast<float2 *>(smem + Ly.d2);
  float4 *csL = reinterpret_cast<float4 *>(smem + Ly.cs);
  float2 *psi = reinterpret_cast<float2 *>(smem + Ly.psi) + (size_t)wave * (N2 + 2) * VW;
  const int q = lane % VW, c0 = lane / VW, total = K2 * ((N + NP - 1) / NP);
  for (int j = lane; j < N2 * VW; j += 64) psi[j] = make_float2(0.f, 0.f);
  init_special_states<VW>(psi, N2, lane);
  build_rot_table<NP>(ixL, P.ab2, N, K2, false);
  const int64_t tj = r_tj(A, s);
  // one doubles step of frame f on every slot (unformed slots hold zeros)
  auto step = [&](int64_t f) {
    const int64_t row = f - P.f_lo;
    __syncthreads();  // the rows of the previous phase are consumed
    stage_step_rows(d2, csL, P.D2 + row * N2, P.cs2 + row * K2, N2, K2);
    __syncthreads();
    for (int sub = 0; sub < P.m; ++sub) {
      dbl_diag<VW>(psi, d2, N2, lane);
      dbl_rotate_tab<NP>(psi, total, ixL, csL, q, c0);
      dbl_diag<VW>(psi, d2, N2, lane);
    }
  };
  // (1) formation and the t1 steps
  for (int64_t f = tj - (t1a + nv - 1); f <= tj; ++f) {
    const int qi = (int)(tj - f) - t1a;  // the slot formed at f
    if (qi >= 0 && qi < nv) {
      for (int j = lane; j < N2; j += 64) psi[j * VW + qi] = make_float2(0.f, 0.f);
      const float *mu = r_mu(A, f, x0);
      for (int a = lane; a < N; a += 64)
        psi[((a * (2 * N - a + 1)) >> 1) * VW + qi] = make_float2(__fmul_rn(kSqrt2, mu[a]), 0.f);
      asm volatile("" ::: "memory");
    }
    if (f < tj) step(f);
  }
  // (2) t1n at tj: the dipole rows mu_x1(tj) and their sqrt2 overtone rows
  if (A.d3) {
    __syncthreads();
    for (int e = t; e < 3 * N; e += blockDim.x) {
      const float v = r_mu(A, tj, e / N)[e % N];
      alL[e] = v, ovL[e] = __fmul_rn(kSqrt2, v);
    }
    __syncthreads();
    if (lane < nv * N) {
      const int c = lane / N, k = lane % N;
      const float ov[3] = {ovL[k], ovL[N + k], ovL[2 * N + k]};
      float2 o[3];
      if ((N & 3) == 0) ddl_multi4<3, VW>(psi + c, N, k, alL, alL + N, alL + 2 * N, ov, o);
      else ddl_multi<3, VW>(psi + c, N, k, alL, alL + N, alL + 2 * N, ov, o);
      for (int x1 = 0; x1 < 3; ++x1) A.t1n[((((int64_t)s * 3 + x0) * 3 + x1) * T1 + t1a + c) * N + k] = o[x1];
    }
  }
  if (!A.d2) return;
  // (3) the t3 steps with diagram 2: the items with this x0
  int mj[7], cj[7], x1j[7], nm = 0;
  for (int m = 0; m < 21; ++m)
    if (c_polar[m][0] == x0 && nm < 7) mj[nm] = m, cj[nm] = raman_pol(c_polar[m][2], c_polar[m][3]), x1j[nm++] = c_polar[m][1];
  // the components these items use (five of the six), and item j's
  // position pj[j] among them
  int comp[6] = {0, 0, 0, 0, 0, 0}, ncomp = 0, pj[7];
  for (int cc = 0; cc < 6; ++cc) {
    bool used = false;
    for (int j = 0; j < nm; ++j) used |= cj[j] == cc;
    if (used) comp[ncomp++] = cc;
  }
  for (int j = 0; j < nm; ++j)
    for (int p = 0; p < ncomp; ++p)
      if (comp[p] == cj[j]) pj[j] = p;
  const int nrow = 3 * T1 + 3;
  const int64_t plane = (int64_t)T3 * T1;
  for (int t3 = 0; t3 < T3; ++t3) {
    const int64_t tl = tj + t3;
    __syncthreads();  // the previous step's rows are consumed
    for (int e = t; e < 6 * N; e += blockDim.x) {
      const float v = r_al(A, tl, e / N)[e % N];
      alL[e] = v, ovL[e] = __fmul_rn(kSqrt2, v);
    }
    for (int e = t; e < 3 * N; e += blockDim.x)
      RnL[e] = gld(A.hist + (((int64_t)s * nrow + 3 * T1 + e / N) * T3 + t3) * N + e % N);
    __syncthreads();
    // lane (c, k): the six contractions, then per batch of items the k-th
    // dot terms; one lane per (c, item) adds them in ascending k
    // o[j]: component comp[j]; only the components this x0's items use
    float2 o[6];
    const int c = lane / N, k = lane % N;
    if (lane < nv * N) {
      auto group = [&](int j0, int n) {  // components comp[j0 .. j0 + n), n <= 3
        const float *r0 = alL + comp[j0] * N, *r1 = alL + comp[min(j0 + 1, ncomp - 1)] * N,
                    *r2 = alL + comp[min(j0 + 2, ncomp - 1)] * N;
        const float ov[3] = {ovL[comp[j0] * N + k], ovL[comp[min(j0 + 1, ncomp - 1)] * N + k],
                             ovL[comp[min(j0 + 2, ncomp - 1)] * N + k]};
        if ((N & 3) == 0) {
          if (n == 3) ddl_multi4<3, VW>(psi + c, N, k, r0, r1, r2, ov, o + j0);
          else if (n == 2) ddl_multi4<2, VW>(psi + c, N, k, r0, r1, r2, ov, o + j0);
          else ddl_multi4<1, VW>(psi + c, N, k, r0, r1, r2, ov, o + j0);
        } else {
          if (n == 3) ddl_multi<3, VW>(psi + c, N, k, r0, r1, r2, ov, o + j0);
          else if (n == 2) ddl_multi<2, VW>(psi + c, N, k, r0, r1, r2, ov, o + j0);
          else ddl_multi<1, VW>(psi + c, N, k, r0, r1, r2, ov, o + j0);
        }
      };
      group(0, min(3, ncomp));
      if (ncomp > 3) group(3, ncomp - 3);
    }
    for (int j0 = 0; j0 < nm; j0 += kRamanItemBatch) {
      const int nb = min(kRamanItemBatch, nm - j0);
      if (lane < nv * N)
        for (int jj = 0; jj < nb; ++jj) {
          const int j = j0 + jj;
          float2 l = o[0];
#pragma unroll
          for (int cc = 1; cc < 6; ++cc)
            if (pj[j] == cc) l = o[cc];
          const float2 Mv = RnL[x1j[j] * N + k];
          const float nr = Mv.x, ni = -Mv.y;
          LvL[(c * kRamanItemBatch + jj) * N + k] = make_float2(__fadd_rn(__fmul_rn(nr, l.y), __fmul_rn(ni, l.x)),
                                                             __fsub_rn(__fmul_rn(nr, l.x), __fmul_rn(ni, l.y)));
        }
      asm volatile("" ::: "memory");
      if (lane < nv * nb) {
        const int cs = lane / nb, jj = lane % nb;
        const float2 *Tm = LvL + (cs * kRamanItemBatch + jj) * N;
        float rr = 0.f, ri = 0.f;
        for (int kk = 0; kk < N; ++kk) {
          const float2 v = Tm[kk];
          rr = __fadd_rn(rr, v.x), ri = __fadd_rn(ri, v.y);
        }
        float *E = A.E + (((int64_t)s * 21 + mj[j0 + jj]) * 6 + 2) * plane + (int64_t)t3 * T1 + t1a + cs;
        E[0] = rr, E[plane] = ri;
      }
      asm volatile("" ::: "memory");
    }
    if (t3 + 1 < T3) step(tl);
  }
}

// Diagram 3 (src/calc_2DIRraman.c:359-402): for every (s, item, t3, t1) the
// sum over sites of alpha_c(tl) t1n(t3), formed as the reference does (leftn3
// = 0 + alpha t1n, then 1 leftn3i + 0 leftn3r against the ground-state right
// side).  Workgroup (s, x0, x1, t3): the items with this (x0, x1) share the
// t1n rows, which are staged through LDS rc rows at a time by coalesced
// loads (row pitch N + 1, so that the threads' walks along their rows fall
// in different banks); thread t1 then sums its row for every such item.
size_t raman_d3_smem(int N, int rc) { return (((size_t)3 * N * 4 + 15) & ~(size_t)15) + (size_t)rc * (N + 1) * 8; }
__global__ void raman_d3_kernel(RamanArgs A, int rc) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, T1 = A.T1, T3 = A.T3;
  const int b = blockIdx.x, s = b / 9, x0 = (b / 3) % 3, x1 = b % 3, t3 = blockIdx.y;
  float *alL = reinterpret_cast<float *>(smem);  // [item][N]
  float2 *rows = reinterpret_cast<float2 *>(smem + (((size_t)3 * N * 4 + 15) & ~(size_t)15));
  int ms[3] = {0, 0, 0}, nm = 0;
  for (int m = 0; m < 21; ++m)
    if (c_polar[m][0] == x0 && c_polar[m][1] == x1 && nm < 3) ms[nm++] = m;
  const int64_t tl = r_tj(A, s) + t3, plane = (int64_t)T3 * T1;
  for (int e = threadIdx.x; e < nm * N; e += blockDim.x) {
    const int j = e / N;
    alL[e] = r_al(A, tl, raman_pol(c_polar[ms[j]][2], c_polar[ms[j]][3]))[e % N];
  }
  // row t1 of (s, x0, x1) at t3
  const float2 *src = A.h3 + ((((int64_t)s * 3 + x0) * 3 + x1) * T1 * T3 + t3) * N;
  const float one = 1.f, zero = 0.f;
  for (int t1c = 0; t1c < T1; t1c += rc) {
    const int nr = min(rc, T1 - t1c);
    __syncthreads();  // the previous chunk's rows are consumed (and the alpha rows written)
    for (int e = threadIdx.x; e < nr * N; e += blockDim.x) {
      const int r = e / N, i = e % N;
      rows[r * (N + 1) + i] = gld(src + ((int64_t)(t1c + r) * T3) * N + i);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nr; r += blockDim.x) {
      const int t1 = t1c + r;
      const float2 *v = rows + r * (N + 1);
      for (int j = 0; j < nm; ++j) {
        const float *al = alL + j * N;
        float rr = 0.f, ri = 0.f;
        for (int i = 0; i < N; ++i) {
          const float a = al[i];
          const float2 p = v[i];
          const float l3r = __fadd_rn(0.f, __fmul_rn(a, p.x)), l3i = __fadd_rn(0.f, __fmul_rn(a, p.y));
          rr = __fadd_rn(rr, __fadd_rn(__fmul_rn(one, l3i), __fmul_rn(zero, l3r)));
          ri = __fadd_rn(ri, __fsub_rn(__fmul_rn(one, l3r), __fmul_rn(zero, l3i)));
        }
        float *E = A.E + (((int64_t)s * 21 + ms[j]) * 6 + 4) * plane + (int64_t)t3 * T1 + t1;
        E[0] = rr, E[plane] = ri;
      }
    }
  }
}

// part[c][4p + q] = sum over the samples of chunk c (kRamanRedS samples)
// and the items, in ascending (s, m) order, of the reference's FP32 products
// E pw, pw = polarweight(p, m) lt(t3, t1): q = 0, 1 diagram 1; q = 2, 3
// diagram 2 minus diagram 3 (src/calc_2DIRraman.c :324-329, :350-355,
// :396-401); FP64 sums.  raman_reduce_final_kernel adds the chunks to R in
// ascending order.
constexpr int kRamanRedS = 16;
__global__ void raman_reduce_kernel(const float *E, int S, int T3, int T1, const float *wpl, const float *lt,
                                    double *part) {
  const int64_t plane = (int64_t)T3 * T1;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y, p = k / 4, q = k % 4, c = blockIdx.z;
  if (e >= plane) return;
  const float l = lt[e];
  double acc = 0.0;
  const int s1 = min(S, (c + 1) * kRamanRedS);
  for (int s = c * kRamanRedS; s < s1; ++s)
    for (int m = 0; m < 21; ++m) {
      const float pw = __fmul_rn(wpl[p * 21 + m], l);
      const float *Em = E + ((int64_t)s * 21 + m) * 6 * plane + e;
      if (q < 2) {
        acc += (double)__fmul_rn(gld(Em + (int64_t)q * plane), pw);
      } else {
        acc += (double)__fmul_rn(gld(Em + (int64_t)q * plane), pw);
        acc -= (double)__fmul_rn(gld(Em + (int64_t)(q + 2) * plane), pw);
      }
    }
  part[((int64_t)c * 12 + k) * plane + e] = acc;
}
__global__ void raman_reduce_final_kernel(const double *part, int nc, int64_t n, double *R) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  double acc = 0.0;
  for (int c = 0; c < nc; ++c) acc += part[(int64_t)c * n + idx];
  R[idx] += acc;
}

}  // namespace nise

using namespace nise;

namespace {
int upr(DevBuf &d, const void *s, size_t n) {
  int rc = d.alloc(n < 16 ? 16 : n);
  if (rc) return rc;
  if (n && s) NISE_HIP_CHECK(hipMemcpy(d.p, s, n, hipMemcpyHostToDevice));
  return NISE_OK;
}
bool tech_eq(const t_non *n, const char *t) { return strcmp(n->technique, t) == 0; }

// Device buffers of the calling thread's Raman runs, kept between calls.
struct RamanBufs {
  DevBuf Hb, mub, Cb, Alb, mk, ab, lay, diag, cs, ab2, D2, cs2, stp, wpl, ltb, Rd;
  DevBuf L0, hist, dbl, t1n, h3, E, part, slots, wgb, slots2, wgb2, slots3, wgb3, gvb, gvc, grow, tvb, tvc, trow;
};

// Groups of at most vw consecutive vectors [first, first + n): begins and
// counts appended
void split_groups(int64_t first, int n, int vw, std::vector<int> &vb, std::vector<int> &vc) {
  for (int o = 0; o < n; o += vw) vb.push_back((int)(first + o)), vc.push_back(std::min(vw, n - o));
}
}  // namespace

// Block conditions of the technique (src/calc_2DIRraman.c): kI singles and
// diagram 1 (k1), kII doubles in t1 (k2), diagram 2 and the t3 steps of ft1
// and rightnr (d2), diagram 3 and t1n (d3), the t3 steps of fr and rightrr
// (p1, :620-629).  Returns 0 for a technique that is not a 2DIR-Raman one.
static int raman_blocks(const t_non *n, int *k1, int *k2, int *d2, int *d3, int *p1) {
  const bool r = tech_eq(n, "2DIRraman"), r1 = tech_eq(n, "2DIRraman1"), r2 = tech_eq(n, "2DIRraman2"),
             r3 = tech_eq(n, "2DIRraman3"), rI = tech_eq(n, "2DIRramanI"), rII = tech_eq(n, "2DIRramanII");
  if (!(r || r1 || r2 || r3 || rI || rII)) return 0;
  *k1 = r || r1 || rI;
  *k2 = r || r2 || r3 || rII;
  *d2 = r || r2 || rII;
  *d3 = r || r3 || rII;
  *p1 = r || r1 || r2;
  return 1;
}

extern "C" int nise_2dirraman_batched(const t_non *non, const float *H_rec, int64_t H_stride, const float *mu_rec,
                                      int64_t mu_stride, int64_t nframes, const float *coupling_rec,
                                      const float *alpha_rec, int where, float *R12) {
  int rc = ensure_device();
  if (rc) return rc;
  const int N = non->singles, T1 = non->tmax1, T3 = non->tmax3;
  if (N < 1 || T1 < 1 || T3 < 1 || non->sample < 1 || non->ts < 1 || !R12 || !alpha_rec)
    return set_error("2DIRraman: bad sizes"), NISE_ERR_ARG;
  int k1, k2, d2, d3, p1;
  if (!raman_blocks(non, &k1, &k2, &d2, &d3, &p1))
    return set_error("2DIRraman: unknown technique %s", non->technique), NISE_ERR_ARG;
  if (non->propagation != 1)
    return set_error("2DIRraman: propagation %d is not provided (1 Coupling)", non->propagation), NISE_ERR_ARG;
  if (non->anharmonicity == 0.f)
    return set_error("2DIRraman: a per-frame anharmonicity is not provided (the reference exits)"), NISE_ERR_ARG;
  if (non->cluster != -1) return set_error("2DIRraman: cluster selection is not provided"), NISE_ERR_ARG;
  const bool ctype = strcmp(non->hamiltonian, "Coupling") == 0;
  if (!ctype && strcmp(non->hamiltonian, "Full") != 0)
    return set_error("2DIRraman: HamiltonianType %s is not provided", non->hamiltonian), NISE_ERR_ARG;
  if (ctype && !coupling_rec) return set_error("2DIRraman: Coupling type needs coupling_rec"), NISE_ERR_ARG;
  const int N2 = N * (N + 1) / 2;
  if (raman_contract_smem(N) > 160 * 1024) return set_error("2DIRraman: N=%d too large for LDS", N), NISE_ERR_ARG;
  const int begin = non->begin;
  int end = non->end;
  if (end == 0) end = (int)((nframes - T1 - non->tmax2 - T3 - 1) / non->sample + 1);  // src/calc_2DIRraman.c:58
  if (end <= begin) return set_error("2DIRraman: no samples (begin %d, end %d)", begin, end), NISE_ERR_RANGE;
  const int64_t f_lo = (int64_t)begin * non->sample;
  const int64_t f_hi = (int64_t)(end - 1) * non->sample + T1 + T3;  // the last frame read is tj + T3 - 1
  if (f_hi > nframes)
    return set_error("2DIRraman: trajectory has %lld frames, the samples need %lld", (long long)nframes,
                     (long long)f_hi), NISE_ERR_RANGE;
  const int64_t nfr = f_hi - f_lo;
  hipStream_t st = default_stream();
  static thread_local RamanBufs B;
  NISE_HIP_CHECK(hipStreamSynchronize(st));

  const float *Hd = H_rec, *mud = mu_rec, *Cd = coupling_rec, *Ald = alpha_rec;
  const int64_t aw = 1 + 6 * (int64_t)N;
  if (where == NISE_HOST) {
    if ((rc = upr(B.Hb, H_rec + f_lo * H_stride, (size_t)(nfr * H_stride * 4)))) return rc;
    Hd = B.Hb.as<float>() - f_lo * H_stride;
    if (ctype) {
      if ((rc = upr(B.mub, mu_rec, (size_t)mu_stride * 4)) || (rc = upr(B.Cb, coupling_rec, (size_t)(1 + N2) * 4)))
        return rc;
      mud = B.mub.as<float>(), Cd = B.Cb.as<float>();
    } else {
      if ((rc = upr(B.mub, mu_rec + f_lo * mu_stride, (size_t)(nfr * mu_stride * 4)))) return rc;
      mud = B.mub.as<float>() - f_lo * mu_stride;
    }
    if ((rc = upr(B.Alb, alpha_rec + f_lo * aw, (size_t)(nfr * aw * 4)))) return rc;
    Ald = B.Alb.as<float>() - f_lo * aw;
  } else if (where != NISE_DEVICE) {
    return set_error("2DIRraman: where must be NISE_HOST or NISE_DEVICE"), NISE_ERR_ARG;
  }
  FrameSource src;
  src.H_rec = Hd, src.H_stride = H_stride, src.coupling_type = ctype, src.C_rec = Cd;
  src.shifte = (non->max1 + non->min1) / 2;  // src/calc_2DIRraman.c:83-86
  src.cut = non->couplingcut, src.apply_shift = 0;

  // singles schedule over the union of kept couplings, per-frame tables of
  // the singles and of the doubles (lexicographic couplings)
  const int64_t npair = (int64_t)N * (N - 1) / 2;
  std::vector<uint8_t> mask(npair);
  if (npair) {
    if ((rc = B.mk.alloc(npair))) return rc;
    if ((rc = launch_union_mask(src, N, f_lo, nfr, B.mk.as<uint8_t>(), st))) return rc;
    NISE_HIP_CHECK(hipMemcpyAsync(mask.data(), B.mk.p, npair, hipMemcpyDeviceToHost, st));
    NISE_HIP_CHECK(hipStreamSynchronize(st));
  }
  Schedule sch;
  build_schedule(N, mask, sch);
  std::vector<uint32_t> lex;
  {
    int64_t p = 0;
    for (int a = 0; a < N; ++a)
      for (int b = a + 1; b < N; ++b, ++p)
        if (mask[p]) lex.push_back((uint32_t)a | ((uint32_t)b << 16));
  }
  const float fac = trotter_factor(non->deltat, 1, non->ts);
  if ((rc = upr(B.ab, sch.ab.data(), sch.ab.size() * 4)) ||
      (rc = upr(B.lay, sch.layer_off.data(), sch.layer_off.size() * 4)) ||
      (rc = B.diag.alloc((size_t)nfr * N * 8)) || (rc = B.cs.alloc((size_t)nfr * std::max(sch.K, 1) * 8)))
    return rc;
  if ((rc = launch_frame_tables(src, N, sch, B.ab.as<uint32_t>(), fac, f_lo, nfr, B.diag.as<float2>(),
                                B.cs.as<float2>(), st)))
    return rc;
  if ((rc = upr(B.ab2, lex.data(), lex.size() * 4)) || (rc = B.D2.alloc((size_t)nfr * N2 * 8)) ||
      (rc = B.cs2.alloc((size_t)nfr * std::max((int)lex.size(), 1) * 16)))
    return rc;
  if ((rc = launch_doubles_tables(src, N, (int)lex.size(), B.ab2.as<uint32_t>(), fac, non->anharmonicity, nullptr,
                                  f_lo, nfr, B.D2.as<float2>(), B.cs2.as<float4>(), st)))
    return rc;
  const int block = 256;
  int vp = 0;
  std::vector<int2> steps;
  for (int v : {64, 32, 16, 8, 4}) {
    build_steps(sch, v, block, steps);
    if (conveyor_smem(N, v, block, 0, (int)steps.size()) <= 160 * 1024) {
      vp = v;
      break;
    }
  }
  if (!vp) return set_error("2DIRraman: N=%d does not fit the conveyor's LDS image", N), NISE_ERR_ARG;
  if ((rc = upr(B.stp, steps.data(), steps.size() * 8))) return rc;

  // weights and lifetime factors (src/calc_2DIRraman.c:100-109; lifetime 0:
  // no decay)
  std::vector<float> wh(63), lth((size_t)T3 * T1);
  for (int p = 0; p < 3; ++p)
    for (int m = 0; m < 21; ++m) wh[p * 21 + m] = polarweight_h(p, m);
  for (int t3 = 0; t3 < T3; ++t3)
    for (int t1 = 0; t1 < T1; ++t1)
      lth[(size_t)t3 * T1 + t1] =
          (non->lifetime > 0) ? (float)exp(-(double)(t1 + t3) * non->deltat / (2 * non->lifetime)) : 1.0f;
  const int64_t plane = (int64_t)T3 * T1;
  if ((rc = upr(B.wpl, wh.data(), wh.size() * 4)) || (rc = upr(B.ltb, lth.data(), lth.size() * 4)) ||
      (rc = B.Rd.alloc((size_t)12 * plane * 8)))
    return rc;
  NISE_HIP_CHECK(hipMemsetAsync(B.Rd.p, 0, (size_t)12 * plane * 8, st));

  const int total = end - begin, nrow = 3 * T1 + 3;
  const double per_sample = 8.0 * 3 * T1 * N + 8.0 * nrow * T3 * N + 8.0 * (3 * T1 + 3) * N2 + 8.0 * 9 * T1 * N +
                            8.0 * 9 * T1 * T3 * N + 4.0 * 126 * plane;
  const int chunk = (int)std::max(1.0, std::min<double>({(double)total, 1024.0, 8.0e9 / per_sample}));
  const int vw = doubles_vw(N);
  // the kII vectors stay in LDS for all their frames (raman_persist_kernel)
  // when their states, rotation table and rows fit
  const int pnp = doubles_np(N);
  const size_t psm = raman_layout(N, 64 / pnp, (int)lex.size()).total;
  const bool pers = k2 && rot_table_fits(N, 64 / pnp, (int)lex.size()) && psm <= 160 * 1024;
  const size_t csm = raman_contract_smem(N);
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)raman_d1_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)csm));
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)raman_d2_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)csm));
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)raman_t1n_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)csm));

  ConveyorArgs a;
  a.N = N, a.K = sch.K, a.nlayers = sch.nlayers(), a.m = non->ts, a.f_lo = f_lo;
  a.ab = B.ab.as<uint32_t>(), a.layer_off = B.lay.as<int>(), a.diag = B.diag.as<float2>(), a.cs = B.cs.as<float2>();
  a.mu_rec = mud, a.mu_stride = ctype ? 0 : mu_stride;
  a.steps = B.stp.as<int2>(), a.nsteps = (int)steps.size();
  DoublesArgs d;
  d.N = N, d.K2 = (int)lex.size(), d.m = non->ts, d.ab2 = B.ab2.as<uint32_t>();
  d.D2 = B.D2.as<float2>(), d.cs2 = B.cs2.as<float4>();

  for (int c0 = 0; c0 < total; c0 += chunk) {
    const int S = std::min(chunk, total - c0);
    RamanArgs A;
    A.N = N, A.T1 = T1, A.T3 = T3, A.S = S, A.s0 = begin + c0, A.sample = non->sample;
    A.k1 = k1, A.k2 = k2, A.d2 = d2, A.d3 = d3, A.p1 = p1;
    A.mu_rec = mud, A.mu_stride = ctype ? 0 : mu_stride, A.al_rec = Ald;
    if ((rc = B.L0.alloc((size_t)S * 3 * T1 * N * 8)) || (rc = B.hist.alloc((size_t)S * nrow * T3 * N * 8)) ||
        (rc = B.dbl.alloc((size_t)S * (3 * T1 + 3) * N2 * 8)) || (rc = B.t1n.alloc((size_t)S * 9 * T1 * N * 8)) ||
        (rc = B.h3.alloc((size_t)S * 9 * T1 * T3 * N * 8)) || (rc = B.E.alloc((size_t)S * 126 * plane * 4)))
      return rc;
    A.L0 = B.L0.as<float2>(), A.hist = B.hist.as<float2>(), A.dbl = B.dbl.as<float2>();
    A.t1n = B.t1n.as<float2>(), A.h3 = B.h3.as<float2>(), A.E = B.E.as<float>();
    d.state = B.dbl.as<float2>();
    NISE_HIP_CHECK(hipMemsetAsync(B.E.p, 0, (size_t)S * 126 * plane * 4, st));
    auto tj_of = [&](int s) { return (int64_t)(begin + c0 + s) * non->sample + T1; };

    // (1) kI seeds L0[x0][t1] (src/calc_2DIRraman.c:197-213) and the history
    //     of L (kI) and M (diagram 2) through t3 from tj
    if (k1) {
      hipLaunchKernelGGL(raman_seed0_kernel, dim3(S, 3), dim3(64), 0, st, A);
      NISE_HIP_CHECK(hipGetLastError());
      std::vector<SlotDesc> sl;
      std::vector<int> wb;
      for (int s = 0; s < S; ++s)
        for (int x = 0; x < 3; ++x) {
          const int first = (int)sl.size();
          for (int t1 = 1; t1 < T1; ++t1)
            sl.push_back(SlotDesc{(int)(tj_of(s) - t1), 0, t1, x, 0, (s * 3 + x) * T1 + t1, 0, -1});
          for (int b = first; b < (int)sl.size(); b += vp) wb.push_back(b);
        }
      wb.push_back((int)sl.size());
      if (!sl.empty()) {
        if ((rc = upr(B.slots, sl.data(), sl.size() * sizeof(SlotDesc))) || (rc = upr(B.wgb, wb.data(), wb.size() * 4)))
          return rc;
        a.slots = B.slots.as<SlotDesc>(), a.wg_slot_begin = B.wgb.as<int>(), a.nwg = (int)wb.size() - 1;
        a.state = B.L0.as<float2>(), a.writeback = 1, a.hist = nullptr;
        if ((rc = launch_conveyor(a, vp, block, st))) return rc;
      }
    }
    if (k1 || d2) {
      std::vector<SlotDesc> sl;
      std::vector<int> wb;
      for (int s = 0; s < S; ++s) {
        const int first = (int)sl.size(), tj = (int)tj_of(s);
        if (k1)
          for (int x = 0; x < 3; ++x)
            for (int t1 = 0; t1 < T1; ++t1)
              sl.push_back(SlotDesc{tj, T3, T3 - 1, x, 1, (s * 3 + x) * T1 + t1, 0, (s * nrow + x * T1 + t1) * T3});
        if (d2)
          for (int x = 0; x < 3; ++x) sl.push_back(SlotDesc{tj, T3, T3 - 1, x, 0, -1, 0, (s * nrow + 3 * T1 + x) * T3});
        for (int b = first; b < (int)sl.size(); b += vp) wb.push_back(b);
      }
      wb.push_back((int)sl.size());
      if ((rc = upr(B.slots2, sl.data(), sl.size() * sizeof(SlotDesc))) || (rc = upr(B.wgb2, wb.data(), wb.size() * 4)))
        return rc;
      a.slots = B.slots2.as<SlotDesc>(), a.wg_slot_begin = B.wgb2.as<int>(), a.nwg = (int)wb.size() - 1;
      a.state = B.L0.as<float2>(), a.writeback = 0, a.hist = B.hist.as<float2>();
      if ((rc = launch_conveyor(a, vp, block, st))) return rc;
    }
    // (2) kII doubles ft1[x0][t1] through frames tj - t1 .. tj - 1: at step r
    //     (frame tj - T1 + r) the vectors of t1 = T1 - r are formed, then
    //     every vector of t1 >= T1 - r takes the step (src/calc_2DIRraman.c
    //     :203-246); fr from mu_x1(tj) (:255-258)
    if (k2 && pers) {
      RamanPersistArgs pa;
      pa.A = A, pa.K2 = (int)lex.size(), pa.m = non->ts, pa.ab2 = B.ab2.as<uint32_t>();
      pa.D2 = B.D2.as<float2>(), pa.cs2 = B.cs2.as<float4>(), pa.f_lo = f_lo;
      const void *fp = pnp == 8 ? (const void *)raman_persist_kernel<8>
                     : pnp == 16 ? (const void *)raman_persist_kernel<16>
                     : pnp == 32 ? (const void *)raman_persist_kernel<32>
                                 : (const void *)raman_persist_kernel<64>;
      NISE_HIP_CHECK(hipFuncSetAttribute(fp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)psm));
      void *kargs[] = {&pa};
      const int nblk = (T1 + 64 / pnp - 1) / (64 / pnp);
      NISE_HIP_CHECK(hipLaunchKernel(fp, dim3((unsigned)(S * nblk)), dim3(192), kargs, psm, st));
    } else if (k2) {
      std::vector<int> vb, vc;
      std::vector<int64_t> rows;
      std::vector<size_t> off(T1 + 1, 0);
      for (int r = 1; r < T1; ++r) {
        off[r] = vb.size();
        for (int s = 0; s < S; ++s) {
          const size_t g0 = vb.size();
          split_groups(((int64_t)s * T1 + T1 - r) * 3, 3 * r, vw, vb, vc);
          for (size_t g = g0; g < vb.size(); ++g) rows.push_back(tj_of(s) - T1 + r - f_lo);
        }
      }
      off[T1] = vb.size();
      if (!vb.empty() && ((rc = upr(B.tvb, vb.data(), vb.size() * 4)) || (rc = upr(B.tvc, vc.data(), vc.size() * 4)) ||
                          (rc = upr(B.trow, rows.data(), rows.size() * 8))))
        return rc;
      for (int r = 1; r < T1; ++r) {
        hipLaunchKernelGGL(raman_inject_kernel, dim3(S * 3), dim3(256), 0, st, A, T1 - r, 0);
        NISE_HIP_CHECK(hipGetLastError());
        DoublesArgs dd = d;
        dd.vec_begin = B.tvb.as<int>() + off[r], dd.vec_count = B.tvc.as<int>() + off[r];
        dd.rows = B.trow.as<int64_t>() + off[r];
        if ((rc = launch_doubles_step(dd, (int)(off[r + 1] - off[r]), st))) return rc;
      }
      hipLaunchKernelGGL(raman_inject_kernel, dim3(S * 3), dim3(256), 0, st, A, 0, 0);
      NISE_HIP_CHECK(hipGetLastError());
    }
    if (k1) {
      hipLaunchKernelGGL(raman_inject_kernel, dim3(S * 3), dim3(256), 0, st, A, 0, 1);
      NISE_HIP_CHECK(hipGetLastError());
    }
    if (d3 && !pers) {
      hipLaunchKernelGGL(raman_t1n_kernel, dim3(S * 3 * T1), dim3(64), csm, st, A);
      NISE_HIP_CHECK(hipGetLastError());
    }
    // (3) per t3: diagrams 1 and 2, then the doubles step of frame tj + t3
    //     for ft1 (d2) and fr (p1)
    std::vector<int> gvb, gvc;
    for (int s = 0; s < S; ++s) {
      if (d2 && !pers) split_groups((int64_t)s * T1 * 3, 3 * T1, vw, gvb, gvc);
      if (p1 && k1) split_groups((int64_t)S * T1 * 3 + s * 3, 3, vw, gvb, gvc);
    }
    const int G = (int)gvb.size();
    std::vector<int64_t> grow((size_t)std::max(T3 - 1, 0) * G);
    {
      int g = 0;
      for (int s = 0; s < S; ++s) {
        const int ng = (d2 && !pers ? (3 * T1 + vw - 1) / vw : 0) + (p1 && k1 ? (3 + vw - 1) / vw : 0);
        for (int j = 0; j < ng; ++j, ++g)
          for (int t3 = 0; t3 + 1 < T3; ++t3) grow[(size_t)t3 * G + g] = tj_of(s) + t3 - f_lo;
      }
    }
    if (G && T3 > 1 &&
        ((rc = upr(B.gvb, gvb.data(), gvb.size() * 4)) || (rc = upr(B.gvc, gvc.data(), gvc.size() * 4)) ||
         (rc = upr(B.grow, grow.data(), grow.size() * 8))))
      return rc;
    for (int t3 = 0; t3 < T3; ++t3) {
      if (k1) {
        hipLaunchKernelGGL(raman_d1_kernel, dim3(S * 3), dim3(64), csm, st, A, t3);
        NISE_HIP_CHECK(hipGetLastError());
      }
      if (d2 && !pers) {
        hipLaunchKernelGGL(raman_d2_kernel, dim3(S * 3 * T1), dim3(64), csm, st, A, t3);
        NISE_HIP_CHECK(hipGetLastError());
      }
      if (t3 + 1 < T3 && G) {
        DoublesArgs dd = d;
        dd.vec_begin = B.gvb.as<int>(), dd.vec_count = B.gvc.as<int>(), dd.rows = B.grow.as<int64_t>() + (size_t)t3 * G;
        if ((rc = launch_doubles_step(dd, G, st))) return rc;
      }
    }
    // (4) diagram 3: t1n through t3 (+1) with its history, then the sums
    if (d3) {
      std::vector<SlotDesc> sl;
      std::vector<int> wb;
      for (int s = 0; s < S; ++s) {
        const int first = (int)sl.size(), tj = (int)tj_of(s);
        for (int j = 0; j < 9 * T1; ++j) {
          const int row = s * 9 * T1 + j;
          sl.push_back(SlotDesc{tj, T3, T3 - 1, 0, 1, row, 0, row * T3});
        }
        for (int b = first; b < (int)sl.size(); b += vp) wb.push_back(b);
      }
      wb.push_back((int)sl.size());
      if ((rc = upr(B.slots3, sl.data(), sl.size() * sizeof(SlotDesc))) || (rc = upr(B.wgb3, wb.data(), wb.size() * 4)))
        return rc;
      a.slots = B.slots3.as<SlotDesc>(), a.wg_slot_begin = B.wgb3.as<int>(), a.nwg = (int)wb.size() - 1;
      a.state = B.t1n.as<float2>(), a.writeback = 0, a.hist = B.h3.as<float2>();
      if ((rc = launch_conveyor(a, vp, block, st))) return rc;
      // t1n rows per LDS chunk: 64 at most, within 64 KB (N <= 200 here gives 40 or more)
      const int rc = (int)std::max<size_t>(1, std::min<size_t>({(size_t)64, (size_t)T1,
                                                                 (65536 - raman_d3_smem(N, 0)) / ((size_t)(N + 1) * 8)}));
      hipLaunchKernelGGL(raman_d3_kernel, dim3(S * 9, T3), dim3(64), raman_d3_smem(N, rc), st, A, rc);
      NISE_HIP_CHECK(hipGetLastError());
    }
    // (5) this chunk into the 12 planes
    const int nred = (S + kRamanRedS - 1) / kRamanRedS;
    if ((rc = B.part.alloc((size_t)nred * 12 * plane * 8))) return rc;
    hipLaunchKernelGGL(raman_reduce_kernel, dim3((unsigned)((plane + 255) / 256), 12, nred), dim3(256), 0, st,
                       (const float *)B.E.as<float>(), S, T3, T1, (const float *)B.wpl.as<float>(),
                       (const float *)B.ltb.as<float>(), B.part.as<double>());
    NISE_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(raman_reduce_final_kernel, dim3((unsigned)((12 * plane + 255) / 256)), dim3(256), 0, st,
                       (const double *)B.part.as<double>(), nred, 12 * plane, B.Rd.as<double>());
    NISE_HIP_CHECK(hipGetLastError());
  }
  std::vector<double> Rh((size_t)12 * plane);
  NISE_HIP_CHECK(hipMemcpyAsync(Rh.data(), B.Rd.p, Rh.size() * 8, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  for (size_t i = 0; i < Rh.size(); ++i) R12[i] = (float)Rh[i];
  return NISE_OK;
}

// src/calc_2DIRraman.c:22-663 (2DIRraman and its variants): rank 0 runs every
// work item on its GPU; other ranks return at once.
extern "C" void calc_2DIRraman(t_non *non, int parentRank, int parentSize, int subRank, int subSize,
                               MPI_Comm subComm, MPI_Comm rootComm) {
  (void)parentSize, (void)subRank, (void)subSize, (void)subComm, (void)rootComm;
  if (parentRank != 0) return;
  const char *name = "calc_2DIRraman";
  if (non->cluster != -1) {
    printf("%s: the GPU path does not implement Cluster\n", name);
    die2d(1, "");
  }
  const int N = non->singles;
  const int totalSampleCount = (non->length - non->tmax1 - non->tmax2 - non->tmax3 - 1) / non->sample + 1;
  if (non->end == 0) non->end = totalSampleCount;
  int sampleCount = non->end - non->begin;
  if (sampleCount <= 0) {
    printf("%s: no samples between Begin and End\n", name);
    die2d(1, "");
  }
  printf("Begin sample: %d, End sample: %d.\n", non->begin, non->end);
  const float shift1 = (non->max1 + non->min1) / 2;
  printf("Rotating Wave Frequency shift %f.\n", shift1);
  non->shifte = shift1;
  non->shiftf = 2 * shift1;
  const bool ctype = strcmp(non->hamiltonian, "Coupling") == 0;
  FILE *probe = fopen(non->energyFName, "rb");
  if (!probe) die2d(1, "Hamiltonian file not found!\n");
  fclose(probe);
  if (!(probe = fopen(non->dipoleFName, "rb"))) {
    printf("Dipole file %s not found!\n", non->dipoleFName);
    die2d(1, "");
  }
  fclose(probe);
  if (!(probe = fopen(non->alphaFName, "rb"))) {
    printf("Raman file %s not found!\n", non->alphaFName);
    die2d(1, "");
  }
  fclose(probe);
  if (non->anharmonicity == 0) die2d(1, "Anharmonicity feature for Raman not implemented yet");
  const int64_t D = non->doubles;
  const int64_t hw = ctype ? 1 + N : 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2;
  const int64_t mw = 1 + 3 * (int64_t)N + 3 * (int64_t)N * D, aw = 1 + 6 * (int64_t)N;
  const int64_t f_lo = (int64_t)non->begin * non->sample;
  const int64_t f_hi = (int64_t)(non->end - 1) * non->sample + non->tmax1 + non->tmax3;
  const int64_t nfr = f_hi - f_lo;
  Pinned H, mu, Cc, Al;
  if (!H.alloc((size_t)(nfr * hw)) || !mu.alloc((size_t)((ctype ? 1 : nfr) * mw)) || !Al.alloc((size_t)(nfr * aw))) {
    printf("%s: could not allocate pinned host memory\n", name);
    die2d(1, "");
  }
  if (!read_records(non->energyFName, hw, f_lo, nfr, H.p))
    die2d(1, "Hamiltonian trajectory file to short, could not fill buffer!!!\n");
  if (!read_records(non->dipoleFName, mw, ctype ? 0 : f_lo, ctype ? 1 : nfr, mu.p))
    die2d(1, "Dipole trajectory file to short, could not fill buffer!!!\n");
  if (!read_records(non->alphaFName, aw, f_lo, nfr, Al.p))
    die2d(1, "Raman trajectory file to short, could not fill buffer!!!\n");
  if (ctype) {
    const int64_t cw = 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2;
    if (!(probe = fopen(non->couplingFName, "rb"))) die2d(1, "Coupling file not found!\n");
    fclose(probe);
    if (!Cc.alloc((size_t)cw) || !read_records(non->couplingFName, cw, 0, 1, Cc.p))
      die2d(1, "Coupling trajectory file to short, could not fill buffer!!!\n");
  }
  // the staged records start at frame f_lo: rebase the samples to begin = 0
  t_non loc = *non;
  loc.begin = 0, loc.end = non->end - non->begin;
  std::vector<float> R((size_t)12 * non->tmax3 * non->tmax1);
  const int rc = nise_2dirraman_batched(&loc, H.p, hw, mu.p, mw, nfr, ctype ? Cc.p : nullptr, Al.p, NISE_HOST,
                                        R.data());
  if (rc) {
    printf("%s: GPU run failed: %s\n", name, nise_last_error());
    die2d(1, "");
  }
  printf("Samples %d\n", sampleCount);
  const size_t pl = (size_t)non->tmax3 * non->tmax1;
  static const char *names[6] = {"RparI.IRraman.dat", "RparII.IRraman.dat", "RperI.IRraman.dat",
                                 "RperII.IRraman.dat", "RcroI.IRraman.dat", "RcroII.IRraman.dat"};
  for (int p = 0; p < 3; ++p)
    for (int k = 0; k < 2; ++k)
      print2D_file(names[2 * p + k], R.data() + (4 * p + 2 * k) * pl, R.data() + (4 * p + 2 * k + 1) * pl, non,
                   sampleCount);
  printf("----------------------------------------\n");
  printf(" 2DIR-Raman calculation succesfully completed\n");
  printf("----------------------------------------\n\n");
}
