// Batched 2DIR response (src/calc_2DIR.c:184-506) for gfx950.
//
// The reference runs 21 molecular-frame polarisation items per sample, each
// repeating propagations that depend on one dipole component only.  Here
// every distinct vector is propagated once per sample:
//   L0[x][t1]  GB seeds: mu_x(tj - t1) propagated t1 frames    (3 x T1)
//   L[x][t1]   L0[x][t1] continued through t3                   (3 x T1)
//   M[x]       mu_x(tk) propagated through t3                    (3)
//   fr[x1][x2] dipole_double(mu_x2(tk), M[x1](tk))               (9, doubles)
//   ft[x0][x2][t1] dipole_double(mu_x2(tk), L0[x0][t1])          (9 x T1)
// with tj = sample*Samplerate + T1 and tk = tj + T2.  Without a waiting
// time (T2 = 0) the SE and EA pathways start from the same vectors (M[x] is
// leftrr, L0 is leftnr).  With T2 > 0 the SE/EA start vectors are mu_x(tj)
// and L0[x][t1] carried through the waiting time by the exact exponential
// (src/calc_2DIR.c:303-326): those are the rows the history records as "M"
// and "L" (started from Linit), and the GB's own M[x] (mu_x(tk) through t3)
// gets three more history rows.  The EA "right" vectors are not
// propagated: the sign -1 step of conj(psi) is, bit for bit, conj of the
// sign +1 step of psi (negation commutes with every rounded product and sum
// of the step), so rightrr[t1](t3) = conj(L[x0][t1](t3)) and rightnr(t3) =
// conj(M[x1](t3)).
//
// Stages: seeds (conveyor, write back), history (conveyor recording L and M
// at every t3), GB/SE responses from the history, doubles setup, then per
// t3 the EA responses and one doubles step, and finally a fixed-order FP64
// reduction over samples and polarisation items into the 12 planes
// [T3][T1] of src/calc_2DIR.c:540-542.  Each dot product keeps the
// reference's element order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "nise_internal.h"
#include "doubles_dev.h"
#include "persist_dev.h"

namespace nise {

struct TwoDirArgs {
  int N, T1, T3, S;         // sites, t1 and t3 points, samples in this chunk
  int T2;                   // waiting time steps: tk = tj + T2
  int nrow;                 // history rows per sample: 3 T1 + 3, plus 3 GB rows when T2 > 0
  int s0, sample;           // first sample (absolute), Samplerate
  int do_gb, do_se, do_ea;
  int es;                   // two-level sites (calc_2DES): no |aa> states, no overtones
  float anh;                // fixed anharmonicity (0: per-frame records)
  const float *mu_rec;      // dipole records (absolute frame index)
  int64_t mu_stride;        // 0: static dipoles
  const float *over_rec;    // overtone dipole records (1 + 3N) when anh == 0
  const float2 *L0;         // [S][3][T1][N]
  const float2 *Linit;      // [S][3T1+3][N]: L[x][t1] and M[x] at tk (T2 > 0), else L0
  int linit_stride;         // rows per sample of Linit (3 T1 + 3, or 3 T1 for L0)
  const float2 *hist;       // [S][nrow][T3][N]: rows of L, then M, then the GB's M (T2 > 0)
  float2 *dbl;              // [S][9 + 9 T1][N2]
  float *E;                 // [S][21][4][T3][T1]
};

__device__ inline const float *mu_at(const TwoDirArgs &A, int64_t f, int x) {
  return A.mu_rec + (A.mu_stride ? f * A.mu_stride : 0) + 1 + (int64_t)x * A.N;
}

__device__ inline int64_t tj_of(const TwoDirArgs &A, int s) {
  return (int64_t)(A.s0 + s) * A.sample + A.T1;
}
__device__ inline int64_t tk_of(const TwoDirArgs &A, int s) { return tj_of(A, s) + A.T2; }

// history row of L[x][t1] / M[x] (the SE and EA vectors) / the GB's M[x]
__device__ inline int64_t hrow_L(const TwoDirArgs &A, int s, int x, int t1) {
  return (((int64_t)s * A.nrow) + x * A.T1 + t1) * A.T3;
}
__device__ inline int64_t hrow_M(const TwoDirArgs &A, int s, int x) {
  return (((int64_t)s * A.nrow) + 3 * A.T1 + x) * A.T3;
}
__device__ inline int64_t hrow_G(const TwoDirArgs &A, int s, int x) {
  return (((int64_t)s * A.nrow) + 3 * A.T1 + (A.T2 > 0 ? 3 : 0) + x) * A.T3;
}

// L0[s][x][0] = mu_x(tj) (no propagation for t1 = 0)
__global__ void seed0_kernel(TwoDirArgs A, float2 *L0) {
  const int s = blockIdx.x, x = blockIdx.y;
  const float *mu = mu_at(A, tj_of(A, s), x);
  for (int i = threadIdx.x; i < A.N; i += blockDim.x)
    L0[(((int64_t)s * 3 + x) * A.T1) * A.N + i] = make_float2(mu[i], 0.f);
}

// acc[a][b] = sum_i mu_b[i] v_a[i] (complex v, real mu) for three rows v_a
// and three rows mu_b, each sum in ascending i with rounded products and
// sums, as the reference's dot loops (the product's operand order does not
// change a rounded product).
__device__ __forceinline__ void dot9(const float2 *v0, const float2 *v1, const float2 *v2, const float *mu,
                                     int mstride, int N, float2 acc[9]) {
#pragma unroll
  for (int j = 0; j < 9; ++j) acc[j] = make_float2(0.f, 0.f);
#pragma unroll 2
  for (int i = 0; i < N; ++i) {
    const float2 v[3] = {gld(v0 + i), gld(v1 + i), gld(v2 + i)};
    const float m[3] = {mu[i], mu[mstride + i], mu[2 * mstride + i]};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        float2 &r = acc[a * 3 + b];
        r.x = __fadd_rn(r.x, __fmul_rn(m[b], v[a].x));
        r.y = __fadd_rn(r.y, __fmul_rn(m[b], v[a].y));
      }
  }
}

// The GB and SE dots that do not depend on both t1 and t3, one workgroup
// per sample:
//   g1[s][x0*3+y][t1] = mu_y(tj) . L0[x0][t1]          GB t1n (src/calc_2DIR.c:244-250)
//   g3[s][t3][a*3+b]  = mu_b(tk + t3) . Mgb[a](t3)     GB t3n (:265-268)
// and for the SE
//   g3P[s][t3][a*3+b] = mu_b(tk + t3) . M[a](t3)       SE t3r (:378-381)
//   g1s[s][x0*3+y][t1] = mu_y(tk) . L[x0][t1](0)       conj: SE t1r (:363-369)
// The reference's SE t1r[x0][x2] is conj(g1s[x0][x2]) and its SE static
// t3n[x1][x2] (:353-357) is conj(g3P[0][x1*3+x2]): the same rounded products
// and sums with the imaginary sign flipped (x - p rounds to -(-x + p)).
// Without a waiting time g3P is g3 and g1s is g1 (the SE vectors are the GB
// ones), and only g1 and g3 are formed.
__global__ void gbse_static_kernel(TwoDirArgs A, float2 *g1, float2 *g3, float2 *g1s, float2 *g3P) {
  extern __shared__ __align__(16) float muj[];  // [3][N] mu(tj), [3][N] mu(tk)
  const int s = blockIdx.x, N = A.N, T1 = A.T1, T3 = A.T3;
  const int64_t tj = tj_of(A, s), tk = tk_of(A, s);
  float *muk = muj + 3 * N;
  for (int e = threadIdx.x; e < 3 * N; e += blockDim.x) {
    muj[e] = mu_at(A, tj, e / N)[e % N];
    muk[e] = mu_at(A, tk, e / N)[e % N];
  }
  __syncthreads();
  const int ne = A.T2 > 0 ? 2 * (T1 + T3) : T1 + T3;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    float2 acc[9];
    const bool se = e >= T1 + T3;  // the SE's own dots (T2 > 0)
    const int r = se ? e - (T1 + T3) : e;
    if (r < T1) {
      const float2 *l0 = se ? A.Linit + ((int64_t)s * A.linit_stride + r) * N : A.L0 + ((int64_t)s * 3 * T1 + r) * N;
      dot9(l0, l0 + (int64_t)T1 * N, l0 + (int64_t)2 * T1 * N, se ? muk : muj, N, N, acc);
      float2 *o = se ? g1s : g1;
#pragma unroll
      for (int j = 0; j < 9; ++j) o[((int64_t)s * 9 + j) * T1 + r] = acc[j];
    } else {
      const int t3 = r - T1;
      const float2 *m0 = A.hist + ((se ? hrow_M(A, s, 0) : hrow_G(A, s, 0)) + t3) * N;
      const int64_t rs = (int64_t)T3 * N;
      dot9(m0, m0 + rs, m0 + 2 * rs, mu_at(A, tk + t3, 0), N, N, acc);
      float2 *o = se ? g3P : g3;
#pragma unroll
      for (int j = 0; j < 9; ++j) o[((int64_t)s * T3 + t3) * 9 + j] = acc[j];
    }
  }
}

// GB and SE terms, one workgroup per (s, t3), one thread per t1: the t3
// dots SE t1n[x0][x3] = L[x0][t1](t3) . mu_x3(tk + t3) (src/calc_2DIR.c
// :384-390) with the static dots of gbse_static_kernel.  Writes
// E = -GB - SE for all 21 items.
__global__ void gbse_kernel(TwoDirArgs A, const float2 *g1, const float2 *g3, const float2 *g1s,
                            const float2 *g3P) {
  extern __shared__ __align__(16) float mul[];  // [3][N] mu(tk + t3)
  constexpr int kPol[21][4] = {
      {0, 0, 0, 0}, {1, 1, 1, 1}, {2, 2, 2, 2}, {0, 0, 1, 1}, {0, 0, 2, 2}, {1, 1, 0, 0},
      {1, 1, 2, 2}, {2, 2, 0, 0}, {2, 2, 1, 1}, {0, 1, 0, 1}, {0, 2, 0, 2}, {1, 0, 1, 0},
      {1, 2, 1, 2}, {2, 0, 2, 0}, {2, 1, 2, 1}, {0, 1, 1, 0}, {0, 2, 2, 0}, {1, 0, 0, 1},
      {1, 2, 2, 1}, {2, 0, 0, 2}, {2, 1, 1, 2}};
  const int s = blockIdx.x, t3 = blockIdx.y, N = A.N, T1 = A.T1;
  const int64_t tl = tk_of(A, s) + t3, plane = (int64_t)A.T3 * T1;
  for (int e = threadIdx.x; e < 3 * N; e += blockDim.x) mul[e] = mu_at(A, tl, e / N)[e % N];
  __syncthreads();
  float2 t3n[9], t3p[9], t3s[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    t3n[j] = cld(g3 + ((int64_t)s * A.T3 + t3) * 9 + j);
    t3p[j] = cld(g3P + ((int64_t)s * A.T3 + t3) * 9 + j);
    t3s[j] = cld(g3P + (int64_t)s * A.T3 * 9 + j);
    t3s[j].y = -t3s[j].y;
  }
  for (int t1 = threadIdx.x; t1 < T1; t1 += blockDim.x) {
    float2 se1[9], gb1[9], gs1[9];
    const float2 *lt = A.hist + (hrow_L(A, s, 0, t1) + t3) * N;
    const int64_t rs = (int64_t)T1 * A.T3 * N;
    dot9(lt, lt + rs, lt + 2 * rs, mul, N, N, se1);
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      gb1[j] = gld(g1 + ((int64_t)s * 9 + j) * T1 + t1);
      gs1[j] = gld(g1s + ((int64_t)s * 9 + j) * T1 + t1);
    }
#pragma unroll
    for (int m = 0; m < 21; ++m) {
      const int x0 = kPol[m][0], x1 = kPol[m][1], x2 = kPol[m][2], x3 = kPol[m][3];
      float e[4] = {0.f, 0.f, 0.f, 0.f};
      if (A.do_gb) {  // src/calc_2DIR.c:272-275
        const float t3nr = t3n[x2 * 3 + x3].x, t3ni = t3n[x2 * 3 + x3].y;
        const float t1nr = gb1[x0 * 3 + x1].x, t1ni = gb1[x0 * 3 + x1].y;
        e[0] -= __fsub_rn(__fmul_rn(t3ni, t1nr), __fmul_rn(t3nr, t1ni));
        e[1] -= __fadd_rn(__fmul_rn(t3nr, t1nr), __fmul_rn(t3ni, t1ni));
        e[2] -= __fadd_rn(__fmul_rn(t3ni, t1nr), __fmul_rn(t3nr, t1ni));
        e[3] -= __fsub_rn(__fmul_rn(t3nr, t1nr), __fmul_rn(t3ni, t1ni));
      }
      if (A.do_se) {  // src/calc_2DIR.c:397-400
        const float t3rr = t3p[x1 * 3 + x3].x, t3ri = t3p[x1 * 3 + x3].y;
        const float t3nr = t3s[x1 * 3 + x2].x, t3ni = t3s[x1 * 3 + x2].y;
        const float t1rr = gs1[x0 * 3 + x2].x, t1ri = -gs1[x0 * 3 + x2].y;
        const float t1nr = se1[x0 * 3 + x3].x, t1ni = se1[x0 * 3 + x3].y;
        e[0] -= __fadd_rn(__fmul_rn(t3rr, t1ri), __fmul_rn(t3ri, t1rr));
        e[1] -= __fsub_rn(__fmul_rn(t3rr, t1rr), __fmul_rn(t3ri, t1ri));
        e[2] -= __fadd_rn(__fmul_rn(t3ni, t1nr), __fmul_rn(t3nr, t1ni));
        e[3] -= __fsub_rn(__fmul_rn(t3nr, t1nr), __fmul_rn(t3ni, t1ni));
      }
      float *E = A.E + (((int64_t)s * 21 + m) * 4) * plane + (int64_t)t3 * T1 + t1;
#pragma unroll
      for (int q = 0; q < 4; ++q) gst(E + q * plane, e[q]);
    }
  }
}

// gbse_kernel with the history rows staged in LDS: per chunk of 64 t1 and
// per x0 the workgroup loads the 64 rows L[x0][t1](t3) with consecutive
// lanes on consecutive sites (row stride N + 1), then each lane forms its
// three dots of that x0 (each sum still in ascending site order).
__global__ void gbse_lds_kernel(TwoDirArgs A, const float2 *g1, const float2 *g3, const float2 *g1s,
                                const float2 *g3P) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int kPol[21][4] = {
      {0, 0, 0, 0}, {1, 1, 1, 1}, {2, 2, 2, 2}, {0, 0, 1, 1}, {0, 0, 2, 2}, {1, 1, 0, 0},
      {1, 1, 2, 2}, {2, 2, 0, 0}, {2, 2, 1, 1}, {0, 1, 0, 1}, {0, 2, 0, 2}, {1, 0, 1, 0},
      {1, 2, 1, 2}, {2, 0, 2, 0}, {2, 1, 2, 1}, {0, 1, 1, 0}, {0, 2, 2, 0}, {1, 0, 0, 1},
      {1, 2, 2, 1}, {2, 0, 0, 2}, {2, 1, 1, 2}};
  const int s = blockIdx.x, t3 = blockIdx.y, N = A.N, T1 = A.T1, t = threadIdx.x, ld = N + 1;
  float *mul = reinterpret_cast<float *>(smem);                                    // [3][N] mu(tk + t3)
  float2 *rowL = reinterpret_cast<float2 *>(smem + ((3 * N * 4 + 15) & ~15));    // [64][N + 1]
  const int64_t tl = tk_of(A, s) + t3, plane = (int64_t)A.T3 * T1;
  for (int e = t; e < 3 * N; e += 64) mul[e] = mu_at(A, tl, e / N)[e % N];
  float2 t3n[9], t3p[9], t3s[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    t3n[j] = cld(g3 + ((int64_t)s * A.T3 + t3) * 9 + j);
    t3p[j] = cld(g3P + ((int64_t)s * A.T3 + t3) * 9 + j);
    t3s[j] = cld(g3P + (int64_t)s * A.T3 * 9 + j);
    t3s[j].y = -t3s[j].y;
  }
  for (int c0 = 0; c0 < T1; c0 += 64) {  // uniform: every lane reaches the barriers
    const int t1 = c0 + t, nr = min(64, T1 - c0);
    float2 se1[9];
#pragma unroll
    for (int x0 = 0; x0 < 3; ++x0) {
      __syncthreads();  // the previous rows are read
      for (int e = t; e < nr * N; e += 64) {
        const int r = e / N, i = e - r * N;
        rowL[r * ld + i] = gld(A.hist + (hrow_L(A, s, x0, c0 + r) + t3) * N + i);
      }
      __syncthreads();
      float2 acc[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
      if (t1 < T1) {
        const float2 *v = rowL + t * ld;
        for (int i = 0; i < N; ++i) {
          const float2 vv = v[i];
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            acc[y].x = __fadd_rn(acc[y].x, __fmul_rn(mul[y * N + i], vv.x));
            acc[y].y = __fadd_rn(acc[y].y, __fmul_rn(mul[y * N + i], vv.y));
          }
        }
      }
#pragma unroll
      for (int y = 0; y < 3; ++y) se1[x0 * 3 + y] = acc[y];
    }
    if (t1 >= T1) continue;
    float2 gb1[9], gs1[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      gb1[j] = gld(g1 + ((int64_t)s * 9 + j) * T1 + t1);
      gs1[j] = gld(g1s + ((int64_t)s * 9 + j) * T1 + t1);
    }
#pragma unroll
    for (int m = 0; m < 21; ++m) {
      const int x0 = kPol[m][0], x1 = kPol[m][1], x2 = kPol[m][2], x3 = kPol[m][3];
      float e[4] = {0.f, 0.f, 0.f, 0.f};
      if (A.do_gb) {  // src/calc_2DIR.c:272-275
        const float t3nr = t3n[x2 * 3 + x3].x, t3ni = t3n[x2 * 3 + x3].y;
        const float t1nr = gb1[x0 * 3 + x1].x, t1ni = gb1[x0 * 3 + x1].y;
        e[0] -= __fsub_rn(__fmul_rn(t3ni, t1nr), __fmul_rn(t3nr, t1ni));
        e[1] -= __fadd_rn(__fmul_rn(t3nr, t1nr), __fmul_rn(t3ni, t1ni));
        e[2] -= __fadd_rn(__fmul_rn(t3ni, t1nr), __fmul_rn(t3nr, t1ni));
        e[3] -= __fsub_rn(__fmul_rn(t3nr, t1nr), __fmul_rn(t3ni, t1ni));
      }
      if (A.do_se) {  // src/calc_2DIR.c:397-400
        const float t3rr = t3p[x1 * 3 + x3].x, t3ri = t3p[x1 * 3 + x3].y;
        const float t3nr = t3s[x1 * 3 + x2].x, t3ni = t3s[x1 * 3 + x2].y;
        const float t1rr = gs1[x0 * 3 + x2].x, t1ri = -gs1[x0 * 3 + x2].y;
        const float t1nr = se1[x0 * 3 + x3].x, t1ni = se1[x0 * 3 + x3].y;
        e[0] -= __fadd_rn(__fmul_rn(t3rr, t1ri), __fmul_rn(t3ri, t1rr));
        e[1] -= __fsub_rn(__fmul_rn(t3rr, t1rr), __fmul_rn(t3ri, t1ri));
        e[2] -= __fadd_rn(__fmul_rn(t3ni, t1nr), __fmul_rn(t3nr, t1ni));
        e[3] -= __fsub_rn(__fmul_rn(t3nr, t1nr), __fmul_rn(t3ni, t1ni));
      }
      float *E = A.E + (((int64_t)s * 21 + m) * 4) * plane + (int64_t)t3 * T1 + t1;
#pragma unroll
      for (int q = 0; q < 4; ++q) gst(E + q * plane, e[q]);
    }
  }
}

// fr and ft = dipole_double(mu_x2(tk), ...) (src/NISE_subs.c:485-509):
// |aa> gets over_a c_a, |ab> gets (0 + mu_a c_b) + mu_b c_a.
// State j of the two-exciton vector vec of sample s: dipole_double of
// mu_x2(tk) on M[x1](tk) (vec < 9, x1 x2 = vec) or on L[x0][t1] at tk.
__device__ __forceinline__ float2 dd_value(const TwoDirArgs &A, int s, int vec, int j) {
  const int N = A.N;
  const int64_t tk = tk_of(A, s);
  int x2;
  const float2 *c;
  if (vec < 9) {
    const int x1 = vec / 3;
    x2 = vec % 3;
    c = A.hist + hrow_M(A, s, x1) * N;  // M[x1] at t3 = 0 (= leftrr at tk)
  } else {
    const int idx = vec - 9, x0 = idx / (3 * A.T1), t1 = idx % A.T1;
    x2 = (idx / A.T1) % 3;
    c = A.Linit + ((int64_t)s * A.linit_stride + x0 * A.T1 + t1) * N;
  }
  const float *dip = mu_at(A, tk, x2);
  int a = 0, rem = j;
  while (rem >= N - a) { rem -= N - a; ++a; }
  const int b = a + rem;
  float2 o;
  if (a == b) {  // two-level sites hold no |aa> (src/NISE_subs.c:533-549)
    const float ov = A.es ? 0.f
                     : (A.anh != 0.f) ? __fmul_rn(kSqrt2, dip[a])
                                      : A.over_rec[tk * (1 + 3 * (int64_t)N) + 1 + (int64_t)x2 * N + a];
    o = A.es ? make_float2(0.f, 0.f) : make_float2(__fmul_rn(ov, c[a].x), __fmul_rn(ov, c[a].y));
  } else {
    o = make_float2(__fadd_rn(__fmul_rn(dip[a], c[b].x), __fmul_rn(dip[b], c[a].x)),
                    __fadd_rn(__fmul_rn(dip[a], c[b].y), __fmul_rn(dip[b], c[a].y)));
  }
  return o;
}

__global__ void dd_kernel(TwoDirArgs A) {
  const int N2 = A.N * (A.N + 1) / 2;
  const int s = blockIdx.y, vec = blockIdx.z;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N2) return;
  A.dbl[((int64_t)s * (9 + 9 * A.T1) + vec) * N2 + j] = dd_value(A, s, vec, j);
}

// dip_x(tl) and the overtone dipoles over_x(tl) for x = 0..2 into LDS
// (src/calc_2DIR.c:450-454; sqrt(2) mu for a fixed anharmonicity,
// src/NISE_subs.c:556-560, else the overtone records).
__device__ inline void load_dip_over(const TwoDirArgs &A, int64_t tl, float *dipL, float *ovL) {
  const int N = A.N;
  for (int e = threadIdx.x; e < 3 * N; e += blockDim.x) {
    const int x = e / N, i = e % N;
    const float m = mu_at(A, tl, x)[i];
    dipL[e] = m;
    // two-level sites: no overtone term (src/NISE_subs.c:578-593); over_k
    // = 0 adds 0 * f_kk = 0 (f_kk stays 0) to the same running sum
    ovL[e] = A.es ? 0.f
             : (A.anh != 0.f) ? __fmul_rn(kSqrt2, m) : A.over_rec[tl * (1 + 3 * (int64_t)N) + 1 + (int64_t)x * N + i];
  }
}

// leftrr_EA at one t3 for every (s, x1, x2, x3): dipole_double_last of
// fr[x1][x2] with the x3 dipoles of frame tl (src/calc_2DIR.c:455-456).
// One workgroup per (s, x1 x2); the doubles vector is staged in LDS when
// kStage.  lrr[s][x1 x2][x3][N].
template <bool kStage>
__global__ void lrr_kernel(TwoDirArgs A, int t3, float2 *lrr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, N2 = N * (N + 1) / 2;
  const int s = blockIdx.x, v = blockIdx.y;
  float *dipL = reinterpret_cast<float *>(smem), *ovL = dipL + 3 * N;
  float2 *fL = reinterpret_cast<float2 *>(ovL + 3 * N);
  const int64_t tl = tk_of(A, s) + t3;
  const float2 *fg = A.dbl + ((int64_t)s * (9 + 9 * A.T1) + v) * N2;
  load_dip_over(A, tl, dipL, ovL);
  if (kStage)
    for (int e = threadIdx.x; e < N2; e += blockDim.x) fL[e] = fg[e];
  __syncthreads();
  const float2 *f = kStage ? fL : fg;
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const float ov[3] = {ovL[k], ovL[N + k], ovL[2 * N + k]};
    float2 o[3];
    ddl_multi<3>(f, N, k, dipL, dipL + N, dipL + 2 * N, ov, o);
    float2 *out = lrr + ((int64_t)s * 9 + v) * 3 * N + k;
    out[0] = o[0], out[N] = o[1], out[2 * N] = o[2];
  }
}

// EA terms at one t3 (src/calc_2DIR.c:448-469).  One workgroup per
// (s, x0 x2, chunk of C t1 values): the C doubles vectors ft[x0][x2][t1] are
// contiguous and staged in LDS with coalesced loads (kStage), each element
// is contracted once for the x3 of every polarisation item with this
// (x0, x2) (NIT = 2 or 3 items), then one thread per (t1, item, part) forms
// the two ordered dot products of that part and adds them into E.
template <int NIT, bool kStage>
__device__ __forceinline__ void ea_block(const TwoDirArgs &A, int t3, int C, int3 mi, const float2 *lrr,
                         unsigned char *smem) {
  const int N = A.N, N2 = N * (N + 1) / 2;
  const int s = blockIdx.z, x0 = blockIdx.y / 3, x2 = blockIdx.y % 3;
  const int t1_0 = blockIdx.x * C, Cc = min(C, A.T1 - t1_0);
  float *dipL = reinterpret_cast<float *>(smem), *ovL = dipL + 3 * N;
  float2 *RnL = reinterpret_cast<float2 *>(ovL + 3 * N);  // [3][N] M[x1](t3)
  float2 *RL = RnL + 3 * N;                               // [C][N] L[x0][t1](t3)
  float2 *lrL = RL + (int64_t)C * N;                      // [NIT][N] leftrr_EA of each item
  float2 *LvL = lrL + NIT * N;                            // [C][NIT][N]
  float2 *ftL = LvL + (int64_t)C * NIT * N;               // [C][N2]
  const int64_t tk = tk_of(A, s), tl = tk + t3;
  const float2 *ftg = A.dbl + ((int64_t)s * (9 + 9 * A.T1) + 9 + (x0 * 3 + x2) * A.T1 + t1_0) * N2;
  const int x3a = c_polar[mi.x][3], x3b = c_polar[mi.y][3], x3c = NIT > 2 ? c_polar[mi.z][3] : 0;
  // dipoles first (into registers when 3N <= 2 blockDim, stored after the
  // main pass), then one pass over the float2 words [M rows | t1 rows |
  // items' leftrr rows | doubles vectors (kStage)]
  const float *dg = mu_at(A, tl, 0);  // 3N contiguous: x N + i
  const float *og = A.over_rec + tl * (1 + 3 * (int64_t)N) + 1;
  const bool dreg = 3 * N <= 2 * (int)blockDim.x, ovr = !A.es && A.anh == 0.f;
  float dv[2] = {0.f, 0.f}, ovv[2] = {0.f, 0.f};
  if (dreg) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = threadIdx.x + u * blockDim.x;
      if (e < 3 * N) {
        dv[u] = gld(dg + e);
        if (ovr) ovv[u] = gld(og + e);
      }
    }
  }
  const float2 *Mrow = A.hist + (hrow_M(A, s, 0) + t3) * N;  // M[x] rows are T3 N apart
  const float2 *Lrow = A.hist + (hrow_L(A, s, x0, t1_0) + t3) * N;  // L[x0][t1] rows likewise
  const int64_t rstride = (int64_t)A.T3 * N;
  const float2 *lrr_s = lrr + (int64_t)s * 27 * N;
  const int lo0 = ((c_polar[mi.x][1] * 3 + x2) * 3 + x3a) * N, lo1 = ((c_polar[mi.y][1] * 3 + x2) * 3 + x3b) * N;
  const int lo2 = NIT > 2 ? ((c_polar[mi.z][1] * 3 + x2) * 3 + x3c) * N : 0;
  const int nsm = (3 + Cc + NIT) * N, fofs = (int)(ftL - RnL);
  gather_lds<16>(RnL, nsm + (kStage ? Cc * N2 : 0), [=](int e) {
    if (e >= nsm) return ftg + (e - nsm);
    const int r = e / N, i = e % N;
    if (r < 3) return Mrow + r * rstride + i;
    if (r < 3 + Cc) return Lrow + (r - 3) * rstride + i;
    const int it = r - 3 - Cc;
    return lrr_s + lo0 + (it >= 1) * (lo1 - lo0) + (it >= 2) * (lo2 - lo1) + i;
  }, [=](int e) {
    if (e >= nsm) return fofs + (e - nsm);  // doubles vectors, contiguous
    return e < (3 + Cc) * N ? e : e + (C - Cc) * N;
  });
  if (dreg) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = threadIdx.x + u * blockDim.x;
      if (e < 3 * N) {
        dipL[e] = dv[u];
        if (ovr) ovL[e] = ovv[u];
      }
    }
  } else {
    stage_lds<4>(dipL, dg, 3 * N);
    if (ovr) stage_lds<4>(ovL, og, 3 * N);
  }
  __syncthreads();
  const int x3v[3] = {x3a, x3b, x3c};
  for (int u = threadIdx.x; u < Cc * N; u += blockDim.x) {
    const int c = u / N, k = u % N;
    float ov[NIT];
    float2 o[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = x3v[it] * N + k;
      ov[it] = A.es ? 0.f : (A.anh != 0.f) ? __fmul_rn(kSqrt2, dipL[e]) : ovL[e];
    }
    const float2 *f = kStage ? ftL + c * N2 : ftg + (int64_t)c * N2;
    ddl_multi<NIT>(f, N, k, dipL + x3a * N, dipL + x3b * N, dipL + x3c * N, ov, o);
#pragma unroll
    for (int it = 0; it < NIT; ++it) LvL[(c * NIT + it) * N + k] = o[it];
  }
  __syncthreads();
  const int64_t plane = (int64_t)A.T3 * A.T1;
  for (int j = threadIdx.x; j < 2 * NIT * Cc; j += blockDim.x) {
    const int part = j & 1, it = (j >> 1) % NIT, c = (j >> 1) / NIT, t1 = t1_0 + c;
    const int m = it == 0 ? mi.x : it == 1 ? mi.y : mi.z, x1 = c_polar[m][1];
    float rr = 0.f, ri = 0.f;
    if (part == 0) {  // rightrr = conj(L[x0][t1](t3)) against leftrr_EA (:465-466)
      const float2 *R = RL + c * N, *lr = lrL + it * N;
      for (int i = 0; i < N; ++i) {
        const float2 l = lr[i], Rv = R[i];
        const float Rr = Rv.x, Ri = -Rv.y;
        rr = __fadd_rn(rr, __fadd_rn(__fmul_rn(l.y, Rr), __fmul_rn(l.x, Ri)));
        ri = __fadd_rn(ri, __fsub_rn(__fmul_rn(l.x, Rr), __fmul_rn(Ri, l.y)));
      }
    } else {  // rightnr = conj(M[x1](t3)) against leftnr_EA[t1] (:457-459, :468-469)
      const float2 *Rn = RnL + x1 * N, *Lv = LvL + (c * NIT + it) * N;
      for (int i = 0; i < N; ++i) {
        const float2 l = Lv[i], Rv = Rn[i];
        const float Nr = Rv.x, Ni = -Rv.y;
        rr = __fadd_rn(rr, __fadd_rn(__fmul_rn(Nr, l.y), __fmul_rn(Ni, l.x)));
        ri = __fadd_rn(ri, __fsub_rn(__fmul_rn(Nr, l.x), __fmul_rn(Ni, l.y)));
      }
    }
    float *E = A.E + (((int64_t)s * 21 + m) * 4 + 2 * part) * plane + (int64_t)t3 * A.T1 + t1;
    E[0] += rr, E[plane] += ri;
  }
}

template <bool kStage>
__global__ void ea_kernel(TwoDirArgs A, int t3, int C, const float2 *lrr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int x0 = blockIdx.y / 3, x2 = blockIdx.y % 3;
  int3 mi = make_int3(0, 0, 0);
  int n = 0;
  for (int m = 0; m < 21; ++m)
    if (c_polar[m][0] == x0 && c_polar[m][2] == x2) {
      if (n == 0) mi.x = m;
      else if (n == 1) mi.y = m;
      else mi.z = m;
      ++n;
    }
  if (n == 3) ea_block<3, kStage>(A, t3, C, mi, lrr, smem);
  else ea_block<2, kStage>(A, t3, C, mi, lrr, smem);
}

// ---------------------------------------------------------------------------
// Persistent EA + doubles kernel: one wave per group of doubles vectors of
// one sample, holding them in LDS for every t3.  At each
// t3 it adds the EA terms its vectors own and then applies the doubles step
// of frame tk + t3, so the two-exciton states never return to device memory.
//   kind 1 (ft vectors of one (x0, x2), t1 = t1_0 + c): the part of the EA
//     response against rightnr = conj(M[x1](t3)) (src/calc_2DIR.c:457-459,
//     :468-469) for every item with that (x0, x2);
//   kind 0 (fr[x1][x2] vectors): leftrr_EA for x3 = 0..2 (:455-456) and the
//     part against rightrr[t1] = conj(L[x0][t1](t3)) (:465-466) for every
//     item with that (x1, x2) and every t1.
// The operations and their order are those of lrr_kernel, ea_kernel and
// doubles_step_kernel, so the results are identical to that path.
// ---------------------------------------------------------------------------
struct DblGroup {
  int s, kind, v0, nv;  // sample, 0 fr / 1 ft, first vector (per-sample index), count
};

struct PersistArgs {
  TwoDirArgs A;
  const DblGroup *groups;
  int K2, m;
  const uint32_t *ab2;   // kept couplings (lexicographic)
  const float2 *D2;      // [frame - f_lo][N2]
  const float4 *cs2;     // [frame - f_lo][K2]
  int64_t f_lo;
  int npw;               // waves per workgroup
};

struct PersistLayout {
  size_t ix, d2, rn, dip, ov, cs, psi, lv, total;  // psi and lv: npw per-wave blocks
};
// Workgroups of npw waves: the waves hold groups of the same sample and
// share the staged frame tables, dipoles and M rows.
// ix: per coupling k, site block cb and lane c0 the rotation of site c =
// cb NP + c0 as two words: the byte offsets (o1 | o2 << 16, o = state index
// x VW x 8) of its two states in a wave's vectors, and (o3 | r << 16): the
// offset of the state that pre-rotates the first (|aa> on the c = b lane,
// else a neutral state) and the byte offset r of its coefficient row in
// cs.  Lanes with no rotation point o1 and o2 at a zero state after the N2
// states of the wave (rotating zeros leaves zeros); one such row ends the
// table.  cs: the frame's coefficient rows, three per coupling.  psi: N2 +
// 2 states of VW vectors per wave (the zero state and the neutral state).
// The EA phase of a t3 (rn, dip, ov, lv) and its doubles step (d2, cs) use
// the same bytes: a barrier separates them.
__host__ __device__ inline PersistLayout persist_layout(int N, int VW, int K2, int npw) {
  const size_t N2 = (size_t)N * (N + 1) / 2;
  const int NP = 64 / VW;
  PersistLayout L;
  size_t o = 0;
  L.ix = o, o += rot_table_bytes(N, NP, K2);
  const size_t u0 = o;
  L.rn = o, o += (size_t)3 * N * 8;
  L.dip = o, o += (size_t)3 * N * 4;
  L.ov = o, o += (size_t)3 * N * 4, o = (o + 15) & ~(size_t)15;
  L.lv = o, o += (size_t)npw * VW * 3 * N * 8;
  const size_t ea_end = o;
  o = u0;
  L.d2 = o, o += N2 * 8, o = (o + 15) & ~(size_t)15;
  L.cs = o, o += (size_t)K2 * 48;
  o = (o > ea_end ? o : ea_end), o = (o + 15) & ~(size_t)15;
  L.psi = o, o += (size_t)npw * (N2 + 2) * VW * 8;
  L.total = o;
  return L;
}

// sum_i (l, conj R) in the reference's order, R read from device memory
// eight elements ahead (part I, :465-466)
__device__ __forceinline__ void dot_part1(const float2 *lr, const float2 *R, int N, float &rr, float &ri) {
  for (int i0 = 0; i0 < N; i0 += 8) {
    float2 rv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) rv[u] = (i0 + u < N) ? gld(R + i0 + u) : make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i0 + u < N) {
        const float2 l = lr[i0 + u];
        const float Rr = rv[u].x, Ri = -rv[u].y;
        rr = __fadd_rn(rr, __fadd_rn(__fmul_rn(l.y, Rr), __fmul_rn(l.x, Ri)));
        ri = __fadd_rn(ri, __fsub_rn(__fmul_rn(l.x, Rr), __fmul_rn(Ri, l.y)));
      }
  }
}

// the slot-th item with (x1, x2) in molecular-polarisation order, or -1
__device__ inline int item_x1x2(int x1, int x2, int slot) {
  for (int m = 0; m < 21; ++m)
    if (c_polar[m][1] == x1 && c_polar[m][2] == x2 && slot-- == 0) return m;
  return -1;
}

// One wave's EA part for ft vectors (lane = threadIdx.x % 64).  Lanes read
// what other lanes of the same wave wrote to LDS: a wave's LDS operations
// run in order, so a compiler fence between the phases suffices.
template <int NP, int NIT>
__device__ __forceinline__ void persist_ft(const TwoDirArgs &A, int lane, int s, int t3, int v0, int nv,
                                           int3 mi, const float2 *psi, const float2 *RnL, float2 *LvL,
                                           const float *dipL, const float *ovL) {
  constexpr int VW = 64 / NP;
  const int N = A.N, T1 = A.T1, t1_0 = (v0 - 9) % T1;
  const int x3a = c_polar[mi.x][3], x3b = c_polar[mi.y][3], x3c = NIT > 2 ? c_polar[mi.z][3] : 0;
  const int x3v[3] = {x3a, x3b, x3c};
  // the lane of (c, k) forms the k-th rounded term of each item's dot
  // (leftnr_EA[t1] against rightnr = conj(M[x1](t3)), src/calc_2DIR.c
  // :457-459, :468-469); one lane per (c, item) then adds the terms in
  // ascending k, the reference's order
  const int x1v[3] = {c_polar[mi.x][1], c_polar[mi.y][1], NIT > 2 ? c_polar[mi.z][1] : 0};
  // the E entries of this lane's (vector, item) (NIT nv <= 24 lanes), read
  // before the contraction so that their latency hides behind it
  const int64_t plane = (int64_t)A.T3 * T1;
  const bool sum_lane = lane < NIT * nv;
  float *Ej = nullptr;
  float e0 = 0.f, e1 = 0.f;
  if (sum_lane) {
    const int it = lane % NIT, c = lane / NIT;
    const int m = it == 0 ? mi.x : it == 1 ? mi.y : mi.z;
    Ej = A.E + (((int64_t)s * 21 + m) * 4 + 2) * plane + (int64_t)t3 * T1 + t1_0 + c;
    e0 = gld(Ej), e1 = gld(Ej + plane);
  }
  for (int u = lane; u < nv * N; u += 64) {
    const int c = u / N, k = u % N;
    float ov[NIT];
    float2 o[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) ov[it] = ovL[x3v[it] * N + k];
    if ((N & 3) == 0)
      ddl_multi4<NIT, VW>(psi + c, N, k, dipL + x3a * N, dipL + x3b * N, dipL + x3c * N, ov, o);
    else
      ddl_multi<NIT, VW>(psi + c, N, k, dipL + x3a * N, dipL + x3b * N, dipL + x3c * N, ov, o);
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const float2 l = o[it], Rv = RnL[x1v[it] * N + k];
      const float Nr = Rv.x, Ni = -Rv.y;
      LvL[(c * NIT + it) * N + k] = make_float2(__fadd_rn(__fmul_rn(Nr, l.y), __fmul_rn(Ni, l.x)),
                                                __fsub_rn(__fmul_rn(Nr, l.x), __fmul_rn(Ni, l.y)));
    }
  }
  asm volatile("" ::: "memory");
  if (sum_lane) {
    const int it = lane % NIT, c = lane / NIT;
    const float2 *Tm = LvL + (c * NIT + it) * N;
    float rr = 0.f, ri = 0.f;
    for (int k = 0; k < N; ++k) {
      const float2 v = Tm[k];
      rr = __fadd_rn(rr, v.x), ri = __fadd_rn(ri, v.y);
    }
    gst(Ej, __fadd_rn(e0, rr)), gst(Ej + plane, __fadd_rn(e1, ri));
  }
  asm volatile("" ::: "memory");
}

// One wave's EA part for fr vectors.
template <int NP>
__device__ __forceinline__ void persist_fr(const TwoDirArgs &A, int lane, int s, int t3, int v0, int nv,
                                           const float2 *psi, float2 *LvL, const float *dipL, const float *ovL) {
  constexpr int VW = 64 / NP;
  const int N = A.N, T1 = A.T1;
  for (int u = lane; u < nv * N; u += 64) {
    const int c = u / N, k = u % N;
    const float ov[3] = {ovL[k], ovL[N + k], ovL[2 * N + k]};
    float2 o[3];
    if ((N & 3) == 0)
      ddl_multi4<3, VW>(psi + c, N, k, dipL, dipL + N, dipL + 2 * N, ov, o);
    else
      ddl_multi<3, VW>(psi + c, N, k, dipL, dipL + N, dipL + 2 * N, ov, o);
    LvL[(c * 3 + 0) * N + k] = o[0], LvL[(c * 3 + 1) * N + k] = o[1], LvL[(c * 3 + 2) * N + k] = o[2];
  }
  asm volatile("" ::: "memory");
  const int64_t plane = (int64_t)A.T3 * T1;
  for (int j = lane; j < nv * 3 * T1; j += 64) {
    const int t1 = j % T1, slot = (j / T1) % 3, c = j / (3 * T1), v = v0 + c;
    const int m = item_x1x2(v / 3, v % 3, slot);
    if (m < 0) continue;
    const int x0 = c_polar[m][0], x3 = c_polar[m][3];
    float rr = 0.f, ri = 0.f;
    dot_part1(LvL + (c * 3 + x3) * N, A.hist + (hrow_L(A, s, x0, t1) + t3) * N, N, rr, ri);
    float *E = A.E + (((int64_t)s * 21 + m) * 4 + 0) * plane + (int64_t)t3 * T1 + t1;
    E[0] += rr, E[plane] += ri;
  }
  asm volatile("" ::: "memory");
}

template <int NP>
__global__ void __launch_bounds__(1024) dbl_ea_kernel(PersistArgs P) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int VW = 64 / NP;
  const TwoDirArgs &A = P.A;
  const int N = A.N, N2 = N * (N + 1) / 2, K2 = P.K2, T1 = A.T1, T3 = A.T3, t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  const int npw = P.npw, nt = blockDim.x;
  const PersistLayout Ly = persist_layout(N, VW, K2, npw);
  uint2 *ixL = reinterpret_cast<uint2 *>(smem + Ly.ix);
  float2 *d2 = reinterpret_cast<float2 *>(smem + Ly.d2);
  float2 *RnL = reinterpret_cast<float2 *>(smem + Ly.rn);
  float *dipL = reinterpret_cast<float *>(smem + Ly.dip);
  float *ovL = reinterpret_cast<float *>(smem + Ly.ov);
  float4 *csL = reinterpret_cast<float4 *>(smem + Ly.cs);
  float2 *psi = reinterpret_cast<float2 *>(smem + Ly.psi) + (size_t)wave * (N2 + 2) * VW;
  float2 *LvL = reinterpret_cast<float2 *>(smem + Ly.lv) + (size_t)wave * VW * 3 * N;
  const DblGroup gd = P.groups[(size_t)blockIdx.x * npw + wave];
  const int s = P.groups[(size_t)blockIdx.x * npw].s;  // every wave of the workgroup: one sample
  const int kind = gd.kind, v0 = gd.v0, nv = gd.nv;
  // this wave's vectors at t3 = 0 (psi[j][w]), formed here from the seeds
  // and the history (dd_value)
  for (int w = 0; w < VW; ++w)
    for (int j = lane; j < N2; j += 64) psi[j * VW + w] = (w < nv) ? dd_value(A, s, v0 + w, j) : make_float2(0.f, 0.f);
  init_special_states<VW>(psi, N2, lane);
  build_rot_table<NP>(ixL, P.ab2, N, K2, A.es != 0);
  // ft groups: the items with this (x0, x2)
  int3 mi = make_int3(0, 0, 0);
  int nit = 0;
  if (kind == 1 && nv > 0) {
    const int idx = v0 - 9, x0 = idx / (3 * T1), x2 = (idx / T1) % 3;
    for (int m = 0; m < 21; ++m)
      if (c_polar[m][0] == x0 && c_polar[m][2] == x2) {
        if (nit == 0) mi.x = m;
        else if (nit == 1) mi.y = m;
        else mi.z = m;
        ++nit;
      }
  }
  const int q = lane % VW, c0 = lane / VW;
  const int64_t tk = tk_of(A, s);
  for (int t3 = 0; t3 < T3; ++t3) {
    const int64_t tl = tk + t3;
    load_dip_over(A, tl, dipL, ovL);
    for (int e = t; e < 3 * N; e += nt) RnL[e] = A.hist[(hrow_M(A, s, e / N) + t3) * N + e % N];
    __syncthreads();
    if (nv > 0) {
      if (kind == 1) {
        if (nit == 3) persist_ft<NP, 3>(A, lane, s, t3, v0, nv, mi, psi, RnL, LvL, dipL, ovL);
        else persist_ft<NP, 2>(A, lane, s, t3, v0, nv, mi, psi, RnL, LvL, dipL, ovL);
      } else {
        persist_fr<NP>(A, lane, s, t3, v0, nv, psi, LvL, dipL, ovL);
      }
    }
    if (t3 + 1 < T3) {  // the doubles step of frame tk + t3 (the state after the last t3 is not used)
      const int64_t row = tk + t3 - P.f_lo;
      __syncthreads();  // d2 and cs take the bytes of the EA phase's rows
      stage_step_rows(d2, csL, P.D2 + row * N2, P.cs2 + row * K2, N2, K2);
      __syncthreads();
      for (int sub = 0; sub < (nv > 0 ? P.m : 0); ++sub) {
        dbl_diag<VW>(psi, d2, N2, lane);
        dbl_rotate_tab<NP>(psi, K2 * ((N + NP - 1) / NP), ixL, csL, q, c0);
        dbl_diag<VW>(psi, d2, N2, lane);
      }
    }
    __syncthreads();  // the staged rows are overwritten at the next t3
  }
}

// part[c][k][t3][t1] = sum over samples s of chunk c (kRedS samples) and
// items m of E[s][m][q][t3][t1] * (w(p, m) lt(t3, t1)), k = 4p + q, in
// ascending (s, m) order and double accumulation; reduce2d_final_kernel adds
// the chunks in ascending order to R.
constexpr int kRedS = 16;
__global__ void reduce2d_part_kernel(const float *E, int S, int T3, int T1, const float *wpl, const float *lt,
                                     double *part) {
  const int64_t plane = (int64_t)T3 * T1;
  const int64_t e0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y, p = k / 4, q = k % 4, c = blockIdx.z;
  if (e0 >= plane) return;
  const float l = lt[e0];
  float pw[21];
#pragma unroll
  for (int m = 0; m < 21; ++m) pw[m] = __fmul_rn(cld(wpl + p * 21 + m), l);
  double acc = 0.0;
  const int s1 = min(S, (c + 1) * kRedS);
  for (int s = c * kRedS; s < s1; ++s) {
    const float *Es = E + ((int64_t)s * 84 + q) * plane + e0;
    float ev[21];
#pragma unroll
    for (int m = 0; m < 21; ++m) ev[m] = gld(Es + (int64_t)m * 4 * plane);
#pragma unroll
    for (int m = 0; m < 21; ++m)
      if (pw[m] != 0.f) acc += (double)__fmul_rn(ev[m], pw[m]);
  }
  part[((int64_t)c * 12 + k) * plane + e0] = acc;
}

__global__ void reduce2d_final_kernel(const double *part, int nc, int64_t n, double *R) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  double acc = 0.0;
  for (int c = 0; c < nc; ++c) acc += part[(int64_t)c * n + idx];
  R[idx] += acc;
}

}  // namespace nise

using namespace nise;

namespace {
int upl(DevBuf &d, const void *s, size_t n) {
  int rc = d.alloc(n < 16 ? 16 : n);
  if (rc) return rc;
  if (n && s) NISE_HIP_CHECK(hipMemcpy(d.p, s, n, hipMemcpyHostToDevice));
  return NISE_OK;
}
bool tech_is(const t_non *n, const char *t) { return strcmp(n->technique, t) == 0; }

// EA launch shape: doubles vectors are staged in LDS when one fits in
// 96 KB; C vectors per workgroup (at most 8), about 40 KB of them.
struct EaShape {
  bool stage;
  int C, nchunk, block, lblock;
  size_t smem, lsmem;
};
EaShape ea_shape(int N, int T1) {
  EaShape e;
  const int N2 = N * (N + 1) / 2;
  e.stage = (size_t)N2 * 8 <= 96 * 1024;
  e.C = e.stage ? std::max(1, std::min({8, T1, 40960 / (N2 * 8)})) : 1;
  e.nchunk = (T1 + e.C - 1) / e.C;
  e.block = std::min(256, std::max(64, (e.C * N + 63) / 64 * 64));
  e.lblock = std::min(256, std::max(64, (N + 63) / 64 * 64));
  // dip, over, M rows, items' leftrr rows; per t1: its L row, contractions
  const size_t fixed = (size_t)6 * N * 4 + (size_t)6 * N * 8;
  e.smem = fixed + (size_t)e.C * 4 * N * 8 + (e.stage ? (size_t)e.C * N2 * 8 : 0);
  e.lsmem = (size_t)6 * N * 4 + (e.stage ? (size_t)N2 * 8 : 0);
  return e;
}
}  // namespace

namespace {
// Device buffers of the calling thread's 2D runs, kept between calls: a run
// of the same shape allocates and frees nothing (the N = 32 ring run holds
// about 3 GB of history, two-exciton vectors and item planes).  Every call
// ends with its stream synchronised.
struct TwoDirBufs {
  DevBuf Hb, mub, Cb, Ab, Ob, mk, ab, lay, diag, cs, ab2, D2, cs2, stp, wpl, ltb, Rd;
  DevBuf L0, hist, dbl, E, part, slots, wgb, slots2, wgb2, vb, rows, lrrb, grp, g1b, g3b;
  // waiting time and the exact-exponential / sparse schemes
  DevBuf Li, g1s, g3P, Ul, Ut2, Ud, jA, jV, jw, jst, el, cptr, cidx, ecnt, dsl, dwb, dsl2, dwb2, dsl3, dwb3;
};

// Groups of DIA slots for launch_dia_slots: consecutive slots with the same
// first frame and step counts, at most max_ns per workgroup.
void dia_groups(const std::vector<DiaSlot> &sl, int max_ns, std::vector<int> &wb) {
  wb.clear();
  for (int i = 0; i < (int)sl.size(); ++i) {
    const bool same = !wb.empty() && i - wb.back() < max_ns && sl[i].fs == sl[wb.back()].fs &&
                      sl[i].nresp == sl[wb.back()].nresp && sl[i].nprop == sl[wb.back()].nprop;
    if (!same) wb.push_back(i);
  }
  wb.push_back((int)sl.size());
}
}  // namespace

extern "C" int nise_2dir_batched(const t_non *non, const float *H_rec, int64_t H_stride,
                                 const float *mu_rec, int64_t mu_stride, int64_t nframes,
                                 const float *coupling_rec, const float *anh_rec,
                                 const float *over_rec, int where, float *R12) {
  int rc = ensure_device();
  if (rc) return rc;
  const int N = non->singles, T1 = non->tmax1, T3 = non->tmax3;
  if (N < 1 || T1 < 1 || T3 < 1 || non->sample < 1 || non->ts < 1 || !R12)
    return set_error("2DIR: bad sizes"), NISE_ERR_ARG;
  if (N > 65535) return set_error("2DIR: N=%d exceeds 65535 sites", N), NISE_ERR_ARG;
  const int prop = non->propagation, T2 = non->tmax2;
  if (prop != 1 && prop != 0)
    return set_error("2DIR: propagation %d is not provided (1 Coupling, 0 Sparse)", prop), NISE_ERR_ARG;
  if (T2 < 0) return set_error("2DIR: tmax2 < 0"), NISE_ERR_ARG;
  // the exact exponential diagonalises N x N frames (FP64 Jacobi in LDS);
  // the sparse two-exciton step holds a frame's element lists in LDS
  if ((T2 > 0 || prop == 0) && N > 100)
    return set_error("2DIR: waiting time and propagation 0 support N <= 100, got %d", N), NISE_ERR_ARG;
  if (prop == 0 && N > 64) return set_error("2DIR: propagation 0 supports N <= 64, got %d", N), NISE_ERR_ARG;
  if (non->cluster != -1) return set_error("2DIR: cluster selection is not provided"), NISE_ERR_ARG;
  // calc_2DES (src/calc_2DES.c:22-586) is this computation for two-level
  // sites: the UVvis techniques, two-exciton states without |aa>
  const int es = tech_is(non, "GBUVvis") || tech_is(non, "SEUVvis") || tech_is(non, "EAUVvis") ||
                 tech_is(non, "2DUVvis") || tech_is(non, "noEAUVvis");
  const int do_gb = tech_is(non, "GBIR") || tech_is(non, "2DIR") || tech_is(non, "noEAIR") ||
                    tech_is(non, "GBUVvis") || tech_is(non, "2DUVvis") || tech_is(non, "noEAUVvis");
  const int do_se = tech_is(non, "SEIR") || tech_is(non, "2DIR") || tech_is(non, "noEAIR") ||
                    tech_is(non, "SEUVvis") || tech_is(non, "2DUVvis") || tech_is(non, "noEAUVvis");
  const int do_ea = tech_is(non, "EAIR") || tech_is(non, "2DIR") || tech_is(non, "EAUVvis") ||
                    tech_is(non, "2DUVvis");
  if (!do_gb && !do_se && !do_ea) return set_error("2DIR: unknown technique %s", non->technique), NISE_ERR_ARG;
  const bool ctype = strcmp(non->hamiltonian, "Coupling") == 0;
  if (!ctype && strcmp(non->hamiltonian, "Full") != 0)
    return set_error("2DIR: HamiltonianType %s is not provided", non->hamiltonian), NISE_ERR_ARG;
  if (ctype && !coupling_rec) return set_error("2DIR: Coupling type needs coupling_rec"), NISE_ERR_ARG;
  if (!es && non->anharmonicity == 0.f && do_ea && (!anh_rec || !over_rec))
    return set_error("2DIR: anharmonicity 0 needs anharmonicity and overtone records"), NISE_ERR_ARG;
  const int begin = non->begin;
  int end = non->end;
  if (end == 0) end = (int)((nframes - T1 - non->tmax2 - T3 - 1) / non->sample + 1);  // src/MPI_subs.c:51,62
  if (end <= begin) return set_error("2DIR: no samples (begin %d, end %d)", begin, end), NISE_ERR_RANGE;
  const int64_t f_lo = (int64_t)begin * non->sample;
  const int64_t f_hi = (int64_t)(end - 1) * non->sample + T1 + T2 + T3;
  if (f_hi > nframes)
    return set_error("2DIR: trajectory has %lld frames, the samples need %lld", (long long)nframes,
                     (long long)f_hi), NISE_ERR_RANGE;
  const int64_t nfr = f_hi - f_lo;
  hipStream_t st = default_stream();
  const int N2 = N * (N + 1) / 2;
  const bool anh_file = do_ea && !es && non->anharmonicity == 0.f;
  static thread_local TwoDirBufs B;
  NISE_HIP_CHECK(hipStreamSynchronize(st));  // a call that returned early may have left work queued

  // inputs on the device (host records are uploaded for [f_lo, f_hi) and
  // indexed by absolute frame through a shifted base pointer)
  DevBuf &Hb = B.Hb, &mub = B.mub, &Cb = B.Cb, &Ab = B.Ab, &Ob = B.Ob;
  const float *Hd = H_rec, *mud = mu_rec, *Cd = coupling_rec, *Ad = anh_rec, *Od = over_rec;
  if (where == NISE_HOST) {
    if ((rc = upl(Hb, H_rec + f_lo * H_stride, (size_t)(nfr * H_stride * 4)))) return rc;
    Hd = Hb.as<float>() - f_lo * H_stride;
    if (ctype) {
      if ((rc = upl(mub, mu_rec, (size_t)mu_stride * 4))) return rc;
      mud = mub.as<float>();
      if ((rc = upl(Cb, coupling_rec, (size_t)(1 + N2) * 4))) return rc;
      Cd = Cb.as<float>();
    } else {
      if ((rc = upl(mub, mu_rec + f_lo * mu_stride, (size_t)(nfr * mu_stride * 4)))) return rc;
      mud = mub.as<float>() - f_lo * mu_stride;
    }
    if (anh_file) {
      if ((rc = upl(Ab, anh_rec + f_lo * (1 + N), (size_t)(nfr * (1 + N) * 4)))) return rc;
      Ad = Ab.as<float>() - f_lo * (1 + N);
      if ((rc = upl(Ob, over_rec + f_lo * (1 + 3 * N), (size_t)(nfr * (1 + 3 * N) * 4)))) return rc;
      Od = Ob.as<float>() - f_lo * (1 + 3 * N);
    }
  } else if (where != NISE_DEVICE) {
    return set_error("2DIR: where must be NISE_HOST or NISE_DEVICE"), NISE_ERR_ARG;
  }
  FrameSource src;
  src.H_rec = Hd, src.H_stride = H_stride, src.coupling_type = ctype, src.C_rec = Cd;
  src.shifte = (non->max1 + non->min1) / 2;  // src/calc_2DIR.c:79-81
  src.cut = non->couplingcut, src.apply_shift = 0;  // read_He applies no single-site shifts

  // singles schedule over the union of kept couplings, per-frame tables
  const int64_t npair = (int64_t)N * (N - 1) / 2;
  std::vector<uint8_t> mask(npair);
  if (npair) {
    DevBuf &mk = B.mk;
    if ((rc = mk.alloc(npair))) return rc;
    if ((rc = launch_union_mask(src, N, f_lo, nfr, mk.as<uint8_t>(), st))) return rc;
    NISE_HIP_CHECK(hipMemcpyAsync(mask.data(), mk.p, npair, hipMemcpyDeviceToHost, st));
    NISE_HIP_CHECK(hipStreamSynchronize(st));
  }
  Schedule sch;
  build_schedule(N, mask, sch);
  std::vector<uint32_t> lex;  // the same couplings in lexicographic order (doubles)
  {
    int64_t p = 0;
    for (int a = 0; a < N; ++a)
      for (int b = a + 1; b < N; ++b, ++p)
        if (mask[p]) lex.push_back((uint32_t)a | ((uint32_t)b << 16));
  }
  const float fac = trotter_factor(non->deltat, 1, non->ts);
  DevBuf &ab = B.ab, &lay = B.lay, &diag = B.diag, &cs = B.cs, &ab2 = B.ab2, &D2 = B.D2, &cs2 = B.cs2, &stp = B.stp;
  if ((rc = upl(ab, sch.ab.data(), sch.ab.size() * 4)) ||
      (rc = upl(lay, sch.layer_off.data(), sch.layer_off.size() * 4)))
    return rc;
  if (prop == 1) {
    if ((rc = diag.alloc((size_t)nfr * N * 8)) || (rc = cs.alloc((size_t)nfr * std::max(sch.K, 1) * 8)))
      return rc;
    if ((rc = launch_frame_tables(src, N, sch, ab.as<uint32_t>(), fac, f_lo, nfr, diag.as<float2>(),
                                  cs.as<float2>(), st)))
      return rc;
  }
  // per-frame exact exponentials: the singles of propagation 0
  // (propagate_vector: DIA when thres is 0 or above 1, else the sparse
  // DIA_S, src/propagate.c:25-33; the EA's right vectors take DIA_S,
  // :1519-1525, which keeps the same elements for thres <= 1), the waiting
  // time (propagate_t2_DIA, no Trotter split, nothing dropped) and the
  // one-exciton operator of the sparse two-exciton step (time_evolution_mat,
  // f / Trotter, elements with |U|^2 <= thres dropped, src/propagate.c:1241-1302)
  volatile float fexp_v = non->deltat * kIcm2ifs;  // src/propagate.c:69
  fexp_v = fexp_v * kTwoPi;
  const float fexp = fexp_v;
  const float thres = non->thres;
  const float thres_l = (thres > 0.f && thres <= 1.f) ? thres : 0.f;
  const float fdbl = fexp / (float)non->ts;  // src/propagate.c:1253
  const float fm_dbl = (float)((double)fexp * 0.5 / non->ts);  // src/propagate.c:1319
  const float2 *Uleft = nullptr, *Utwo = nullptr, *Udbl = nullptr;
  int64_t t2_lo = 0;
  {
    DevBuf &jst = B.jst;
    if ((rc = jst.alloc((size_t)3 * nfr * 4))) return rc;
    std::vector<int64_t> jn;  // frames per table, for the convergence check
    if (prop == 0) {
      if ((rc = B.Ul.alloc((size_t)nfr * N * N * 8))) return rc;
      if ((rc = launch_dia_tables(src, N, f_lo, nfr, fexp, thres_l, B.Ul.as<float2>(), B.jA, B.jV, B.jw,
                                  B.jst.as<int>(), st)))
        return rc;
      Uleft = B.Ul.as<float2>();
      jn.push_back(nfr);
      if (do_ea) {
        if (non->ts == 1 && thres_l == thres) {
          Udbl = Uleft;
        } else {
          if ((rc = B.Ud.alloc((size_t)nfr * N * N * 8))) return rc;
          if ((rc = launch_dia_tables(src, N, f_lo, nfr, fdbl, thres, B.Ud.as<float2>(), B.jA, B.jV, B.jw,
                                      B.jst.as<int>() + nfr, st)))
            return rc;
          Udbl = B.Ud.as<float2>();
          jn.push_back(nfr);
        }
        const int64_t NN = (int64_t)N * N;
        if ((rc = B.el.alloc((size_t)nfr * NN * 16)) || (rc = B.cidx.alloc((size_t)nfr * NN * 4)) ||
            (rc = B.cptr.alloc((size_t)nfr * (N + 1) * 4)) || (rc = B.ecnt.alloc((size_t)nfr * 4)))
          return rc;
        if ((rc = launch_sparse_lists(Udbl, N, nfr, thres, B.el.as<float4>(), B.cptr.as<int>(), B.cidx.as<int>(),
                                      B.ecnt.as<int>(), st)))
          return rc;
      }
    }
    if (T2 > 0) {  // frames tj .. tj + T2 - 1 of every sample
      t2_lo = f_lo + T1;
      const int64_t n2 = f_hi - T3 - t2_lo;
      if (prop == 0 && thres_l == 0.f) {
        Utwo = Uleft + (t2_lo - f_lo) * N * N;
      } else {
        if ((rc = B.Ut2.alloc((size_t)n2 * N * N * 8))) return rc;
        if ((rc = launch_dia_tables(src, N, t2_lo, n2, fexp, 0.f, B.Ut2.as<float2>(), B.jA, B.jV, B.jw,
                                    B.jst.as<int>() + 2 * nfr, st)))
          return rc;
        Utwo = B.Ut2.as<float2>();
        jn.push_back(-n2);  // third slot of the status words
      }
    }
    if (!jn.empty()) {  // every frame's eigensystem converged
      std::vector<int> sts((size_t)3 * nfr, 0);
      NISE_HIP_CHECK(hipMemcpyAsync(sts.data(), jst.p, sts.size() * 4, hipMemcpyDeviceToHost, st));
      NISE_HIP_CHECK(hipStreamSynchronize(st));
      for (size_t k = 0; k < jn.size(); ++k) {
        const int64_t base = jn[k] < 0 ? 2 * nfr : (int64_t)k * nfr, n = jn[k] < 0 ? -jn[k] : jn[k];
        for (int64_t i = 0; i < n; ++i)
          if (sts[base + i] < 0) return set_error("2DIR: Jacobi did not converge for frame %lld", (long long)i), NISE_ERR_NUMERIC;
      }
    }
  }
  if (do_ea && prop == 1) {
    if ((rc = upl(ab2, lex.data(), lex.size() * 4))) return rc;
    if ((rc = D2.alloc((size_t)nfr * N2 * 8)) ||
        (rc = cs2.alloc((size_t)nfr * std::max((int)lex.size(), 1) * 16)))
      return rc;
    if ((rc = launch_doubles_tables(src, N, (int)lex.size(), ab2.as<uint32_t>(), fac,
                                    es ? 0.f : non->anharmonicity, es ? nullptr : Ad, f_lo, nfr,
                                    D2.as<float2>(), cs2.as<float4>(), st)))
      return rc;
  }
  // conveyor shape: the largest slot count whose LDS image fits
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
  if (!vp) return set_error("2DIR: N=%d does not fit the conveyor's LDS image", N), NISE_ERR_ARG;
  if (do_ea && std::max(ea_shape(N, T1).smem, ea_shape(N, T1).lsmem) > 160 * 1024)
    return set_error("2DIR: N=%d does not fit the EA kernel's LDS", N), NISE_ERR_ARG;
  if ((rc = upl(stp, steps.data(), steps.size() * 8))) return rc;

  // response weights: polarweight x lifetime factor (src/calc_2DIR.c:105-115)
  std::vector<float> wh(63), lth((size_t)T3 * T1);
  for (int p = 0; p < 3; ++p)
    for (int m = 0; m < 21; ++m) wh[p * 21 + m] = polarweight_h(p, m);
  for (int t3 = 0; t3 < T3; ++t3)
    for (int t1 = 0; t1 < T1; ++t1)
      lth[(size_t)t3 * T1 + t1] =
          (non->lifetime > 0) ? (float)exp(-(double)(t1 + t3) * non->deltat / (2 * non->lifetime)) : 1.0f;
  DevBuf &wpl = B.wpl, &ltb = B.ltb, &Rd = B.Rd;
  const int64_t plane = (int64_t)T3 * T1;
  if ((rc = upl(wpl, wh.data(), wh.size() * 4)) || (rc = upl(ltb, lth.data(), lth.size() * 4))) return rc;
  if ((rc = Rd.alloc((size_t)12 * plane * 8))) return rc;
  NISE_HIP_CHECK(hipMemsetAsync(Rd.p, 0, (size_t)12 * plane * 8, st));

  // samples per chunk, sized to a device memory budget
  const int total = end - begin;
  const int nvec = 9 + 9 * T1;  // doubles vectors per sample
  const int nrow = 3 * T1 + 3 + (T2 > 0 ? 3 : 0);  // history rows per sample
  const double per_sample = 8.0 * 3 * T1 * N + 8.0 * nrow * T3 * N + (T2 > 0 ? 8.0 * (3 * T1 + 3) * N : 0.0) +
                            (do_ea ? 8.0 * nvec * N2 : 0.0) + 4.0 * 84 * plane;
  const int chunk = (int)std::max(1.0, std::min<double>({(double)total, 1024.0, 8.0e9 / per_sample}));
  const int vw = doubles_vw(N);
  const int gps = (nvec + vw - 1) / vw;  // doubles groups per sample

  // the seed and history conveyors have their own slot buffers: a host copy
  // into a buffer that a queued kernel still reads would not wait for it
  DevBuf &L0 = B.L0, &hist = B.hist, &dbl = B.dbl, &E = B.E, &slots = B.slots, &wgb = B.wgb, &vb = B.vb,
         &slots2 = B.slots2, &wgb2 = B.wgb2,
         &rows = B.rows, &lrrb = B.lrrb, &grp = B.grp, &g1b = B.g1b, &g3b = B.g3b, &part = B.part;
  const EaShape ea = ea_shape(N, T1);
  ConveyorArgs a;
  a.N = N, a.K = sch.K, a.nlayers = sch.nlayers(), a.m = non->ts, a.f_lo = f_lo;
  a.ab = ab.as<uint32_t>(), a.layer_off = lay.as<int>(), a.diag = diag.as<float2>(), a.cs = cs.as<float2>();
  a.mu_rec = mud, a.mu_stride = ctype ? 0 : mu_stride;
  a.steps = stp.as<int2>(), a.nsteps = (int)steps.size();

  for (int c0 = 0; c0 < total; c0 += chunk) {
    const int S = std::min(chunk, total - c0);
    TwoDirArgs A;
    A.N = N, A.T1 = T1, A.T3 = T3, A.S = S, A.s0 = begin + c0, A.sample = non->sample;
    A.T2 = T2, A.nrow = nrow;
    A.do_gb = do_gb, A.do_se = do_se, A.do_ea = do_ea, A.es = es, A.anh = non->anharmonicity;
    A.mu_rec = mud, A.mu_stride = ctype ? 0 : mu_stride, A.over_rec = Od;
    if ((rc = L0.alloc((size_t)S * 3 * T1 * N * 8)) ||
        (rc = hist.alloc((size_t)S * nrow * T3 * N * 8)) ||
        (rc = E.alloc((size_t)S * 84 * plane * 4)) || (rc = g1b.alloc((size_t)S * 9 * T1 * 8)) ||
        (rc = g3b.alloc((size_t)S * 9 * T3 * 8)) ||
        (rc = part.alloc((size_t)((S + kRedS - 1) / kRedS) * 12 * plane * 8)))
      return rc;
    if (T2 > 0 && ((rc = B.Li.alloc((size_t)S * (3 * T1 + 3) * N * 8)) || (rc = B.g1s.alloc((size_t)S * 9 * T1 * 8)) ||
                   (rc = B.g3P.alloc((size_t)S * 9 * T3 * 8))))
      return rc;
    A.L0 = L0.as<float2>(), A.hist = hist.as<float2>(), A.dbl = dbl.as<float2>(), A.E = E.as<float>();
    A.Linit = T2 > 0 ? B.Li.as<float2>() : L0.as<float2>();
    A.linit_stride = T2 > 0 ? 3 * T1 + 3 : 3 * T1;
    const float2 *g1s = T2 > 0 ? B.g1s.as<float2>() : g1b.as<float2>();
    const float2 *g3P = T2 > 0 ? B.g3P.as<float2>() : g3b.as<float2>();
    DiaSlotArgs ds;
    ds.N = N, ds.f_lo = f_lo, ds.U = Uleft, ds.mu_rec = mud, ds.mu_stride = ctype ? 0 : mu_stride;
    const int dmax = dia_slots_max(N);

    // (1) GB seeds L0[x][t1]: mu_x(tj - t1) propagated through frames
    //     tj-t1 .. tj-1 (src/calc_2DIR.c:215-239); t1 = 0 is mu_x(tj)
    hipLaunchKernelGGL(seed0_kernel, dim3(S, 3), dim3(64), 0, st, A, L0.as<float2>());
    NISE_HIP_CHECK(hipGetLastError());
    std::vector<SlotDesc> sl;
    std::vector<int> wb;
    if (prop == 0) {  // the exact exponential / sparse scheme, three polarisations per (s, t1)
      std::vector<DiaSlot> dl;
      for (int s = 0; s < S; ++s) {
        const int tj = (begin + c0 + s) * non->sample + T1;
        for (int t1 = 1; t1 < T1; ++t1)
          for (int x = 0; x < 3; ++x) dl.push_back(DiaSlot{tj - t1, 0, t1, x, -1, (s * 3 + x) * T1 + t1, -1});
      }
      dia_groups(dl, dmax, wb);
      if (!dl.empty()) {
        if ((rc = upl(B.dsl, dl.data(), dl.size() * sizeof(DiaSlot))) || (rc = upl(B.dwb, wb.data(), wb.size() * 4)))
          return rc;
        ds.slots = B.dsl.as<DiaSlot>(), ds.wg_begin = B.dwb.as<int>(), ds.nwg = (int)wb.size() - 1;
        ds.in = nullptr, ds.out = L0.as<float2>(), ds.hist = nullptr;
        if ((rc = launch_dia_slots(ds, dmax, st))) return rc;
      }
      wb.clear();
    } else {
      for (int s = 0; s < S; ++s) {
        const int tj = (begin + c0 + s) * non->sample + T1;
        for (int x = 0; x < 3; ++x) {
          const int first = (int)sl.size();
          for (int t1 = 1; t1 < T1; ++t1)
            sl.push_back(SlotDesc{tj - t1, 0, t1, x, 0, (s * 3 + x) * T1 + t1, 0, -1});
          for (int b = first; b < (int)sl.size(); b += vp) wb.push_back(b);
        }
      }
      wb.push_back((int)sl.size());
    }
    if (prop == 1 && !sl.empty()) {
      if ((rc = upl(slots, sl.data(), sl.size() * sizeof(SlotDesc))) || (rc = upl(wgb, wb.data(), wb.size() * 4)))
        return rc;
      a.slots = slots.as<SlotDesc>(), a.wg_slot_begin = wgb.as<int>(), a.nwg = (int)wb.size() - 1;
      a.state = L0.as<float2>(), a.writeback = 1, a.hist = nullptr;
      if ((rc = launch_conveyor(a, vp, block, st))) return rc;
    }
    // (1b) waiting time: L0[x][t1] and mu_x(tj) through frames tj .. tj+T2-1
    //      by the exact exponential (src/calc_2DIR.c:303-326) into Linit
    if (T2 > 0) {
      std::vector<DiaSlot> dl;
      for (int s = 0; s < S; ++s) {
        const int tj = (begin + c0 + s) * non->sample + T1, base = s * (3 * T1 + 3);
        for (int x = 0; x < 3; ++x)
          for (int t1 = 0; t1 < T1; ++t1)
            dl.push_back(DiaSlot{tj, 0, T2, x, (s * 3 + x) * T1 + t1, base + x * T1 + t1, -1});
        for (int x = 0; x < 3; ++x) dl.push_back(DiaSlot{tj, 0, T2, x, -1, base + 3 * T1 + x, -1});
      }
      dia_groups(dl, dmax, wb);
      if ((rc = upl(B.dsl2, dl.data(), dl.size() * sizeof(DiaSlot))) || (rc = upl(B.dwb2, wb.data(), wb.size() * 4)))
        return rc;
      DiaSlotArgs d2 = ds;
      d2.U = Utwo, d2.f_lo = t2_lo;
      d2.slots = B.dsl2.as<DiaSlot>(), d2.wg_begin = B.dwb2.as<int>(), d2.nwg = (int)wb.size() - 1;
      d2.in = L0.as<float2>(), d2.out = B.Li.as<float2>(), d2.hist = nullptr;
      if ((rc = launch_dia_slots(d2, dmax, st))) return rc;
    }
    // (2) history at t3 = 0..T3-1 from tk: L[x][t1] (from Linit), M[x] (from
    //     mu_x(tk), or from Linit after a waiting time) and, after a waiting
    //     time, the GB's M[x] from mu_x(tk)
    sl.clear(), wb.clear();
    std::vector<DiaSlot> hl;
    for (int s = 0; s < S; ++s) {
      const int tk = (begin + c0 + s) * non->sample + T1 + T2;
      const int base = s * nrow, lb = s * A.linit_stride;
      for (int x = 0; x < 3; ++x)
        for (int t1 = 0; t1 < T1; ++t1) {
          const int lrow = lb + x * T1 + t1, h = (base + x * T1 + t1) * T3;
          sl.push_back(SlotDesc{tk, T3, T3 - 1, x, 1, lrow, 0, h});
          hl.push_back(DiaSlot{tk, T3, T3 - 1, x, lrow, -1, h});
        }
      for (int x = 0; x < 3; ++x) {
        const int h = (base + 3 * T1 + x) * T3;
        if (T2 > 0) {
          sl.push_back(SlotDesc{tk, T3, T3 - 1, x, 1, lb + 3 * T1 + x, 0, h});
          hl.push_back(DiaSlot{tk, T3, T3 - 1, x, lb + 3 * T1 + x, -1, h});
        } else {
          sl.push_back(SlotDesc{tk, T3, T3 - 1, x, 0, -1, 0, h});
          hl.push_back(DiaSlot{tk, T3, T3 - 1, x, -1, -1, h});
        }
      }
      if (T2 > 0)
        for (int x = 0; x < 3; ++x) {
          const int h = (base + 3 * T1 + 3 + x) * T3;
          sl.push_back(SlotDesc{tk, T3, T3 - 1, x, 0, -1, 0, h});
          hl.push_back(DiaSlot{tk, T3, T3 - 1, x, -1, -1, h});
        }
      for (int b = 0; b < nrow; b += vp) wb.push_back(base + b);
    }
    wb.push_back((int)sl.size());
    if (prop == 1) {
      if ((rc = upl(slots2, sl.data(), sl.size() * sizeof(SlotDesc))) || (rc = upl(wgb2, wb.data(), wb.size() * 4)))
        return rc;
      a.slots = slots2.as<SlotDesc>(), a.wg_slot_begin = wgb2.as<int>(), a.nwg = (int)wb.size() - 1;
      a.state = const_cast<float2 *>(A.Linit), a.writeback = 0, a.hist = hist.as<float2>();
      if ((rc = launch_conveyor(a, vp, block, st))) return rc;
    } else {
      dia_groups(hl, dmax, wb);
      if ((rc = upl(B.dsl3, hl.data(), hl.size() * sizeof(DiaSlot))) || (rc = upl(B.dwb3, wb.data(), wb.size() * 4)))
        return rc;
      DiaSlotArgs d3 = ds;
      d3.slots = B.dsl3.as<DiaSlot>(), d3.wg_begin = B.dwb3.as<int>(), d3.nwg = (int)wb.size() - 1;
      d3.in = A.Linit, d3.out = nullptr, d3.hist = hist.as<float2>();
      if ((rc = launch_dia_slots(d3, dmax, st))) return rc;
    }
    // (3) GB and SE terms from the history
    hipLaunchKernelGGL(gbse_static_kernel, dim3(S), dim3(256), 6 * N * sizeof(float), st, A, g1b.as<float2>(),
                       g3b.as<float2>(), const_cast<float2 *>(g1s), const_cast<float2 *>(g3P));
    NISE_HIP_CHECK(hipGetLastError());
    const size_t gsm = ((size_t)3 * N * 4 + 15) / 16 * 16 + (size_t)64 * (N + 1) * 8;
    if (gsm <= 64 * 1024) {
      NISE_HIP_CHECK(hipFuncSetAttribute((const void *)gbse_lds_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)gsm));
      hipLaunchKernelGGL(gbse_lds_kernel, dim3(S, T3), dim3(64), gsm, st, A, (const float2 *)g1b.as<float2>(),
                         (const float2 *)g3b.as<float2>(), g1s, g3P);
    } else {
      hipLaunchKernelGGL(gbse_kernel, dim3(S, T3), dim3(64), 3 * N * sizeof(float), st, A,
                         (const float2 *)g1b.as<float2>(), (const float2 *)g3b.as<float2>(), g1s, g3P);
    }
    NISE_HIP_CHECK(hipGetLastError());
    // (4) EA: doubles vectors at tk, then per t3 the responses and one
    //     doubles step with frame tk + t3
    const int K2 = (int)lex.size(), np = doubles_np(N), pvw = 64 / np;
    // waves per workgroup: the most resident waves per CU (160 KB of LDS),
    // counting the padding of each sample's groups to whole workgroups; at
    // most 4 (larger workgroups wait longer at the per-t3 barriers: 7 waves
    // of the N = 32 ring ran 15% slower than 4 at 2 more waves per CU)
    const int gps_p = (9 + pvw - 1) / pvw + 9 * ((T1 + pvw - 1) / pvw);
    int npw = 0;
    double best = 0;
    for (int w = 1; w <= 4; ++w) {
      const size_t L = persist_layout(N, pvw, K2, w).total;
      if (L > 160 * 1024) break;
      const double score = (double)(160 * 1024 / L) * w * gps_p / (double)((gps_p + w - 1) / w * w);
      if (score > best * 1.0001) best = score, npw = w;
    }
    const size_t psmem = npw ? persist_layout(N, pvw, K2, npw).total : 0;
    if (do_ea && prop == 1 && npw && rot_table_fits(N, pvw, K2)) {
      // every doubles vector stays in the LDS of its workgroup for all t3
      // (and is formed there: no doubles vectors in device memory)
      std::vector<DblGroup> gh;
      for (int s = 0; s < S; ++s) {
        for (int v0 = 0; v0 < 9; v0 += pvw) gh.push_back(DblGroup{s, 0, v0, std::min(pvw, 9 - v0)});
        for (int p = 0; p < 9; ++p)
          for (int t1 = 0; t1 < T1; t1 += pvw) gh.push_back(DblGroup{s, 1, 9 + p * T1 + t1, std::min(pvw, T1 - t1)});
        while (gh.size() % npw) gh.push_back(DblGroup{s, 0, 0, 0});  // workgroups never mix samples
      }
      if ((rc = upl(grp, gh.data(), gh.size() * sizeof(DblGroup)))) return rc;
      PersistArgs P;
      P.A = A, P.groups = grp.as<DblGroup>(), P.K2 = K2, P.m = non->ts, P.ab2 = ab2.as<uint32_t>();
      P.D2 = D2.as<float2>(), P.cs2 = cs2.as<float4>(), P.f_lo = f_lo, P.npw = npw;
      const void *fp = np == 8 ? (const void *)dbl_ea_kernel<8>
                     : np == 16 ? (const void *)dbl_ea_kernel<16>
                     : np == 32 ? (const void *)dbl_ea_kernel<32>
                                : (const void *)dbl_ea_kernel<64>;
      NISE_HIP_CHECK(hipFuncSetAttribute(fp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)psmem));
      void *kargs[] = {&P};
      NISE_HIP_CHECK(hipLaunchKernel(fp, dim3((unsigned)(gh.size() / npw)), dim3(64 * npw), kargs, psmem, st));
    } else if (do_ea) {
      // per t3: leftrr, the EA responses from the doubles vectors in device
      // memory, then one doubles step of every vector
      if ((rc = dbl.alloc((size_t)S * nvec * N2 * 8))) return rc;
      A.dbl = dbl.as<float2>();
      hipLaunchKernelGGL(dd_kernel, dim3((N2 + 127) / 128, S, nvec), dim3(128), 0, st, A);
      NISE_HIP_CHECK(hipGetLastError());
      std::vector<int> vbh((size_t)S * gps + 1);
      for (int s = 0; s < S; ++s)
        for (int g = 0; g < gps; ++g) vbh[(size_t)s * gps + g] = s * nvec + g * vw;
      vbh.back() = S * nvec;
      std::vector<int64_t> rowh((size_t)T3 * S * gps);
      for (int t3 = 0; t3 < T3; ++t3)
        for (int s = 0; s < S; ++s)
          for (int g = 0; g < gps; ++g)
            rowh[((size_t)t3 * S + s) * gps + g] = (int64_t)(begin + c0 + s) * non->sample + T1 + T2 + t3 - f_lo;
      if ((rc = upl(vb, vbh.data(), vbh.size() * 4)) || (rc = upl(rows, rowh.data(), rowh.size() * 8))) return rc;
      if ((rc = lrrb.alloc((size_t)S * 27 * N * 8))) return rc;
      const void *fl = ea.stage ? (const void *)lrr_kernel<true> : (const void *)lrr_kernel<false>;
      const void *fe = ea.stage ? (const void *)ea_kernel<true> : (const void *)ea_kernel<false>;
      NISE_HIP_CHECK(hipFuncSetAttribute(fl, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ea.lsmem));
      NISE_HIP_CHECK(hipFuncSetAttribute(fe, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ea.smem));
      DoublesArgs d;
      d.N = N, d.K2 = (int)lex.size(), d.m = non->ts, d.ab2 = ab2.as<uint32_t>();
      d.D2 = D2.as<float2>(), d.cs2 = cs2.as<float4>(), d.state = dbl.as<float2>(), d.es = es;
      d.vec_begin = vb.as<int>();
      for (int t3 = 0; t3 < T3; ++t3) {
        if (ea.stage) {
          hipLaunchKernelGGL(lrr_kernel<true>, dim3(S, 9), dim3(ea.lblock), ea.lsmem, st, A, t3, lrrb.as<float2>());
          NISE_HIP_CHECK(hipGetLastError());
          hipLaunchKernelGGL(ea_kernel<true>, dim3(ea.nchunk, 9, S), dim3(ea.block), ea.smem, st, A, t3, ea.C,
                             (const float2 *)lrrb.as<float2>());
        } else {
          hipLaunchKernelGGL(lrr_kernel<false>, dim3(S, 9), dim3(ea.lblock), ea.lsmem, st, A, t3, lrrb.as<float2>());
          NISE_HIP_CHECK(hipGetLastError());
          hipLaunchKernelGGL(ea_kernel<false>, dim3(ea.nchunk, 9, S), dim3(ea.block), ea.smem, st, A, t3, ea.C,
                             (const float2 *)lrrb.as<float2>());
        }
        NISE_HIP_CHECK(hipGetLastError());
        if (t3 + 1 < T3) {  // the state after the last t3 is not used
          if (prop == 1) {
            d.rows = rows.as<int64_t>() + (size_t)t3 * S * gps;
            if ((rc = launch_doubles_step(d, S * gps, st))) return rc;
          } else {  // src/propagate.c:1486-1509 (and :1606-1629 for two-level sites)
            SparseDblArgs sp;
            sp.N = N, sp.m = non->ts, sp.es = es, sp.anh = non->anharmonicity, sp.anh_rec = es ? nullptr : Ad;
            sp.fm = fm_dbl, sp.f_lo = f_lo, sp.el = B.el.as<float4>(), sp.colptr = B.cptr.as<int>();
            sp.colidx = B.cidx.as<int>(), sp.cnt = B.ecnt.as<int>(), sp.state = dbl.as<float2>();
            sp.vec_begin = vb.as<int>(), sp.rows = rows.as<int64_t>() + (size_t)t3 * S * gps;
            if ((rc = launch_sparse_doubles(sp, S * gps, st))) return rc;
          }
        }
      }
    }
    // (5) fixed-order reduction of this chunk into the 12 planes
    const int nred = (S + kRedS - 1) / kRedS;
    hipLaunchKernelGGL(reduce2d_part_kernel, dim3((unsigned)((plane + 255) / 256), 12, nred), dim3(256), 0, st,
                       E.as<float>(), S, T3, T1, wpl.as<float>(), ltb.as<float>(), part.as<double>());
    NISE_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(reduce2d_final_kernel, dim3((unsigned)((12 * plane + 255) / 256)), dim3(256), 0, st,
                       (const double *)part.as<double>(), nred, 12 * plane, Rd.as<double>());
    NISE_HIP_CHECK(hipGetLastError());
    NISE_HIP_CHECK(hipStreamSynchronize(st));  // host slot vectors are rebuilt next chunk
  }
  std::vector<double> Rh((size_t)12 * plane);
  NISE_HIP_CHECK(hipMemcpyAsync(Rh.data(), Rd.p, Rh.size() * 8, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  for (size_t i = 0; i < Rh.size(); ++i) R12[i] = (float)Rh[i];
  return NISE_OK;
}

// ---------------------------------------------------------------------------
// File-level drop-in
// ---------------------------------------------------------------------------
[[noreturn]] void nise::die2d(int code, const char *msg) {
  printf("%s", msg);
  fflush(stdout);
  exit(code);
}

// src/MPI_subs.c:20-34
void nise::print2D_file(const char *name, float *R, float *I, const t_non *non, int sampleCount) {
  FILE *out = fopen(name, "w");
  if (!out) return;
  const int T1 = non->tmax1;
  for (int t1 = 0; t1 < non->tmax1; t1 += non->dt1) {
    const int t2 = non->tmax2;
    for (int t3 = 0; t3 < non->tmax3; t3 += non->dt3) {
      float &r = R[(size_t)t3 * T1 + t1], &i = I[(size_t)t3 * T1 + t1];
      r /= sampleCount;
      i /= sampleCount;
      if (t3 == 0) r /= 2, i /= 2;
      fprintf(out, "%f %f %f %e %e\n", t1 * non->deltat, t2 * non->deltat, t3 * non->deltat, r, i);
    }
  }
  fclose(out);
}

namespace {
// The file flow of src/calc_2DIR.c:22-629 and src/calc_2DES.c:22-586 (the
// same program for anharmonic oscillators and for two-level sites): rank 0
// runs every work item on its GPU; other ranks have nothing to add to the
// sums and return.
void run_2d_files(t_non *non, int parentRank, const char *name, bool es) {
  if (parentRank != 0) return;
  if (non->cluster != -1) {
    printf("%s: the GPU path does not implement Cluster\n", name);
    die2d(1, "");
  }
  const int N = non->singles;
  const int totalSampleCount = (non->length - non->tmax1 - non->tmax2 - non->tmax3 - 1) / non->sample + 1;
  if (totalSampleCount > 0) printf("Making %d samples!\n", totalSampleCount);
  else die2d(1, "Insufficient data to calculate spectrum.\nPlease, lower max times or provide longer\ntrajectory.\n");
  if (non->end == 0) non->end = totalSampleCount;
  int sampleCount = non->end - non->begin;
  if (sampleCount == 0) sampleCount = 1;
  const float shift1 = (non->max1 + non->min1) / 2;
  printf("Rotating Wave Frequency shift %f.\n", shift1);
  non->shifte = shift1;
  non->shiftf = 2 * shift1;
  const bool ctype = strcmp(non->hamiltonian, "Coupling") == 0;
  const bool ea = tech_is(non, "EAIR") || tech_is(non, "2DIR");
  const int64_t D = non->doubles;
  const int64_t hw = ctype ? 1 + N : 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2;
  const int64_t mw = 1 + 3 * (int64_t)N + 3 * (int64_t)N * D;
  const int64_t f_lo = (int64_t)non->begin * non->sample;
  const int64_t f_hi = (int64_t)(non->end - 1) * non->sample + non->tmax1 + non->tmax2 + non->tmax3;
  if (non->end <= non->begin) {
    printf("%s: no samples between Begin and End\n", name);
    die2d(1, "");
  }
  const int64_t nfr = f_hi - f_lo;
  FILE *probe = fopen(non->energyFName, "rb");
  if (!probe) die2d(1, "Hamiltonian file not found!\n");
  fclose(probe);
  probe = fopen(non->dipoleFName, "rb");
  if (!probe) {
    printf("Dipole file %s not found!\n", non->dipoleFName);
    die2d(1, "");
  }
  fclose(probe);
  // src/calc_2DIR.c:132-146; calc_2DES reads no anharmonicities
  const bool anh_file = !es && non->anharmonicity == 0;
  if (anh_file) {
    if (!(probe = fopen(non->anharFName, "rb"))) {
      printf("Anharmonicity file %s not found!\n", non->anharFName);
      die2d(1, "");
    }
    fclose(probe);
    if (!(probe = fopen(non->overdipFName, "rb"))) {
      printf("Overtone dipole file %s not found!\n", non->overdipFName);
      die2d(1, "");
    }
    fclose(probe);
  }
  Pinned H, mu, Cc, Ah, Oh;
  if (!H.alloc((size_t)(nfr * hw)) || !mu.alloc((size_t)((ctype ? 1 : nfr) * mw))) {
    printf("%s: could not allocate pinned host memory\n", name);
    die2d(1, "");
  }
  if (!read_records(non->energyFName, hw, f_lo, nfr, H.p))
    die2d(1, "Hamiltonian trajectory file to short, could not fill buffer!!!\n");
  if (!read_records(non->dipoleFName, mw, ctype ? 0 : f_lo, ctype ? 1 : nfr, mu.p))
    die2d(1, "Dipole trajectory file to short, could not fill buffer!!!\n");
  if (ctype) {
    const int64_t cw = 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2;
    if (!(probe = fopen(non->couplingFName, "rb"))) die2d(1, "Coupling file not found!\n");
    fclose(probe);
    if (!Cc.alloc((size_t)cw) || !read_records(non->couplingFName, cw, 0, 1, Cc.p))
      die2d(1, "Coupling trajectory file to short, could not fill buffer!!!\n");
  }
  if (anh_file && ea) {
    if (!Ah.alloc((size_t)(nfr * (1 + N))) || !Oh.alloc((size_t)(nfr * (1 + 3 * N)))) {
      printf("%s: could not allocate pinned host memory\n", name);
      die2d(1, "");
    }
    if (!read_records(non->anharFName, 1 + N, f_lo, nfr, Ah.p) ||
        !read_records(non->overdipFName, 1 + 3 * (int64_t)N, f_lo, nfr, Oh.p))
      die2d(1, "Anharmonicity or overtone dipole trajectory file to short, could not fill buffer!!!\n");
  }
  // the staged records start at frame f_lo: rebase the samples to begin = 0
  t_non loc = *non;
  loc.begin = 0, loc.end = non->end - non->begin;
  std::vector<float> R((size_t)12 * non->tmax3 * non->tmax1);
  const int rc = nise_2dir_batched(&loc, H.p, hw, mu.p, mw, nfr, ctype ? Cc.p : nullptr,
                                   Ah.p, Oh.p, NISE_HOST, R.data());
  if (rc) {
    printf("%s: GPU run failed: %s\n", name, nise_last_error());
    die2d(1, "");
  }
  printf("Samples %d\n", sampleCount);
  const size_t pl = (size_t)non->tmax3 * non->tmax1;
  static const char *names[6] = {"RparI.dat", "RparII.dat", "RperI.dat", "RperII.dat", "RcroI.dat", "RcroII.dat"};
  for (int p = 0; p < 3; ++p)
    for (int k = 0; k < 2; ++k)
      print2D_file(names[2 * p + k], R.data() + (4 * p + 2 * k) * pl, R.data() + (4 * p + 2 * k + 1) * pl, non,
                   sampleCount);
  printf("----------------------------------------\n");
  printf(" %s calculation succesfully completed\n", es ? "2DES" : "2DIR");
  printf("----------------------------------------\n\n");
}
}  // namespace

// src/calc_2DIR.c:22-629
extern "C" void calc_2DIR(t_non *non, int parentRank, int parentSize, int subRank, int subSize,
                          MPI_Comm subComm, MPI_Comm rootComm) {
  (void)parentSize, (void)subRank, (void)subSize, (void)subComm, (void)rootComm;
  run_2d_files(non, parentRank, "calc_2DIR", false);
}

// src/calc_2DES.c:22-586 (2DUVvis and its pathway techniques)
extern "C" void calc_2DES(t_non *non, int parentRank, int parentSize, int subRank, int subSize,
                          MPI_Comm subComm, MPI_Comm rootComm) {
  (void)parentSize, (void)subRank, (void)subSize, (void)subComm, (void)rootComm;
  run_2d_files(non, parentRank, "calc_2DES", true);
}
