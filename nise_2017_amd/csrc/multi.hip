// Batched absorption of many independent systems of uneven size.
//
// Systems with propagation 1 (Trotter coupling scheme) run on the row
// conveyor in one launch: every workgroup reads its own system descriptor,
// so waves of different N share the launch.  Systems with propagation 0
// (exact exponential, src/propagate.c:60-125) and N <= 64 run on the DIA
// conveyor in a second launch: per frame U = exp(-i H f) is formed once
// (FP64 Taylor series with scaling and squaring, rounded to FP32) and each
// workgroup applies its frame's U, staged in LDS, to all its live vectors at
// that frame.
//
// Each workgroup records its start and end on the device wall clock; a
// system's completion latency is the end of its last workgroup measured from
// the first start of the run.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "nise_internal.h"

namespace nise {

int launch_jacobi(const double *A, int N, int nmat, double *w, double *V, int *status, hipStream_t st);

// Coupling as read_He leaves it (|J| < cut -> 0, src/read_trajectory.c:305-311);
// the DIA path uses every remaining element.
__device__ inline float he_coupling(const FrameSource &s, int N, int64_t f, int a, int b) {
  const float J = raw_coupling(s, N, f, a, b);
  return fabsf(J) < s.cut ? 0.f : J;
}

// Dense double Hamiltonian of frames f_lo .. (src/propagate.c:79-85).
__global__ void dia_dense_kernel(FrameSource s, int N, int64_t f_lo, double *A) {
  const int64_t m = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * N) return;
  const int a = i / N, b = i % N;
  const int64_t f = f_lo + m;
  const float h = a == b ? site_energy(s, N, f, a) : he_coupling(s, N, f, a < b ? a : b, a < b ? b : a);
  A[m * N * N + i] = (double)h;
}

// U[a][b] = sum_c V[a][c] V[b][c] (cos, -sin)(e_c f) (src/propagate.c:88-106),
// e_c rounded to float and multiplied by f in float as the reference does.
// With thres > 0 (the sparse scheme, src/propagate.c:185-258) an element is
// kept only where its FP32 |U_ab|^2 exceeds thres (:242), else it is 0.
__global__ void dia_u_kernel(const double *w, const double *V, int N, float fac, float thres, float2 *U) {
  const int64_t m = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * N) return;
  const int a = i / N, b = i % N;
  const double *Vm = V + m * N * N;
  double ur = 0.0, ui = 0.0;
  for (int c = 0; c < N; ++c) {
    const float ef = __fmul_rn((float)w[m * N + c], fac);
    const double p = Vm[a * N + c] * Vm[b * N + c];
    ur += p * cos((double)ef);
    ui -= p * sin((double)ef);
  }
  float2 u = make_float2((float)ur, (float)ui);
  if (thres > 0.f && !(__fadd_rn(__fmul_rn(u.x, u.x), __fmul_rn(u.y, u.y)) > thres)) u = make_float2(0.f, 0.f);
  U[m * N * N + i] = u;
}

int launch_dia_u(const double *w, const double *V, int N, int64_t nmat, float fac, float thres, float2 *U,
                 hipStream_t st) {
  for (int64_t d0 = 0; d0 < nmat; d0 += 65535) {
    const int64_t nd = std::min<int64_t>(65535, nmat - d0);
    hipLaunchKernelGGL(dia_u_kernel, dim3((N * N + 255) / 256, (unsigned)nd), dim3(256), 0, st, w + d0 * N,
                       V + d0 * N * N, N, fac, thres, U + d0 * N * N);
    NISE_HIP_CHECK(hipGetLastError());
  }
  return NISE_OK;
}

// U = exp(-i fac H) of one frame per workgroup (N <= kExpmMaxN), the same
// operator as V exp(-i E fac) V^T (src/propagate.c:60-125) without an
// eigensystem: A = fac H in FP64 is scaled by 2^-s until its 1-norm is at
// most 1/2, exp(-i A) = cos A - i sin A is summed as the Taylor series up to
// the power K whose term theta^K / K! falls below 1e-15, and the result is
// squared s times.  With B = A^2, cos A = sum_k (-1)^k B^k / (2k)! and
// sin A = A sum_k (-1)^k B^k / (2k+1)! are evaluated by the Paterson-
// Stockmeyer scheme in B and B^2 (Horner in B^2 over chunks c_j + c_j' B):
// 2 + K/2 + 1 real products instead of K for the plain power series.
// Products run from LDS; each thread keeps its own 16 output elements in
// registers.  Elements are rounded to FP32 once; with thres > 0 (the sparse
// scheme) an element is kept only where its FP32 |U_ab|^2 exceeds thres,
// else it is 0.
constexpr int kExpmMaxN = 64;

// Every N x N matrix sits in LDS with row pitch N + 1 (odd); thread t holds
// the elements (ry + 16 i, cx + 16 j), i, j < 4, ry = t / 16, cx = t % 16, so
// that the 16 lanes of a row group read 16 consecutive entries of a row of
// B and the 4 row groups of a wave read rows in different banks.
struct ExpmTile {
  int N, L, ry, cx;
  int S;  // stride between a thread's rows / columns (16 for 256 threads, 8 for 64)
  __device__ bool in(int i, int j) const { return ry + S * i < N && cx + S * j < N; }
  __device__ int at(int i, int j) const { return (ry + S * i) * L + cx + S * j; }
};

// q = P B for the thread's elements.
__device__ __forceinline__ void tile_mm(const double *P, const double *B, const ExpmTile &g, double q[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) q[i][j] = 0.0;
  for (int c = 0; c < g.N; ++c) {
    double pr[4], bc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) pr[i] = g.ry + g.S * i < g.N ? P[(g.ry + g.S * i) * g.L + c] : 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) bc[j] = g.cx + g.S * j < g.N ? B[c * g.L + g.cx + g.S * j] : 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) q[i][j] += pr[i] * bc[j];
  }
}

__device__ __forceinline__ void tile_store(double *X, const double q[4][4], const ExpmTile &g) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (g.in(i, j)) X[g.at(i, j)] = q[i][j];
}

// 1 / n! for n < 32, by the divisions 1 / 2 / 3 / ... / n in turn (the
// values the series coefficients had when each thread divided at run time).
struct InvFact {
  double v[32];
  constexpr InvFact() : v() {
    for (int n = 0; n < 32; ++n) {
      double c = 1.0;
      for (int d = 2; d <= n; ++d) c /= d;
      v[n] = c;
    }
  }
};
__constant__ constexpr InvFact kInvFact{};

// Coefficient k of the cos (odd = 0) or sin (odd = 1) series in B = A^2:
// (-1)^k / (2k + odd)!, zero past the last power kept (K < 30, so
// 2k + odd < 32).
__device__ __forceinline__ double trig_coef(int k, int odd, int kmax) {
  if (k > kmax) return 0.0;
  const double c = kInvFact.v[2 * k + odd];
  return (k & 1) ? -c : c;
}

// R = p(B) = sum_{k <= kmax} coef_k B^k for the thread's elements, by Horner
// in B^2 over the chunks coef_2j I + coef_2j+1 B (the top chunk also takes
// B^2); Bv: the thread's elements of B; B2 (LDS) the product operand; X an
// LDS buffer for the running value.  Every iteration ends with a barrier.
__device__ void trig_poly(const double Bv[4][4], const double *B2, double *X, const ExpmTile &g, int odd, int kmax,
                          double R[4][4]) {
  const int J = kmax >= 1 ? (kmax - 1) / 2 : 0;  // top chunk: powers 2J .. kmax (at most 3)
  auto chunk = [&](int j, int top, double Y[4][4]) {
    const double a0 = trig_coef(2 * j, odd, kmax), a1 = trig_coef(2 * j + 1, odd, kmax),
                 a2 = top ? trig_coef(2 * j + 2, odd, kmax) : 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const bool in = g.in(i, l);
        const int e = in ? g.at(i, l) : 0;
        Y[i][l] = (g.ry + g.S * i == g.cx + g.S * l ? a0 : 0.0) + (in ? a1 * Bv[i][l] + a2 * B2[e] : 0.0);
      }
  };
  chunk(J, 1, R);
  for (int j = J - 1; j >= 0; --j) {
    tile_store(X, R, g);
    __syncthreads();
    double q[4][4], Y[4][4];
    tile_mm(B2, X, g, q);
    chunk(j, 0, Y);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int l = 0; l < 4; ++l) R[i][l] = Y[i][l] + q[i][l];
    __syncthreads();
  }
}

// pairs = 0: U[m][a][b] row-major; pairs = 1: U[m][b / 2][a] as float4
// (U_a,b, U_a,b+1), the second half zero past the last column (the DIA
// conveyor's layout: one 16-byte read gives a site two columns).
// One system's frames for expm_jobs_kernel; b is set by launch_expm_jobs.
struct ExpmJob {
  FrameSource src;
  int N = 0, pairs = 0;
  int64_t f_first = 0, nfr = 0, b = 0;
  float fac = 0.f, thres = 0.f;
  float2 *U = nullptr;
};
__host__ __device__ inline int64_t expm_per(int N, int pairs) {
  return pairs ? (int64_t)2 * ((N + 1) >> 1) * N : (int64_t)N * N;
}
__host__ __device__ inline size_t expm_smem(int N) { return (size_t)2 * N * (N + 1) * 8 + 64; }
// U of frame f into Um (the frame's block of U), xs: expm_smem(N) bytes.
// NT threads: 256 (a 16 x 16 grid, N <= 64) or 64 (8 x 8, N <= 32); each
// thread holds the elements (ry + S i, cx + S j), i, j < 4.
template <int NT>
__device__ __forceinline__ void expm_frame(const FrameSource &src, int N, int64_t f, float fac, float thres,
                                           int pairs, float2 *Um, double *xs) {
  // Two N x N matrices in LDS (M0, M1): A, then B = A^2 (the thread's
  // elements kept in registers), B^2 and the running value of the series,
  // then A again (reloaded from the records) for sin A = A (sin A / A).
  // Four matrices took 133 KB at N = 64, one workgroup per CU.
  const int t = threadIdx.x, L = N + 1, NL = N * L;
  double *M0 = xs, *M1 = M0 + NL;
  double *red = M1 + NL;  // [4]
  constexpr int S = NT == 64 ? 8 : 16;
  const ExpmTile g{N, L, t / S, t % S, S};
  // M[a][b] = (H_ab fac) mul
  auto fill = [&](double *M, double mul) {
    if (!src.onfly) {
      // couplings and energies from the records: every load of the thread
      // in flight before the first use (per element, with the on-the-fly
      // coupling code beside it, each load's latency showed), indexes past
      // the matrix clamped to its last element and not stored
      constexpr int kLd = (4 * S) * (4 * S) / NT;
      const float *rec = src.H_rec + f * src.H_stride + 1;
      const float *crec = src.coupling_type ? src.C_rec + 1 : rec;
      float hv[kLd];
#pragma unroll
      for (int u = 0; u < kLd; ++u) {
        const int e = min(t + NT * u, N * N - 1), a = e / N, b = e - a * N;
        hv[u] = a == b ? gld(rec + (src.coupling_type ? a : sindex(a, a, N)))
                       : gld(crec + sindex(a < b ? a : b, a < b ? b : a, N));
      }
#pragma unroll
      for (int u = 0; u < kLd; ++u) {
        const int e = t + NT * u, a = e / N, b = e - a * N;
        if (e < N * N) {
          const float h = a == b ? shift_diag(src, a, hv[u]) : (fabsf(hv[u]) < src.cut ? 0.f : hv[u]);
          M[a * L + b] = ((double)h * (double)fac) * mul;
        }
      }
    } else {
      for (int e = t; e < N * N; e += NT) {
        const int a = e / N, b = e - a * N;
        const float h = a == b ? site_energy(src, N, f, a) : he_coupling(src, N, f, a < b ? a : b, a < b ? b : a);
        M[a * L + b] = ((double)h * (double)fac) * mul;
      }
    }
  };
  fill(M0, 1.0);
  __syncthreads();
  // 1-norm (largest column sum of |A|)
  double cm = 0.0;
  for (int c = t; c < N; c += NT) {
    double sum = 0.0;
    for (int r = 0; r < N; ++r) sum += fabs(M0[r * L + c]);
    cm = fmax(cm, sum);
  }
  for (int o = 32; o > 0; o >>= 1) cm = fmax(cm, __shfl_xor(cm, o));
  if ((t & 63) == 0) red[t >> 6] = cm;
  __syncthreads();
  double theta = red[0];
  for (int w = 1; w < NT / 64; ++w) theta = fmax(theta, red[w]);
  int sq = 0;
  while (theta > 0.5) theta *= 0.5, ++sq;
  int K = 0;
  for (double term = 1.0; term >= 1e-15 && K < 30; ++K) term *= theta / (K + 1);
  const double scale = ldexp(1.0, -sq);
  __syncthreads();  // every thread has read red
  for (int e = t; e < N * N; e += NT) {
    const int a = e / N, b = e - a * N;
    M0[a * L + b] *= scale;
  }
  __syncthreads();
  double q[4][4], bv[4][4];
  tile_mm(M0, M0, g, bv), tile_store(M1, bv, g);  // B
  __syncthreads();
  tile_mm(M1, M1, g, q);  // B^2
  __syncthreads();        // every thread has read M0 (A) and M1 (B)
  tile_store(M0, q, g);
  __syncthreads();
  double ur[4][4], ui[4][4];
  trig_poly(bv, M0, M1, g, 0, K / 2, ur);        // cos A
  trig_poly(bv, M0, M1, g, 1, (K - 1) / 2, ui);  // sin A / A
  __syncthreads();  // the products have read M0 (B^2) and M1
  fill(M0, scale);  // A
  tile_store(M1, ui, g);
  __syncthreads();
  tile_mm(M0, M1, g, q);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) ui[i][j] = -q[i][j];  // exp(-i A) = cos A - i sin A
  double *P = M0, *Q = M1;
  for (int it = 0; it < sq; ++it) {  // U <- U U (complex)
    __syncthreads();  // the previous products have read P, Q (and X)
    tile_store(P, ur, g), tile_store(Q, ui, g);
    __syncthreads();
    double rr[4][4], ii[4][4], ri[4][4], ir[4][4];
    tile_mm(P, P, g, rr);
    tile_mm(Q, Q, g, ii);
    tile_mm(P, Q, g, ri);
    tile_mm(Q, P, g, ir);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) ur[i][j] = rr[i][j] - ii[i][j], ui[i][j] = ri[i][j] + ir[i][j];
  }
  const int Pr = (N + 1) >> 1;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!g.in(i, j)) continue;
      const int a = g.ry + S * i, b = g.cx + S * j;
      float2 u = make_float2((float)ur[i][j], (float)ui[i][j]);
      if (thres > 0.f && !(__fadd_rn(__fmul_rn(u.x, u.x), __fmul_rn(u.y, u.y)) > thres)) u = make_float2(0.f, 0.f);
      Um[pairs ? ((b >> 1) * N + a) * 2 + (b & 1) : a * N + b] = u;
    }
  if (pairs && (N & 1))
    for (int a = t; a < N; a += NT) Um[((Pr - 1) * N + a) * 2 + 1] = make_float2(0.f, 0.f);
}

__global__ void __launch_bounds__(256) expm_kernel(FrameSource src, int N, int64_t f_first, float fac, float thres,
                                                   int pairs, float2 *U) {
  extern __shared__ __align__(16) double xs[];
  const int64_t m = blockIdx.x;
  expm_frame<256>(src, N, f_first + m, fac, thres, pairs, U + m * expm_per(N, pairs), xs);
}

// The frames of several systems in one launch: workgroup b0 + blockIdx.x
// computes frame b - jobs[j].b of job j, the last job with jobs[j].b <= b
// (NT threads: 64 for N <= 32).
template <int NT>
__global__ void __launch_bounds__(NT) expm_jobs_kernel(const ExpmJob *jobs, int njobs, int64_t b0) {
  extern __shared__ __align__(16) double xs[];
  const int64_t b = b0 + blockIdx.x;
  int lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (jobs[mid].b <= b) lo = mid;
    else hi = mid;
  }
  const ExpmJob J = jobs[lo];
  const int64_t m = b - J.b;
  expm_frame<NT>(J.src, J.N, J.f_first + m, J.fac, J.thres, J.pairs, J.U + m * expm_per(J.N, J.pairs), xs);
}

int launch_expm_tables(const FrameSource &src, int N, int64_t f_first, int64_t nfr, float fac, float thres,
                       int pairs, float2 *U, hipStream_t st) {
  if (N < 1 || N > kExpmMaxN) return set_error("expm: N=%d outside 1..%d", N, kExpmMaxN), NISE_ERR_ARG;
  const size_t smem = expm_smem(N);
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)expm_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem));
  for (int64_t d0 = 0; d0 < nfr; d0 += 65535) {
    const int64_t nd = std::min<int64_t>(65535, nfr - d0);
    const int64_t per = expm_per(N, pairs);
    hipLaunchKernelGGL(expm_kernel, dim3((unsigned)nd), dim3(256), smem, st, src, N, f_first + d0, fac, thres,
                       pairs, U + d0 * per);
    NISE_HIP_CHECK(hipGetLastError());
  }
  return NISE_OK;
}

// The jobs' frames in one launch per size class (N <= 16, 32, 48, 64: each
// launch takes the LDS of its largest N), larger N first within a class.
// host and dev hold the job table between the upload and the launches.
int launch_expm_jobs(std::vector<ExpmJob> &jobs, std::vector<ExpmJob> &host, DevBuf &dev, hipStream_t st) {
  host.clear();
  std::vector<std::pair<int, int>> cls;  // (class, first job of the class in host)
  for (int c = 3; c >= 0; --c) {
    std::vector<ExpmJob> in;
    for (const ExpmJob &J : jobs) {
      if (J.N < 1 || J.N > kExpmMaxN) return set_error("expm: N=%d outside 1..%d", J.N, kExpmMaxN), NISE_ERR_ARG;
      if ((J.N - 1) / 16 == c) in.push_back(J);
    }
    if (in.empty()) continue;
    std::stable_sort(in.begin(), in.end(), [](const ExpmJob &x, const ExpmJob &y) { return x.N > y.N; });
    cls.push_back({c, (int)host.size()});
    host.insert(host.end(), in.begin(), in.end());
  }
  if (host.empty()) return NISE_OK;
  cls.push_back({-1, (int)host.size()});
  for (size_t k = 0; k + 1 < cls.size(); ++k) {  // b: the job's first workgroup within its class
    int64_t b = 0;
    for (int j = cls[k].second; j < cls[k + 1].second; ++j) host[j].b = b, b += host[j].nfr;
  }
  int rc = dev.alloc(host.size() * sizeof(ExpmJob));
  if (rc) return rc;
  NISE_HIP_CHECK(hipMemcpyAsync(dev.p, host.data(), host.size() * sizeof(ExpmJob), hipMemcpyHostToDevice, st));
  for (size_t k = 0; k + 1 < cls.size(); ++k) {
    const int j0 = cls[k].second, nj = cls[k + 1].second - j0;
    const size_t smem = expm_smem(16 * (cls[k].first + 1));
    // N <= 32: one wave per frame (eight frames per CU instead of two at
    // the kernel's register count)
    const bool one_wave = cls[k].first <= 1;
    const void *fn = one_wave ? (const void *)expm_jobs_kernel<64> : (const void *)expm_jobs_kernel<256>;
    NISE_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    const int64_t nb = host[j0 + nj - 1].b + host[j0 + nj - 1].nfr;
    for (int64_t d0 = 0; d0 < nb; d0 += 65535) {
      const ExpmJob *jd = (const ExpmJob *)dev.as<ExpmJob>() + j0;
      void *kargs[] = {&jd, (void *)&nj, &d0};
      NISE_HIP_CHECK(hipLaunchKernel(fn, dim3((unsigned)std::min<int64_t>(65535, nb - d0)), dim3(one_wave ? 64 : 256),
                                     kargs, smem, st));
    }
  }
  return NISE_OK;
}

// Per-frame propagators U (FP32 [nfr][N][N]) of frames f_first .. for the
// exact exponential with factor fac (thres > 0: the sparse scheme's dropped
// elements zeroed), through FP64 work arrays of at most 4096 frames; one
// Jacobi convergence word per frame in status (negative: not converged).
int launch_dia_tables(const FrameSource &src, int N, int64_t f_first, int64_t nfr, float fac, float thres,
                      float2 *U, DevBuf &A, DevBuf &V, DevBuf &w, int *status, hipStream_t st) {
  if (nfr <= 0) return NISE_OK;
  if (N <= kExpmMaxN) {  // no eigensystem needed: the series, and nothing to converge
    NISE_HIP_CHECK(hipMemsetAsync(status, 0, (size_t)nfr * sizeof(int), st));
    return launch_expm_tables(src, N, f_first, nfr, fac, thres, 0, U, st);
  }
  const int64_t chunk = std::min<int64_t>(nfr, 4096);
  int rc;
  if ((rc = A.alloc((size_t)chunk * N * N * 8)) || (rc = V.alloc((size_t)chunk * N * N * 8)) ||
      (rc = w.alloc((size_t)chunk * N * 8)))
    return rc;
  for (int64_t c0 = 0; c0 < nfr; c0 += chunk) {
    const int64_t nc = std::min(chunk, nfr - c0);
    for (int64_t d0 = 0; d0 < nc; d0 += 65535) {
      const int64_t nd = std::min<int64_t>(65535, nc - d0);
      hipLaunchKernelGGL(dia_dense_kernel, dim3((N * N + 255) / 256, (unsigned)nd), dim3(256), 0, st, src, N,
                         f_first + c0 + d0, A.as<double>() + d0 * N * N);
      NISE_HIP_CHECK(hipGetLastError());
    }
    if ((rc = launch_jacobi(A.as<double>(), N, (int)nc, w.as<double>(), V.as<double>(), status + c0, st)))
      return rc;
    if ((rc = launch_dia_u(w.as<double>(), V.as<double>(), N, nc, fac, thres, U + c0 * N * N, st))) return rc;
  }
  return NISE_OK;
}

// One workgroup (256 threads): vectors v0 .. v0+vp-1 of one system (vp a
// multiple of 3, so that the three polarisations of a sample share it).
// Thread (site a, group sg) holds site a of vectors sg + j * (256 / Np),
// j < kDiaVec, in registers while it applies U; each 16-byte LDS read gives
// it two columns of U or two entries of one vector.  At frame f only the
// vectors whose window [fs, fs + tmax) holds f do any work.
constexpr int kDiaVec = 8;
// vector rows of Np + 2 (even: 16-byte aligned) entries, so that the rows
// read at once by the groups of a wave start in different banks
__host__ __device__ inline int dia_pitch(int Np) { return (Np + 3) & ~1; }

// psB[v] = U psA[v] at site a for the C vectors v = vb, vb + vs, ...: one
// 16-byte read of U (two columns) serves all C of them.
template <int C>
__device__ __forceinline__ void dia_apply(const float4 *Ut, const float2 *psA, float2 *psB, int N, int Pr, int Ps,
                                          int a, int vb, int vs) {
  float2 y[C];
#pragma unroll
  for (int j = 0; j < C; ++j) y[j] = make_float2(0.f, 0.f);
  for (int bp = 0; bp < Pr; ++bp) {
    const float4 u = Ut[bp * N + a];
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const float4 c = *reinterpret_cast<const float4 *>(psA + (vb + j * vs) * Ps + 2 * bp);
      y[j].x = fmaf(u.x, c.x, y[j].x);
      y[j].x = fmaf(-u.y, c.y, y[j].x);
      y[j].x = fmaf(u.z, c.z, y[j].x);
      y[j].x = fmaf(-u.w, c.w, y[j].x);
      y[j].y = fmaf(u.x, c.y, y[j].y);
      y[j].y = fmaf(u.y, c.x, y[j].y);
      y[j].y = fmaf(u.z, c.w, y[j].y);
      y[j].y = fmaf(u.w, c.z, y[j].y);
    }
  }
#pragma unroll
  for (int j = 0; j < C; ++j) psB[(vb + j * vs) * Ps + a] = y[j];
}

__global__ void __launch_bounds__(256) dia_conveyor_kernel(const DiaSys *sys, const int2 *wmap,
                                                           unsigned long long *times) {
  extern __shared__ __align__(16) unsigned char smem[];
  const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
  const int2 wm = wmap[blockIdx.x];
  const DiaSys &S = sys[wm.x];
  const int N = S.N, Np = S.Np, T = S.tmax, v0 = wm.y, Pr = (N + 1) >> 1, Ps = dia_pitch(Np);
  const int nv = min(S.vp, S.nvec - v0);
  const int groups = 256 / Np;
  float4 *Ut = reinterpret_cast<float4 *>(smem);                // [Pr][N]
  float2 *psA = reinterpret_cast<float2 *>(Ut + (size_t)Pr * N); // [vp][Np + 2]
  float2 *psB = psA + (size_t)S.vp * Ps;                         // [vp][Np + 2]
  double2 *acc = reinterpret_cast<double2 *>(psB + (size_t)S.vp * Ps);  // [T]
  float2 *red = reinterpret_cast<float2 *>(acc + T);             // [vp]
  const int t = threadIdx.x, a = t % Np, sg = t / Np;
  for (int i = t; i < T; i += 256) acc[i] = make_double2(0.0, 0.0);
  // entries past N stay zero: U's pad column meets a zero, never garbage
  for (int i = t; i < 2 * S.vp * Ps; i += 256) psA[i] = make_float2(0.f, 0.f);
  const int64_t s0 = S.begin + v0 / 3;  // first sample of the workgroup
  const int nsamp = (nv + 2) / 3;
  const int64_t F0 = s0 * S.sample, F1 = (s0 + nsamp - 1) * S.sample + T;
  // U and the dipoles of the next frame in flight in registers while this
  // frame works
  const int nU = Pr * N;
  constexpr int kUPer = (32 * 64 + 255) / 256;
  float4 ureg[kUPer];
  float mreg[3] = {0.f, 0.f, 0.f};
  auto load_u = [&](int64_t f) {
    const float4 *Uf = reinterpret_cast<const float4 *>(S.U) + (f - S.f_lo) * nU;
#pragma unroll
    for (int u = 0; u < kUPer; ++u)
      if (t + u * 256 < nU) ureg[u] = gld(Uf + t + u * 256);
    const float *mu = S.mu_rec + (S.mu_stride ? f * S.mu_stride : 0) + 1;
    if (a < N)
#pragma unroll
      for (int x = 0; x < 3; ++x) mreg[x] = gld(mu + x * N + a);
  };
  load_u(F0);
  __syncthreads();
  for (int64_t f = F0; f < F1; ++f) {
#pragma unroll
    for (int u = 0; u < kUPer; ++u)
      if (t + u * 256 < nU) Ut[t + u * 256] = ureg[u];
    const float m3[3] = {mreg[0], mreg[1], mreg[2]};
    if (f + 1 < F1) load_u(f + 1);
    // vectors live at f: samples s0 + k with 0 <= f - (s0 + k) sample < T
    const int64_t kf = (f - F0) / S.sample;  // last sample started
    const int64_t klo = f - F0 - T < 0 ? 0 : (f - F0 - T) / S.sample + 1;
    const int vlo = (int)(3 * klo), vhi = (int)min<int64_t>(nv, 3 * (kf + 1));
    // this thread's live vectors: j in [jl, jh)
    int jl = kDiaVec, jh = 0;
#pragma unroll
    for (int j = 0; j < kDiaVec; ++j) {
      const int v = sg + j * groups;
      if (v >= vlo && v < vhi) jl = min(jl, j), jh = j + 1;
    }
    // responses of the live vectors (and the dipole start of new ones); the
    // site sums of all of them go through the lane butterfly together
    float r[kDiaVec], q[kDiaVec];
#pragma unroll
    for (int j = 0; j < kDiaVec; ++j) {
      r[j] = q[j] = 0.f;
      const int v = sg + j * groups;
      if (j < jl || j >= jh || a >= N) continue;
      const float m = m3[v % 3];
      float2 p;
      if (f == (s0 + v / 3) * S.sample) {
        p = make_float2(m, 0.f);
        psA[v * Ps + a] = p;
      } else {
        p = psA[v * Ps + a];
      }
      r[j] = m * p.x, q[j] = m * p.y;
    }
    for (int off = 1; off < Np; off <<= 1)
#pragma unroll
      for (int j = 0; j < kDiaVec; ++j)
        if (j >= jl && j < jh) r[j] += __shfl_xor(r[j], off), q[j] += __shfl_xor(q[j], off);
    if (a == 0)
#pragma unroll
      for (int j = 0; j < kDiaVec; ++j)
        if (j >= jl && j < jh) red[sg + j * groups] = make_float2(r[j], q[j]);
    lds_barrier();
    // responses into the accumulator, one thread per sample in polarisation
    // order (samples of one workgroup sit at different delays at any frame)
    if (t < nsamp) {
      const int64_t rel = f - (s0 + t) * S.sample;
      if (rel >= 0 && rel < T)
        for (int x = 0; x < 3; ++x) {
          const int v = 3 * t + x;
          if (v < nv) acc[rel].x += (double)red[v].x, acc[rel].y += (double)red[v].y;
        }
    }
    // psi <- U psi for the live vectors that go on (src/propagate.c:109-119)
    int jlo = kDiaVec, jhi = 0;
#pragma unroll
    for (int j = 0; j < kDiaVec; ++j) {
      const int v = sg + j * groups;
      const bool go = v >= vlo && v < vhi && f - (s0 + v / 3) * S.sample < T - 1;
      if (go) jlo = min(jlo, j), jhi = j + 1;
    }
    if (a < N && jlo < jhi) {
      const int vb = sg + jlo * groups;
      switch (jhi - jlo) {
        case 1: dia_apply<1>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
        case 2: dia_apply<2>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
        case 3: dia_apply<3>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
        case 4: dia_apply<4>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
        case 5: dia_apply<5>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
        case 6: dia_apply<6>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
        case 7: dia_apply<7>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
        default: dia_apply<8>(Ut, psA, psB, N, Pr, Ps, a, vb, groups); break;
      }
    }
    lds_barrier();
    float2 *tmp = psA;
    psA = psB, psB = tmp;
  }
  const int w = v0 / S.vp;
  for (int i = t; i < T; i += 256) gst(S.partial + (int64_t)w * T + i, acc[i]);
  if (t == 0 && times) {
    times[2 * blockIdx.x] = t_start;
    times[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
  }
}

size_t dia_conveyor_smem(int N, int Np, int vp, int T) {
  return (size_t)((N + 1) / 2) * N * 16 + (size_t)2 * vp * dia_pitch(Np) * 8 + (size_t)T * 16 + (size_t)vp * 8;
}

// Vectors per workgroup: kDiaVec per thread, whole samples, at most one
// sample per thread (the accumulation step).
int dia_vectors_per_group(int Np, int nvec) {
  const int v = std::min(3 * 256, 3 * ((kDiaVec * (256 / Np)) / 3));
  return std::max(3, std::min(v, 3 * ((nvec + 2) / 3)));
}

namespace {
int up(DevBuf &d, const void *s, size_t n) {
  int rc = d.alloc(n < 16 ? 16 : n);
  if (rc) return rc;
  if (n && s) NISE_HIP_CHECK(hipMemcpy(d.p, s, n, hipMemcpyHostToDevice));
  return NISE_OK;
}

struct SysState {
  int N = 0, T = 0, dia = 0, nvec = 0, nwg = 0;  // nwg: rows of partial responses
  float thres = 0.f;  // sparse DIA threshold (0: full DIA)
  int64_t f_lo = 0, f_hi = 0;
  FrameSource src;
  DevBuf H, mu, C, shs, sha, ab, diag, cs, steps, U, partial;
  const float *mu_dev = nullptr;
  int64_t mu_stride = 0;
  Schedule sch_asap, sch;
  std::vector<uint8_t> mask;
  // the schedules of the previous call, reused while their inputs repeat
  // (the row schedule's step table and frame-table order stay on the device)
  std::vector<uint8_t> asap_mask, rows_mask;
  int asap_N = 0, rows_R = 0, rows_N = 0;
  // everything but the device buffers and the cached schedules, which the
  // next call reuses
  void reset() {
    N = T = dia = nvec = nwg = 0;
    thres = 0.f;
    f_lo = f_hi = 0;
    src = FrameSource();
    mu_dev = nullptr;
    mu_stride = 0;
    mask.clear();  // the schedules and their inputs stay for the next call
  }
};

// Device buffers of the calling thread's batches, kept between calls: a
// batch of the same shape then allocates and frees nothing (hipMalloc and
// hipFree of the cfg4 batch's buffers took about 55 ms of host time per
// call).  Every call ends with both streams synchronised.
struct MultiCache {
  std::vector<std::unique_ptr<SysState>> sys;
  DevBuf mk, rsys, rwm, dsys, dwm, times, out, ejobs;
  std::vector<ExpmJob> ejobs_h;  // the job table of the last upload
  // the row count nise_absorption_multi_tune chose for a batch of this shape
  std::vector<int64_t> tune_sig;
  // 0: none; > 0 that count for every system; -1 each its own (capped at
  // kMultiRowCap); < -1 each its own capped at -tune_R
  int tune_R = 0;
};
// Cap of a system's own row count: 32 took the 96-system batch from 71.0 to
// 65.5 ms against 64 (the largest systems, N = 900-1000, then take 32 rows
// and two vectors per wave instead of one: twice the layers, half the
// waves, and half the VMEM loads per layer-vector).
constexpr int kMultiRowCap = 32;
thread_local MultiCache g_multi;

// The shape of a batch: per system N, tmax, samples, spacing, scheme, Trotter.
void batch_signature(const nise_system *systems, int nsys, std::vector<int64_t> &sig) {
  sig.clear();
  for (int k = 0; k < nsys; ++k) {
    const t_non *n = systems[k].non;
    sig.insert(sig.end(), {(int64_t)n->singles, (int64_t)n->tmax1, (int64_t)n->begin, (int64_t)n->end,
                           (int64_t)n->sample, (int64_t)n->propagation, (int64_t)n->ts});
  }
}
}  // namespace

}  // namespace nise

using namespace nise;

// One batch; R_force > 0 runs every Trotter system with that row count,
// R_force = -1 each with its own (capped at kMultiRowCap), R_force < -1 each
// with its own capped at -R_force; R_force = 0 takes the tuned choice for a
// batch of this shape, else each system's own count.
static int multi_run(const nise_system *systems, int nsys, int where, float *re_S, float *im_S, double *done_ms,
                     int R_force) {
  int rc = ensure_device();
  if (rc) return rc;
  if (nsys < 1 || !systems || !re_S || !im_S) return set_error("multi: bad arguments"), NISE_ERR_ARG;
  // Trotter systems on the default stream, DIA systems (per-frame Jacobi,
  // U and their conveyor) on a second one, so that the two overlap
  hipStream_t st = default_stream(), sd = aux_stream();
  if (!sd) return set_error("multi: no second stream"), NISE_ERR_HIP;
  MultiCache &cache = g_multi;
  // a call that returned early may have left work queued on the buffers
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  NISE_HIP_CHECK(hipStreamSynchronize(sd));
  std::vector<std::unique_ptr<SysState>> &S = cache.sys;
  if ((int)S.size() < nsys) S.resize(nsys);
  // 1. per-system inputs, structure and validation
  for (int k = 0; k < nsys; ++k) {
    if (!S[k]) S[k].reset(new SysState());
    S[k]->reset();
    SysState &s = *S[k];
    const nise_system &y = systems[k];
    const t_non *non = y.non;
    if (!non || !y.H_rec || !y.mu_rec) return set_error("multi: system %d has null inputs", k), NISE_ERR_ARG;
    if (strcmp(non->hamiltonian, "TransitionDipole") == 0 || strcmp(non->hamiltonian, "ExtendedDipole") == 0)
      return set_error("multi: system %d: on-the-fly couplings (%s) need nise_abs_plan_create_onfly", k,
                       non->hamiltonian), NISE_ERR_ARG;
    const int N = non->singles, T = non->tmax1;
    if (N < 1 || N > 65535 || T < 1 || non->sample < 1 || non->ts < 1)
      return set_error("multi: system %d has bad sizes", k), NISE_ERR_ARG;
    s.N = N, s.T = T;
    if (non->propagation == 0) {
      if (N > 64) return set_error("multi: DIA systems need N <= 64 (system %d has %d)", k, N), NISE_ERR_ARG;
      s.dia = 1;
      // src/propagate.c:25-28: a threshold in (0, 1] selects the sparse scheme
      s.thres = (non->thres > 0 && non->thres <= 1) ? non->thres : 0.f;
    } else if (non->propagation != 1) {
      return set_error("multi: system %d: propagation %d not provided", k, non->propagation), NISE_ERR_ARG;
    }
    int end = non->end ? non->end : (int)((y.nframes - T - 1) / non->sample + 1);
    if (end <= non->begin) return set_error("multi: system %d has no samples", k), NISE_ERR_RANGE;
    s.nvec = 3 * (end - non->begin);
    s.f_lo = (int64_t)non->begin * non->sample;
    s.f_hi = (int64_t)(end - 1) * non->sample + T;
    if (s.f_hi > y.nframes) return set_error("multi: system %d trajectory too short", k), NISE_ERR_RANGE;
    const bool ctype = strcmp(non->hamiltonian, "Coupling") == 0;
    if (ctype && !y.coupling_rec) return set_error("multi: system %d needs coupling_rec", k), NISE_ERR_ARG;
    const int64_t nfr = s.f_hi - s.f_lo;
    const float *Hd = y.H_rec, *mud = y.mu_rec, *Cd = y.coupling_rec;
    if (where == NISE_HOST) {
      if ((rc = up(s.H, y.H_rec + s.f_lo * y.H_stride, (size_t)(nfr * y.H_stride * 4)))) return rc;
      Hd = s.H.as<float>() - s.f_lo * y.H_stride;
      if (ctype) {
        if ((rc = up(s.mu, y.mu_rec, (size_t)y.mu_stride * 4))) return rc;
        mud = s.mu.as<float>();
        if ((rc = up(s.C, y.coupling_rec, (size_t)(1 + (int64_t)N * (N + 1) / 2) * 4))) return rc;
        Cd = s.C.as<float>();
      } else {
        if ((rc = up(s.mu, y.mu_rec + s.f_lo * y.mu_stride, (size_t)(nfr * y.mu_stride * 4)))) return rc;
        mud = s.mu.as<float>() - s.f_lo * y.mu_stride;
      }
    }
    s.mu_dev = mud, s.mu_stride = ctype ? 0 : y.mu_stride;
    s.src.H_rec = Hd, s.src.H_stride = y.H_stride, s.src.coupling_type = ctype, s.src.C_rec = Cd;
    s.src.shifte = non->shifte, s.src.cut = non->couplingcut, s.src.apply_shift = 1;
    s.src.nshift = non->SingleShiftSites;
    if (s.src.nshift > 0) {
      if ((rc = up(s.shs, non->SingleShiftSite, s.src.nshift * 4)) ||
          (rc = up(s.sha, non->SingleShift, 2 * s.src.nshift * 4)))
        return rc;
      s.src.shift_site = s.shs.as<int>(), s.src.shift_ab = s.sha.as<float>();
    }
  }
  // 2. the DIA systems' U per frame (the matrix exponential of each frame,
  //    expm_jobs_kernel: one launch per size class for all systems) on the
  //    second stream, queued before the host builds the Trotter schedules so
  //    that the two overlap
  std::vector<RowsArgs> rows;
  std::vector<int> rows_of;
  std::vector<DiaSys> dias;
  std::vector<int> dia_of;
  std::vector<ExpmJob> jobs;
  for (int k = 0; k < nsys; ++k) {
    SysState &s = *S[k];
    const t_non *non = systems[k].non;
    const int N = s.N;
    const int64_t nfr = s.f_hi - s.f_lo;
    if (!s.dia) continue;
      if ((rc = s.U.alloc((size_t)nfr * ((N + 1) / 2) * N * 16))) return rc;
      volatile float fv = non->deltat * kIcm2ifs;  // src/propagate.c:69
      fv = fv * kTwoPi;
      const float fac = fv;
      ExpmJob J;
      J.src = s.src, J.N = N, J.pairs = 1, J.f_first = s.f_lo, J.nfr = nfr, J.fac = fac, J.thres = s.thres;
      J.U = s.U.as<float2>();
      jobs.push_back(J);
      DiaSys d;
      d.N = N, d.Np = 1;
      while (d.Np < N) d.Np <<= 1;
      d.vp = dia_vectors_per_group(d.Np, s.nvec);
      d.tmax = s.T, d.begin = non->begin, d.sample = non->sample, d.nvec = s.nvec, d.f_lo = s.f_lo;
      d.U = s.U.as<float2>(), d.mu_rec = s.mu_dev, d.mu_stride = s.mu_stride;
      s.nwg = (s.nvec + d.vp - 1) / d.vp;
      if ((rc = s.partial.alloc((size_t)s.nwg * s.T * sizeof(double2)))) return rc;
      d.partial = s.partial.as<double2>();
      dias.push_back(d);
      dia_of.push_back(k);
  }
  if ((rc = launch_expm_jobs(jobs, cache.ejobs_h, cache.ejobs, sd))) return rc;
  // the coupling masks of all Trotter systems in one device buffer, one
  // copy back (no synchronisation per system)
  {
    std::vector<int64_t> moff(nsys + 1, 0);
    for (int k = 0; k < nsys; ++k)
      moff[k + 1] = moff[k] + (S[k]->dia ? 0 : (int64_t)S[k]->N * (S[k]->N - 1) / 2);
    if (moff[nsys] > 0) {
      DevBuf &mk = cache.mk;
      if ((rc = mk.alloc((size_t)moff[nsys]))) return rc;
      for (int k = 0; k < nsys; ++k) {
        SysState &s = *S[k];
        if (moff[k + 1] > moff[k] &&
            (rc = launch_union_mask(s.src, s.N, s.f_lo, s.f_hi - s.f_lo, mk.as<uint8_t>() + moff[k], st)))
          return rc;
      }
      std::vector<uint8_t> all((size_t)moff[nsys]);
      NISE_HIP_CHECK(hipMemcpyAsync(all.data(), mk.p, all.size(), hipMemcpyDeviceToHost, st));
      NISE_HIP_CHECK(hipStreamSynchronize(st));
      for (int k = 0; k < nsys; ++k)
        if (!S[k]->dia) S[k]->mask.assign(all.begin() + moff[k], all.begin() + moff[k + 1]);
    }
    for (int k = 0; k < nsys; ++k) {
      SysState &s = *S[k];
      if (s.dia || (s.asap_N == s.N && s.asap_mask == s.mask)) continue;
      build_schedule(s.N, s.mask, s.sch_asap);
      s.asap_mask = s.mask, s.asap_N = s.N;
    }
  }
  // 3. row counts of the Trotter launch
  // R > 0: every Trotter system with R rows; 0: each system its own count,
  // the power of two nearest its mean layer width up to 64 (systems with
  // many couplings per layer take wide layers, which shortens the chain of
  // layers each of their workgroups walks)
  int R = 0, cap = kMultiRowCap;
  std::vector<int64_t> sig;
  if (R_force == 0 && cache.tune_R != 0) batch_signature(systems, nsys, sig);
  if (R_force > 0) R = R_force;
  else if (R_force < -1) cap = -R_force;
  else if (R_force == 0 && cache.tune_R != 0 && sig == cache.tune_sig) {
    if (cache.tune_R > 0) R = cache.tune_R;
    else if (cache.tune_R < -1) cap = -cache.tune_R;
  }
  // 4. tables of the Trotter systems
  for (int k = 0; k < nsys; ++k) {
    SysState &s = *S[k];
    const t_non *non = systems[k].non;
    const int N = s.N;
    const int64_t nfr = s.f_hi - s.f_lo;
    if (s.dia) continue;
      const int Rk = R > 0 ? R : rows_for(N, s.sch_asap.Kreal, s.sch_asap.nlayers(), cap);
      if (s.rows_N != N || s.rows_R != Rk || s.rows_mask != s.mask) {
        s.rows_N = 0;  // until the tables below are on the device
        build_schedule_rows(N, s.mask, Rk, s.sch);
        std::vector<uint2> tab;
        rows_step_table(s.sch, Rk, tab);
        std::vector<uint32_t> abd;  // the frame tables follow the step table's order
        rows_device_order(s.sch, Rk, abd);
        if ((rc = up(s.ab, abd.data(), abd.size() * 4)) ||
            (rc = up(s.steps, tab.data(), tab.size() * sizeof(uint2))))
          return rc;
        s.rows_mask = s.mask, s.rows_R = Rk, s.rows_N = N;
      }
      if ((rc = s.diag.alloc((size_t)nfr * N * 8)) || (rc = s.cs.alloc((size_t)nfr * s.sch.K * 8))) return rc;
      if ((rc = launch_frame_tables(s.src, N, s.sch, s.ab.as<uint32_t>(), trotter_factor(non->deltat, 1, non->ts),
                                    s.f_lo, nfr, s.diag.as<float2>(), s.cs.as<float2>(), st)))
        return rc;
      s.nwg = s.nvec;  // partial rows: one per vector
      if ((rc = s.partial.alloc((size_t)s.nwg * s.T * sizeof(double2)))) return rc;
      RowsArgs r;
      r.N = N, r.tmax = s.T, r.m = non->ts, r.nlayers = s.sch.nlayers(), r.K = s.sch.K, r.f_lo = s.f_lo;
      r.rows = Rk;
      r.nvec = s.nvec, r.begin = non->begin, r.sample = non->sample;
      r.steps = s.steps.as<uint2>(), r.cs = s.cs.as<float2>(), r.diag = s.diag.as<float2>();
      r.mu_rec = s.mu_dev, r.mu_stride = s.mu_stride, r.partial = s.partial.as<double2>();
      rows.push_back(r);
      rows_of.push_back(k);
  }
  // 5. the two launches (row conveyor for all Trotter systems, DIA conveyor
  //    for all DIA systems), timed on the device clock
  std::vector<int2> rmap, dmap;
  size_t rsmem = 0, dsmem = 0;
  bool rpf = false;
  if (!rows.empty() && (rc = rows_prepare(rows.data(), (int)rows.size(), R, rmap, rsmem, rpf))) return rc;
  // longest workgroups first (layers x steps x Trotter substeps of their
  // system), so that the short ones fill the end of the launch
  std::stable_sort(rmap.begin(), rmap.end(), [&](const int2 &x, const int2 &y) {
    const RowsArgs &a = rows[x.x], &b = rows[y.x];
    return (int64_t)a.nlayers * a.tmax * a.m > (int64_t)b.nlayers * b.tmax * b.m;
  });
  for (size_t i = 0; i < dias.size(); ++i) {
    dsmem = std::max(dsmem, dia_conveyor_smem(dias[i].N, dias[i].Np, dias[i].vp, dias[i].tmax));
    for (int v = 0; v < dias[i].nvec; v += dias[i].vp) dmap.push_back(make_int2((int)i, v));
  }
  if (dsmem > 160 * 1024) return set_error("multi: DIA conveyor needs %zu B of LDS", dsmem), NISE_ERR_ARG;
  DevBuf &rsys = cache.rsys, &rwm = cache.rwm, &dsys = cache.dsys, &dwm = cache.dwm, &times = cache.times;
  const size_t nt = rmap.size() + dmap.size();
  if ((rc = up(rsys, rows.data(), rows.size() * sizeof(RowsArgs))) ||
      (rc = up(rwm, rmap.data(), rmap.size() * sizeof(int2))) ||
      (rc = up(dsys, dias.data(), dias.size() * sizeof(DiaSys))) ||
      (rc = up(dwm, dmap.data(), dmap.size() * sizeof(int2))) || (rc = times.alloc(nt * 16 + 16)))
    return rc;
  unsigned long long *tm = times.as<unsigned long long>();
  // The row conveyor starts after the matrix exponentials: its long-lived
  // waves fill the CUs' LDS, and an exponential launch queued behind them
  // (75-133 KB of LDS per workgroup) waited for them to drain (the N = 33-48
  // class took 50 ms instead of 0.6; the batch 65.6 -> 64.2 ms with the wait).
  {
    hipEvent_t ev;
    NISE_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e1 = hipEventRecord(ev, sd), e2 = e1 == hipSuccess ? hipStreamWaitEvent(st, ev, 0) : e1;
    (void)hipEventDestroy(ev);
    NISE_HIP_CHECK(e2);
  }
  if (!rmap.empty() && (rc = launch_rows(rsys.as<RowsArgs>(), rwm.as<int2>(), (int)rmap.size(), R, rpf, rsmem, st, tm)))
    return rc;
  if (!dmap.empty()) {
    NISE_HIP_CHECK(hipFuncSetAttribute((const void *)dia_conveyor_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)dsmem));
    hipLaunchKernelGGL(dia_conveyor_kernel, dim3((unsigned)dmap.size()), dim3(256), dsmem, sd,
                       dsys.as<DiaSys>(), dwm.as<int2>(), tm + 2 * rmap.size());
    NISE_HIP_CHECK(hipGetLastError());
  }
  // the reductions and copies below wait for the DIA stream
  {
    hipEvent_t ev;
    NISE_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e1 = hipEventRecord(ev, sd), e2 = e1 == hipSuccess ? hipStreamWaitEvent(st, ev, 0) : e1;
    (void)hipEventDestroy(ev);
    NISE_HIP_CHECK(e2);
  }
  // 6. per-system reductions and results
  std::vector<float> hr, hi;
  size_t off = 0;
  DevBuf &out = cache.out;
  int64_t Tsum = 0;
  for (int k = 0; k < nsys; ++k) Tsum += S[k]->T;
  if ((rc = out.alloc((size_t)Tsum * 8))) return rc;
  for (int k = 0; k < nsys; ++k) {
    SysState &s = *S[k];
    float *o = out.as<float>() + 2 * off;
    if ((rc = launch_reduce_partials(s.partial.as<double2>(), s.nwg, s.T, o, o + s.T, st))) return rc;
    off += s.T;
  }
  std::vector<float> ho((size_t)2 * Tsum);
  std::vector<unsigned long long> ht(2 * nt);
  NISE_HIP_CHECK(hipMemcpyAsync(ho.data(), out.p, ho.size() * 4, hipMemcpyDeviceToHost, st));
  if (nt) NISE_HIP_CHECK(hipMemcpyAsync(ht.data(), times.p, ht.size() * 8, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  NISE_HIP_CHECK(hipStreamSynchronize(sd));
  off = 0;
  for (int k = 0; k < nsys; ++k) {
    std::memcpy(re_S + off, ho.data() + 2 * off, S[k]->T * 4);
    std::memcpy(im_S + off, ho.data() + 2 * off + S[k]->T, S[k]->T * 4);
    off += S[k]->T;
  }
  if (done_ms && nt) {
    int khz = 0;
    NISE_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
    unsigned long long t0 = ~0ull;
    for (size_t i = 0; i < nt; ++i) t0 = std::min(t0, ht[2 * i]);
    std::vector<unsigned long long> done(nsys, t0);
    for (size_t i = 0; i < rmap.size(); ++i) {
      const int k = rows_of[rmap[i].x];
      done[k] = std::max(done[k], ht[2 * i + 1]);
    }
    for (size_t i = 0; i < dmap.size(); ++i) {
      const int k = dia_of[dmap[i].x];
      done[k] = std::max(done[k], ht[2 * (rmap.size() + i) + 1]);
    }
    for (int k = 0; k < nsys; ++k) done_ms[k] = khz > 0 ? (double)(done[k] - t0) / khz : -1.0;
  }
  return NISE_OK;
}

extern "C" int nise_absorption_multi(const nise_system *systems, int nsys, int where, float *re_S,
                                     float *im_S, double *done_ms) {
  return multi_run(systems, nsys, where, re_S, im_S, done_ms, 0);
}

// Times the batch with each system's own row count (capped at 32, 16 and
// 64: codes 0, -16, -64) and with every common row count of the row
// conveyor (1, 2, ..., 64; the best of two runs each, after a warm run) and
// keeps the fastest for later nise_absorption_multi calls on a batch of the
// same shape in this thread.
extern "C" int nise_absorption_multi_tune(const nise_system *systems, int nsys, int where, int *R_best, float *ms,
                                          int *codes, int n) {
  int rc = ensure_device();
  if (rc) return rc;
  if (nsys < 1 || !systems) return set_error("multi tune: bad arguments"), NISE_ERR_ARG;
  int64_t Tsum = 0;
  for (int k = 0; k < nsys; ++k) Tsum += systems[k].non ? systems[k].non->tmax1 : 0;
  std::vector<float> re((size_t)std::max<int64_t>(Tsum, 1)), im(re.size());
  g_multi.tune_R = 0;
  if ((rc = multi_run(systems, nsys, where, re.data(), im.data(), nullptr, 0))) return rc;
  int best = -2, cnt = 0;
  double best_ms = 0;
  for (int R : {0, -16, -64, 1, 2, 4, 8, 16, 32, 64}) {  // <= 0: each system its own count
    double t = -1;
    for (int rep = 0; rep < 2; ++rep) {
      const auto t0 = std::chrono::steady_clock::now();
      const int r = multi_run(systems, nsys, where, re.data(), im.data(), nullptr, R ? R : -1);  // -1: cap 32
      const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (r) {  // a row count the batch cannot take (LDS): skipped
        t = -1;
        break;
      }
      t = t < 0 ? dt : std::min(t, dt);
    }
    if (t < 0) continue;
    if (ms && codes && cnt < n) ms[cnt] = (float)t, codes[cnt] = R;
    ++cnt;
    if (best == -2 || t < best_ms) best = R, best_ms = t;
  }
  if (best == -2) return set_error("multi tune: no row count ran"), NISE_ERR_ARG;
  batch_signature(systems, nsys, g_multi.tune_sig);
  g_multi.tune_R = best ? best : -1;  // -16, -64: those caps
  if (R_best) *R_best = best;
  // one run with the kept choice, so that the next call finds its schedules
  // and tables in the cache (the last candidate timed left its own there)
  if ((rc = multi_run(systems, nsys, where, re.data(), im.data(), nullptr, 0))) return rc;
  return cnt;
}
