// Exact-exponential (DIA) propagation for gfx950.
//
// Reference: propagate_vec_DIA, src/propagate.c:60-125, which builds the
// dense Hamiltonian, diagonalises it with LAPACK ssyev_ (diagonalizeLPD,
// src/propagate.c:1185-1221) and applies U = V exp(-i E f) V^T with
// f = deltat*icm2ifs*twoPi*sign (no Trotter division).
//
// Here one workgroup per matrix runs a parallel cyclic Jacobi method in
// double precision with both A and V resident in LDS: every round rotates
// M/2 disjoint index pairs (round-robin tournament order), so the angles of
// a round can all be computed from the same A.  Sweeps continue until the
// off-diagonal Frobenius norm falls below 1e-28 of the total (at most 40).
// U is then formed once per frame and applied to every vector that shares
// the frame.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nise_internal.h"

namespace nise {

constexpr int kJacBlock = 256;
constexpr int kJacMaxN = 100;  // 2 N^2 doubles must fit 160 KiB of LDS

// Pair i of round r of the round-robin schedule on M (even) positions.
__device__ __forceinline__ void rr_pair(int M, int r, int i, int &p, int &q) {
  auto pos = [&](int k) { return k == 0 ? 0 : 1 + (k - 1 + r) % (M - 1); };
  p = pos(i), q = pos(M - 1 - i);
  if (p > q) { int t = p; p = q; q = t; }
}

// LDS row stride of the Jacobi matrices: odd (in doubles) when the padded
// matrices fit, so that a wave walking down a column spreads over the banks
// (a stride of N = 64 doubles put every lane on one bank).
__host__ __device__ inline int jacobi_ld(int N) {
  const int ld = N | 1;
  return (size_t)2 * N * ld * sizeof(double) + 4096 <= 160 * 1024 ? ld : N;
}

// task = i N + k for 0 <= task < (M / 2) N, without an integer division.
__device__ __forceinline__ void split_task(int task, int N, float invN, int &i, int &k) {
  i = (int)((float)task * invN);
  k = task - i * N;
  if (k >= N) ++i, k -= N;
  else if (k < 0) --i, k += N;
}

// A: row-major symmetric N x N (double), one matrix per workgroup.
// Outputs: w[N] ascending eigenvalues, V[N][N] with V[i*N+j] = component i of
// eigenvector j.  status[b] = sweeps used, or -1 without convergence.
// In LDS both matrices have the row stride ld = jacobi_ld(N).
__global__ void __launch_bounds__(kJacBlock) jacobi_kernel(const double *Ain, int N, double *w_out,
                                                           double *V_out, int *status) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ld = jacobi_ld(N);
  double *A = reinterpret_cast<double *>(smem);
  double *V = A + N * ld;
  double *cs = V + N * ld;                   // [M/2][2]
  int *pq = reinterpret_cast<int *>(cs + N + 2);  // [M/2][2], even length
  double *red = reinterpret_cast<double *>(pq + ((N + 3) & ~1));  // [kJacBlock/64][2]
  const int t = threadIdx.x;
  const int64_t b = blockIdx.x;
  const int M = N + (N & 1);
  const float invN = 1.0f / (float)N;
  for (int e = t; e < N * N; e += kJacBlock) {
    int r, c;
    split_task(e, N, invN, r, c);
    A[r * ld + c] = Ain[b * N * N + e];
    V[r * ld + c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  int sweep = 0, converged = 0;
  for (; sweep < 40; ++sweep) {
    // convergence test: off-diagonal vs total Frobenius norm
    double off = 0, tot = 0;
    for (int e = t; e < N * N; e += kJacBlock) {
      int r, c;
      split_task(e, N, invN, r, c);
      double a = A[r * ld + c] * A[r * ld + c];
      tot += a;
      if (r != c) off += a;
    }
    for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o), tot += __shfl_xor(tot, o);
    if ((t & 63) == 0) red[2 * (t >> 6)] = off, red[2 * (t >> 6) + 1] = tot;
    __syncthreads();
    off = 0, tot = 0;
    for (int k = 0; k < kJacBlock / 64; ++k) off += red[2 * k], tot += red[2 * k + 1];
    __syncthreads();
    if (off <= 1e-28 * tot || off == 0.0) { converged = 1; break; }
    for (int r = 0; r < M - 1; ++r) {
      for (int i = t; i < M / 2; i += kJacBlock) {
        int p, q;
        rr_pair(M, r, i, p, q);
        double c = 1.0, s = 0.0;
        if (q < N) {
          double apq = A[p * ld + q];
          if (apq != 0.0) {
            double tau = (A[q * ld + q] - A[p * ld + p]) / (2.0 * apq);
            double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tt * tt);
            s = tt * c;
          }
        } else {
          q = p;  // dummy partner: identity
        }
        cs[2 * i] = c, cs[2 * i + 1] = s;
        pq[2 * i] = p, pq[2 * i + 1] = q;
      }
      __syncthreads();
      // A <- A J and V <- V J (columns p, q of each pair)
      for (int task = t; task < (M / 2) * N; task += kJacBlock) {
        int i, k;
        split_task(task, N, invN, i, k);
        int p = pq[2 * i], q = pq[2 * i + 1];
        if (p == q) continue;
        double c = cs[2 * i], s = cs[2 * i + 1];
        double akp = A[k * ld + p], akq = A[k * ld + q];
        A[k * ld + p] = c * akp - s * akq;
        A[k * ld + q] = s * akp + c * akq;
        double vkp = V[k * ld + p], vkq = V[k * ld + q];
        V[k * ld + p] = c * vkp - s * vkq;
        V[k * ld + q] = s * vkp + c * vkq;
      }
      __syncthreads();
      // A <- J^T A (rows p, q of each pair)
      for (int task = t; task < (M / 2) * N; task += kJacBlock) {
        int i, k;
        split_task(task, N, invN, i, k);
        int p = pq[2 * i], q = pq[2 * i + 1];
        if (p == q) continue;
        double c = cs[2 * i], s = cs[2 * i + 1];
        double apk = A[p * ld + k], aqk = A[q * ld + k];
        A[p * ld + k] = c * apk - s * aqk;
        A[q * ld + k] = s * apk + c * aqk;
      }
      __syncthreads();
    }
  }
  // eigenvalues ascending (the order ssyev_ returns), eigenvectors follow
  if (t == 0) {
    int *perm = pq;  // reuse: N entries
    for (int i = 0; i < N; ++i) perm[i] = i;
    for (int i = 1; i < N; ++i) {  // insertion sort on the diagonal
      int k = perm[i];
      double d = A[k * ld + k];
      int j = i - 1;
      while (j >= 0 && A[perm[j] * ld + perm[j]] > d) { perm[j + 1] = perm[j]; --j; }
      perm[j + 1] = k;
    }
    status[b] = converged ? sweep : -1;
  }
  __syncthreads();
  for (int j = t; j < N; j += kJacBlock) w_out[b * N + j] = A[pq[j] * ld + pq[j]];
  for (int e = t; e < N * N; e += kJacBlock) {
    int r, j;
    split_task(e, N, invN, r, j);
    V_out[b * N * N + e] = V[r * ld + pq[j]];
  }
}

size_t jacobi_smem(int N) {
  return (size_t)2 * N * jacobi_ld(N) * sizeof(double) + (size_t)(N + 2) * sizeof(double) +
         (size_t)((N + 3) & ~1) * sizeof(int) + (kJacBlock / 64) * 2 * sizeof(double) + 16;
}

int launch_jacobi(const double *A, int N, int nmat, double *w, double *V, int *status, hipStream_t st) {
  if (N < 1 || N > kJacMaxN) return set_error("DIA: N=%d outside 1..%d", N, kJacMaxN), NISE_ERR_ARG;
  size_t smem = jacobi_smem(N);
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)jacobi_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(jacobi_kernel, dim3(nmat), dim3(kJacBlock), smem, st, A, N, w, V, status);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

// Dense double matrix from a packed float triangle (src/propagate.c:79-85).
__global__ void dense_from_triangle(const float *Htri, int N, double *A) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * N) return;
  int a = i / N, b = i % N;
  A[i] = (double)Htri[sindex(a, b, N)];
}

// U = V exp(-i E f) V^T (src/propagate.c:87-106) and psi <- U psi for nvec
// column vectors (src/propagate.c:109-119).  One workgroup per vector.
__global__ void dia_apply_kernel(const double *w, const double *V, int N, float fac, float2 *psi, int nvec) {
  extern __shared__ __align__(16) unsigned char smem[];
  double *yr = reinterpret_cast<double *>(smem);  // V^T psi, then phase
  double *yi = yr + N;
  const int v = blockIdx.x, t = threadIdx.x;
  if (v >= nvec) return;
  float2 *p = psi + (int64_t)v * N;
  for (int j = t; j < N; j += blockDim.x) {
    double sr = 0, si = 0;
    for (int a = 0; a < N; ++a) sr += V[a * N + j] * p[a].x, si += V[a * N + j] * p[a].y;
    // reference: re_U = cos(e f), im_U = -sin(e f) with e, f float
    float ef = __fmul_rn((float)w[j], fac);
    double c = cos((double)ef), s = -sin((double)ef);
    yr[j] = sr * c - si * s;
    yi[j] = si * c + sr * s;
  }
  __syncthreads();
  for (int a = t; a < N; a += blockDim.x) {
    double sr = 0, si = 0;
    for (int j = 0; j < N; ++j) sr += V[a * N + j] * yr[j], si += V[a * N + j] * yi[j];
    p[a] = make_float2((float)sr, (float)si);
  }
}

// psi <- U psi with a (thresholded) propagator in FP32, the kept elements
// accumulated in the reference's order (src/propagate.c:238-247).  One
// workgroup per vector.
__global__ void dense_apply_kernel(const float2 *U, int N, float2 *psi, int nvec) {
  extern __shared__ __align__(16) unsigned char smem[];
  float2 *c = reinterpret_cast<float2 *>(smem);
  const int v = blockIdx.x, t = threadIdx.x;
  if (v >= nvec) return;
  float2 *p = psi + (int64_t)v * N;
  for (int b = t; b < N; b += blockDim.x) c[b] = p[b];
  __syncthreads();
  for (int a = t; a < N; a += blockDim.x) {
    float yr = 0.f, yi = 0.f;
    for (int b = 0; b < N; ++b) {
      const float2 u = U[a * N + b];
      if (u.x == 0.f && u.y == 0.f) continue;  // dropped element
      yr = __fadd_rn(yr, __fsub_rn(__fmul_rn(u.x, c[b].x), __fmul_rn(u.y, c[b].y)));
      yi = __fadd_rn(yi, __fadd_rn(__fmul_rn(u.x, c[b].y), __fmul_rn(u.y, c[b].x)));
    }
    p[a] = make_float2(yr, yi);
  }
}

namespace {
struct DiaCtx {
  DevBuf H, A, w, V, st, psi, U;
  std::vector<float2> sbuf;
};
DiaCtx &dctx() {
  static thread_local DiaCtx c;
  return c;
}
int up(DevBuf &d, const void *s, size_t n, hipStream_t st) {
  int rc = d.alloc(n < 16 ? 16 : n);
  if (rc) return rc;
  if (n && s) NISE_HIP_CHECK(hipMemcpyAsync(d.p, s, n, hipMemcpyHostToDevice, st));
  return NISE_OK;
}
}  // namespace

// nvec vectors stored consecutively (stride N) in cr/ci, one frame H.
int dia_propagate_host(const t_non *non, const float *H, float *cr, float *ci, int sign, int nvec,
                       hipStream_t st) {
  int rc = ensure_device();
  if (rc) return rc;
  const int N = non->singles;
  if (N < 1 || N > kJacMaxN) return set_error("DIA on the GPU supports 1..%d sites, got %d", kJacMaxN, N), NISE_ERR_ARG;
  DiaCtx &c = dctx();
  const int nn2 = N * (N + 1) / 2;
  if ((rc = up(c.H, H, nn2 * 4, st))) return rc;
  if ((rc = c.A.alloc((size_t)N * N * 8)) || (rc = c.V.alloc((size_t)N * N * 8)) ||
      (rc = c.w.alloc(N * 8)) || (rc = c.st.alloc(4)))
    return rc;
  hipLaunchKernelGGL(dense_from_triangle, dim3((N * N + 255) / 256), dim3(256), 0, st, c.H.as<float>(), N,
                     c.A.as<double>());
  if ((rc = launch_jacobi(c.A.as<double>(), N, 1, c.w.as<double>(), c.V.as<double>(), c.st.as<int>(), st)))
    return rc;
  c.sbuf.resize((size_t)nvec * N);
  for (size_t i = 0; i < c.sbuf.size(); ++i) c.sbuf[i] = make_float2(cr[i], ci[i]);
  if ((rc = up(c.psi, c.sbuf.data(), c.sbuf.size() * sizeof(float2), st))) return rc;
  volatile float fac = non->deltat * kIcm2ifs;
  fac = fac * kTwoPi;
  fac = fac * (float)sign;
  hipLaunchKernelGGL(dia_apply_kernel, dim3(nvec), dim3(64), 2 * N * sizeof(double), st,
                     c.w.as<double>(), c.V.as<double>(), N, (float)fac, c.psi.as<float2>(), nvec);
  NISE_HIP_CHECK(hipGetLastError());
  int status = 0;
  NISE_HIP_CHECK(hipMemcpyAsync(&status, c.st.p, 4, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipMemcpyAsync(c.sbuf.data(), c.psi.p, c.sbuf.size() * sizeof(float2),
                                hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  if (status < 0) return set_error("Jacobi eigensolver did not converge"), NISE_ERR_NUMERIC;
  for (size_t i = 0; i < c.sbuf.size(); ++i) cr[i] = c.sbuf[i].x, ci[i] = c.sbuf[i].y;
  return NISE_OK;
}

// The sparse scheme (src/propagate.c:185-258): U as for DIA, elements with
// |U_ab|^2 <= thres dropped, psi <- U psi; *elements counts the kept ones.
int dia_s_propagate_host(const t_non *non, const float *H, float *cr, float *ci, int sign, int nvec,
                         hipStream_t st, int *elements) {
  int rc = ensure_device();
  if (rc) return rc;
  const int N = non->singles;
  if (N < 1 || N > kJacMaxN) return set_error("DIA on the GPU supports 1..%d sites, got %d", kJacMaxN, N), NISE_ERR_ARG;
  DiaCtx &c = dctx();
  const int nn2 = N * (N + 1) / 2;
  if ((rc = up(c.H, H, nn2 * 4, st))) return rc;
  if ((rc = c.A.alloc((size_t)N * N * 8)) || (rc = c.V.alloc((size_t)N * N * 8)) ||
      (rc = c.w.alloc(N * 8)) || (rc = c.st.alloc(4)) || (rc = c.U.alloc((size_t)N * N * 8)))
    return rc;
  hipLaunchKernelGGL(dense_from_triangle, dim3((N * N + 255) / 256), dim3(256), 0, st, c.H.as<float>(), N,
                     c.A.as<double>());
  if ((rc = launch_jacobi(c.A.as<double>(), N, 1, c.w.as<double>(), c.V.as<double>(), c.st.as<int>(), st)))
    return rc;
  volatile float fac = non->deltat * kIcm2ifs;  // src/propagate.c:197, no /m
  fac = fac * kTwoPi;
  fac = fac * (float)sign;
  if ((rc = launch_dia_u(c.w.as<double>(), c.V.as<double>(), N, 1, (float)fac, non->thres, c.U.as<float2>(), st)))
    return rc;
  c.sbuf.resize((size_t)nvec * N);
  for (size_t i = 0; i < c.sbuf.size(); ++i) c.sbuf[i] = make_float2(cr[i], ci[i]);
  if ((rc = up(c.psi, c.sbuf.data(), c.sbuf.size() * sizeof(float2), st))) return rc;
  hipLaunchKernelGGL(dense_apply_kernel, dim3(nvec), dim3(64), N * sizeof(float2), st, c.U.as<float2>(), N,
                     c.psi.as<float2>(), nvec);
  NISE_HIP_CHECK(hipGetLastError());
  std::vector<float2> hU((size_t)N * N);
  int status = 0;
  NISE_HIP_CHECK(hipMemcpyAsync(&status, c.st.p, 4, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipMemcpyAsync(hU.data(), c.U.p, hU.size() * sizeof(float2), hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipMemcpyAsync(c.sbuf.data(), c.psi.p, c.sbuf.size() * sizeof(float2),
                                hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  if (status < 0) return set_error("Jacobi eigensolver did not converge"), NISE_ERR_NUMERIC;
  int kept = 0;
  for (const float2 &u : hU) kept += (u.x != 0.f || u.y != 0.f);
  if (elements) *elements = kept;
  for (size_t i = 0; i < c.sbuf.size(); ++i) cr[i] = c.sbuf[i].x, ci[i] = c.sbuf[i].y;
  return NISE_OK;
}

}  // namespace nise

using namespace nise;

// src/propagate.c:60-125
extern "C" void propagate_vec_DIA(t_non *non, float *Hamiltonian_i, float *cr, float *ci, int sign) {
  if (dia_propagate_host(non, Hamiltonian_i, cr, ci, sign, 1, default_stream())) {
    // src/propagate.c:1206-1209 prints and exits on a failed diagonalisation
    printf("Something went wrong trying to diagonalize a matrix...\n(%s)\n", get_error());
    exit(0);
  }
}

// src/propagate.c:1185-1221
extern "C" void diagonalizeLPD(float *H, float *v, int N) {
  hipStream_t st = default_stream();
  int rc = ensure_device();
  DevBuf A, w, V, s;
  std::vector<double> hA((size_t)N * N), hw(N), hV((size_t)N * N);
  for (size_t i = 0; i < hA.size(); ++i) hA[i] = H[i];
  int status = -1;
  if (!rc) rc = A.alloc(hA.size() * 8);
  if (!rc) rc = w.alloc(N * 8);
  if (!rc) rc = V.alloc(hV.size() * 8);
  if (!rc) rc = s.alloc(4);
  if (!rc && hipMemcpyAsync(A.p, hA.data(), hA.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess) rc = NISE_ERR_HIP;
  if (!rc) rc = launch_jacobi(A.as<double>(), N, 1, w.as<double>(), V.as<double>(), s.as<int>(), st);
  if (!rc && (hipMemcpyAsync(hw.data(), w.p, N * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipMemcpyAsync(hV.data(), V.p, hV.size() * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipMemcpyAsync(&status, s.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess))
    rc = NISE_ERR_HIP;
  if (rc || status < 0) {
    printf("Something went wrong trying to diagonalize a matrix...\n");
    exit(0);
  }
  for (int i = 0; i < N; ++i) v[i] = (float)hw[i];
  for (size_t i = 0; i < hV.size(); ++i) H[i] = (float)hV[i];
}
