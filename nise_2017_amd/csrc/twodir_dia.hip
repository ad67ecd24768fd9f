// Exact-exponential and sparse propagation for the 2D response
// (src/calc_2DIR.c:184-506 with propagation 0 or a waiting time):
//
//  - dia_slots_kernel: singles vectors stepped through per-frame propagators
//    U(f) = V exp(-i E f) V^T (FP32 tables from the FP64 Jacobi solver),
//    the exact exponential of propagate_vec_DIA (src/propagate.c:60-125),
//    of the sparse scheme propagate_vec_DIA_S (:185-258, dropped elements
//    zeroed in the table) and of the waiting-time step propagate_t2_DIA
//    (:130-180).  Slots of one workgroup share their first frame and step
//    counts; a step is psi_a <- sum_b U_ab psi_b, summed in double in
//    ascending b and rounded once.
//  - sparse_lists_kernel: per frame, the kept elements of the one-exciton
//    propagator in time_evolution_mat's order (src/propagate.c:1287-1296,
//    row a outer, column b inner) with the caller's swap of the row and
//    column arrays (:1496): element e carries (s, t) = (b, a), and per
//    column s the ascending list of its elements.
//  - sparse_doubles_kernel: the sparse two-exciton step of
//    propagate_double_sparce / _ES (src/propagate.c:1304-1473).  Output
//    state (p, q) gathers the reference's terms U_e1 U_e2 g f[Sindex(t1,t2)]
//    over the element pairs e2 <= e1 with {s1, s2} = {p, q}, in the
//    reference's order (e1 ascending, then e2 ascending), each term with the
//    reference's FP32 roundings.
#include <algorithm>
#include <cmath>

#include "nise_internal.h"

namespace nise {

__global__ void __launch_bounds__(256) dia_slots_kernel(DiaSlotArgs a) {
  extern __shared__ __align__(16) float2 dsm[];
  const int b0 = a.wg_begin[blockIdx.x], b1 = a.wg_begin[blockIdx.x + 1], ns = b1 - b0, N = a.N;
  const int t = threadIdx.x, nt = blockDim.x;
  float2 *U = dsm;                 // [N][N] the frame's propagator
  float2 *pa = U + (size_t)N * N;  // [ns][N]
  float2 *pb = pa + (size_t)ns * N;
  const DiaSlot d0 = a.slots[b0];  // first frame and step counts of the group
  for (int e = t; e < ns * N; e += nt) {
    const int v = e / N, i = e - v * N;
    const DiaSlot d = a.slots[b0 + v];
    if (d.in_row >= 0) {
      pa[e] = a.in[(int64_t)d.in_row * N + i];
    } else {
      const float *mu = a.mu_rec + (a.mu_stride ? (int64_t)d.fs * a.mu_stride : 0) + 1 + (int64_t)d.x * N;
      pa[e] = make_float2(mu[i], 0.f);
    }
  }
  const int steps = max(d0.nresp, d0.nprop);
  for (int k = 0; k < steps; ++k) {
    if (k < d0.nresp && a.hist)
      for (int e = t; e < ns * N; e += nt) {
        const int v = e / N, i = e - v * N;
        const int h = a.slots[b0 + v].hist_row;
        if (h >= 0) a.hist[((int64_t)h + k) * N + i] = pa[e];
      }
    if (k >= d0.nprop) continue;
    const float2 *Uf = a.U + ((int64_t)d0.fs + k - a.f_lo) * N * N;
    __syncthreads();  // pa is complete; U is free
    for (int e = t; e < N * N; e += nt) U[e] = Uf[e];
    __syncthreads();
    for (int e = t; e < ns * N; e += nt) {
      const int v = e / N, i = e - v * N;
      const float2 *u = U + (size_t)i * N, *p = pa + (size_t)v * N;
      double sr = 0.0, si = 0.0;
      for (int b = 0; b < N; ++b) {
        const float2 ub = u[b], pv = p[b];
        sr += (double)ub.x * pv.x - (double)ub.y * pv.y;
        si += (double)ub.x * pv.y + (double)ub.y * pv.x;
      }
      pb[e] = make_float2((float)sr, (float)si);
    }
    float2 *tmp = pa;
    pa = pb, pb = tmp;
  }
  __syncthreads();
  if (a.out)
    for (int e = t; e < ns * N; e += nt) {
      const int v = e / N, i = e - v * N;
      const int o = a.slots[b0 + v].out_row;
      if (o >= 0) a.out[(int64_t)o * N + i] = pa[e];
    }
}

size_t dia_slots_smem(int N, int ns) { return (size_t)N * N * 8 + (size_t)2 * ns * N * 8; }

int launch_dia_slots(const DiaSlotArgs &a, int max_ns, hipStream_t st) {
  if (a.nwg == 0) return NISE_OK;
  const size_t smem = dia_slots_smem(a.N, max_ns);
  if (smem > 160 * 1024) return set_error("DIA slots: %zu B of LDS", smem), NISE_ERR_ARG;
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)dia_slots_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem));
  hipLaunchKernelGGL(dia_slots_kernel, dim3((unsigned)a.nwg), dim3(256), smem, st, a);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

// Largest slot count per workgroup whose LDS image fits.
int dia_slots_max(int N) {
  const int64_t room = 160 * 1024 - (int64_t)N * N * 8;
  return room <= 0 ? 0 : (int)std::min<int64_t>(256, room / (16 * (int64_t)N));
}

// One workgroup per frame row.  Element e of row m: el[m][e] = (s, t as
// int bits, Re U, Im U); colptr[m][s] .. colptr[m][s+1] index colidx[m][..],
// the elements of column s in ascending e.
__global__ void __launch_bounds__(256) sparse_lists_kernel(const float2 *U, int N, float thres, float4 *el,
                                                           int *colptr, int *colidx, int *cnt) {
  extern __shared__ __align__(16) int lsm[];
  int *rank = lsm;                     // [N][N] rank of a kept element in its row, -1 dropped
  int *rowoff = rank + N * N;          // [N + 1]
  int *cp = rowoff + N + 1;            // [N + 1]
  const int64_t m = blockIdx.x;
  const float2 *Um = U + m * N * N;
  const int t = threadIdx.x;
  for (int a = t; a < N; a += blockDim.x) {  // src/propagate.c:1287-1296, row a
    int r = 0;
    for (int b = 0; b < N; ++b) {
      const float2 u = Um[a * N + b];
      const bool keep = __fadd_rn(__fmul_rn(u.x, u.x), __fmul_rn(u.y, u.y)) > thres;
      rank[a * N + b] = keep ? r++ : -1;
    }
    rowoff[a + 1] = r;
  }
  __syncthreads();
  if (t == 0) {
    rowoff[0] = 0;
    for (int a = 0; a < N; ++a) rowoff[a + 1] += rowoff[a];
    cnt[m] = rowoff[N];
  }
  __syncthreads();
  float4 *elm = el + m * N * N;
  for (int i = t; i < N * N; i += blockDim.x) {
    const int a = i / N, b = i % N, r = rank[i];
    if (r >= 0) {
      const float2 u = Um[i];
      elm[rowoff[a] + r] = make_float4(__int_as_float(b), __int_as_float(a), u.x, u.y);
    }
  }
  for (int s = t; s < N; s += blockDim.x) {
    int c = 0;
    for (int a = 0; a < N; ++a) c += rank[a * N + s] >= 0;
    cp[s + 1] = c;
  }
  __syncthreads();
  if (t == 0) {
    cp[0] = 0;
    for (int s = 0; s < N; ++s) cp[s + 1] += cp[s];
  }
  __syncthreads();
  int *cpm = colptr + m * (N + 1), *cim = colidx + m * N * N;
  for (int s = t; s <= N; s += blockDim.x) cpm[s] = cp[s];
  for (int s = t; s < N; s += blockDim.x) {
    int k = cp[s];
    for (int a = 0; a < N; ++a)
      if (rank[a * N + s] >= 0) cim[k++] = rowoff[a] + rank[a * N + s];
  }
}

int launch_sparse_lists(const float2 *U, int N, int64_t nrows, float thres, float4 *el, int *colptr, int *colidx,
                        int *cnt, hipStream_t st) {
  if (nrows <= 0) return NISE_OK;
  const size_t smem = ((size_t)N * N + 2 * (N + 1)) * 4;
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)sparse_lists_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem));
  for (int64_t r0 = 0; r0 < nrows; r0 += 65535) {
    const int64_t nr = std::min<int64_t>(65535, nrows - r0);
    hipLaunchKernelGGL(sparse_lists_kernel, dim3((unsigned)nr), dim3(256), smem, st, U + r0 * N * N, N, thres,
                       el + r0 * N * N, colptr + r0 * (N + 1), colidx + r0 * N * N, cnt + r0);
    NISE_HIP_CHECK(hipGetLastError());
  }
  return NISE_OK;
}

__device__ __forceinline__ int sidx(int a, int b, int N) {  // src/NISE_subs.h:54-65
  return a > b ? a + ((b * (2 * N - b - 1)) >> 1) : b + ((a * (2 * N - a - 1)) >> 1);
}

// One term of src/propagate.c:1370-1373 (g as :1364-1368, or :1451-1455
// for two-level sites) added to acc in the reference's statement order.
__device__ __forceinline__ void sparse_term(float2 &acc, const float4 a, const float4 b, const float2 *f, int N,
                                            int es) {
  const int s = __float_as_int(a.x), t = __float_as_int(a.y), u = __float_as_int(b.x), v = __float_as_int(b.y);
  float g = 1.f;
  if (s != u) g = es ? 0.f : kSqrt2;
  if (v != t) g = es ? 0.f : kSqrt2;
  if (s != u && v != t) g = 1.f;
  const float2 fb = f[sidx(t, v, N)];
  const float re = __fsub_rn(__fmul_rn(a.z, b.z), __fmul_rn(a.w, b.w));
  const float im = __fadd_rn(__fmul_rn(a.w, b.z), __fmul_rn(a.z, b.w));
  acc.x = __fadd_rn(acc.x, __fmul_rn(__fmul_rn(re, fb.x), g));
  acc.x = __fsub_rn(acc.x, __fmul_rn(__fmul_rn(im, fb.y), g));
  acc.y = __fadd_rn(acc.y, __fmul_rn(__fmul_rn(re, fb.y), g));
  acc.y = __fadd_rn(acc.y, __fmul_rn(__fmul_rn(im, fb.x), g));
}

// The anharmonic half phase on |aa> (src/propagate.c:1345-1349, 1385-1389):
// co f - si f rotated by +fm * Delta_a.
__device__ __forceinline__ float2 anh_phase(float2 v, float co, float si) {
  return make_float2(__fsub_rn(__fmul_rn(co, v.x), __fmul_rn(si, v.y)),
                     __fadd_rn(__fmul_rn(co, v.y), __fmul_rn(si, v.x)));
}

__global__ void __launch_bounds__(256) sparse_doubles_kernel(SparseDblArgs A) {
  extern __shared__ __align__(16) unsigned char ssm[];
  const int N = A.N, N2 = N * (N + 1) / 2, t = threadIdx.x, nt = blockDim.x, g = blockIdx.x;
  const int64_t row = A.rows[g];
  const int ne = A.cnt[row];
  float4 *elL = reinterpret_cast<float4 *>(ssm);   // [N*N]
  float2 *fL = reinterpret_cast<float2 *>(elL + N * N);  // [N2]
  float2 *vL = fL + N2;                            // [N2]
  float *coL = reinterpret_cast<float *>(vL + N2); // [N]
  float *siL = coL + N;                            // [N]
  int *cpL = reinterpret_cast<int *>(siL + N);     // [N + 1]
  int *ciL = cpL + N + 1;                          // [N*N]
  uint32_t *pq = reinterpret_cast<uint32_t *>(ciL + N * N);  // [N2] (p | q << 16), p <= q
  for (int e = t; e < ne; e += nt) elL[e] = A.el[row * N * N + e], ciL[e] = A.colidx[row * N * N + e];
  for (int s = t; s <= N; s += nt) cpL[s] = A.colptr[row * (N + 1) + s];
  for (int a = t; a < N; a += nt) {
    if (!A.es) {  // src/propagate.c:1325-1335
      const float d = A.anh != 0.f ? A.anh : A.anh_rec[(A.f_lo + row) * (1 + (int64_t)N) + 1 + a];
      const float x = __fmul_rn(A.fm, d);
      coL[a] = (float)cos((double)x), siL[a] = (float)sin((double)x);
    }
    const int r0 = (a * (2 * N - a - 1)) >> 1;
    for (int b = a; b < N; ++b) pq[b + r0] = (uint32_t)a | ((uint32_t)b << 16);
  }
  for (int v = A.vec_begin[g]; v < A.vec_begin[g + 1]; ++v) {
    float2 *fg = A.state + (int64_t)v * N2;
    __syncthreads();  // the previous vector's write-back has read fL
    for (int j = t; j < N2; j += nt) fL[j] = fg[j];
    for (int sub = 0; sub < A.m; ++sub) {
      __syncthreads();
      if (!A.es)
        for (int a = t; a < N; a += nt) {
          const int k = sidx(a, a, N);
          fL[k] = anh_phase(fL[k], coL[a], siL[a]);
        }
      __syncthreads();
      for (int j = t; j < N2; j += nt) {
        const int p = (int)(pq[j] & 0xffffu), q = (int)(pq[j] >> 16);
        float2 acc = make_float2(0.f, 0.f);
        if (p == q) {
          for (int i1 = cpL[p]; i1 < cpL[p + 1]; ++i1) {
            const float4 e1 = elL[ciL[i1]];
            for (int i2 = cpL[p]; i2 <= i1; ++i2) sparse_term(acc, e1, elL[ciL[i2]], fL, N, A.es);
          }
        } else {
          int i = cpL[p], jq = cpL[q];
          const int ip = cpL[p + 1], iq = cpL[q + 1];
          while (i < ip || jq < iq) {
            if (jq >= iq || (i < ip && ciL[i] < ciL[jq])) {  // e1 in column p: e2 in column q before it
              const float4 e1 = elL[ciL[i]];
              for (int k = cpL[q]; k < jq; ++k) sparse_term(acc, e1, elL[ciL[k]], fL, N, A.es);
              ++i;
            } else {  // e1 in column q: e2 in column p before it
              const float4 e1 = elL[ciL[jq]];
              for (int k = cpL[p]; k < i; ++k) sparse_term(acc, e1, elL[ciL[k]], fL, N, A.es);
              ++jq;
            }
          }
        }
        vL[j] = acc;
      }
      __syncthreads();
      for (int j = t; j < N2; j += nt) {
        const int p = (int)(pq[j] & 0xffffu), q = (int)(pq[j] >> 16);
        fL[j] = (p == q && !A.es) ? anh_phase(vL[j], coL[p], siL[p]) : vL[j];
      }
    }
    __syncthreads();
    for (int j = t; j < N2; j += nt) fg[j] = fL[j];
  }
}

size_t sparse_doubles_smem(int N) {
  const size_t N2 = (size_t)N * (N + 1) / 2;
  return (size_t)N * N * 16 + 2 * N2 * 8 + 2 * (size_t)N * 4 + (size_t)(N + 1) * 4 + (size_t)N * N * 4 + N2 * 4;
}

int launch_sparse_doubles(const SparseDblArgs &a, int ngroups, hipStream_t st) {
  if (ngroups == 0) return NISE_OK;
  const size_t smem = sparse_doubles_smem(a.N);
  if (smem > 160 * 1024) return set_error("sparse doubles: N=%d too large for LDS", a.N), NISE_ERR_ARG;
  NISE_HIP_CHECK(hipFuncSetAttribute((const void *)sparse_doubles_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(sparse_doubles_kernel, dim3((unsigned)ngroups), dim3(256), smem, st, a);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

}  // namespace nise
