// Adjoint (reverse-mode) gradient of the batched linear absorption response
// with respect to per-site static energy offsets delta_a and a global scale
// c of the kept couplings, for gfx950.
//
// Forward (src/absorption.c:92-146 with the parameters applied to each
// frame): psi(0) = mu_x(ti), S(t) += mu_x(ti+t).psi(t), psi(t+1) = U_t psi(t),
// U_t = [D R_K ... R_1 D]^m, D = exp(-i f h/2), R_k = exp(-i f c J_k sx).
// Given g(t) = dL/dRe S(t) + i dL/dIm S(t), the adjoint obeys
//   lambda(T-1) = g(T-1) mu(T-1),  lambda(t) = U_t^dagger lambda(t+1) + g(t) mu(t)
// and dL/dtheta = sum_t Re <lambda(t+1), (dU_t/dtheta) psi(t)>.  Walking one
// step backwards from (phi, chi) = (psi(t+1), lambda(t+1)):
//   D:   dL/d delta_a += (f/2) Im(conj(chi_a) phi_a); phi, chi <- conj(D) .
//   R_k: dL/dc       += f J_k Im(conj(chi_a) phi_b + conj(chi_b) phi_a);
//        phi, chi <- R_k^dagger .   (rotations in reverse order)
//   D:   as above, giving (psi(t), U_t^dagger lambda(t+1)).
// The forward states are rebuilt by the inverse steps (U_t^-1 = U_t^dagger)
// and replaced by the forward checkpoint psi(rel) whenever rel is a multiple
// of K, so the rounding drift of the inversion spans at most K steps.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nise_internal.h"

namespace nise {


__device__ __forceinline__ float2 cmul_conj_d(float2 p, float2 d) {
  // p * conj(d), d = (cos, -sin): the inverse of the reference's
  // (cr*re - ci*im, ci*re + cr*im) product with d
  return make_float2(__fadd_rn(__fmul_rn(p.x, d.x), __fmul_rn(p.y, d.y)),
                     __fsub_rn(__fmul_rn(p.y, d.x), __fmul_rn(p.x, d.y)));
}

__device__ __forceinline__ void rotate_back(float2 &pa, float2 &pb, float co, float si) {
  // R^dagger: the forward rotation of src/propagate.c:883-887 with -si
  const float ar = __fadd_rn(__fmul_rn(co, pa.x), __fmul_rn(si, pb.y));
  const float ai = __fsub_rn(__fmul_rn(co, pa.y), __fmul_rn(si, pb.x));
  const float br = __fadd_rn(__fmul_rn(co, pb.x), __fmul_rn(si, pa.y));
  const float bi = __fsub_rn(__fmul_rn(co, pb.y), __fmul_rn(si, pa.x));
  pa = make_float2(ar, ai);
  pb = make_float2(br, bi);
}

// Im(conj(x) y)
__device__ __forceinline__ float im_cj(float2 x, float2 y) { return x.x * y.y - x.y * y.x; }

// One wave per workgroup: VP vectors, 64/VP rotations of a layer at a time.
// LDS operations of one wave execute in order, so layers need no barrier.
template <int VP>
__global__ void __launch_bounds__(64) adjoint_kernel(AdjArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int R = 64 / VP;
  const int N = A.N, T = A.T;
  double *gd = reinterpret_cast<double *>(smem);                      // [N]
  float2 *phi = reinterpret_cast<float2 *>(smem + (size_t)N * 8);     // [N][VP]
  float2 *chi = phi + (size_t)N * VP;                                 // [N][VP]
  float2 *Dsh = chi + (size_t)N * VP;                                 // [N]
  float2 *gsh = Dsh + N;                                              // [T]
  float *mush = reinterpret_cast<float *>(gsh + T);                   // [3N]
  const int g = blockIdx.x, t = threadIdx.x, v = t % VP, q = t / VP;
  const int s0 = A.wg_begin[g], ns = A.wg_begin[g + 1] - s0;
  for (int i = t; i < N; i += 64) gd[i] = 0.0;
  for (int i = t; i < T; i += 64) gsh[i] = A.g[i];
  const bool has = v < ns;
  const SlotDesc sd = has ? A.slots[s0 + v] : SlotDesc{0, 0, 0, 0, 0, -1, 0, -1};
  int64_t F0 = INT64_MAX, F1 = INT64_MIN;
  for (int k = 0; k < ns; ++k) {
    const int64_t fs = A.slots[s0 + k].fs;
    F0 = fs < F0 ? fs : F0;
    F1 = fs + T > F1 ? fs + T : F1;
  }
  const float hf = 0.5f * A.fac;
  double cg = 0.0;
  // per-site offset gradient: sum over the VP slots of a site, added by the
  // group's v == 0 lane (sites of one q are disjoint from other q's)
  auto add_site = [&](int i, float c) {
#pragma unroll
    for (int off = 1; off < VP; off <<= 1) c += __shfl_xor(c, off);
    if (v == 0) gd[i] += (double)c;
  };
  for (int64_t f = F1 - 1; f >= F0; --f) {
    const int64_t row = f - A.f_lo;
    const int64_t rel = f - sd.fs;
    const bool init = has && rel == T - 1;
    const bool step = has && rel >= 0 && rel < T - 1;
    // frame f's phases and dipoles (one wave: the LDS writes are seen by the
    // reads below in program order)
    const float *mu = A.mu_rec + (A.mu_stride ? f * A.mu_stride : 0) + 1;
    for (int i = t; i < N; i += 64) Dsh[i] = A.diag[row * N + i];
    for (int i = t; i < 3 * N; i += 64) mush[i] = mu[i];
    const float *mx = mush + sd.x * N;
    const float2 gv = (has && rel >= 0 && rel < T) ? gsh[rel] : make_float2(0.f, 0.f);
    const bool reset = step && (rel % A.every == 0);
    const int64_t ckrow = (int64_t)sd.hist + rel / A.every;
    if (init) {
      const int64_t last = (int64_t)sd.hist + A.nck - 1;
      for (int i = q; i < N; i += R) {
        phi[i * VP + v] = A.ck[last * N + i];
        chi[i * VP + v] = make_float2(__fmul_rn(gv.x, mx[i]), __fmul_rn(gv.y, mx[i]));
      }
    }
    for (int sub = 0; sub < A.m; ++sub) {
      // second half-phase of the substep
      for (int i = q; i < N; i += R) {
        float c = 0.f;
        if (step) {
          float2 p = phi[i * VP + v], x = chi[i * VP + v];
          c = hf * im_cj(x, p);
          const float2 d = Dsh[i];
          phi[i * VP + v] = cmul_conj_d(p, d);
          chi[i * VP + v] = cmul_conj_d(x, d);
        }
        add_site(i, c);
      }
      // rotations in reverse order
      float cf = 0.f;
      for (int s = A.nsteps - 1; s >= 0; --s) {
        const int2 st = A.steps[s];
        const int r = st.x + q;
        const uint32_t pk = (step && r < st.y) ? A.ab[r] : 0u;
        const int a = (int)(pk & 0xffffu), b = (int)(pk >> 16);
        if (step && r < st.y && a < N) {  // a >= N: row-schedule padding
          const float2 c = A.cs[row * A.K + r];
          const float j = A.jf[row * A.K + r];
          float2 pa = phi[a * VP + v], pb = phi[b * VP + v];
          float2 xa = chi[a * VP + v], xb = chi[b * VP + v];
          cf += j * (im_cj(xa, pb) + im_cj(xb, pa));
          rotate_back(pa, pb, c.x, c.y);
          rotate_back(xa, xb, c.x, c.y);
          phi[a * VP + v] = pa, phi[b * VP + v] = pb;
          chi[a * VP + v] = xa, chi[b * VP + v] = xb;
        }
      }
      cg += (double)cf;
      // first half-phase, then (last substep) the response term of psi(rel)
      const bool fin = sub == A.m - 1;
      for (int i = q; i < N; i += R) {
        float c = 0.f;
        if (step) {
          float2 p = phi[i * VP + v], x = chi[i * VP + v];
          c = hf * im_cj(x, p);
          const float2 d = Dsh[i];
          p = cmul_conj_d(p, d);
          x = cmul_conj_d(x, d);
          if (fin) {
            x.x = __fadd_rn(x.x, __fmul_rn(gv.x, mx[i]));
            x.y = __fadd_rn(x.y, __fmul_rn(gv.y, mx[i]));
            if (reset) p = A.ck[ckrow * N + i];
          }
          phi[i * VP + v] = p;
          chi[i * VP + v] = x;
        }
        add_site(i, c);
      }
    }
  }
  // per-workgroup partials (fixed order)
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) cg += __shfl_xor(cg, off);
  for (int i = t; i < N; i += 64) A.part_delta[(int64_t)g * N + i] = gd[i];
  if (t == 0) A.part_c[g] = cg;
}

__global__ void adjoint_reduce_kernel(const double *pd, const double *pc, int nwg, int N,
                                      double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    double s = 0.0;
    for (int g = 0; g < nwg; ++g) s += pd[(int64_t)g * N + i];
    out[i] = s;
  } else if (i == N) {
    double s = 0.0;
    for (int g = 0; g < nwg; ++g) s += pc[g];
    out[N] = s;
  }
}

// Row variant (used with the row conveyor's schedule and tables): one wave
// holds G = 64/R vectors, lane q*G + g = (row q, vector g), each vector at
// its own frame (time skew, as in rows_kernel), walking t = tmax-1 .. 0 in
// lock step.  Rotation entries are the row schedule's (padding pairs act on
// the zeroed scratch sites N..N+R-1 with jf = 0).  Per-site offset
// gradients are summed over the wave's vectors with shuffles and added to an
// FP64 LDS accumulator by the g == 0 lane of each row.
template <int R>
__global__ void __launch_bounds__(64) adjoint_rows_kernel(AdjRowsArgs A) {
  constexpr int G = 64 / R;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, T = A.T, nl = A.nlayers, NS = N + R;
  double *gd = reinterpret_cast<double *>(smem);                  // [N]
  float2 *phi = reinterpret_cast<float2 *>(gd + N);                // [N+R][G]
  float2 *chi = phi + (size_t)NS * G;                              // [N+R][G]
  const int lane = threadIdx.x, q = lane / G, g = lane % G;
  const int v = blockIdx.x * G + g;
  const bool has = v < A.nvec;
  const int vv = has ? v : blockIdx.x * G;
  const int s = vv / 3, x = vv % 3;
  const int64_t fs = (int64_t)(A.begin + s) * A.sample;
  const int64_t ck0 = (int64_t)vv * A.nck;
  const float hf = 0.5f * A.fac;
  for (int i = lane; i < N; i += 64) gd[i] = 0.0;
  for (int i = lane; i < NS * G; i += 64) phi[i] = chi[i] = make_float2(0.f, 0.f);
  double cg = 0.0;
  auto site_pass = [&](const float2 *D, int t, bool fin) {
    // half-phase conj(D) on phi and chi with the offset gradient; on the
    // last substep also the response term g(t) mu(t) and the checkpoint
    const float2 gv = A.g[t];
    const float *mu = A.mu_rec + (A.mu_stride ? (fs + t) * A.mu_stride : 0) + 1 + (int64_t)x * N;
    const bool reset = fin && (t % A.every == 0);
    const int64_t ckr = ck0 + t / A.every;
    for (int i = q; i < N; i += R) {
      float2 p = phi[i * G + g], c = chi[i * G + g];
      float cc = hf * im_cj(c, p);
      const float2 d = D[i];
      p = cmul_conj_d(p, d);
      c = cmul_conj_d(c, d);
      if (fin) {
        const float m = mu[i];
        c.x = __fadd_rn(c.x, __fmul_rn(gv.x, m));
        c.y = __fadd_rn(c.y, __fmul_rn(gv.y, m));
        if (reset) p = A.ck[ckr * N + i];
      }
      phi[i * G + g] = p;
      chi[i * G + g] = c;
      if (!has) cc = 0.f;
#pragma unroll
      for (int off = 1; off < G; off <<= 1) cc += __shfl_xor(cc, off);
      if (g == 0) gd[i] += (double)cc;
    }
  };
  for (int t = T - 1; t >= 0; --t) {
    const int64_t row = fs + t - A.f_lo;
    if (t == T - 1) {  // start: psi(T-1) and lambda(T-1) = g(T-1) mu(T-1)
      const float2 gv = A.g[t];
      const float *mu = A.mu_rec + (A.mu_stride ? (fs + t) * A.mu_stride : 0) + 1 + (int64_t)x * N;
      for (int i = q; i < N; i += R) {
        phi[i * G + g] = A.ck[(ck0 + A.nck - 1) * N + i];
        chi[i * G + g] = make_float2(__fmul_rn(gv.x, mu[i]), __fmul_rn(gv.y, mu[i]));
      }
      continue;
    }
    const float2 *D = A.diag + row * N;
    // entry q of layer l in device order (rows_device_order)
    const float2 *cs = A.cs + row * A.K + 4 * q;
    const float *jf = A.jf + row * A.K + 4 * q;
    const uint2 *stq = A.steps + 4 * q;
    for (int sub = 0; sub < A.m; ++sub) {
      site_pass(D, t, false);
      float cf = 0.f;
      for (int l = nl - 1; l >= 0; --l) {
        const int le = (l >> 2) * 4 * R + (l & 3);
        const uint2 e = stq[le];
        const float2 c = cs[le];
        const float j = jf[le];
        float2 *pa_ = reinterpret_cast<float2 *>(reinterpret_cast<unsigned char *>(phi) + e.x + g * 8);
        float2 *pb_ = reinterpret_cast<float2 *>(reinterpret_cast<unsigned char *>(phi) + e.y + g * 8);
        float2 *xa_ = reinterpret_cast<float2 *>(reinterpret_cast<unsigned char *>(chi) + e.x + g * 8);
        float2 *xb_ = reinterpret_cast<float2 *>(reinterpret_cast<unsigned char *>(chi) + e.y + g * 8);
        float2 pa = *pa_, pb = *pb_, xa = *xa_, xb = *xb_;
        cf += j * (im_cj(xa, pb) + im_cj(xb, pa));
        rotate_back(pa, pb, c.x, c.y);
        rotate_back(xa, xb, c.x, c.y);
        *pa_ = pa, *pb_ = pb, *xa_ = xa, *xb_ = xb;
      }
      cg += has ? (double)cf : 0.0;
      site_pass(D, t, sub == A.m - 1);
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) cg += __shfl_xor(cg, off);
  for (int i = lane; i < N; i += 64) A.part_delta[(int64_t)blockIdx.x * N + i] = gd[i];
  if (lane == 0) A.part_c[blockIdx.x] = cg;
}

size_t adjoint_rows_smem(int N, int R) { return (size_t)N * 8 + (size_t)2 * (N + R) * (64 / R) * 8; }

int launch_adjoint_rows(const AdjRowsArgs &a, int R, double *out, hipStream_t st) {
  const size_t smem = adjoint_rows_smem(a.N, R);
  if (smem > 160 * 1024) return set_error("adjoint: N=%d needs %zu B of LDS", a.N, smem), NISE_ERR_ARG;
  const void *fn = R == 1 ? (const void *)adjoint_rows_kernel<1>
                 : R == 2 ? (const void *)adjoint_rows_kernel<2>
                 : R == 4 ? (const void *)adjoint_rows_kernel<4>
                 : R == 8 ? (const void *)adjoint_rows_kernel<8>
                 : R == 16 ? (const void *)adjoint_rows_kernel<16> : nullptr;
  if (!fn) return set_error("adjoint: unsupported rows %d", R), NISE_ERR_ARG;
  const int nwg = (a.nvec + 64 / R - 1) / (64 / R);
  if (nwg > 0) {
    NISE_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    AdjRowsArgs args = a;
    void *kargs[] = {&args};
    NISE_HIP_CHECK(hipLaunchKernel(fn, dim3(nwg), dim3(64), kargs, smem, st));
  }
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((a.N + 1 + 127) / 128), dim3(128), 0, st,
                     a.part_delta, a.part_c, nwg, a.N, out);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

size_t adjoint_smem(int N, int vp, int T) {
  return (size_t)N * 8 + (size_t)2 * N * vp * 8 + (size_t)N * 8 + (size_t)T * 8 + (size_t)3 * N * 4;
}

int launch_adjoint(const AdjArgs &a, int vp, double *out, hipStream_t st) {
  const size_t smem = adjoint_smem(a.N, vp, a.T);
  if (smem > 160 * 1024)
    return set_error("adjoint: N=%d with %d vectors per workgroup needs %zu B of LDS", a.N, vp, smem),
           NISE_ERR_ARG;
  const void *fn = vp == 4 ? (const void *)adjoint_kernel<4>
                 : vp == 8 ? (const void *)adjoint_kernel<8>
                 : vp == 16 ? (const void *)adjoint_kernel<16>
                 : vp == 32 ? (const void *)adjoint_kernel<32>
                 : vp == 64 ? (const void *)adjoint_kernel<64> : nullptr;
  if (!fn) return set_error("adjoint: unsupported vp=%d", vp), NISE_ERR_ARG;
  if (a.nwg > 0) {
    NISE_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    AdjArgs args = a;
    void *kargs[] = {&args};
    NISE_HIP_CHECK(hipLaunchKernel(fn, dim3(a.nwg), dim3(64), kargs, smem, st));
  }
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((a.N + 1 + 127) / 128), dim3(128), 0, st,
                     a.part_delta, a.part_c, a.nwg, a.N, out);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

// Vectors per workgroup: one rotation lane per rotation of an average layer,
// the rest of the wave's lanes for vectors, within the LDS budget.
int adjoint_vp(const Schedule &s, int N, int T) {
  const double per_layer = s.nlayers() ? (double)s.K / s.nlayers() : 1.0;
  int vp = 64;
  while (vp > 4 && 64 / vp < per_layer) vp >>= 1;
  while (vp > 4 && adjoint_smem(N, vp, T) > 80 * 1024) vp >>= 1;
  return vp;
}

// Steps of at most 64/vp rotations, each inside one layer.
void adjoint_steps(const Schedule &s, int vp, std::vector<int2> &steps) {
  const int R = 64 / vp;
  steps.clear();
  for (int l = 0; l < s.nlayers(); ++l)
    for (int r = s.layer_off[l]; r < s.layer_off[l + 1]; r += R)
      steps.push_back(make_int2(r, std::min(r + R, s.layer_off[l + 1])));
}

}  // namespace nise
