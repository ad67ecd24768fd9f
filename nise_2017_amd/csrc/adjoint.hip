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
#include <climits>
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
// holds G = 64/R vectors, lane q*G + g = (row q, vector g).  As in the
// forward row conveyor, each quad of vectors (g / 4) is frame-synchronous
// and quads are time-skewed: at wave time tau quad h is at frame
// Fmax(h) - tau (Fmax: its last vector's last frame), vector g at relative
// time t = frame - fs_g.  A vector starts (t = tmax-1) from its last
// checkpoint and lambda = g mu, then steps back while t >= 0; lanes of
// vectors outside their window are rotated too but contribute nothing.
// Rotation entries (device order) are loaded one per lane per 4-layer block
// with a quad broadcast, through a register ring walking the layers
// backwards.  Padding pairs act on the scratch sites N..N+R-1 with jf = 0.
// Per-site offset gradients are summed over the wave's vectors with
// shuffles and added to an FP64 LDS accumulator by the g == 0 lane of each
// row.
constexpr int kAdjRingG = 2;              // ring slots per lane, 4 layers each
// Sites per batch of HBM loads in the site pass, whose gradient butterflies
// also run interleaved: a lane walks N / R sites per frame, so waves with
// few rows take larger batches.
template <int R>
constexpr int adj_site_batch() {
  return R >= 4 ? 4 : R == 2 ? 8 : 16;
}

// Layer j of a ring block, backwards (src/propagate.c:883-887 inverted):
// the coupling-scale gradient term and R^dagger on phi and chi.
template <int j>
__device__ __forceinline__ void adj_slot(unsigned char *phib, unsigned char *chib, uint32_t gofs, uint2 e,
                                         float2 c, float jv, float &cf) {
  const uint32_t ea = quad_bcast<j>(e.x) + gofs, eb = quad_bcast<j>(e.y) + gofs;
  const float co = quad_bcast<j>(c.x), si = quad_bcast<j>(c.y), jj = quad_bcast<j>(jv);
  float2 *pa_ = reinterpret_cast<float2 *>(phib + ea), *pb_ = reinterpret_cast<float2 *>(phib + eb);
  float2 *xa_ = reinterpret_cast<float2 *>(chib + ea), *xb_ = reinterpret_cast<float2 *>(chib + eb);
  float2 pa = *pa_, pb = *pb_, xa = *xa_, xb = *xb_;
  cf += jj * (im_cj(xa, pb) + im_cj(xb, pa));
  rotate_back(pa, pb, co, si);
  rotate_back(xa, xb, co, si);
  *pa_ = pa, *pb_ = pb, *xa_ = xa, *xb_ = xb;
}

template <int R>
__global__ void __launch_bounds__(64) adjoint_rows_kernel(AdjRowsArgs A) {
  constexpr int G = 64 / R;
  static_assert(G >= 4, "a row needs a whole quad of lanes");
  constexpr int kAdjSiteU = adj_site_batch<R>();
  // Site pass without per-site branches for one and two rows (a lane walks
  // N / R sites per frame; the N = 128 chain's adjoint ran 15.4 ms with the
  // branches); wider rows keep the branches, which take fewer registers
  // (at 4-16 rows 140-162 VGPRs against 177-198).
  constexpr bool kAdjSelect = R <= 2;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, T = A.T, nl = A.nlayers, NS = N + R, m = A.m;
  const int64_t K = A.K;
  double *gd = reinterpret_cast<double *>(smem);                  // [N]
  float2 *phi = reinterpret_cast<float2 *>(gd + N);                // [N+R][G]
  float2 *chi = phi + (size_t)NS * G;                              // [N+R][G]
  unsigned char *phib = reinterpret_cast<unsigned char *>(phi), *chib = reinterpret_cast<unsigned char *>(chi);
  const int lane = threadIdx.x, q = lane / G, g = lane % G;
  const int v0 = blockIdx.x * G, v = v0 + g;
  const bool has = v < A.nvec;
  const int x = has ? v % 3 : 0;
  auto start_of = [&](int vec) { return (int64_t)(A.begin + vec / 3) * A.sample; };
  const int64_t fs = has ? start_of(v) : LLONG_MAX / 4;
  const int64_t ck0 = (int64_t)v * A.nck;
  const uint32_t gofs = (uint32_t)g * 8u;
  const float hf = 0.5f * A.fac;
  // quads: frames run from Fmax(h) down to F0(h)
  const int vlast = min(v0 + G, A.nvec) - 1;
  int64_t span = 0;
  for (int h = 0; 4 * h < G && v0 + 4 * h <= vlast; ++h)
    span = max(span, start_of(min(v0 + 4 * h + 3, vlast)) + T - start_of(v0 + 4 * h));
  const int vq = min(v0 + (g & ~3), vlast);
  const int64_t F0 = start_of(vq), Fmax = start_of(min(vq + 3, vlast)) + T - 1;

  for (int i = lane; i < N; i += 64) gd[i] = 0.0;
  for (int i = lane; i < NS * G; i += 64) phi[i] = chi[i] = make_float2(0.f, 0.f);
  double cg = 0.0;

  // Site pass before sweep (tau, sub), per element of the lane's vector:
  //   B: the closing half-phase conj(D) of the previous sweep with its
  //      offset gradient -- of frame f+1 (then also lambda += g mu and the
  //      checkpoint reset) when sub == 0 and the vector stepped at t+1, of
  //      frame f when sub > 0 and it steps at t;
  //   A: the opening half-phase conj(D(f)) of this sweep with its offset
  //      gradient, when the vector steps at t.
  // Each phase's gradient term uses (phi, chi) just before the phase.  A
  // vector's start (t = tmax-1) is written after the frame's sweeps, which
  // rotate every column.
  auto site_pass = [&](int64_t t, int sub, int64_t row, int64_t f) {
    const bool step = t >= 0 && t < T - 1;
    const bool bflag = sub == 0 ? (t + 1 >= 0 && t + 1 < T - 1) : step;
    const bool bfin = sub == 0 && bflag;
    const int64_t tb = t + 1;
    const bool reset = bfin && (tb % A.every == 0);
    // response and checkpoint of the closing phase of step t+1
    const float2 gv = bfin ? gld(A.g + tb) : make_float2(0.f, 0.f);
    // (f + 1 clamped to the quad's frames F0..Fmax: lanes outside their
    // window load valid rows they do not use)
    const int64_t fmu = max(min(f + 1, Fmax), F0);
    const float *mu = A.mu_rec + (A.mu_stride ? fmu * A.mu_stride : 0) + 1 + (int64_t)x * N;
    const int64_t ckr = ck0 + (reset ? tb / A.every : 0);
    const float2 *Da = A.diag + row * N;
    const float2 *Db = sub == 0 ? A.diag + (max(min(f + 1, Fmax), F0) - A.f_lo) * N : Da;
    for (int i0 = q; i0 < N; i0 += kAdjSiteU * R) {
      float2 da[kAdjSiteU], db[kAdjSiteU], cv[kAdjSiteU];
      float mv[kAdjSiteU];
#pragma unroll
      for (int u = 0; u < kAdjSiteU; ++u) {
        const int i = i0 + u * R, ii = i < N ? i : 0;
        da[u] = gld(Da + ii);
        db[u] = gld(Db + ii);
        mv[u] = gld(mu + ii);
        cv[u] = reset ? gld(A.ck + ckr * N + ii) : make_float2(0.f, 0.f);  // padding lanes have no rows
      }
      __builtin_amdgcn_sched_barrier(0);
      float ccu[kAdjSiteU];
      if constexpr (kAdjSelect) {
        // no branch per site: both phases are computed and selected, so the
        // LDS reads of a batch are in flight together (the selected values
        // are those of the branches: (0 + x) and x agree, -0 included)
#pragma unroll
        for (int u = 0; u < kAdjSiteU; ++u) {
          const int i = i0 + u * R;
          const bool ok = i < N;
          const int is = ok ? i : q;
          const float2 p0 = phi[is * G + g], c0 = chi[is * G + g];
          // B: closing phase of the previous sweep
          const float ccb = hf * im_cj(c0, p0);
          const float2 pb = cmul_conj_d(p0, db[u]);
          float2 cb = cmul_conj_d(c0, db[u]);
          const float2 cbf = make_float2(__fadd_rn(cb.x, __fmul_rn(gv.x, mv[u])), __fadd_rn(cb.y, __fmul_rn(gv.y, mv[u])));
          cb = bfin ? cbf : cb;
          const float2 p1 = bflag ? (reset ? cv[u] : pb) : p0;
          const float2 c1 = bflag ? cb : c0;
          const float cc1 = bflag ? ccb : 0.f;
          // A: opening phase of this sweep
          const float cca = hf * im_cj(c1, p1);
          const float2 pa = cmul_conj_d(p1, da[u]), ca = cmul_conj_d(c1, da[u]);
          const float2 p = step ? pa : p1, c = step ? ca : c1;
          ccu[u] = step ? cc1 + cca : cc1;
          if (ok) phi[is * G + g] = p, chi[is * G + g] = c;
        }
      } else {  // per-site branches
#pragma unroll
        for (int u = 0; u < kAdjSiteU; ++u) {
          const int i = i0 + u * R;
          const bool ok = i < N;
          const int is = ok ? i : q;
          float2 p = phi[is * G + g], c = chi[is * G + g];
          float cc = 0.f;
          if (bflag) {
            cc = hf * im_cj(c, p);
            p = cmul_conj_d(p, db[u]);
            c = cmul_conj_d(c, db[u]);
            if (bfin) {
              c.x = __fadd_rn(c.x, __fmul_rn(gv.x, mv[u]));
              c.y = __fadd_rn(c.y, __fmul_rn(gv.y, mv[u]));
              if (reset) p = cv[u];
            }
          }
          if (step) {
            cc += hf * im_cj(c, p);
            p = cmul_conj_d(p, da[u]);
            c = cmul_conj_d(c, da[u]);
          }
          if (ok) phi[is * G + g] = p, chi[is * G + g] = c;
          ccu[u] = cc;
        }
      }
      // each site's sum over the wave's vectors, the butterflies of the
      // batch's sites interleaved (each site's own order unchanged)
#pragma unroll
      for (int off = 1; off < G; off <<= 1)
#pragma unroll
        for (int u = 0; u < kAdjSiteU; ++u) ccu[u] += __shfl_xor(ccu[u], off);
#pragma unroll
      for (int u = 0; u < kAdjSiteU; ++u) {
        const int i = i0 + u * R;
        if (i < N && g == 0) gd[i] += (double)ccu[u];
      }
    }
  };

  const uint32_t loff = (uint32_t)(4 * q + (lane & 3));
  const int nb = nl / 4;  // 4-layer blocks (a multiple of kAdjRingG)
  uint2 re[kAdjRingG];
  float2 rc[kAdjRingG];
  float rj[kAdjRingG];
  int64_t ring_key = -1;  // sweep (tau * m + sub) whose top blocks the ring holds

  // tau runs one past the longest span: the last closing half-phase
  for (int64_t tau = 0; tau <= span; ++tau) {
    const int64_t f = Fmax - tau;  // this lane's frame
    const int64_t t = f - fs;
    const bool live = t >= -1 && t <= T - 1;  // start, step, or the closing B of step 0
    if (!__ballot(live)) {  // no vector of the wave is in its window: jump
      int64_t nt = span + 1;
      for (int k = 0; k < G && v0 + k <= vlast; ++k) {
        const int64_t fq = start_of(min(v0 + (k & ~3) + 3, vlast)) + T - 1;  // quad's Fmax
        const int64_t o = fq - (start_of(v0 + k) + T - 1);
        if (o > tau) nt = min(nt, o);
      }
      tau = nt - 1;
      continue;
    }
    const bool step = t >= 0 && t < T - 1;
    const int64_t row = max(f, F0) - A.f_lo;  // clamped for finished quads
    const bool sweep = __ballot(step) != 0;
    for (int sub = 0; sub < (sweep ? m : 1); ++sub) {
      site_pass(t, sub, row, f);
      if (!sweep) break;
      const float2 *csf = A.cs + row * K + loff;
      const float *jff = A.jf + row * K + loff;
      if (ring_key != tau * m + sub) {
#pragma unroll
        for (int k = 0; k < kAdjRingG; ++k) {
          const int64_t o = (int64_t)(nb - 1 - k) * 4 * R;
          re[k] = gld(A.steps + loff + o), rc[k] = gld(csf + o), rj[k] = gld(jff + o);
        }
      }
      float cf = 0.f;
      // blocks nb-1 .. 0; the last group refills from the top blocks of
      // the next sweep (this frame again, or the frame before)
      const int64_t nrow = sub + 1 < m ? row : max(f - 1, F0) - A.f_lo;
      for (int b0 = nb - 1; b0 >= 0; b0 -= kAdjRingG) {
        const bool tail = b0 - kAdjRingG < 0;
#pragma unroll
        for (int k = 0; k < kAdjRingG; ++k) {
          adj_slot<3>(phib, chib, gofs, re[k], rc[k], rj[k], cf);
          adj_slot<2>(phib, chib, gofs, re[k], rc[k], rj[k], cf);
          adj_slot<1>(phib, chib, gofs, re[k], rc[k], rj[k], cf);
          adj_slot<0>(phib, chib, gofs, re[k], rc[k], rj[k], cf);
          const int64_t o = (int64_t)(tail ? nb - 1 - k : b0 - kAdjRingG - k) * 4 * R;
          const int64_t r2 = tail ? nrow : row;
          re[k] = gld(A.steps + loff + o);
          rc[k] = gld(A.cs + r2 * K + loff + o);
          rj[k] = gld(A.jf + r2 * K + loff + o);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      ring_key = tau * m + sub + 1;
      cg += step ? (double)cf : 0.0;
    }
    if (t == T - 1) {  // start: psi(T-1) and lambda(T-1) = g(T-1) mu(T-1)
      const float2 gv = gld(A.g + (T - 1));
      const float *mu = A.mu_rec + (A.mu_stride ? f * A.mu_stride : 0) + 1 + (int64_t)x * N;
      for (int i = q; i < N; i += R) {
        const float mm = gld(mu + i);
        phi[i * G + g] = gld(A.ck + (ck0 + A.nck - 1) * N + i);
        chi[i * G + g] = make_float2(__fmul_rn(gv.x, mm), __fmul_rn(gv.y, mm));
      }
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
