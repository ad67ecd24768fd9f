// Device pieces of the two-exciton (doubles) Trotter step shared by the
// doubles step kernel and the persistent 2DIR kernel.
// Reference: propagate_vec_coupling_S_doubles, src/propagate.c:902-1048, and
// the two-level form propagate_vec_coupling_S_doubles_ES, :1051-1181.
#pragma once

#include "nise_internal.h"

namespace nise {

__device__ __forceinline__ float2 cmul_diag2(float2 c, float2 u) {
  // src/propagate.c:960-961: (cr*re - ci*im, cr*im + ci*re)
  return make_float2(__fsub_rn(__fmul_rn(c.x, u.x), __fmul_rn(c.y, u.y)),
                     __fadd_rn(__fmul_rn(c.x, u.y), __fmul_rn(c.y, u.x)));
}

__device__ __forceinline__ void rot2(float2 &p1, float2 &p2, float co, float si) {
  // src/propagate.c:986-990 (same form for all five c ranges)
  float cr1 = __fsub_rn(__fmul_rn(co, p1.x), __fmul_rn(si, p2.y));
  float ci1 = __fadd_rn(__fmul_rn(co, p1.y), __fmul_rn(si, p2.x));
  float cr2 = __fsub_rn(__fmul_rn(co, p2.x), __fmul_rn(si, p1.y));
  float ci2 = __fadd_rn(__fmul_rn(co, p2.y), __fmul_rn(si, p1.x));
  p1 = make_float2(cr1, ci1);
  p2 = make_float2(cr2, ci2);
}

// Couplings staged in LDS per pass over the coupling list: (co, si, co2,
// si2) rows and pair words of at most kDblCh couplings.
constexpr int kDblCh = 512;
__host__ __device__ inline int dbl_chunk(int K2) {
  return K2 <= 64 ? 64 : K2 < kDblCh ? (K2 + 63) / 64 * 64 : kDblCh;
}
// Vectors per lane.  Two (one index computation for the same pair of
// states in two vectors) doubles the LDS image of a wave; the lower
// occupancy made the N = 32 step slower, so every shape uses one.
__host__ __device__ constexpr int dbl_vpl(int) { return 1; }
// The rotations of couplings [0, nk) (pair words abL, angle rows csL, both
// in LDS) on the states psi[j * VW + w] of the vectors w < nv, VW = 64 / NP
// * VPL: lane group q owns vectors q VPL .. q VPL + VPL - 1 and lane c0 the
// rotations of sites c = c0, c0 + NP, ... (the caller maps lanes to (q,
// c0)).  rs0 is the row start of |c0 c0> in the packed basis.  es: two-level
// sites, no rotation for c == a or c == b.
template <int NP, int VPL>
__device__ __forceinline__ void dbl_rotate(float2 *psi, int N, int nv, int nk, const uint32_t *abL,
                                           const float4 *csL, int es, int q, int c0, int rs0) {
  constexpr int VW = 64 / NP * VPL;
  for (int k = 0; k < nk; ++k) {
    // the coupling is wave-uniform: pair and angles in scalar registers
    const uint32_t p = __builtin_amdgcn_readfirstlane(abL[k]);
    const int a = (int)(p & 0xffffu), b = (int)(p >> 16);
    const int rs_a = (a * (2 * N - a - 1)) >> 1, rs_b = (b * (2 * N - b - 1)) >> 1;
    const float4 cv = csL[k];
    const float co = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cv.x)));
    const float si = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cv.y)));
    const float co2 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cv.z)));
    const float si2 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cv.w)));
    for (int cb = 0; cb < N; cb += NP) {
      const int c = cb + c0;
      // every c except b, ascending; c == a (sqrt(2) angle) also does the
      // c == b rotation (|ab>, |bb>): it shares |ab>, and no rotation of
      // a c in between touches |ab> or |bb>
      if (c < N && c != b && !(es && c == a)) {
        const int rs_c = cb == 0 ? rs0 : (c * (2 * N - c - 1)) >> 1;
        const int i1 = (c <= a ? a + rs_c : c + rs_a) * VW + q * VPL;
        const int i2 = (c <= b ? b + rs_c : c + rs_b) * VW + q * VPL;
        const int i3 = (b + rs_b) * VW + q * VPL;
        const bool ca = c == a;
        const float cl = ca ? co2 : co, sl = ca ? si2 : si;
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
          if (q * VPL + v < nv) {
            float2 p1 = psi[i1 + v], p2 = psi[i2 + v], p3 = make_float2(0.f, 0.f);
            if (ca) p3 = psi[i3 + v];  // read with the pair: one LDS round trip
            rot2(p1, p2, cl, sl);
            psi[i1 + v] = p1;
            if (ca) {
              rot2(p2, p3, co2, si2);
              psi[i3 + v] = p3;
            }
            psi[i2 + v] = p2;
          }
        }
      }
      asm volatile("" ::: "memory");
    }
  }
}

// Element k of dipole_double_last(dip, f) (src/NISE_subs.c:552-575) for NIT
// dipole sets at once, in the reference's order: sum over i < k of
// dip_i f_ik, then over_k f_kk, then dip_j f_kj for j > k, i.e. one pass over
// i = 0..N-1 with the weight over_k at i = k.  Each element f is read once
// for all NIT sets.  f_ik with i <= k sits at k + rs(i), rs(i) = i(2N-i-1)/2,
// and f_ki with i > k at rs(k) + i (src/NISE_subs.h:54-65).  Element j of
// f is read at f[j * FS].
template <int NIT, int FS = 1>
__device__ __forceinline__ void ddl_multi(const float2 *f, int N, int k, const float *dip0, const float *dip1,
                                 const float *dip2, const float ov[NIT], float2 out[NIT]) {
  const float *dip[3] = {dip0, dip1, dip2};
  float r[NIT], q[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) r[it] = 0.f, q[it] = 0.f;
  // state (i, k): index k + rs(i) for i <= k, rs(k) + i above, rs(i) =
  // i (2N - i - 1) / 2; stepped from i = 0 (index k)
  const char *fb = reinterpret_cast<const char *>(f);
  int ob = k * FS * 8;
  for (int i = 0; i < N; ++i) {
    const float2 v = *reinterpret_cast<const float2 *>(fb + ob);
    const bool d = i == k;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      // dip[i] is wave-uniform: read it once into a scalar register (a
      // convergent read, so it is not turned into a masked per-lane load)
      const float dv = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(dip[it][i])));
      const float w = d ? ov[it] : dv;
      r[it] = __fadd_rn(r[it], __fmul_rn(w, v.x)), q[it] = __fadd_rn(q[it], __fmul_rn(w, v.y));
    }
    ob += (i < k) ? (N - i - 1) * FS * 8 : FS * 8;
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) out[it] = make_float2(r[it], q[it]);
}

// ddl_multi for N % 4 == 0 with 16-byte aligned dipole rows in LDS: one
// broadcast 128-bit read per item covers four i, and the state address is
// stepped from i = 0 (index k; + N - i - 1 while i < k, + 1 after).
template <int NIT, int FS>
__device__ __forceinline__ void ddl_multi4(const float2 *f, int N, int k, const float *dip0, const float *dip1,
                                           const float *dip2, const float ov[NIT], float2 out[NIT]) {
  const float *dip[3] = {dip0, dip1, dip2};
  float r[NIT], q[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) r[it] = 0.f, q[it] = 0.f;
  const char *fb = reinterpret_cast<const char *>(f);
  int ob = k * FS * 8;
  for (int i0 = 0; i0 < N; i0 += 4) {
    float4 d4[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) d4[it] = *reinterpret_cast<const float4 *>(dip[it] + i0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u;
      const float2 v = *reinterpret_cast<const float2 *>(fb + ob);
      const bool d = i == k;
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const float dv = u == 0 ? d4[it].x : u == 1 ? d4[it].y : u == 2 ? d4[it].z : d4[it].w;
        const float w = d ? ov[it] : dv;
        r[it] = __fadd_rn(r[it], __fmul_rn(w, v.x)), q[it] = __fadd_rn(q[it], __fmul_rn(w, v.y));
      }
      ob += (i < k) ? (N - i - 1) * FS * 8 : FS * 8;
    }
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) out[it] = make_float2(r[it], q[it]);
}

}  // namespace nise
