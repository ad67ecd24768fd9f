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
// * VPL: lane group q = t / NP owns vectors q VPL .. q VPL + VPL - 1 and
// lane c0 = t % NP the rotations of sites c = c0, c0 + NP, ...  rs0 is the
// row start of |c0 c0> in the packed basis.  es: two-level sites, no
// rotation for c == a or c == b.
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

}  // namespace nise
