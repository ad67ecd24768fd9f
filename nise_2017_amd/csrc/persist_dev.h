// Device pieces of the persistent two-exciton kernels (twodir.hip,
// twodir_raman.hip): a wave holds VW = 64 / NP doubles vectors in LDS as
// psi[state][VW] for many frames and steps them with a per-lane rotation
// table built once per launch.  Lane (c0, q) = (lane / VW, lane % VW) is
// site c0 of a rotation and vector q.  A wave's states are followed by a
// zero state (N2) and a neutral state (N2 + 1).
#pragma once

#include "nise_internal.h"
#include "doubles_dev.h"

namespace nise {

// The rotations of one substep (as dbl_rotate), each lane's from its table
// entries (read one rotation ahead).  Coupling (a, b) rotates (|ca>, |cb>)
// for every c, and |ab> twice in the reference's order: with |aa> (c = a),
// then with |bb> (c = b), both with the sqrt2 angles.  When a and b fall in
// the same site block, the c = b lane does its rotation in the same round
// as the c = a lane: it first pre-rotates |ab> with |aa> itself, forming the
// c = a lane's |ab> output from the same inputs and operations, and the
// c = a lane's copy is overwritten: the rounds store the second state first.
// Every other lane's pre-rotation is an exact identity (coefficients (1, 0)
// on a neutral state (-0, +0): x - 0 * 0 and y + 0 * (-0) return x and y,
// signed zeros included).  In different blocks the c = a lane's block
// writes |ab> and the c = b lane's later block rotates it with |bb>.
// Coefficient rows per coupling: (1, 0, co, si) for the plain rotations,
// (1, 0, co2, si2) for the c = a lane and a c = b lane in another block,
// (co2, si2, co2, si2) for a c = b lane in the same block.
__device__ __forceinline__ void rot_entry(char *base, const char *cs, uint2 e) {
  float2 *a1 = reinterpret_cast<float2 *>(base + (e.x & 0xffffu));
  float2 *a2 = reinterpret_cast<float2 *>(base + (e.x >> 16));
  const float2 q = *reinterpret_cast<const float2 *>(base + (e.y & 0xffffu));
  const float4 cf = *reinterpret_cast<const float4 *>(cs + (e.y >> 16));
  float2 p1 = *a1, p2 = *a2;
  p1 = make_float2(__fsub_rn(__fmul_rn(cf.x, p1.x), __fmul_rn(cf.y, q.y)),
                   __fadd_rn(__fmul_rn(cf.x, p1.y), __fmul_rn(cf.y, q.x)));
  rot2(p1, p2, cf.z, cf.w);
  *a2 = p2;
  *a1 = p1;
}
template <int NP>
__device__ __forceinline__ void dbl_rotate_tab(float2 *psi, int total, const uint2 *ix, const float4 *csL, int q,
                                               int c0) {
  if (total == 0) return;
  char *base = reinterpret_cast<char *>(psi) + q * 8;
  const char *cs = reinterpret_cast<const char *>(csL);
  const uint2 *e = ix + c0;
  uint2 e0 = e[0];
  int j = 0;
  for (; j + 1 < total; j += 2) {
    const uint2 e1 = e[(j + 1) * NP];
    rot_entry(base, cs, e0);
    e0 = e[(j + 2) * NP];
    rot_entry(base, cs, e1);
  }
  if (j < total) rot_entry(base, cs, e0);
  asm volatile("" ::: "memory");
}

// psi[j * VW + w] *= d2[j] for the states j < N2 of the VW vectors, two
// entries per lane access when VW > 1 (src/propagate.c:960-961).
template <int VW>
__device__ __forceinline__ void dbl_diag(float2 *psi, const float2 *d2, int N2, int lane) {
  if constexpr (VW == 1) {
    for (int j = lane; j < N2; j += 64) psi[j] = cmul_diag2(psi[j], d2[j]);
  } else {
    float4 *p4 = reinterpret_cast<float4 *>(psi);
    for (int j = lane; j < N2 * (VW / 2); j += 64) {
      const float2 d = d2[j / (VW / 2)];
      const float4 v = p4[j];
      const float2 lo = cmul_diag2(make_float2(v.x, v.y), d), hi = cmul_diag2(make_float2(v.z, v.w), d);
      p4[j] = make_float4(lo.x, lo.y, hi.x, hi.y);
    }
  }
  asm volatile("" ::: "memory");
}

// LDS bytes of the rotation table (K2 nb + 1 rows of NP two-word entries).
__host__ __device__ inline size_t rot_table_bytes(int N, int NP, int K2) {
  const int nb = (N + NP - 1) / NP;
  return (((size_t)K2 * nb + 1) * NP * 8 + 15) & ~(size_t)15;
}
// The table's 16-bit offsets reach every state of a wave and every
// coefficient row.
__host__ __device__ inline bool rot_table_fits(int N, int VW, int K2) {
  const size_t N2 = (size_t)N * (N + 1) / 2;
  return (N2 + 2) * VW * 8 <= 65536 && (size_t)K2 * 48 <= 65536;
}

// The rotation table (first two words per coupling k, site block cb and lane
// c0): every c, ascending; two-level sites (es) have no c = a and no c = b
// rotation.  Coefficient row offsets count three rows of 16 bytes per
// coupling (stage_step_rows).
template <int NP>
__device__ inline void build_rot_table(uint2 *ixL, const uint32_t *ab2, int N, int K2, bool es) {
  constexpr int VW = 64 / NP;
  const int N2 = N * (N + 1) / 2, nb = (N + NP - 1) / NP;
  const uint32_t zs = (uint32_t)(N2 * VW * 8), ns = (uint32_t)((N2 + 1) * VW * 8);
  for (int e = threadIdx.x; e < (K2 * nb + 1) * NP; e += blockDim.x) {
    const int k = e / (nb * NP), c = e % (nb * NP);
    const uint32_t p = k < K2 ? ab2[k] : 0u;
    const int a = (int)(p & 0xffffu), b = (int)(p >> 16);
    uint2 v = make_uint2(zs * 0x10001u, ns | ((uint32_t)(k < K2 ? k * 48 : 0) << 16));
    if (k < K2 && c < N && !(es && (c == a || c == b))) {
      const int rs_c = (c * (2 * N - c - 1)) >> 1;
      const int rs_a = (a * (2 * N - a - 1)) >> 1, rs_b = (b * (2 * N - b - 1)) >> 1;
      const int i1 = c <= a ? a + rs_c : c + rs_a, i2 = c <= b ? b + rs_c : c + rs_b;
      const bool same = a / NP == b / NP;
      uint32_t o3 = ns, row = 0;
      if (!es && c == a) row = 1;
      if (!es && c == b) row = same ? 2 : 1, o3 = same ? (uint32_t)((a + rs_a) * VW * 8) : ns;
      v = make_uint2((uint32_t)(i1 * VW * 8) | ((uint32_t)(i2 * VW * 8) << 16),
                     o3 | ((uint32_t)(k * 48 + row * 16) << 16));
    }
    ixL[e] = v;
  }
}

// The zero and neutral states of a wave's vectors (lane < VW writes slot
// lane).
template <int VW>
__device__ __forceinline__ void init_special_states(float2 *psi, int N2, int lane) {
  if (lane < VW) psi[N2 * VW + lane] = make_float2(0.f, 0.f);
  if (lane < VW) psi[(N2 + 1) * VW + lane] = make_float2(-0.f, 0.f);
}

// A frame's doubles-step rows in LDS: the D2 row, and per coupling e the
// coefficient rows (1, 0, co, si), (1, 0, co2, si2), (co2, si2, co2, si2).
__device__ inline void stage_step_rows(float2 *d2, float4 *csL, const float2 *D2row, const float4 *csrow, int N2,
                                       int K2) {
  stage_lds<8>(d2, D2row, N2);
  for (int e = threadIdx.x; e < K2; e += blockDim.x) {
    const float4 c = gld(csrow + e);
    csL[3 * e] = make_float4(1.f, 0.f, c.x, c.y);
    csL[3 * e + 1] = make_float4(1.f, 0.f, c.z, c.w);
    csL[3 * e + 2] = make_float4(c.z, c.w, c.z, c.w);
  }
}

}  // namespace nise
