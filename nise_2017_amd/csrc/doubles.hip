// Two-exciton (doubles) Trotter propagation for gfx950.
//
// Reference: propagate_vec_coupling_S_doubles, src/propagate.c:902-1048.
// Basis: the N(N+1)/2 states |ab>, a <= b, packed with Sindex.  One Trotter
// substep is D2 . R_kmax ... R_1 . D2 with
//   D2     = exp(-i f H0/2),  H0_ab = e_a + e_b - delta_ab * anharmonicity
//   R_k    = for c = 0..N-1 the 2x2 rotation of (|ca>, |cb>) with angle
//            f J_k, or sqrt(2) f J_k when c is a or b, in ascending c.
// The N rotations of one coupling touch disjoint pairs except c = a and
// c = b, which share |ab>: they run as two ordered rounds.  Couplings follow
// each other strictly (any two share a state).  The reference evaluates the
// trigonometry in float libm (cosf, sinf, sqrtf); here each is computed in
// double and rounded once, the correctly rounded float result.
//
// One wave propagates 64/NP vectors (NP = lanes per vector, N <= NP <= 64,
// or NP = 64 with a loop over site chunks when N > 64); the wave's vectors
// share one frame, live in LDS as psi[state][vector], and LDS operations of
// a wave execute in order, so rounds need no barrier.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nise_internal.h"
#include "doubles_dev.h"

namespace nise {

// Per-frame doubles tables: D2[row][N2] and (co, si, co2, si2)[row][K2]
// for the kept couplings in lexicographic order (ab2).  anh_rec != null:
// per-site anharmonicities from records (1 + N words per frame).
__global__ void doubles_tables_kernel(FrameSource s, int N, int K2, const uint32_t *ab2, float fac,
                                      float anh_fixed, const float *anh_rec, int64_t f_lo,
                                      float2 *D2, float4 *cs2) {
  const int64_t row = blockIdx.y;
  const int64_t f = f_lo + row;
  const int N2 = N * (N + 1) / 2;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < N2) {
    // invert the packed index j -> (a, b), a <= b
    int a = 0, rem = j;
    while (rem >= N - a) { rem -= N - a; ++a; }
    const int b = a + rem;
    const float ea = __fsub_rn(raw_diag(s, N, f, a), s.shifte);
    const float eb = __fsub_rn(raw_diag(s, N, f, b), s.shifte);
    float h0 = __fadd_rn(ea, eb);
    if (a == b) {
      // no records and no fixed value: two-level sites, no anharmonicity
      const float d = (anh_fixed == 0.f) ? (anh_rec ? anh_rec[f * (1 + (int64_t)N) + 1 + a] : 0.f) : anh_fixed;
      h0 = __fsub_rn(h0, d);
    }
    const float arg = __fmul_rn(__fmul_rn(0.5f, h0), fac);
    D2[row * N2 + j] = make_float2((float)cos((double)arg), (float)(-sin((double)arg)));
  } else if (j < N2 + K2) {
    const int k = j - N2;
    const uint32_t p = ab2[k];
    const int a = (int)(p & 0xffffu), b = (int)(p >> 16);
    float J = kept_coupling(s, N, f, a, b);
    float4 out = make_float4(1.f, 0.f, 1.f, 0.f);
    if (J != 0.f) {
      const float Jf = __fmul_rn(J, fac);
      const float si = (float)(-sin((double)Jf));
      const float co = (float)sqrt((double)__fsub_rn(1.0f, __fmul_rn(si, si)));
      const float si2 = (float)(-sin((double)__fmul_rn(Jf, kSqrt2)));
      const float co2 = (float)sqrt((double)__fsub_rn(1.0f, __fmul_rn(si2, si2)));
      out = make_float4(co, si, co2, si2);
    }
    cs2[row * K2 + k] = out;
  }
}

int launch_doubles_tables(const FrameSource &src, int N, int K2, const uint32_t *ab2, float fac,
                          float anh_fixed, const float *anh_rec, int64_t f_lo, int64_t nfr,
                          float2 *D2, float4 *cs2, hipStream_t st) {
  const int N2 = N * (N + 1) / 2;
  const int n = N2 + K2;
  int64_t done = 0;
  while (done < nfr) {
    const int64_t chunk = std::min<int64_t>(nfr - done, 65535);
    hipLaunchKernelGGL(doubles_tables_kernel, dim3((n + 255) / 256, (unsigned)chunk), dim3(256), 0, st,
                       src, N, K2, ab2, fac, anh_fixed, anh_rec, f_lo + done, D2 + done * N2,
                       cs2 + done * K2);
    NISE_HIP_CHECK(hipGetLastError());
    done += chunk;
  }
  return NISE_OK;
}

// LDS bytes of a doubles step: states [N2][vectors per wave], the D2 row
// when d2_lds, then the coupling chunk.  The D2 row is staged when a wave
// holds four or more vectors (N <= 16) and everything fits; with one or two
// vectors it would take a half or a third of the workgroup's LDS, and reading
// it through the cache keeps more waves resident (the N = 32 ring's 2DIR-
// Raman run: 72.2 -> 69.1 ms).
__host__ __device__ inline size_t doubles_smem(int N, int np, int K2, bool d2_lds) {
  const size_t N2 = (size_t)N * (N + 1) / 2, vw = (size_t)(64 / np) * dbl_vpl(np);
  return ((N2 * (vw + (d2_lds ? 1 : 0)) * 8 + 15) & ~(size_t)15) + (size_t)dbl_chunk(K2) * 20;
}
__host__ __device__ inline bool doubles_d2_lds(int N, int np, int K2) {
  return (64 / np) * dbl_vpl(np) >= 4 && doubles_smem(N, np, K2, true) <= 160 * 1024;
}

// One frame of m Trotter substeps for groups of vectors.  Group g (one
// wave) holds vectors vec_begin[g] .. vec_begin[g+1]-1 of the state buffer
// (at most 64/NP x VPL of them) and uses table row rows[g].  The states,
// the frame's D2 row and its coupling rows are staged in LDS first (loads in
// flight together), so the ordered rotation chain waits on LDS only.  Lane
// group q = t % (64 / NP) owns vectors q VPL .. q VPL + VPL - 1, and c0 =
// t / (64 / NP) the rotation of site c: the vectors of one state sit in
// adjacent lanes, so a round's contiguous states fill the LDS banks.
template <int NP, int VPL>
__global__ void __launch_bounds__(64) doubles_step_kernel(DoublesArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int VW = 64 / NP * VPL;  // vectors per wave
  const int N = A.N, N2 = N * (N + 1) / 2, K2 = A.K2, ch = dbl_chunk(K2);
  const bool d2l = doubles_d2_lds(N, NP, K2);
  float2 *psi = reinterpret_cast<float2 *>(smem);  // [N2][VW]
  float4 *csL = reinterpret_cast<float4 *>(smem + (((size_t)N2 * (VW + d2l) * 8 + 15) & ~(size_t)15));  // [ch]
  uint32_t *abL = reinterpret_cast<uint32_t *>(csL + ch);                                              // [ch]
  const int g = blockIdx.x, t = threadIdx.x;
  const int v0 = A.vec_begin[g], nv = A.vec_count ? A.vec_count[g] : A.vec_begin[g + 1] - v0;
  const int64_t row = A.rows[g];
  const float4 *cs2 = A.cs2 + row * K2;
  const float2 *d2L = psi + N2 * VW, *d2g = A.D2 + row * N2;  // the D2 row in LDS (d2l) or device memory
  // one diagonal half-step on the LDS states (LDS and global reads kept
  // apart: a pointer that may be either compiles to flat loads)
  auto diag = [&]() {
    if (d2l) {
#pragma unroll 4
      for (int j = t; j < N2 * VW; j += 64) psi[j] = cmul_diag2(psi[j], d2L[j / VW]);
    } else {
      for (int j = t; j < N2 * VW; j += 64) psi[j] = cmul_diag2(psi[j], gld(d2g + j / VW));
    }
    asm volatile("" ::: "memory");
  };
  // Staging in one pass over the float2 words [states of the nv vectors |
  // D2 row | coupling rows]: state[v0 + w][j] -> psi[j][w]; the pair words
  // (at most kDblCh, 8 per lane) are loaded before it and stored after, so
  // all of a thread's loads are in flight together.
  const bool once = K2 <= ch;
  for (int w = nv; w < VW; ++w)
    for (int j = t; j < N2; j += 64) psi[j * VW + w] = make_float2(0.f, 0.f);
  uint32_t abv[kDblCh / 64];
  if (once) {
#pragma unroll
    for (int u = 0; u < kDblCh / 64; ++u)
      if (t + 64 * u < K2) abv[u] = gld(A.ab2 + t + 64 * u);
  }
  {
    const int nS = nv * N2, nD = d2l ? N2 : 0, nC = once ? 2 * K2 : 0;
    const float2 *sg = A.state + (int64_t)v0 * N2, *cg = reinterpret_cast<const float2 *>(cs2);
    const int cofs = (int)(reinterpret_cast<float2 *>(csL) - psi);
    gather_lds<32>(psi, nS + nD + nC,
                   [=](int e) { return e < nS ? sg + e : e < nS + nD ? d2g + (e - nS) : cg + (e - nS - nD); },
                   [=](int e) {
                     if (e < nS) {
                       int w = 0;
#pragma unroll
                       for (int u = 1; u < VW; ++u) w += e >= u * N2;
                       return (e - w * N2) * VW + w;
                     }
                     return e < nS + nD ? N2 * VW + (e - nS) : cofs + (e - nS - nD);
                   });
  }
  if (once) {
#pragma unroll
    for (int u = 0; u < kDblCh / 64; ++u)
      if (t + 64 * u < K2) abL[t + 64 * u] = abv[u];
  }
  __syncthreads();
  const int q = t % (64 / NP), c0 = t / (64 / NP);
  const int rs0 = (c0 * (2 * N - c0 - 1)) >> 1;  // row start of |c0 c0> in the packed basis
  for (int sub = 0; sub < A.m; ++sub) {
    diag();
    for (int k0 = 0; k0 < K2; k0 += ch) {
      const int nk = min(ch, K2 - k0);
      if (!once) {
        __syncthreads();
        stage_lds<4>(csL, cs2 + k0, nk);
        stage_lds<4>(abL, A.ab2 + k0, nk);
        __syncthreads();
      }
      dbl_rotate<NP, VPL>(psi, N, nv, nk, abL, csL, A.es, q, c0, rs0);
    }
    if (sub + 1 < A.m) diag();
  }
  __syncthreads();
  // the last diagonal half-step, fused into the store
  for (int ww = 0; ww < nv; ++ww)
    for (int j = t; j < N2; j += 64)
      gst(A.state + (int64_t)(v0 + ww) * N2 + j,
          cmul_diag2(psi[j * VW + ww], d2l ? d2L[j] : gld(d2g + j)));
}

int doubles_np(int N) { return N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64; }
int doubles_vw(int N) { return 64 / doubles_np(N) * dbl_vpl(doubles_np(N)); }

int launch_doubles_step(const DoublesArgs &a, int ngroups, hipStream_t st) {
  if (ngroups == 0) return NISE_OK;
  const int np = doubles_np(a.N);
  const size_t smem = doubles_smem(a.N, np, a.K2, doubles_d2_lds(a.N, np, a.K2));
  if (smem > 160 * 1024) return set_error("doubles: N=%d too large for LDS", a.N), NISE_ERR_ARG;
  const void *fn = np == 8 ? (const void *)doubles_step_kernel<8, dbl_vpl(8)>
                 : np == 16 ? (const void *)doubles_step_kernel<16, dbl_vpl(16)>
                 : np == 32 ? (const void *)doubles_step_kernel<32, dbl_vpl(32)>
                            : (const void *)doubles_step_kernel<64, dbl_vpl(64)>;
  NISE_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  DoublesArgs args = a;
  void *kargs[] = {&args};
  NISE_HIP_CHECK(hipLaunchKernel(fn, dim3(ngroups), dim3(64), kargs, smem, st));
  return NISE_OK;
}

// Lexicographic list of kept couplings of one packed frame (host side).
void kept_pairs(int N, const float *H, float cut, std::vector<uint32_t> &ab) {
  ab.clear();
  for (int a = 0; a < N; ++a)
    for (int b = a + 1; b < N; ++b)
      if (std::fabs(H[sindex(a, b, N)]) > cut) ab.push_back((uint32_t)a | ((uint32_t)b << 16));
}

namespace {
struct DoublesCtx {
  DevBuf H, ab, D2, cs2, state, vb, rows, anh;
  std::vector<float2> sbuf;
};
DoublesCtx &dctx2() {
  static thread_local DoublesCtx c;
  return c;
}
int up2(DevBuf &d, const void *s, size_t n) {
  int rc = d.alloc(n < 16 ? 16 : n);
  if (rc) return rc;
  if (n && s) NISE_HIP_CHECK(hipMemcpy(d.p, s, n, hipMemcpyHostToDevice));
  return NISE_OK;
}
}  // namespace

// nvec doubles vectors (stride N(N+1)/2) through one frame H (packed,
// diagonal already shifted), m substeps.
// es: two-level sites (src/propagate.c:1051-1181): no anharmonicity, no
// rotations for c == a or c == b; Anh and anharmonicity are not used.
int doubles_host(int N, const float *H, float *const *cr, float *const *ci, int nvec, int m,
                 const float *Anh, float anharmonicity, float deltat, float cut, int es) {
  int rc = ensure_device();
  if (rc) return rc;
  if (N < 1 || m < 1 || nvec < 1) return set_error("doubles: bad arguments"), NISE_ERR_ARG;
  hipStream_t st = default_stream();
  DoublesCtx &c = dctx2();
  const int N2 = N * (N + 1) / 2;
  std::vector<float> rec(1 + N2);
  rec[0] = 0.f;
  std::memcpy(rec.data() + 1, H, N2 * 4);
  std::vector<uint32_t> ab;
  kept_pairs(N, H, cut, ab);
  const int K2 = (int)ab.size();
  const int vw = doubles_vw(N);
  const int ng = (nvec + vw - 1) / vw;
  std::vector<int> vb(ng + 1);
  for (int g = 0; g <= ng; ++g) vb[g] = std::min(nvec, g * vw);
  c.sbuf.resize((size_t)nvec * N2);
  for (int v = 0; v < nvec; ++v)
    for (int j = 0; j < N2; ++j) c.sbuf[(size_t)v * N2 + j] = make_float2(cr[v][j], ci[v][j]);
  if ((rc = up2(c.H, rec.data(), rec.size() * 4)) || (rc = up2(c.ab, ab.data(), ab.size() * 4)) ||
      (rc = up2(c.vb, vb.data(), vb.size() * 4)) ||
      (rc = up2(c.state, c.sbuf.data(), c.sbuf.size() * sizeof(float2))))
    return rc;
  std::vector<int64_t> rows64(ng, 0);
  if ((rc = up2(c.rows, rows64.data(), rows64.size() * 8))) return rc;
  if (es) anharmonicity = 0.f, Anh = nullptr;
  if (anharmonicity == 0.f && Anh) {
    std::vector<float> arec(1 + N);
    arec[0] = 0.f;
    std::memcpy(arec.data() + 1, Anh, N * 4);
    if ((rc = up2(c.anh, arec.data(), arec.size() * 4))) return rc;
  }
  if ((rc = c.D2.alloc((size_t)N2 * sizeof(float2))) || (rc = c.cs2.alloc((size_t)std::max(K2, 1) * sizeof(float4))))
    return rc;
  FrameSource src;
  src.H_rec = c.H.as<float>(), src.H_stride = 1 + N2, src.shifte = 0.f, src.cut = cut, src.apply_shift = 0;
  const float fac = trotter_factor(deltat, 1, m);
  if ((rc = launch_doubles_tables(src, N, K2, c.ab.as<uint32_t>(), fac, anharmonicity,
                                  (anharmonicity == 0.f && Anh) ? c.anh.as<float>() : nullptr, 0, 1,
                                  c.D2.as<float2>(), c.cs2.as<float4>(), st)))
    return rc;
  DoublesArgs a;
  a.N = N, a.K2 = K2, a.m = m, a.ab2 = c.ab.as<uint32_t>();
  a.D2 = c.D2.as<float2>(), a.cs2 = c.cs2.as<float4>();
  a.state = c.state.as<float2>(), a.vec_begin = c.vb.as<int>(), a.rows = c.rows.as<int64_t>();
  a.es = es;
  if ((rc = launch_doubles_step(a, ng, st))) return rc;
  NISE_HIP_CHECK(hipMemcpyAsync(c.sbuf.data(), c.state.p, c.sbuf.size() * sizeof(float2),
                                hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  for (int v = 0; v < nvec; ++v)
    for (int j = 0; j < N2; ++j) {
      cr[v][j] = c.sbuf[(size_t)v * N2 + j].x;
      ci[v][j] = c.sbuf[(size_t)v * N2 + j].y;
    }
  return NISE_OK;
}

int coupling_S_host(int N, const float *H, float *cr, float *ci, int nvec, int m, int sign,
                    float deltat, float cut);

}  // namespace nise

using namespace nise;

// src/propagate.c:902-1048
extern "C" void propagate_vec_coupling_S_doubles(t_non *non, float *Hamiltonian_i, float *cr,
                                                 float *ci, int m, float *Anh) {
  float *r[1] = {cr}, *i[1] = {ci};
  if (doubles_host(non->singles, Hamiltonian_i, r, i, 1, m, Anh, non->anharmonicity, non->deltat,
                   non->couplingcut, 0)) {
    printf("propagate_vec_coupling_S_doubles: GPU propagation failed: %s\n", get_error());
    exit(1);
  }
}

// src/propagate.c:1051-1181 (two-level sites, calc_2DES)
extern "C" void propagate_vec_coupling_S_doubles_ES(t_non *non, float *Hamiltonian_i, float *cr,
                                                    float *ci, int m) {
  float *r[1] = {cr}, *i[1] = {ci};
  if (doubles_host(non->singles, Hamiltonian_i, r, i, 1, m, nullptr, 0.f, non->deltat,
                   non->couplingcut, 1)) {
    printf("propagate_vec_coupling_S_doubles_ES: GPU propagation failed: %s\n", get_error());
    exit(1);
  }
}

namespace {
// The Coupling scheme of the doubles control routines: the doubles vectors
// fr and ft1r[0..tmax1) in one launch, the singles vectors rightnr and
// rightrr[0..tmax1) with sign -1 in another.
void doubles_control(const char *who, t_non *non, float *Hamil_i_e, float **ft1r, float **ft1i, float *fr,
                     float *fi, float **rightrr, float **rightri, float *rightnr, float *rightni, float *Anh,
                     int es) {
  if (non->propagation != 1) {
    printf("%s: propagation scheme %d is not provided by the GPU library\n", who, non->propagation);
    exit(1);
  }
  const int T1 = non->tmax1, N = non->singles;
  std::vector<float *> r(T1 + 1), i(T1 + 1);
  r[0] = fr, i[0] = fi;
  for (int t = 0; t < T1; ++t) r[t + 1] = ft1r[t], i[t + 1] = ft1i[t];
  if (doubles_host(N, Hamil_i_e, r.data(), i.data(), T1 + 1, non->ts, Anh, non->anharmonicity, non->deltat,
                   non->couplingcut, es)) {
    printf("%s: GPU propagation failed: %s\n", who, get_error());
    exit(1);
  }
  std::vector<float> sr((size_t)(T1 + 1) * N), si((size_t)(T1 + 1) * N);
  std::memcpy(sr.data(), rightnr, N * 4);
  std::memcpy(si.data(), rightni, N * 4);
  for (int t = 0; t < T1; ++t) {
    std::memcpy(sr.data() + (size_t)(t + 1) * N, rightrr[t], N * 4);
    std::memcpy(si.data() + (size_t)(t + 1) * N, rightri[t], N * 4);
  }
  if (coupling_S_host(N, Hamil_i_e, sr.data(), si.data(), T1 + 1, non->ts, -1, non->deltat,
                      non->couplingcut)) {
    printf("%s: GPU propagation failed: %s\n", who, get_error());
    exit(1);
  }
  std::memcpy(rightnr, sr.data(), N * 4);
  std::memcpy(rightni, si.data(), N * 4);
  for (int t = 0; t < T1; ++t) {
    std::memcpy(rightrr[t], sr.data() + (size_t)(t + 1) * N, N * 4);
    std::memcpy(rightri[t], si.data() + (size_t)(t + 1) * N, N * 4);
  }
}
}  // namespace

// src/propagate.c:1476-1591, Coupling scheme (non->propagation == 1)
extern "C" void propagate_double_2DIR_control(t_non *non, float *Hamil_i_e, float **ft1r,
                                              float **ft1i, float *fr, float *fi, float **rightrr,
                                              float **rightri, float *rightnr, float *rightni,
                                              int currentSample, int molPol, int t3, float *Anh) {
  (void)currentSample, (void)molPol, (void)t3;
  doubles_control("propagate_double_2DIR_control", non, Hamil_i_e, ft1r, ft1i, fr, fi, rightrr, rightri,
                  rightnr, rightni, Anh, 0);
}

// src/propagate.c:1594-1709, Coupling scheme (non->propagation == 1): the
// two-level doubles of calc_2DES
extern "C" void propagate_double_2DES_control(t_non *non, float *Hamil_i_e, float **ft1r,
                                              float **ft1i, float *fr, float *fi, float **rightrr,
                                              float **rightri, float *rightnr, float *rightni,
                                              int currentSample, int molPol, int t3) {
  (void)currentSample, (void)molPol, (void)t3;
  doubles_control("propagate_double_2DES_control", non, Hamil_i_e, ft1r, ft1i, fr, fi, rightrr, rightri,
                  rightnr, rightni, nullptr, 1);
}
