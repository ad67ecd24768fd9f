// Internal declarations shared by the HIP translation units of libnise_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nise_abi.h"

namespace nise {

// Constants exactly as the reference declares them (src/types.h:5-12):
// float variables initialised from double expressions.
constexpr float kCv = 2.99792458e-5;
constexpr float kTwoPi = 2 * 3.14159265359;
constexpr float kIcm2ifs = 2.99792458e-5;
constexpr float kSqrt2 = 1.41421356237;

// Packed upper-triangle index, src/NISE_subs.h:54-65.
__host__ __device__ inline int sindex(int a, int b, int N) {
  return (a > b) ? a + b * ((N << 1) - b - 1) / 2 : b + a * ((N << 1) - a - 1) / 2;
}

// Error state for the nise_* entry points (thread local message).
void set_error(const char *fmt, ...);
const char *get_error();

#define NISE_HIP_CHECK(expr)                                                  \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::nise::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,            \
                        hipGetErrorString(e_));                               \
      return NISE_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

// Device buffer with RAII release.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  int alloc(size_t n);
  void release();
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Frame-independent rotation schedule of one Trotter sweep.
//   ab[r]      : (a | b << 16) of the r-th rotation in schedule order
//   layer_off  : rotations of layer l are [layer_off[l], layer_off[l+1])
// Rotations inside a layer touch disjoint sites; for every site the layers
// keep the lexicographic (a<b) order of src/propagate.c:848-888, so the
// layered sweep performs, site by site, the same floating point operations
// in the same order as the reference's sequential sweep.
struct Schedule {
  int N = 0;
  int K = 0;      // entries per sweep (rotations, plus padding in row schedules)
  int Kreal = 0;  // rotations of the union coupling structure
  std::vector<uint32_t> ab;
  std::vector<int> layer_off;
  int nlayers() const { return (int)layer_off.size() - 1; }
};

// Build the layered schedule from a kept-coupling mask over the strict upper
// triangle, enumerated in lexicographic (a<b) order.
void build_schedule(int N, const std::vector<uint8_t> &mask, Schedule &s);
// Row schedule: layers of exactly R entries (rotations, then dummy pairs
// (N+q, N+q) in the unused rows q), for the row conveyor.
void build_schedule_rows(int N, const std::vector<uint8_t> &mask, int R, Schedule &s);

// Singles parameters shared by the frame-table kernels.
struct FrameSource {
  const float *H_rec = nullptr;  // device, record 0 header
  int64_t H_stride = 0;          // words per record
  int coupling_type = 0;         // 1: H_rec holds N diagonal values
  const float *C_rec = nullptr;  // device, Full record with couplings
  float shifte = 0.f;
  float cut = 0.f;
  int nshift = 0;                // single-site shifts, applied in order
  const int *shift_site = nullptr;   // device
  const float *shift_ab = nullptr;   // device, (a, b) pairs
  int apply_shift = 1;
  const float *delta = nullptr;  // device, per-site static offsets added to
                                 // the shifted diagonal (gradient runs)
  float cscale = 1.f;            // scale of the kept couplings
  // Couplings built per frame from geometry (src/read_trajectory.c:181-303):
  // 1 TransitionDipole, 2 ExtendedDipole; H_rec then holds N diagonal values.
  int onfly = 0;
  const float *pos_rec = nullptr;  // position records (int + 3N), one per
  int64_t pos_stride = 0;          // frame (two per frame for ExtendedDipole)
  const float *dip_rec = nullptr;  // transition dipole records (int + 3N)
  int64_t dip_stride = 0;
  const float *box_rec = nullptr;  // periodic box records (int + 3), or none
};

// ---------------------------------------------------------------------------
// Frame element access.
// ---------------------------------------------------------------------------
__device__ inline float raw_diag(const FrameSource &s, int N, int64_t f, int a) {
  const float *rec = s.H_rec + f * s.H_stride + 1;
  return (s.coupling_type || s.onfly) ? rec[a] : rec[sindex(a, a, N)];
}

// Minimum image along axis x when a box is given (pbc1, src/NISE_subs.c:641-648).
__device__ inline float pbc_image(float r, int x, const float *box) {
  if (box && box[0] > 0.f) {
    if (r > box[x] / 2) r = __fsub_rn(r, box[x]);
    if (r < -box[x] / 2) r = __fadd_rn(r, box[x]);
  }
  return r;
}

// |p - q| with the minimum image, in the reference's FP32 order.
__device__ inline float pbc_dist(const float *p, const float *q, int i, int j, int N, const float *box) {
  const float rx = pbc_image(__fsub_rn(p[i], q[j]), 0, box);
  const float ry = pbc_image(__fsub_rn(p[N + i], q[N + j]), 1, box);
  const float rz = pbc_image(__fsub_rn(p[2 * N + i], q[2 * N + j]), 2, box);
  return __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(rx, rx), __fmul_rn(ry, ry)), __fmul_rn(rz, rz)));
}

// Coupling of sites i < j at frame f built from geometry, FP32 in the
// reference's order.  TransitionDipole (src/read_trajectory.c:198-216):
// J = A (m_i.m_j - 3 (m_i.R)(m_j.R)/R^2) / R^3.  ExtendedDipole (:244-297):
// charges q = |m|/d on the two position records, four charge-pair
// distances; the reference takes q_j from site i's dipole (:266), kept so
// that results are identical.
__device__ inline float onfly_coupling(const FrameSource &s, int N, int64_t f, int i, int j) {
  const float A = 5034.11861687f;  // Debye^2/Angstrom^3 to cm^-1
  const float *box = s.box_rec ? s.box_rec + f * 4 + 1 : nullptr;
  const float *mu = s.dip_rec + f * s.dip_stride + 1;
  const float m1x = mu[i], m1y = mu[N + i], m1z = mu[2 * N + i];
  if (s.onfly == 1) {
    const float *R = s.pos_rec + f * s.pos_stride + 1;
    const float m2x = mu[j], m2y = mu[N + j], m2z = mu[2 * N + j];
    const float rx = pbc_image(__fsub_rn(R[i], R[j]), 0, box);
    const float ry = pbc_image(__fsub_rn(R[N + i], R[N + j]), 1, box);
    const float rz = pbc_image(__fsub_rn(R[2 * N + i], R[2 * N + j]), 2, box);
    const float dist = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(rx, rx), __fmul_rn(ry, ry)), __fmul_rn(rz, rz)));
    const float idist = __fdiv_rn(1.f, dist);
    const float idist3 = __fmul_rn(__fmul_rn(idist, idist), idist);
    const float f1 = __fadd_rn(__fadd_rn(__fmul_rn(m1x, m2x), __fmul_rn(m1y, m2y)), __fmul_rn(m1z, m2z));
    const float d1 = __fadd_rn(__fadd_rn(__fmul_rn(m1x, rx), __fmul_rn(m1y, ry)), __fmul_rn(m1z, rz));
    const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(m2x, rx), __fmul_rn(m2y, ry)), __fmul_rn(m2z, rz));
    const float f2 = __fmul_rn(__fmul_rn(__fmul_rn(__fmul_rn(-3.f, d1), d2), idist), idist);
    return __fmul_rn(__fmul_rn(A, __fadd_rn(f1, f2)), idist3);
  }
  const float *R = s.pos_rec + 2 * f * s.pos_stride + 1;  // charges at R ...
  const float *R2 = R + s.pos_stride;                     // ... and at R2
  // q = sqrt(|m|^2) / d: the square root and the division in double, rounded
  // once when stored to float (C's sqrt of a float argument is double)
  const double m1 = __dsqrt_rn((double)__fadd_rn(__fadd_rn(__fmul_rn(m1x, m1x), __fmul_rn(m1y, m1y)), __fmul_rn(m1z, m1z)));
  const float q1 = (float)__ddiv_rn(m1, (double)pbc_dist(R2, R, i, i, N, box));
  const float q2 = (float)__ddiv_rn(m1, (double)pbc_dist(R2, R, j, j, N, box));
  float J = __fdiv_rn(1.f, pbc_dist(R, R, i, j, N, box));
  J = __fsub_rn(J, __fdiv_rn(1.f, pbc_dist(R, R2, i, j, N, box)));
  J = __fsub_rn(J, __fdiv_rn(1.f, pbc_dist(R2, R, i, j, N, box)));
  J = __fadd_rn(J, __fdiv_rn(1.f, pbc_dist(R2, R2, i, j, N, box)));
  return __fmul_rn(__fmul_rn(__fmul_rn(A, q1), q2), J);
}

__device__ inline float raw_coupling(const FrameSource &s, int N, int64_t f, int a, int b) {
  if (s.onfly) return onfly_coupling(s, N, f, a, b);
  const float *rec = s.coupling_type ? (s.C_rec + 1) : (s.H_rec + f * s.H_stride + 1);
  return rec[sindex(a, b, N)];
}

// Diagonal element as read_Hamiltonian leaves it: minus shifte
// (src/read_trajectory.c:153,176,337), then the single-site shifts
// (src/hamshift.c:59-66) in list order.
__device__ inline float shift_diag(const FrameSource &s, int a, float raw) {
  float h = __fsub_rn(raw, s.shifte);
  if (s.apply_shift) {
    for (int i = 0; i < s.nshift; ++i)
      if (s.shift_site[i] == a) h = __fadd_rn(__fmul_rn(h, s.shift_ab[2 * i]), s.shift_ab[2 * i + 1]);
  }
  return h;
}
__device__ inline float site_energy(const FrameSource &s, int N, int64_t f, int a) {
  return shift_diag(s, a, raw_diag(s, N, f, a));
}

// Coupling as kept by read_He (|J| < cut -> 0, src/read_trajectory.c:305-311)
// and the propagator (|J| > cut, src/propagate.c:853).  Returns 0 when not kept.
__device__ inline float kept_coupling(const FrameSource &s, int N, int64_t f, int a, int b) {
  float J = raw_coupling(s, N, f, a, b);
  if (fabsf(J) < s.cut) J = 0.f;
  return (fabsf(J) > s.cut) ? J : 0.f;
}

// Global-memory access through a pointer read from a descriptor in memory.
// The compiler cannot place such a pointer in an address space and emits
// flat instructions, which also count against the LDS wait counter: every
// wait for an LDS result would then wait for outstanding HBM loads too.
// The access goes through a plain integer (or integer vector) type of the
// same size: the address-space qualifier does not survive on class types.
template <int B> struct GWord;
template <> struct GWord<4> { typedef unsigned type; };
template <> struct GWord<8> { typedef unsigned long long type; };
template <> struct GWord<16> { typedef unsigned type __attribute__((ext_vector_type(4))); };
template <class T>
__device__ __forceinline__ T gld(const T *p) {
  typedef typename GWord<sizeof(T)>::type W;
  const W w = *(const __attribute__((address_space(1))) W *)p;
  T v;
  __builtin_memcpy(&v, &w, sizeof(T));
  return v;
}
template <class T>
__device__ __forceinline__ void gst(T *p, T v) {
  typedef typename GWord<sizeof(T)>::type W;
  W w;
  __builtin_memcpy(&w, &v, sizeof(T));
  *(__attribute__((address_space(1))) W *)p = w;
}

// Workgroup barrier for LDS data: this wave's LDS traffic complete, then
// s_barrier.  Global loads already issued stay outstanding (__syncthreads
// would wait for them too, which undoes a prefetch); use it only where no
// global store of this workgroup must be visible to another wave.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Load of wave-uniform, read-only data through the constant address space,
// which the compiler turns into scalar loads (p must be the same in every
// lane and the data must not be written while the kernel runs).
template <class T>
__device__ __forceinline__ T cld(const T *p) {
  typedef typename GWord<sizeof(T)>::type W;
  const W w = *(const __attribute__((address_space(4))) W *)p;
  T v;
  __builtin_memcpy(&v, &w, sizeof(T));
  return v;
}

// dst[map(i)] = *src(i) for i < n, device memory to LDS, by the whole
// workgroup with U loads in flight per thread before the first LDS store (a
// plain copy loop waits for each load before it issues the next).
struct IdentityMap {
  __device__ int operator()(int i) const { return i; }
};
template <int U, class T, class Src, class Map = IdentityMap>
__device__ __forceinline__ void gather_lds(T *dst, int n, Src src, Map map = Map()) {
  const int nt = blockDim.x;
  int i = threadIdx.x;
  for (; i + (U - 1) * nt < n; i += U * nt) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = gld(src(i + u * nt));
#pragma unroll
    for (int u = 0; u < U; ++u) dst[map(i + u * nt)] = v[u];
  }
  if (i < n) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i + u * nt < n) v[u] = gld(src(i + u * nt));
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i + u * nt < n) dst[map(i + u * nt)] = v[u];
  }
}
template <int U, class T, class Map = IdentityMap>
__device__ __forceinline__ void stage_lds(T *dst, const T *src, int n, Map map = Map()) {
  gather_lds<U>(dst, n, [=](int i) { return src + i; }, map);
}

// Value of quad lane j in every lane of the quad (DPP quad_perm; the
// compiler folds it into the VALU instruction that uses the value).
template <int j>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, j * 0x55, 0xF, 0xF, false);
}
template <int j>
__device__ __forceinline__ float quad_bcast(float x) {
  return __uint_as_float(quad_bcast<j>(__float_as_uint(x)));
}

// Complex helpers with the reference's operation order.
// src/propagate.c:871-872 / :892-893: (cr*re - ci*im, ci*re + cr*im)
__device__ __forceinline__ float2 cmul_diag(float2 c, float2 u) {
  return make_float2(__fsub_rn(__fmul_rn(c.x, u.x), __fmul_rn(c.y, u.y)),
                     __fadd_rn(__fmul_rn(c.y, u.x), __fmul_rn(c.x, u.y)));
}

// src/propagate.c:883-887
__device__ __forceinline__ void rotate_pair(float2 &pa, float2 &pb, float co, float si) {
  float cr1 = __fsub_rn(__fmul_rn(co, pa.x), __fmul_rn(si, pb.y));
  float ci1 = __fadd_rn(__fmul_rn(co, pa.y), __fmul_rn(si, pb.x));
  float cr2 = __fsub_rn(__fmul_rn(co, pb.x), __fmul_rn(si, pa.y));
  float ci2 = __fadd_rn(__fmul_rn(co, pb.y), __fmul_rn(si, pa.x));
  pa = make_float2(cr1, ci1);
  pb = make_float2(cr2, ci2);
}

// Reference time factor f = deltat*icm2ifs*twoPi*sign/m (src/propagate.c:836).
inline float trotter_factor(float deltat, int sign, int m) {
  volatile float f = deltat * kIcm2ifs;
  f = f * kTwoPi;
  f = f * (float)sign;
  f = f / (float)m;
  return f;
}

// Launch wrappers (trotter.hip).
int launch_union_mask(const FrameSource &src, int N, int64_t f_lo, int64_t nfr,
                      uint8_t *mask_dev, hipStream_t st);
// jf_dev (optional, [rows][K]): fac * J of each kept coupling before the
// coupling scale, 0 for absent ones (angle derivative for gradients).
int launch_frame_tables(const FrameSource &src, int N, const Schedule &sch,
                        const uint32_t *ab_dev, float fac, int64_t f_lo,
                        int64_t nfr, float2 *diag_dev, float2 *cs_dev,
                        hipStream_t st, float *jf_dev = nullptr);

// One vector slot of the conveyor kernel.
struct SlotDesc {
  int fs;     // first frame (absolute frame index)
  int nresp;  // response steps (0 = none)
  int nprop;  // propagation steps
  int x;      // dipole component for init/response
  int init;   // 0: psi = mu_x(fs), 1: psi from state buffer
  int out;    // state buffer row (init 1 or write back), -1 none
  int grp;    // response group leader: number of consecutive slots (same fs)
              // whose responses it sums; 0 for members
  int hist;   // history row: psi at frames fs .. fs+nresp-1 (after the
              // pending half-phase, before the next) goes to rows hist+rel
              // of ConveyorArgs::hist (hist_every 1), or only the
              // checkpoints rel = 0, K, 2K, ... and nresp-1 to rows
              // hist + 0, 1, 2, ... (hist_every K); -1 none
};

struct ConveyorArgs {
  int N = 0, K = 0, nlayers = 0, m = 1;
  int64_t f_lo = 0;          // absolute frame index of table row 0
  const uint32_t *ab = nullptr;
  const int *layer_off = nullptr;
  const float2 *diag = nullptr;  // [rows][N]
  const float2 *cs = nullptr;    // [rows][K]
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;         // 0: static dipole (Coupling type)
  const SlotDesc *slots = nullptr;
  const int *wg_slot_begin = nullptr;  // [nwg+1]
  int nwg = 0;
  int tmax_acc = 0;              // response length (absorption tmax)
  double2 *partial = nullptr;    // [nwg][tmax_acc]
  float2 *state = nullptr;       // [rows][N] for init/write back
  int writeback = 0;
  const int2 *steps = nullptr;   // rotation steps (build_steps)
  int nsteps = 0;
  float2 *hist = nullptr;        // [rows][N] recorded states
  int hist_every = 1;            // record every step, or checkpoints every K
};

int launch_conveyor(const ConveyorArgs &a, int vp, int block, hipStream_t st);
void build_steps(const Schedule &s, int vp, int block, std::vector<int2> &steps);
size_t conveyor_smem(int N, int vp, int block, int tmax_acc, int nsteps);
int launch_reduce_partials(const double2 *partial, int nwg, int tmax,
                           float *re, float *im, hipStream_t st);

// Plain one-sample-per-launch kernel (baseline).
struct PlainArgs {
  int N = 0, m = 1, tmax = 0;
  float fac = 0.f;
  FrameSource src;
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;
  int64_t ti = 0;            // start frame of the sample
  double2 *resp = nullptr;   // [tmax] accumulated (3 polarisations)
  const uint32_t *ab = nullptr;     // layered schedule of the union structure
  const int *layer_off = nullptr;
  int K = 0, nlayers = 0;
  int cs_lds = 1;            // rotation angles in LDS (else in ws_cs)
  float2 *ws_cs = nullptr;   // [K] workspace when they do not fit
};
size_t plain_smem(int N, int K, bool cs_lds);
int launch_plain_sample(const PlainArgs &a, hipStream_t st);

// Doubles propagation (doubles.hip).
struct DoublesArgs {
  int N = 0, K2 = 0, m = 1;
  const uint32_t *ab2 = nullptr;   // kept couplings, lexicographic
  const float2 *D2 = nullptr;      // [rows][N2]
  const float4 *cs2 = nullptr;     // [rows][K2] (co, si, co2, si2)
  float2 *state = nullptr;         // [vectors][N2]
  const int *vec_begin = nullptr;  // [groups + 1], or [groups] with vec_count
  const int *vec_count = nullptr;  // [groups] vectors of each group (groups need
                                   // not be adjacent), or null: vec_begin[g+1] - vec_begin[g]
  const int64_t *rows = nullptr;   // [groups] table row of each group
  int es = 0;                      // two-level sites (src/propagate.c:1051-1181):
                                   // no rotation for c == a or c == b
};
int launch_doubles_tables(const FrameSource &src, int N, int K2, const uint32_t *ab2, float fac,
                          float anh_fixed, const float *anh_rec, int64_t f_lo, int64_t nfr,
                          float2 *D2, float4 *cs2, hipStream_t st);
int launch_doubles_step(const DoublesArgs &a, int ngroups, hipStream_t st);
int doubles_np(int N);
int doubles_vw(int N);  // vectors per doubles group (one wave)

// Row conveyor (rows.hip), one descriptor per system.  Vector v = 3 s + x
// is polarisation x of sample begin + s, starting at frame (begin + s) *
// sample.  partial holds one row of tmax per vector of the system.
struct RowsArgs {
  int N = 0, tmax = 0, m = 1, nlayers = 0;  // nlayers: multiple of 8
  int rows = 0;                  // this system's row count (launches with R = 0)
  int64_t K = 0;                 // entries per sweep (= nlayers * R)
  int64_t f_lo = 0;
  int nvec = 0, begin = 0, sample = 1;
  const uint2 *steps = nullptr;  // [nlayers * R] LDS byte offsets (a, b)
  const float2 *cs = nullptr;    // [rows][K] (cos, sin) in row-schedule order
  const float2 *diag = nullptr;  // [rows][N]
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;
  double2 *partial = nullptr;    // [vectors][tmax]
  float2 *hist = nullptr;        // checkpoints (hist_every, nck per vector)
  int hist_every = 1, nck = 0;
};
int rows_prepare(const RowsArgs *sys, int nsys, int R, std::vector<int2> &wmap, size_t &smem, bool &pf);
// times (optional): per workgroup (start, end) on the device wall clock.
int launch_rows(const RowsArgs *sys, const int2 *wmap, int nwg, int R, bool pf, size_t smem, hipStream_t st,
                unsigned long long *times = nullptr);

// DIA conveyor (multi.hip): one descriptor per system, vp consecutive
// vectors per 256-thread workgroup, Np = N rounded up to a power of two.
struct DiaSys {
  int N = 0, Np = 1, vp = 1, tmax = 0, begin = 0, sample = 1, nvec = 0;
  int64_t f_lo = 0;
  const float2 *U = nullptr;     // [rows][N][N] per-frame propagators
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;
  double2 *partial = nullptr;    // [workgroups][tmax]
};
// U = V exp(-i E f) V^T for nmat eigensystems (multi.hip); thres > 0 drops
// the elements with FP32 |U_ab|^2 <= thres (the sparse scheme).
int launch_dia_u(const double *w, const double *V, int N, int64_t nmat, float fac, float thres, float2 *U,
                 hipStream_t st);
// Singles vectors through per-frame exact-exponential propagators
// (twodir_dia.hip).  The slots of one workgroup share fs, nresp and nprop.
struct DiaSlot {
  int fs;        // first frame (absolute)
  int nresp;     // history steps: psi before each of frames fs .. fs+nresp-1
  int nprop;     // propagation steps
  int x;         // dipole component of the initial state when in_row < 0
  int in_row;    // initial state row of DiaSlotArgs::in, or -1: mu_x(fs)
  int out_row;   // final state row of DiaSlotArgs::out, or -1
  int hist_row;  // first history row of DiaSlotArgs::hist, or -1
};
struct DiaSlotArgs {
  int N = 0;
  int64_t f_lo = 0;                // absolute frame of table row 0
  const float2 *U = nullptr;       // [rows][N][N]
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;           // 0: static dipoles
  const DiaSlot *slots = nullptr;
  const int *wg_begin = nullptr;   // [nwg + 1]
  int nwg = 0;
  const float2 *in = nullptr;
  float2 *out = nullptr;
  float2 *hist = nullptr;
};
int launch_dia_slots(const DiaSlotArgs &a, int max_ns, hipStream_t st);
int dia_slots_max(int N);

// The sparse two-exciton step (twodir_dia.hip): per-frame element lists of
// the one-exciton propagator and the step on groups of vectors.
int launch_sparse_lists(const float2 *U, int N, int64_t nrows, float thres, float4 *el, int *colptr, int *colidx,
                        int *cnt, hipStream_t st);
struct SparseDblArgs {
  int N = 0, m = 1, es = 0;
  float anh = 0.f;                 // fixed anharmonicity (0: per-frame records)
  const float *anh_rec = nullptr;  // (1 + N) words per frame, absolute frame index
  float fm = 0.f;                  // the half-phase factor f * 0.5 / m
  int64_t f_lo = 0;                // absolute frame of list row 0
  const float4 *el = nullptr;      // [rows][N*N]
  const int *colptr = nullptr;     // [rows][N + 1]
  const int *colidx = nullptr;     // [rows][N*N]
  const int *cnt = nullptr;        // [rows]
  float2 *state = nullptr;         // [vectors][N2]
  const int *vec_begin = nullptr;  // [groups + 1]
  const int64_t *rows = nullptr;   // [groups] list row of each group
};
int launch_sparse_doubles(const SparseDblArgs &a, int ngroups, hipStream_t st);

// Per-frame exact-exponential propagators of frames f_first .. f_first+nfr
// (multi.hip): dense H, FP64 Jacobi, U = V exp(-i E fac) V^T in FP32.
int launch_dia_tables(const FrameSource &src, int N, int64_t f_first, int64_t nfr, float fac, float thres,
                      float2 *U, DevBuf &A, DevBuf &V, DevBuf &w, int *status, hipStream_t st);
size_t rows_smem(int N, int R);
int rows_for(int N, int Kreal, int nlayers_asap, int cap = 16);
void rows_device_order(const Schedule &s, int R, std::vector<uint32_t> &ab);
void rows_step_table(const Schedule &s, int R, std::vector<uint2> &tab);

// Adjoint of the batched absorption (adjoint.hip).
struct AdjArgs {
  int N = 0, K = 0, m = 1, T = 0, every = 1, nck = 0;
  float fac = 0.f;
  int64_t f_lo = 0;
  const uint32_t *ab = nullptr;
  const float2 *cs = nullptr;   // [rows][K] (co, si) with the coupling scale
  const float *jf = nullptr;    // [rows][K] f * J before the scale
  const float2 *diag = nullptr; // [rows][N] half-step phases with the offsets
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;
  const SlotDesc *slots = nullptr;  // fs, x, hist = checkpoint row base
  const int *wg_begin = nullptr;
  int nwg = 0;
  const int2 *steps = nullptr;  // rotation steps in forward order
  int nsteps = 0;
  const float2 *ck = nullptr;   // forward checkpoints
  const float2 *g = nullptr;    // [T]
  double *part_delta = nullptr; // [nwg][N]
  double *part_c = nullptr;     // [nwg]
};
int launch_adjoint(const AdjArgs &a, int vp, double *out, hipStream_t st);

// Row variant of the adjoint (the row conveyor's schedule and tables).
struct AdjRowsArgs {
  int N = 0, T = 0, m = 1, nlayers = 0, every = 1, nck = 0, nvec = 0, begin = 0, sample = 1;
  int64_t K = 0, f_lo = 0;
  float fac = 0.f;
  const uint2 *steps = nullptr;  // row-schedule LDS byte offsets (a, b) for 64/R vectors
  const float2 *cs = nullptr;
  const float *jf = nullptr;
  const float2 *diag = nullptr;
  const float *mu_rec = nullptr;
  int64_t mu_stride = 0;
  const float2 *ck = nullptr;
  const float2 *g = nullptr;
  double *part_delta = nullptr;
  double *part_c = nullptr;
};
int launch_adjoint_rows(const AdjRowsArgs &a, int R, double *out, hipStream_t st);
int adjoint_vp(const Schedule &s, int N, int T);
void adjoint_steps(const Schedule &s, int vp, std::vector<int2> &steps);

// Library-wide stream and device setup.
int ensure_device();
hipStream_t default_stream();
hipStream_t aux_stream();

// src/polar.c:36-105 (directions) and :12-33 (weights)
static __constant__ int c_polar[21][4] = {
    {0, 0, 0, 0}, {1, 1, 1, 1}, {2, 2, 2, 2}, {0, 0, 1, 1}, {0, 0, 2, 2}, {1, 1, 0, 0},
    {1, 1, 2, 2}, {2, 2, 0, 0}, {2, 2, 1, 1}, {0, 1, 0, 1}, {0, 2, 0, 2}, {1, 0, 1, 0},
    {1, 2, 1, 2}, {2, 0, 2, 0}, {2, 1, 2, 1}, {0, 1, 1, 0}, {0, 2, 2, 0}, {1, 0, 0, 1},
    {1, 2, 2, 1}, {2, 0, 0, 2}, {2, 1, 1, 2}};

inline float polarweight_h(int pol, int x) {
  const float ithirty = 1.0 / 30;
  float w = 0;
  if (pol == 0) w = (x <= 2) ? 6 * ithirty : 2 * ithirty;
  if (pol == 1) w = (x <= 2) ? 2 * ithirty : (x <= 8) ? 4 * ithirty : -1 * ithirty;
  if (pol == 2) w = (x <= 8) ? 0.0f : (x <= 14) ? (float)(2.5 * ithirty) : (float)(-2.5 * ithirty);
  return w;
}

// Host file helpers (driver.hip).
// Read records [first, first+count) of a GROBIN file into dst; false if the
// file holds fewer records.
bool read_records(const char *name, int64_t rec_words, int64_t first, int64_t count, float *dst);

// The 2D file drop-ins (twodir.hip): print the message and exit with code;
// write one print2D file (src/MPI_subs.c:20-34), dividing R and I in place.
[[noreturn]] void die2d(int code, const char *msg);
void print2D_file(const char *name, float *R, float *I, const t_non *non, int sampleCount);

// Pinned host staging buffer (hipHostMalloc) with RAII release.
struct Pinned {
  float *p = nullptr;
  Pinned() = default;
  Pinned(const Pinned &) = delete;
  Pinned &operator=(const Pinned &) = delete;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  bool alloc(size_t words) { return hipHostMalloc((void **)&p, words * 4, hipHostMallocDefault) == hipSuccess; }
};

}  // namespace nise
