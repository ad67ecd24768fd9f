// Batched absorption of many independent systems of uneven size.
//
// Systems with propagation 1 (Trotter coupling scheme) run on the row
// conveyor in one launch: every workgroup reads its own system descriptor,
// so waves of different N share the launch.  Systems with propagation 0
// (exact exponential, src/propagate.c:60-125) and N <= 64 run on the DIA
// conveyor in a second launch: per frame U = V exp(-i E f) V^T is formed once
// (batched FP64 Jacobi, then U in double rounded to FP32) and each workgroup
// applies its frame's U, staged in LDS, to all its vectors at that frame.
//
// Each workgroup records its start and end on the device wall clock; a
// system's completion latency is the end of its last workgroup measured from
// the first start of the run.
#include <algorithm>
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

// Per-frame propagators U (FP32 [nfr][N][N]) of frames f_first .. for the
// exact exponential with factor fac (thres > 0: the sparse scheme's dropped
// elements zeroed), through FP64 work arrays of at most 4096 frames; one
// Jacobi convergence word per frame in status (negative: not converged).
int launch_dia_tables(const FrameSource &src, int N, int64_t f_first, int64_t nfr, float fac, float thres,
                      float2 *U, DevBuf &A, DevBuf &V, DevBuf &w, int *status, hipStream_t st) {
  if (nfr <= 0) return NISE_OK;
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

// One workgroup (256 threads): vectors v0 .. v0+vp-1 of one system.  Thread
// (site a, group sg) handles sites a of vectors sg, sg + 256/Np, ...
__global__ void __launch_bounds__(256) dia_conveyor_kernel(const DiaSys *sys, const int2 *wmap,
                                                           unsigned long long *times) {
  extern __shared__ __align__(16) unsigned char smem[];
  const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
  const int2 wm = wmap[blockIdx.x];
  const DiaSys &S = sys[wm.x];
  const int N = S.N, Np = S.Np, T = S.tmax, v0 = wm.y;
  const int nv = min(S.vp, S.nvec - v0);
  const int groups = 256 / Np;
  float2 *U = reinterpret_cast<float2 *>(smem);                 // [N][N]
  float2 *psA = U + (size_t)N * N;                              // [vp][Np]
  float2 *psB = psA + (size_t)S.vp * Np;                        // [vp][Np]
  double2 *acc = reinterpret_cast<double2 *>(psB + (size_t)S.vp * Np);  // [T]
  float2 *red = reinterpret_cast<float2 *>(acc + T);            // [vp]
  const int t = threadIdx.x, a = t % Np, sg = t / Np;
  for (int i = t; i < T; i += 256) acc[i] = make_double2(0.0, 0.0);
  auto fs_of = [&](int v) { return (int64_t)(S.begin + (v0 + v) / 3) * S.sample; };
  const int64_t F0 = fs_of(0), F1 = fs_of(nv - 1) + T;
  __syncthreads();
  for (int64_t f = F0; f < F1; ++f) {
    const int64_t row = f - S.f_lo;
    const float2 *Uf = S.U + row * N * N;
    for (int i = t; i < N * N; i += 256) U[i] = gld(Uf + i);
    const float *mu = S.mu_rec + (S.mu_stride ? f * S.mu_stride : 0) + 1;
    __syncthreads();
    for (int v = sg; v < nv; v += groups) {
      const int64_t rel = f - fs_of(v);
      const int x = (v0 + v) % 3;
      const bool act = rel >= 0 && rel < T;
      float r = 0.f, q = 0.f;
      float2 p = make_float2(0.f, 0.f);
      if (act && a < N) {
        const float m = gld(mu + x * N + a);
        p = rel == 0 ? make_float2(m, 0.f) : psA[v * Np + a];
        if (rel == 0) psA[v * Np + a] = p;
        r = __fmul_rn(m, p.x);
        q = __fmul_rn(m, p.y);
      }
      // response of this vector: sum over its Np lanes (one wave slice)
      for (int off = 1; off < Np; off <<= 1) r += __shfl_xor(r, off), q += __shfl_xor(q, off);
      if (a == 0) red[v] = act ? make_float2(r, q) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    // psi <- U psi for vectors that propagate at this frame (src/propagate.c:109-119)
    for (int v = sg; v < nv; v += groups) {
      const int64_t rel = f - fs_of(v);
      if (rel >= 0 && rel < T - 1 && a < N) {
        float yr = 0.f, yi = 0.f;
        const float2 *ps = psA + v * Np;
        for (int b = 0; b < N; ++b) {
          const float2 u = U[a * N + b], c = ps[b];
          yr = __fadd_rn(yr, __fsub_rn(__fmul_rn(u.x, c.x), __fmul_rn(u.y, c.y)));
          yi = __fadd_rn(yi, __fadd_rn(__fmul_rn(u.x, c.y), __fmul_rn(u.y, c.x)));
        }
        psB[v * Np + a] = make_float2(yr, yi);
      }
    }
    if (t == 0)  // responses in vector order
      for (int v = 0; v < nv; ++v) {
        const int64_t rel = f - fs_of(v);
        if (rel >= 0 && rel < T) acc[rel].x += (double)red[v].x, acc[rel].y += (double)red[v].y;
      }
    __syncthreads();
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
  return (size_t)N * N * 8 + (size_t)2 * vp * Np * 8 + (size_t)T * 16 + (size_t)vp * 8;
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
  // the row schedule of the previous call, reused while its inputs repeat
  std::vector<uint8_t> rows_mask;
  int rows_R = 0, rows_N = 0;
  // everything but the device buffers, which the next call reuses
  void reset() {
    N = T = dia = nvec = nwg = 0;
    thres = 0.f;
    f_lo = f_hi = 0;
    src = FrameSource();
    mu_dev = nullptr;
    mu_stride = 0;
    sch_asap = Schedule();
    mask.clear();
  }
};

// Device buffers of the calling thread's batches, kept between calls: a
// batch of the same shape then allocates and frees nothing (hipMalloc and
// hipFree of the cfg4 batch's buffers took about 55 ms of host time per
// call).  Every call ends with both streams synchronised.
struct MultiCache {
  std::vector<std::unique_ptr<SysState>> sys;
  DevBuf mk, A, V, w, stt, rsys, rwm, dsys, dwm, times, out;
};
}  // namespace

}  // namespace nise

using namespace nise;

extern "C" int nise_absorption_multi(const nise_system *systems, int nsys, int where, float *re_S,
                                     float *im_S, double *done_ms) {
  int rc = ensure_device();
  if (rc) return rc;
  if (nsys < 1 || !systems || !re_S || !im_S) return set_error("multi: bad arguments"), NISE_ERR_ARG;
  // Trotter systems on the default stream, DIA systems (per-frame Jacobi,
  // U and their conveyor) on a second one, so that the two overlap
  hipStream_t st = default_stream(), sd = aux_stream();
  if (!sd) return set_error("multi: no second stream"), NISE_ERR_HIP;
  static thread_local MultiCache cache;
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
  // 2. the DIA systems' U per frame on the second stream, queued before
  //    the host builds the Trotter schedules so that the two overlap.  The
  //    DIA systems share FP64 work arrays sized for the largest of them and
  //    one status word per frame, checked once after every launch is queued.
  std::vector<int64_t> soff(nsys + 1, 0);
  int64_t wmax = 0;
  for (int k = 0; k < nsys; ++k) {
    const int64_t nfr = S[k]->f_hi - S[k]->f_lo;
    soff[k + 1] = soff[k] + (S[k]->dia ? nfr : 0);
    if (S[k]->dia) wmax = std::max(wmax, std::min<int64_t>(nfr, 4096) * S[k]->N * S[k]->N);
  }
  DevBuf &A = cache.A, &V = cache.V, &w = cache.w, &stt = cache.stt;
  if (wmax > 0 && ((rc = A.alloc((size_t)wmax * 8)) || (rc = V.alloc((size_t)wmax * 8)) ||
                   (rc = w.alloc((size_t)wmax * 8)) || (rc = stt.alloc((size_t)soff[nsys] * 4))))
    return rc;
  std::vector<RowsArgs> rows;
  std::vector<int> rows_of;
  std::vector<DiaSys> dias;
  std::vector<int> dia_of;
  for (int k = 0; k < nsys; ++k) {
    SysState &s = *S[k];
    const t_non *non = systems[k].non;
    const int N = s.N;
    const int64_t nfr = s.f_hi - s.f_lo;
    if (!s.dia) continue;
      // U per frame, in chunks of frames for the FP64 work arrays
      if ((rc = s.U.alloc((size_t)nfr * N * N * 8))) return rc;
      const int64_t chunk = std::min<int64_t>(nfr, 4096);
      volatile float fv = non->deltat * kIcm2ifs;  // src/propagate.c:69
      fv = fv * kTwoPi;
      const float fac = fv;
      for (int64_t c0 = 0; c0 < nfr; c0 += chunk) {
        const int64_t nc = std::min(chunk, nfr - c0);
        for (int64_t d0 = 0; d0 < nc; d0 += 65535) {
          const int64_t nd = std::min<int64_t>(65535, nc - d0);
          hipLaunchKernelGGL(dia_dense_kernel, dim3((N * N + 255) / 256, (unsigned)nd), dim3(256), 0, sd, s.src, N,
                             s.f_lo + c0 + d0, A.as<double>() + d0 * N * N);
          NISE_HIP_CHECK(hipGetLastError());
        }
        if ((rc = launch_jacobi(A.as<double>(), N, (int)nc, w.as<double>(), V.as<double>(),
                                stt.as<int>() + soff[k] + c0, sd)))
          return rc;
        if ((rc = launch_dia_u(w.as<double>(), V.as<double>(), N, nc, fac, s.thres, s.U.as<float2>() + c0 * N * N,
                               sd)))
          return rc;
      }
      DiaSys d;
      d.N = N, d.Np = 1;
      while (d.Np < N) d.Np <<= 1;
      d.vp = std::min(32, std::max(1, 256 / d.Np) * 8);
      d.tmax = s.T, d.begin = non->begin, d.sample = non->sample, d.nvec = s.nvec, d.f_lo = s.f_lo;
      d.U = s.U.as<float2>(), d.mu_rec = s.mu_dev, d.mu_stride = s.mu_stride;
      s.nwg = (s.nvec + d.vp - 1) / d.vp;
      if ((rc = s.partial.alloc((size_t)s.nwg * s.T * sizeof(double2)))) return rc;
      d.partial = s.partial.as<double2>();
      dias.push_back(d);
      dia_of.push_back(k);
  }
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
    for (int k = 0; k < nsys; ++k)
      if (!S[k]->dia) build_schedule(S[k]->N, S[k]->mask, S[k]->sch_asap);
  }
  // 3. one row count for the Trotter launch: the work-weighted median
  std::vector<std::pair<int, double>> rw;
  for (int k = 0; k < nsys; ++k)
    if (const SysState *p = S[k].get(); !p->dia)
      rw.push_back({rows_for(p->N, p->sch_asap.Kreal, p->sch_asap.nlayers()),
                    (double)p->nvec * p->T * std::max(p->sch_asap.Kreal, 1)});
  int R = 16;
  if (!rw.empty()) {
    std::sort(rw.begin(), rw.end());
    double tot = 0, acc = 0;
    for (auto &e : rw) tot += e.second;
    for (auto &e : rw) {
      acc += e.second;
      if (acc >= 0.5 * tot) { R = e.first; break; }
    }
  }
  // 4. tables of the Trotter systems
  for (int k = 0; k < nsys; ++k) {
    SysState &s = *S[k];
    const t_non *non = systems[k].non;
    const int N = s.N;
    const int64_t nfr = s.f_hi - s.f_lo;
    if (s.dia) continue;
      if (s.rows_N != N || s.rows_R != R || s.rows_mask != s.mask) {
        build_schedule_rows(N, s.mask, R, s.sch);
        s.rows_mask = s.mask, s.rows_R = R, s.rows_N = N;
      }
      std::vector<uint2> tab;
      rows_step_table(s.sch, R, tab);
      std::vector<uint32_t> abd;  // the frame tables follow the step table's order
      rows_device_order(s.sch, R, abd);
      if ((rc = up(s.ab, abd.data(), abd.size() * 4)) ||
          (rc = up(s.steps, tab.data(), tab.size() * sizeof(uint2))) ||
          (rc = s.diag.alloc((size_t)nfr * N * 8)) || (rc = s.cs.alloc((size_t)nfr * s.sch.K * 8)))
        return rc;
      if ((rc = launch_frame_tables(s.src, N, s.sch, s.ab.as<uint32_t>(), trotter_factor(non->deltat, 1, non->ts),
                                    s.f_lo, nfr, s.diag.as<float2>(), s.cs.as<float2>(), st)))
        return rc;
      s.nwg = s.nvec;  // partial rows: one per vector
      if ((rc = s.partial.alloc((size_t)s.nwg * s.T * sizeof(double2)))) return rc;
      RowsArgs r;
      r.N = N, r.tmax = s.T, r.m = non->ts, r.nlayers = s.sch.nlayers(), r.K = s.sch.K, r.f_lo = s.f_lo;
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
  if (!rows.empty() && (rc = rows_prepare(rows.data(), (int)rows.size(), R, rmap, rsmem))) return rc;
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
  if (!rmap.empty() && (rc = launch_rows(rsys.as<RowsArgs>(), rwm.as<int2>(), (int)rmap.size(), R, rsmem, st, tm)))
    return rc;
  if (!dmap.empty()) {
    NISE_HIP_CHECK(hipFuncSetAttribute((const void *)dia_conveyor_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)dsmem));
    hipLaunchKernelGGL(dia_conveyor_kernel, dim3((unsigned)dmap.size()), dim3(256), dsmem, sd,
                       dsys.as<DiaSys>(), dwm.as<int2>(), tm + 2 * rmap.size());
    NISE_HIP_CHECK(hipGetLastError());
  }
  // the reductions and copies below wait for the DIA stream
  std::vector<int> sts((size_t)soff[nsys]);
  if (!sts.empty()) NISE_HIP_CHECK(hipMemcpyAsync(sts.data(), stt.p, sts.size() * 4, hipMemcpyDeviceToHost, sd));
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
  for (int k = 0; k < nsys; ++k)
    for (int64_t i = soff[k]; i < soff[k + 1]; ++i)
      if (sts[i] < 0) return set_error("multi: Jacobi did not converge (system %d)", k), NISE_ERR_NUMERIC;
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
