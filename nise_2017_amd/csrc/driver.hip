// File-level drop-ins: absorption() (src/absorption.c:18-202), do_1DFFT()
// (src/1DFFT.c:144-274) and calc_S1() (src/absorption.c:205-212), plus the
// GROBIN readers they share with calc_2DIR().
//
// absorption() reads the frames its samples need from the files named in
// t_non into pinned host memory, runs the batched frame-major kernels and
// writes the reference's output files.  do_1DFFT() applies the
// apodisation and the discrete Fourier transform in double precision on the
// GPU and writes the file on the host.  Messages and exit codes on bad input
// follow the reference.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "nise_internal.h"

namespace nise {

[[noreturn]] static void die(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vprintf(fmt, ap);
  va_end(ap);
  fflush(stdout);
  exit(code);
}

// src/hamshift.c:13-47: "count" then count lines "site a b".
void read_shift_file(t_non *non) {
  if (non->singleShiftFName[0] == '\0') {
    non->SingleShiftSites = 0;
    return;
  }
  FILE *f = fopen(non->singleShiftFName, "r");
  if (!f) {
    printf("Single shift file %s not found!\n", non->singleShiftFName);
    return;
  }
  if (fscanf(f, "%d", &non->SingleShiftSites) != 1)
    die(-1, "Could not read number of single shifts site from %s\n", non->singleShiftFName);
  non->SingleShiftSite = (int *)calloc(non->SingleShiftSites, sizeof(int));
  non->SingleShift = (float *)calloc(2 * non->SingleShiftSites, sizeof(float));
  for (int i = 0; i < non->SingleShiftSites; i++)
    if (fscanf(f, "%d %f %f", &non->SingleShiftSite[i], &non->SingleShift[i * 2],
               &non->SingleShift[i * 2 + 1]) != 3)
      break;
  fclose(f);
}

// Read records [first, first+count) of a GROBIN file into dst (pinned).
// Returns false if the file holds fewer records.
bool read_records(const char *name, int64_t rec_words, int64_t first, int64_t count, float *dst) {
  FILE *f = fopen(name, "rb");
  if (!f) return false;
  if (fseeko(f, (off_t)(first * rec_words * 4), SEEK_SET) != 0) {
    fclose(f);
    return false;
  }
  const size_t want = (size_t)(count * rec_words);
  const size_t got = fread(dst, 4, want, f);
  fclose(f);
  return got == want;
}

static bool is_coupling(const t_non *non) { return strcmp(non->hamiltonian, "Coupling") == 0; }


// ---------------------------------------------------------------------------
// do_1DFFT on the device
// ---------------------------------------------------------------------------
// in[i] = (im_S[i]/samples, re_S[i]/samples) x apodisation, in[0] halved;
// out[k] = sum_n in[n] exp(-2 pi i n k / fft) with the index product
// reduced modulo fft so the twiddle argument stays exact.
__global__ void fft1d_kernel(const float *re_S, const float *im_S, int tmax1, int tmax, int fft,
                             int samples, float deltat, float lifetime, float homogen,
                             float inhomogen, int window, float *spec_r, float *spec_i) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= fft) return;
  double xr = 0, xi = 0;
  for (int n = 0; n < tmax1; ++n) {
    // src/1DFFT.c:176-197 (float divisions, double exponentials)
    double ar = im_S[n] / (float)samples, ai = re_S[n] / (float)samples;
    double fac = 1.0;
    if (lifetime > 0.0f) fac *= exp((double)(-n * deltat / (2 * lifetime)));
    if (homogen > 0.0f) fac *= exp((double)(-n * deltat / (2 * homogen)));
    if (inhomogen > 0.0f)
      fac *= exp((double)(-n * deltat * n * deltat / (2 * inhomogen * inhomogen)));
    if (window == 1) {
      const double c = cos((double)(n * kTwoPi / (4 * tmax - 4)));
      fac *= c * c;
    }
    ar *= fac, ai *= fac;
    if (n == 0) ar *= 0.5, ai *= 0.5;
    const int64_t j = ((int64_t)n * k) % fft;
    const double th = 2.0 * M_PI * (double)j / (double)fft;
    const double c = cos(th), s = sin(th);
    xr += ar * c + ai * s;
    xi += ai * c - ar * s;
  }
  // src/1DFFT.c:203-206: spec_r = Im(out) 2 dt c_v, spec_i = Re(out) 2 dt c_v
  spec_r[k] = (float)(xi * 2 * (double)deltat * (double)kCv);
  spec_i[k] = (float)(xr * 2 * (double)deltat * (double)kCv);
}

int fft1d_device(const t_non *non, const float *re_S, const float *im_S, int samples, int fft,
                 std::vector<float> &sr, std::vector<float> &si) {
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t st = default_stream();
  DevBuf in, out;
  const int T = non->tmax1;
  if ((rc = in.alloc((size_t)2 * T * 4)) || (rc = out.alloc((size_t)2 * fft * 4))) return rc;
  NISE_HIP_CHECK(hipMemcpy(in.p, re_S, T * 4, hipMemcpyHostToDevice));
  NISE_HIP_CHECK(hipMemcpy(in.as<float>() + T, im_S, T * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(fft1d_kernel, dim3((fft + 127) / 128), dim3(128), 0, st, in.as<float>(),
                     in.as<float>() + T, T, non->tmax, fft, samples, non->deltat, non->lifetime,
                     non->homogen, non->inhomogen, non->window, out.as<float>(),
                     out.as<float>() + fft);
  NISE_HIP_CHECK(hipGetLastError());
  sr.resize(fft), si.resize(fft);
  NISE_HIP_CHECK(hipMemcpyAsync(sr.data(), out.p, fft * 4, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipMemcpyAsync(si.data(), out.as<float>() + fft, fft * 4, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  return NISE_OK;
}

// ---------------------------------------------------------------------------
// calc_S1 on the device: the reference's running FP32 sum, term by term.
// ---------------------------------------------------------------------------
__global__ void calc_s1_kernel(float *acc, const float *cr, const float *ci, const float *mu, int N) {
  if (threadIdx.x != 0) return;
  float r = acc[0], q = acc[1];
  for (int i = 0; i < N; ++i) {
    r = __fadd_rn(r, __fmul_rn(mu[i], cr[i]));
    q = __fadd_rn(q, __fmul_rn(mu[i], ci[i]));
  }
  acc[0] = r, acc[1] = q;
}

}  // namespace nise

using namespace nise;

extern "C" void calc_S1(float *re_S_1, float *im_S_1, int t1, t_non *non, float *cr, float *ci,
                        float *mu) {
  const int N = non->singles;
  static thread_local DevBuf buf;
  std::vector<float> h(2 + 3 * (size_t)N);
  h[0] = re_S_1[t1], h[1] = im_S_1[t1];
  std::memcpy(h.data() + 2, cr, N * 4);
  std::memcpy(h.data() + 2 + N, ci, N * 4);
  std::memcpy(h.data() + 2 + 2 * N, mu, N * 4);
  if (ensure_device() || buf.alloc(h.size() * 4) ||
      hipMemcpy(buf.p, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    die(1, "calc_S1: GPU unavailable: %s\n", get_error());
  float *d = buf.as<float>();
  hipLaunchKernelGGL(calc_s1_kernel, dim3(1), dim3(64), 0, default_stream(), d, d + 2, d + 2 + N,
                     d + 2 + 2 * N, N);
  if (hipMemcpyAsync(h.data(), buf.p, 8, hipMemcpyDeviceToHost, default_stream()) != hipSuccess ||
      hipStreamSynchronize(default_stream()) != hipSuccess)
    die(1, "calc_S1: GPU call failed\n");
  re_S_1[t1] = h[0], im_S_1[t1] = h[1];
}

// src/1DFFT.c:144-274 (projection segments handled as in the reference:
// pro_dim consecutive blocks of tmax values).
extern "C" void do_1DFFT(t_non *non, char fname[], float *re_S_1, float *im_S_1, int samples) {
  const float shift1 = non->shifte;
  int fft = non->fft;
  if (fft < non->tmax1) fft = non->tmax1;
  int pro_dim = 1;
  if (non->Npsites == non->singles && non->Npsites > 0) {
    int mx = 0;
    for (int i = 0; i < non->singles; ++i) mx = non->psites[i] > mx ? non->psites[i] : mx;
    pro_dim = mx + 1;
  }
  std::vector<std::vector<float>> SR(pro_dim), SI(pro_dim);
  for (int ip = 0; ip < pro_dim; ++ip)
    if (fft1d_device(non, re_S_1 + non->tmax * ip, im_S_1 + non->tmax * ip, samples, fft, SR[ip], SI[ip]))
      die(1, "do_1DFFT: GPU transform failed: %s\n", get_error());
  FILE *out = fopen(fname, "w");
  if (!out) die(1, "Could not open %s\n", fname);
  for (int i = fft / 2; i <= fft - 1; i++) {
    const float w = -((fft - i) / non->deltat / kCv / fft - shift1);
    if (w > non->min1 && w < non->max1) {
      fprintf(out, "%f ", w);
      for (int ip = 0; ip < pro_dim; ++ip) fprintf(out, "%e %e ", SR[ip][i], SI[ip][i]);
      fprintf(out, "\n");
    }
  }
  for (int i = 0; i <= fft / 2 - 1; i++) {
    const float w = -((-i) / non->deltat / kCv / fft - shift1);
    if (w > non->min1 && w < non->max1) {
      fprintf(out, "%f ", w);
      for (int ip = 0; ip < pro_dim; ++ip) fprintf(out, "%e %e ", SR[ip][i], SI[ip][i]);
      fprintf(out, "\n");
    }
  }
  fclose(out);
}

// ---------------------------------------------------------------------------
// Streamed absorption: the samples in windows, each window's frames read and
// uploaded while the previous window runs
// ---------------------------------------------------------------------------
namespace {
struct StreamWindow {
  int b = 0, e = 0;            // samples [b, e)
  int64_t f_lo = 0, nfr = 0;   // their frames
};

// Frames of a window and the bytes of one frame's records.
int64_t window_frames(int samples, int sample, int tmax) { return (int64_t)(samples - 1) * sample + tmax; }
}  // namespace

extern "C" int nise_absorption_stream(const t_non *non, int window_samples, float *re_S, float *im_S) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!non || !re_S || !im_S) return set_error("stream: null argument"), NISE_ERR_ARG;
  const int N = non->singles, T = non->tmax1, sr = non->sample;
  if (N < 1 || T < 1 || sr < 1) return set_error("stream: bad sizes"), NISE_ERR_ARG;
  const bool ctype = is_coupling(non);
  if (!ctype && strcmp(non->hamiltonian, "Full") != 0)
    return set_error("stream: HamiltonianType %s is not streamed (Full and Coupling are)", non->hamiltonian),
           NISE_ERR_ARG;
  const int begin = non->begin, end = non->end;
  if (end <= begin) return set_error("stream: no samples (begin %d, end %d)", begin, end), NISE_ERR_RANGE;
  const int64_t D = non->doubles;
  const int64_t hw = ctype ? 1 + N : 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2;
  const int64_t mw = 1 + 3 * (int64_t)N + 3 * (int64_t)N * D;
  const int64_t cw = 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2;
  // window: the given sample count, else what a quarter of the free device
  // memory holds in the two frame buffers and the run's per-frame tables
  int W = window_samples;
  if (W <= 0) {
    size_t fr = 0, tot = 0;
    NISE_HIP_CHECK(hipMemGetInfo(&fr, &tot));
    const double per_frame = 4.0 * (2 * hw + (ctype ? 0 : 2 * mw)) + 8.0 * N * (N + 1);
    const int64_t F = (int64_t)((double)fr / 4 / per_frame);
    W = F < T ? 1 : (int)std::min<int64_t>(end - begin, (F - T) / sr + 1);
  }
  W = std::max(1, std::min(W, end - begin));
  std::vector<StreamWindow> win;
  for (int b = begin; b < end; b += W) {
    StreamWindow w;
    w.b = b, w.e = std::min(end, b + W);
    w.f_lo = (int64_t)b * sr, w.nfr = window_frames(w.e - b, sr, T);
    win.push_back(w);
  }
  const int64_t fmax = window_frames(W, sr, T);
  // static dipoles and couplings of the Coupling type, once
  Pinned hmu[2], hH[2], hC, hmu0;
  DevBuf dH[2], dmu[2], dC, dmu0;
  for (int i = 0; i < 2 && i < (int)win.size(); ++i) {
    if (!hH[i].alloc((size_t)(fmax * hw)) || (!ctype && !hmu[i].alloc((size_t)(fmax * mw))))
      return set_error("stream: could not allocate pinned host memory"), NISE_ERR_NOMEM;
    if ((rc = dH[i].alloc((size_t)(fmax * hw * 4))) || (!ctype && (rc = dmu[i].alloc((size_t)(fmax * mw * 4)))))
      return rc;
  }
  if (ctype) {
    if (!hC.alloc((size_t)cw) || !hmu0.alloc((size_t)mw))
      return set_error("stream: could not allocate pinned host memory"), NISE_ERR_NOMEM;
    if (!read_records(non->couplingFName, cw, 0, 1, hC.p))
      return set_error("stream: coupling file %s missing or too short", non->couplingFName), NISE_ERR_RANGE;
    if (!read_records(non->dipoleFName, mw, 0, 1, hmu0.p))
      return set_error("stream: dipole file %s missing or too short", non->dipoleFName), NISE_ERR_RANGE;
    if ((rc = dC.alloc((size_t)cw * 4)) || (rc = dmu0.alloc((size_t)mw * 4))) return rc;
    NISE_HIP_CHECK(hipMemcpy(dC.p, hC.p, (size_t)cw * 4, hipMemcpyHostToDevice));
    NISE_HIP_CHECK(hipMemcpy(dmu0.p, hmu0.p, (size_t)mw * 4, hipMemcpyHostToDevice));
  }
  hipStream_t cp = nullptr;
  NISE_HIP_CHECK(hipStreamCreate(&cp));
  // read window k into buffer k % 2 and upload it on the copy stream; the
  // message of a failure is returned, not set (it runs on a reader thread)
  auto fetch = [&](int k, std::string &err) {
    const StreamWindow &w = win[k];
    const int i = k % 2;
    if (!read_records(non->energyFName, hw, w.f_lo, w.nfr, hH[i].p)) {
      err = std::string("Hamiltonian trajectory file ") + non->energyFName + " too short for frames up to " +
            std::to_string(w.f_lo + w.nfr);
      return;
    }
    if (!ctype && !read_records(non->dipoleFName, mw, w.f_lo, w.nfr, hmu[i].p)) {
      err = std::string("Dipole trajectory file ") + non->dipoleFName + " too short for frames up to " +
            std::to_string(w.f_lo + w.nfr);
      return;
    }
    if (hipMemcpyAsync(dH[i].p, hH[i].p, (size_t)(w.nfr * hw * 4), hipMemcpyHostToDevice, cp) != hipSuccess ||
        (!ctype &&
         hipMemcpyAsync(dmu[i].p, hmu[i].p, (size_t)(w.nfr * mw * 4), hipMemcpyHostToDevice, cp) != hipSuccess) ||
        hipStreamSynchronize(cp) != hipSuccess)
      err = "upload of a window failed";
  };
  std::vector<double> acc_r(T, 0.0), acc_i(T, 0.0);
  std::vector<float> wr(T), wi(T);
  std::string err;
  fetch(0, err);
  for (size_t k = 0; k < win.size() && err.empty(); ++k) {
    std::string next_err;
    std::thread reader;
    if (k + 1 < win.size()) reader = std::thread([&, k] { fetch((int)k + 1, next_err); });
    const StreamWindow &w = win[k];
    const int i = (int)(k % 2);
    t_non loc = *non;  // the window's records start at its first frame
    loc.begin = 0, loc.end = w.e - w.b;
    const int r = nise_absorption_batched(&loc, dH[i].as<float>(), hw, ctype ? dmu0.as<float>() : dmu[i].as<float>(),
                                          mw, w.nfr, ctype ? dC.as<float>() : nullptr, NISE_DEVICE, 0, wr.data(),
                                          wi.data());
    if (reader.joinable()) reader.join();
    if (r) {
      (void)hipStreamDestroy(cp);
      return r;
    }
    for (int t = 0; t < T; ++t) acc_r[t] += wr[t], acc_i[t] += wi[t];
    err = next_err;
  }
  (void)hipStreamDestroy(cp);
  if (!err.empty()) return set_error("stream: %s", err.c_str()), NISE_ERR_RANGE;
  for (int t = 0; t < T; ++t) re_S[t] = (float)acc_r[t], im_S[t] = (float)acc_i[t];
  return NISE_OK;
}

static void write_absorption(t_non *non, float *re, float *im);

// src/absorption.c:18-202
extern "C" void absorption(t_non *non) {
  const int N = non->singles;
  const float shift1 = (non->max1 + non->min1) / 2;  // :57-59
  printf("Frequency shift %f.\n", shift1);
  non->shifte = shift1;
  if (non->Npsites != 0) die(1, "absorption: the GPU path does not implement Projection\n");
  if (non->cluster != -1) die(1, "absorption: the GPU path does not implement Cluster\n");
  const bool ctype = is_coupling(non);
  // on-the-fly couplings from positions (src/read_trajectory.c:181-303)
  const int onfly = strcmp(non->hamiltonian, "TransitionDipole") == 0 ? 1
                  : strcmp(non->hamiltonian, "ExtendedDipole") == 0 ? 2 : 0;
  if (!ctype && !onfly && strcmp(non->hamiltonian, "Full") != 0)
    die(1, "absorption: HamiltonianType %s is not provided by the GPU path\n", non->hamiltonian);
  if (onfly && non->propagation != 1)
    die(1, "absorption: %s couplings run with propagation Coupling on the GPU path\n", non->hamiltonian);
  // the files are opened first (src/absorption.c:76, open_files
  // src/read_trajectory.c:15-27), then the samples counted
  FILE *probe = fopen(non->energyFName, "rb");
  if (!probe) die(1, "Hamiltonian file not found!\n");
  fclose(probe);
  probe = fopen(non->dipoleFName, "rb");
  if (!probe) die(1, "Dipole file %s not found!\n", non->dipoleFName);
  fclose(probe);
  read_shift_file(non);
  // src/NISE_subs.c:197-221
  const int N_samples = (non->length - non->tmax1 - 1) / non->sample + 1;
  if (N_samples > 0) printf("Making %d samples!\n", N_samples);
  else die(1, "Insufficient data to calculate spectrum.\nPlease, lower max times or provide longer\ntrajectory.\n");
  if (non->end == 0) non->end = N_samples;
  if (non->end > N_samples)
    die(0, "Endpoint larger than number of samples was specified.\nEndpoint was %d but cannot be larger than %d.\n",
        non->end, N_samples);
  const int64_t D = non->doubles;
  const int64_t hw = (ctype || onfly) ? 1 + N : 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2;
  const int64_t mw = 1 + 3 * (int64_t)N + 3 * (int64_t)N * D;
  const int64_t f_lo = (int64_t)non->begin * non->sample;
  const int64_t f_hi = (int64_t)(non->end - 1) * non->sample + non->tmax1;
  const int64_t nfr = f_hi - f_lo;
  // a frame range beyond a third of the free device memory runs in windows
  // of samples, streamed from the files (Full and Coupling types)
  size_t dfree = 0, dtot = 0;
  if (!onfly && hipMemGetInfo(&dfree, &dtot) == hipSuccess &&
      (double)nfr * (hw + (ctype ? 0 : mw)) * 4 > (double)dfree / 3) {
    std::vector<float> re(non->tmax1), im(non->tmax1);
    if (nise_absorption_stream(non, 0, re.data(), im.data()))
      die(1, "absorption: GPU run failed: %s\n", nise_last_error());
    write_absorption(non, re.data(), im.data());
    return;
  }
  Pinned H, mu, Cc;
  if (!H.alloc((size_t)(nfr * hw)) || !mu.alloc((size_t)((ctype ? 1 : nfr) * mw)))
    die(1, "absorption: could not allocate pinned host memory\n");
  if (!read_records(non->energyFName, hw, f_lo, nfr, H.p))
    die(1, "Hamiltonian trajectory file to short, could not fill buffer!!!\n");
  if (!read_records(non->dipoleFName, mw, ctype ? 0 : f_lo, ctype ? 1 : nfr, mu.p))
    die(1, "Dipole trajectory file to short, could not fill buffer!!!\n");
  if (ctype) {
    if (!Cc.alloc((size_t)(1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2)))
      die(1, "absorption: could not allocate pinned host memory\n");
    probe = fopen(non->couplingFName, "rb");
    if (!probe) die(1, "Coupling file not found!\n");
    fclose(probe);
    if (!read_records(non->couplingFName, 1 + (int64_t)N * (N + 1) / 2 + D * (D + 1) / 2, 0, 1, Cc.p))
      die(1, "Coupling trajectory file to short, could not fill buffer!!!\n");
  }
  Pinned pos, box;
  const bool pbc = strlen(non->pbcFName) > 0;
  if (onfly) {  // positions: read_mue records, two per frame for ExtendedDipole; boxes: int + 3 floats
    const int64_t per = onfly == 2 ? 2 : 1;
    probe = fopen(non->positionFName, "rb");
    if (!probe) die(1, "Atom position file for TDC on the fly not found!\n");
    fclose(probe);
    if (!pos.alloc((size_t)(per * nfr * mw))) die(1, "absorption: could not allocate pinned host memory\n");
    if (!read_records(non->positionFName, mw, per * f_lo, per * nfr, pos.p))
      die(1, "Dipole trajectory file to short, could not fill buffer!!!\n");
    if (pbc) {
      probe = fopen(non->pbcFName, "rb");
      if (!probe) die(1, "Periodic Boundary Condition file not found!\n");
      fclose(probe);
      if (!box.alloc((size_t)(nfr * 4)) || !read_records(non->pbcFName, 4, f_lo, nfr, box.p))
        die(1, "Periodic Boundary Condition file to short, could not fill buffer!!!\n");
    }
  }
  std::vector<float> re(non->tmax1), im(non->tmax1);
  // the records start at frame f_lo: the run counts frames from there
  t_non loc = *non;
  loc.begin = 0, loc.end = non->end - non->begin;
  int rc;
  if (onfly) {
    nise_abs_plan *plan = nullptr;
    rc = nise_abs_plan_create_onfly(&plan, &loc, H.p, hw, mu.p, mw, nfr, pos.p, mw, pbc ? box.p : nullptr,
                                    NISE_HOST, 0);
    if (!rc) rc = nise_abs_plan_execute(plan, re.data(), im.data(), 0);
    if (plan) nise_abs_plan_destroy(plan);
  } else {
    rc = nise_absorption_batched(&loc, H.p, hw, mu.p, mw, nfr, ctype ? Cc.p : nullptr, NISE_HOST, 0, re.data(),
                                 im.data());
  }
  if (rc) die(1, "absorption: GPU run failed: %s\n", nise_last_error());
  write_absorption(non, re.data(), im.data());
}

// the output files of src/absorption.c:157-201
static void write_absorption(t_non *non, float *re, float *im) {
  const int samples = non->end;  // src/absorption.c:185,192: loop exit value
  FILE *log = fopen("NISE.log", "a");
  if (log) {
    fprintf(log, "Finished Calculating Response!\nWriting to file!\n");
    fclose(log);
  }
  FILE *out = fopen("TD_Absorption.dat", "w");
  if (!out) die(1, "Could not open TD_Absorption.dat\n");
  for (int t1 = 0; t1 < non->tmax1; t1 += non->dt1) {
    fprintf(out, "%f ", t1 * non->deltat);
    fprintf(out, "%e %e ", re[t1] / samples, im[t1] / samples);
    fprintf(out, "\n");
  }
  fclose(out);
  do_1DFFT(non, (char *)"Absorption.dat", re, im, samples);
  printf("----------------------------------------------\n");
  printf(" Absorption calculation succesfully completed\n");
  printf("----------------------------------------------\n\n");
}
