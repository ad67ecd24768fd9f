// The 2DFFT post-processing program (src/2DFFT.c:20-830) as an in-process
// entry point: response files R{par,per,cro}{I,II}.dat (written by print2D,
// src/MPI_subs.c:20-34) -> apodised, edge-halved, zero-padded fft x fft
// grids -> forward 2D DFT in double precision on the GPU -> Rw*.I/II.dat
// -> rephasing and non-rephasing sums 2D.{par,per,cro}.dat and the
// broadband pump-probe projections PP.*.dat.
//
// The file handling follows the program step by step, including its text
// round trips (the frequency files are written with %e and read back as
// float to form the sums), its FP32 arithmetic on times, axes and sums in
// source order, the waiting-time loop over Waitingtime.dat and the 2DIRraman
// file names.  The transform is a direct separable DFT with the index
// product reduced modulo fft, so every twiddle is one of fft exact table
// values; FFTW's forward transform computes the same sums with a different
// rounding (~1e-15 relative).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "nise_internal.h"

namespace nise {

// X[p][a][j] = sum_{b < ncols} in[p][a][b] tw[(b j) mod fft], rows a < nrows
__global__ void dft2d_cols_kernel(const double2 *in, int fft, int nrows, int ncols, const double2 *tw,
                                  double2 *X) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y, p = blockIdx.z;
  if (j >= fft) return;
  const double2 *row = in + ((int64_t)p * fft + a) * fft;
  double xr = 0.0, xi = 0.0;
  int m = 0;  // (b j) mod fft
  for (int b = 0; b < ncols; ++b) {
    const double2 v = row[b], w = tw[m];
    xr += v.x * w.x - v.y * w.y;
    xi += v.x * w.y + v.y * w.x;
    m += j;
    if (m >= fft) m -= fft;
  }
  X[((int64_t)p * nrows + a) * fft + j] = make_double2(xr, xi);
}

// out[p][i][j] = sum_{a < nrows} X[p][a][j] tw[(a i) mod fft]
__global__ void dft2d_rows_kernel(const double2 *X, int fft, int nrows, const double2 *tw, double2 *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, p = blockIdx.z;
  if (j >= fft) return;
  double xr = 0.0, xi = 0.0;
  int m = 0;  // (a i) mod fft
  for (int a = 0; a < nrows; ++a) {
    const double2 v = X[((int64_t)p * nrows + a) * fft + j], w = tw[m];
    xr += v.x * w.x - v.y * w.y;
    xi += v.x * w.y + v.y * w.x;
    m += i;
    if (m >= fft) m -= fft;
  }
  out[((int64_t)p * fft + i) * fft + j] = make_double2(xr, xi);
}

}  // namespace nise

using namespace nise;

// Forward 2D DFT of P grids fft x fft (complex interleaved, host arrays):
// out[p][k0][k1] = sum in[p][j0][j1] exp(-2 pi i (j0 k0 + j1 k1) / fft), as
// FFTW_FORWARD.  Only rows < nrows and columns < ncols of the input may be
// nonzero.
extern "C" int nise_dft2d(const double *in, int P, int fft, int nrows, int ncols, double *out) {
  int rc = ensure_device();
  if (rc) return rc;
  if (P < 1 || P > 65535 || fft < 1 || fft > 65535 || nrows < 0 || nrows > fft || ncols < 0 || ncols > fft ||
      !in || !out)
    return set_error("dft2d: bad arguments"), NISE_ERR_ARG;
  hipStream_t st = default_stream();
  const size_t grid = (size_t)fft * fft;
  std::vector<double> twh(2 * (size_t)fft);
  for (int m = 0; m < fft; ++m) {
    const double th = 2.0 * M_PI * (double)m / (double)fft;
    twh[2 * m] = cos(th), twh[2 * m + 1] = -sin(th);
  }
  DevBuf din, dX, dout, dtw;
  const int nr = nrows > 0 ? nrows : 1;
  if ((rc = din.alloc((size_t)P * grid * 16)) || (rc = dX.alloc((size_t)P * nr * fft * 16)) ||
      (rc = dout.alloc((size_t)P * grid * 16)) || (rc = dtw.alloc((size_t)fft * 16)))
    return rc;
  NISE_HIP_CHECK(hipMemcpyAsync(din.p, in, (size_t)P * grid * 16, hipMemcpyHostToDevice, st));
  NISE_HIP_CHECK(hipMemcpyAsync(dtw.p, twh.data(), (size_t)fft * 16, hipMemcpyHostToDevice, st));
  if (nrows == 0) {
    NISE_HIP_CHECK(hipMemsetAsync(dout.p, 0, (size_t)P * grid * 16, st));
  } else {
    const dim3 blk(128), g1((fft + 127) / 128, nrows, P), g2((fft + 127) / 128, fft, P);
    hipLaunchKernelGGL(dft2d_cols_kernel, g1, blk, 0, st, din.as<double2>(), fft, nrows, ncols,
                       dtw.as<double2>(), dX.as<double2>());
    NISE_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(dft2d_rows_kernel, g2, blk, 0, st, dX.as<double2>(), fft, nrows, dtw.as<double2>(),
                       dout.as<double2>());
    NISE_HIP_CHECK(hipGetLastError());
  }
  NISE_HIP_CHECK(hipMemcpyAsync(out, dout.p, (size_t)P * grid * 16, hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  return NISE_OK;
}

namespace {

constexpr float kTwoPi2d = 2 * 3.14159265359;  // src/types.h twoPi
constexpr float kCv2d = 2.99792458e-5;         // src/2DFFT.c:46 c_v

// the program's round() macro (src/2DFFT.c:11): 0.5 is a double
long round_ref(float x) { return x > 0 ? (long)(x + 0.5) : (long)(x - 0.5); }

struct Grid {  // one response function transformed
  bool present = false;
  std::vector<double> in;  // [fft][fft] complex interleaved
  int maxrow = -1, maxcol = -1;
};

// Read one print2D file into the zero-padded grid (src/2DFFT.c:318-362 for
// kI, :466-515 for kII).  kII adds into the grid and takes its inhomogeneous
// factor on (t1 + t3).
void read_response(FILE *input, const t_non *non, int fft, bool kII, Grid &g) {
  g.in.assign(2 * (size_t)fft * fft, 0.0);
  const float deltat = non->deltat, homo = non->homogen, inhomo = non->inhomogen;
  float ti[4] = {0, 0, 0, 0}, rr = 0, ir = 0;
  for (int i = 0; i < non->tmax1; i++) {
    for (int k = 0; k < non->tmax3; k++) {
      if (fscanf(input, "%f %f %f %e %e", &ti[1], &ti[2], &ti[3], &rr, &ir) != 5) {
        // the program does not check: the previous values are used again
      }
      // kI: (fft t1) / deltat; kII: fft (t1 / deltat) (src/2DFFT.c:335, :484-488)
      const long index = kII ? round_ref(ti[3] / (deltat * 1) + fft * (ti[1] / (deltat * 1)))
                             : round_ref(ti[3] / (deltat * 1) + fft * ti[1] / (deltat * 1));
      if (homo > 0.0f) {
        const float x = -(ti[1] + ti[3]) / 2 / homo;
        rr = (float)(rr * exp((double)x));
        ir = (float)(ir * exp((double)x));
      }
      if (inhomo > 0.0f) {
        const float d = kII ? ti[1] + ti[3] : ti[1] - ti[3];
        const float x = -d * d / 2 / inhomo / inhomo;
        rr = (float)(rr * exp((double)x));
        ir = (float)(ir * exp((double)x));
      }
      if (non->window == 1) {
        const double c1 = cos((double)(ti[1] * kTwoPi2d / (float)(4 * non->tmax1 - 4) / deltat));
        const double c3 = cos((double)(ti[3] * kTwoPi2d / (float)(4 * non->tmax3 - 4) / deltat));
        rr = (float)(rr * c1 * c1);
        rr = (float)(rr * c3 * c3);
        ir = (float)(ir * c1 * c1);
        ir = (float)(ir * c3 * c3);
      }
      if (i == 0) rr = rr / 2, ir = ir / 2;
      if (k == 0) rr = rr / 2, ir = ir / 2;
      if (index < (long)fft * fft && index >= 0) {
        double *c = &g.in[2 * (size_t)index];
        if (kII) c[0] += rr, c[1] += ir;
        else c[0] = rr, c[1] = ir;
        const int row = (int)(index / fft), col = (int)(index % fft);
        if (row > g.maxrow) g.maxrow = row;
        if (col > g.maxcol) g.maxcol = col;
      }
    }
  }
}

// Transform the present grids on the device; results replace g.in.
int transform(std::vector<Grid *> &gs, int fft) {
  int P = 0, nrows = 0, ncols = 0;
  for (Grid *g : gs)
    if (g->present) {
      ++P;
      nrows = std::max(nrows, g->maxrow + 1);
      ncols = std::max(ncols, g->maxcol + 1);
    }
  if (P == 0) return NISE_OK;
  const size_t grid = 2 * (size_t)fft * fft;
  std::vector<double> in((size_t)P * grid), out((size_t)P * grid);
  int p = 0;
  for (Grid *g : gs)
    if (g->present) std::memcpy(in.data() + (size_t)p++ * grid, g->in.data(), grid * 8);
  int rc = nise_dft2d(in.data(), P, fft, nrows, ncols, out.data());
  if (rc) return rc;
  p = 0;
  for (Grid *g : gs)
    if (g->present) std::memcpy(g->in.data(), out.data() + (size_t)p++ * grid, grid * 8);
  return NISE_OK;
}

struct Names {
  std::string time[2][3], freq[2][3], twoD[3], twoD2[3], pp[3], pp2[3];
};

// File names of src/2DFFT.c:292-305, :442-456, :659-683.
Names names_for(bool raman, bool wfile, const char *wt) {
  static const char *pn[3] = {"par", "per", "cro"};
  Names n;
  char b[512];
  for (int pol = 0; pol < 3; ++pol) {
    for (int k = 0; k < 2; ++k) {
      const char *kk = k == 0 ? "I" : "II";
      if (raman) {
        snprintf(b, sizeof b, "R%s%s.IRraman.dat", pn[pol], kk), n.time[k][pol] = b;
        snprintf(b, sizeof b, "Rw%sIRraman.%s.dat", pn[pol], kk), n.freq[k][pol] = b;
      } else if (wfile) {
        snprintf(b, sizeof b, "R%s%s_%sfs.dat", pn[pol], kk, wt), n.time[k][pol] = b;
        snprintf(b, sizeof b, "Rw%s.%s.dat", pn[pol], kk), n.freq[k][pol] = b;
      } else {
        snprintf(b, sizeof b, "R%s%s.dat", pn[pol], kk), n.time[k][pol] = b;
        snprintf(b, sizeof b, "Rw%s.%s.dat", pn[pol], kk), n.freq[k][pol] = b;
      }
    }
    if (raman) {
      snprintf(b, sizeof b, "2DIRraman.I.%s.dat", pn[pol]), n.twoD[pol] = b;
      snprintf(b, sizeof b, "2DIRraman.II.%s.dat", pn[pol]), n.twoD2[pol] = b;
      snprintf(b, sizeof b, "PPIRraman.I.%s.dat", pn[pol]), n.pp[pol] = b;
      snprintf(b, sizeof b, "PPIRraman.II.%s.dat", pn[pol]), n.pp2[pol] = b;
    } else if (wfile) {
      snprintf(b, sizeof b, "2D.%s_%sfs.dat", pn[pol], wt), n.twoD[pol] = b;
      snprintf(b, sizeof b, "PP.%s_%sfs.dat", pn[pol], wt), n.pp[pol] = b;
    } else {
      snprintf(b, sizeof b, "2D.%s.dat", pn[pol]), n.twoD[pol] = b;
      snprintf(b, sizeof b, "PP.%s.dat", pn[pol]), n.pp[pol] = b;
    }
  }
  return n;
}

// The point at grid index (i, j) of the transform, and its output.
void put(FILE *out, int form, float w1, float w3, const double *v) {
  if (form == 0 || form == 2) fprintf(out, "%f %f %e %e\n", w1, w3, v[0], v[1]);
  if (form == 1) fprintf(out, "%e ", v[1]);
}

// src/2DFFT.c:367-439: kI written over w1 in (-max1, -min1), w3 in (min3, max3).
void write_kI(FILE *out, const Grid &g, int fft, float deltat, float shift1, float shift3, const t_non *non,
              int form) {
  auto w1_of = [&](int i) { return i >= fft / 2 ? (-fft + i) / deltat / kCv2d / fft - shift1 : i / deltat / kCv2d / fft - shift1; };
  auto w3_of = [&](int j) { return j < fft / 2 ? j / deltat / kCv2d / fft - shift3 : (-fft + j) / deltat / kCv2d / fft - shift3; };
  auto row = [&](int i) {
    float w1 = 0.f;
    for (int half = 0; half < 2; ++half) {
      const int j0 = half == 0 ? fft / 2 : 0, j1 = half == 0 ? fft : fft / 2;
      for (int j = j0; j < j1; j++) {
        w1 = w1_of(i);
        const float w3 = w3_of(j);
        if (w1 > -non->max1 && w3 < non->max3)
          if (w1 < -non->min1 && w3 > non->min3) put(out, form, w1, w3, &g.in[2 * ((size_t)j + (size_t)fft * i)]);
      }
    }
    if ((form == 1 || form == 2) && w1 > -non->max1 && w1 < -non->min1) fprintf(out, "\n");
  };
  for (int i = fft / 2; i < fft; i++) row(i);
  for (int i = 0; i < fft / 2; i++) row(i);
}

// src/2DFFT.c:519-626: kII written over w1 in (min1, max1), w3 in (min3, max3),
// tracking the extent of the written axes.
void write_kII(FILE *out, FILE *axis, const Grid &g, int fft, float deltat, float shift1, float shift3,
               float sw, const t_non *non, int form, bool raman, float &rwmi1, float &rwma1, float &rwmi3,
               float &rwma3) {
  auto w1_of = [&](int i) { return i >= fft / 2 ? i / deltat / kCv2d / fft + shift1 - sw : (fft + i) / deltat / kCv2d / fft + shift1 - sw; };
  auto w3_of = [&](int j) { return j < fft / 2 ? (fft + j) / deltat / kCv2d / fft + shift3 - sw : j / deltat / kCv2d / fft + shift3 - sw; };
  auto row = [&](int i, bool second) {
    float w1 = 0.f;
    for (int half = 0; half < 2; ++half) {
      const int j0 = half == 0 ? fft / 2 : 0, j1 = half == 0 ? fft : fft / 2;
      for (int j = j0; j < j1; j++) {
        w1 = w1_of(i);
        const float w3 = w3_of(j);
        if (w1 < non->max1 && w3 < non->max3) {
          if (w1 > non->min1 && w3 > non->min3) {
            if (w1 < rwmi1) rwmi1 = w1;
            if (w1 > rwma1) rwma1 = w1;
            if (w3 < rwmi3) rwmi3 = w3;
            if (w3 > rwma3) rwma3 = w3;
            // the program shifts w1 by shift3 in its third block (i < fft/2,
            // j >= fft/2) and by shift1 elsewhere
            if (raman) w1 = w1 + ((second && half == 0) ? shift3 : shift1);
            put(out, form, w1, w3, &g.in[2 * ((size_t)j + (size_t)fft * i)]);
          }
        }
      }
    }
    if ((form == 1 || form == 2) && w1 < non->max1 && w1 > non->min1) {
      fprintf(out, "\n");
      if (form == 1 && axis) fprintf(axis, "%f\n", w1);
    }
  };
  for (int i = fft / 2; i < fft; i++) row(i, false);
  for (int i = 0; i < fft / 2; i++) row(i, true);
}

typedef std::vector<std::vector<float>> Mat;  // [1..p][1..q]

// fscanf("%f %f %e %e") of p x q points into four matrices (src/2DFFT.c:680-712);
// entries past the end of the file keep zero.
void read_freq(FILE *in, int p, int q, Mat &a, Mat &b, Mat &c, Mat &d) {
  a.assign(p + 1, std::vector<float>(q + 1, 0.f)), b = a, c = a, d = a;
  for (int i = 1; i <= p; i++)
    for (int j = 1; j <= q; j++)
      if (fscanf(in, "%f %f %e %e", &a[i][j], &b[i][j], &c[i][j], &d[i][j]) != 4) {
      }
}

}  // namespace

// The 2DFFT program (src/2DFFT.c:20-830) in process, on the files of the
// current directory.  format: 0 Dislin / Python, 1 Matlab, 2 Gnuplot.
extern "C" int nise_2dfft(const t_non *non, int format) {
  const int form = (format == 1 || format == 2) ? format : 0;
  const float deltat = non->deltat;
  const int fft = non->fft;
  if (non->min1 < 0 || non->min2 < 0 || non->min3 < 0 || non->max1 < 0 || non->max2 < 0 || non->max3 < 0) {
    printf("All field frequencies must be positive!\n");
    return set_error("2dfft: negative field frequency"), NISE_ERR_ARG;
  }
  if (non->min1 > non->max1 || non->min2 > non->max2 || non->min3 > non->max3) {
    printf("The max frequency must be larger than the min frequency!\n");
    return set_error("2dfft: min frequency above max"), NISE_ERR_ARG;
  }
  if (fft <= 0) {
    printf("Fatal error.\nFFT keyword not specified!\n");
    return set_error("2dfft: FFT not set"), NISE_ERR_ARG;
  }
  const char *t = non->technique;
  const bool raman = !strcmp(t, "2DIRraman") || !strcmp(t, "2DIRraman1") || !strcmp(t, "2DIRraman2") ||
                     !strcmp(t, "2DIRraman3") || !strcmp(t, "2DIRramanI") || !strcmp(t, "2DIRramanII");
  // waiting-time loop (src/2DFFT.c:259-283)
  FILE *WTime = nullptr;
  if (access("Waitingtime.dat", F_OK) != -1) {
    printf("Waitingtime file found.\n");
    WTime = fopen("Waitingtime.dat", "r");
  }
  const bool wfile = WTime != nullptr;
  int rc = NISE_OK;
  for (;;) {
    char waittime[256] = "";
    if (wfile) {
      if (fscanf(WTime, "%255s", waittime) != 1) break;
      printf("Doing 2DFFT for %s fs\n", waittime);
    }
    const Names nm = names_for(raman, wfile, waittime);
    const float shift1 = 0.5 * (non->min1 + non->max1);
    float shift3 = -0.5 * (non->min3 + non->max3);
    // kI and kII: read all present files, transform on the device, write
    Grid g[2][3];
    float rwmi1 = shift1, rwma1 = shift1, rwmi3 = -shift3, rwma3 = -shift3;
    for (int k = 0; k < 2; ++k) {
      for (int pol = 0; pol < 3; ++pol) {
        FILE *input = fopen(nm.time[k][pol].c_str(), "r");
        if (!input) {
          printf("The response function file %s was not found!\n", nm.time[k][pol].c_str());
          printf("Warning: Skipping this response function polarization!\n");
          continue;
        }
        printf(k == 0 ? "Reading response function (kI) file!\n" : "Reading response function file (kII)!\n");
        read_response(input, non, fft, k == 1, g[k][pol]);
        g[k][pol].present = true;
        fclose(input);
      }
      std::vector<Grid *> gs = {&g[k][0], &g[k][1], &g[k][2]};
      if ((rc = transform(gs, fft))) break;
      if (k == 1) shift3 = -shift3;
      const float sw = 1 / deltat / kCv2d;
      FILE *axisFH = nullptr;
      for (int pol = 0; pol < 3; ++pol) {
        if (!g[k][pol].present) continue;
        FILE *out = fopen(nm.freq[k][pol].c_str(), "w");
        if (!out) {
          set_error("2dfft: cannot write %s", nm.freq[k][pol].c_str());
          rc = NISE_ERR_ARG;
          break;
        }
        if (k == 0) {
          write_kI(out, g[k][pol], fft, deltat, shift1, shift3, non, form);
        } else {
          rwma1 = shift1, rwmi1 = shift1, rwma3 = shift3, rwmi3 = shift3;
          if (form == 1) axisFH = fopen("waxis.dat", "w");
          write_kII(out, axisFH, g[k][pol], fft, deltat, shift1, shift3, sw, non, form, raman, rwmi1, rwma1,
                    rwmi3, rwma3);
          if (axisFH) fclose(axisFH), axisFH = nullptr;
        }
        fclose(out);
      }
      if (rc) break;
    }
    if (rc) break;
    if (form == 1) break;  // Matlab: the user adds the parts (src/2DFFT.c:628-629)
    // rephasing + non-rephasing (src/2DFFT.c:631-826)
    const float dw1 = 1 / deltat / kCv2d / fft, dw3 = 1 / deltat / kCv2d / fft;
    const int p1 = (int)round_ref((rwma1 - rwmi1) / dw1 + 1), p3 = (int)round_ref((rwma3 - rwmi3) / dw3 + 1);
    if (!raman) printf("Adding rephasing and nonrephasing contributions!\n");
    for (int pol = 0; pol < 3; ++pol) {
      FILE *in = fopen(nm.freq[0][pol].c_str(), "r");
      if (!in) {
        printf("The response function file %s was not found!\n", nm.freq[0][pol].c_str());
        printf("Warning: Skipping this response function polarization!\n");
        continue;
      }
      Mat kI1, kI3, kIi, kI, kII1, kII3, kIIi, kII;
      read_freq(in, p1, p3, kI1, kI3, kIi, kI);
      fclose(in);
      for (int i = 1; i <= p1; i++)
        for (int j = 1; j <= p3; j++) kI1[i][j] = -kI1[i][j];
      in = fopen(nm.freq[1][pol].c_str(), "r");
      if (!in) {
        // the program reads from the missing file regardless; skip instead
        printf("The response function file %s was not found!\n", nm.freq[1][pol].c_str());
        printf("Warning: Skipping this response function polarization!\n");
        continue;
      }
      read_freq(in, p1, p3, kII1, kII3, kIIi, kII);
      fclose(in);
      Mat IR3(p1 + 1, std::vector<float>(p3 + 1, 0.f)), IR = IR3, IRi = IR3;
      for (int k = 1; k <= p1; k++)
        for (int l = 1; l <= p3; l++) IR3[k][l] = rwmi3 + (rwma3 - rwmi3) * l / (p3);
      auto pump_probe = [&](const char *name, const Mat &S, const Mat &axis) {
        FILE *o = fopen(name, "w");
        if (!o) {
          printf("Name for broadband pump probe not given.\nSkipping pump probe calculation.\n");
          return;
        }
        fprintf(o, "### Broadband pump probe spectrum\n");
        for (int l = 1; l <= p3; l++) {
          float pumpProbe = 0;
          for (int k = 1; k <= p1; k++) pumpProbe += S[k][l];
          fprintf(o, "%f %e\n", axis[1][l], pumpProbe);
        }
        fclose(o);
      };
      auto write2d = [&](const char *name, const Mat &a, const Mat &b, const Mat &c, const Mat &d) {
        FILE *o = fopen(name, "w");
        if (!o) return;
        for (int k = 1; k <= p1; k++) {
          for (int l = 1; l <= p3; l++) fprintf(o, "%f %f %e %e\n", a[k][l], b[k][l], c[k][l], d[k][l]);
          if (form == 2) fprintf(o, "\n");
        }
        fclose(o);
      };
      if (raman) {
        write2d(nm.twoD[pol].c_str(), kI1, kI3, kIi, kI);
        write2d(nm.twoD2[pol].c_str(), kII1, kII3, kIIi, kII);
        pump_probe(nm.pp[pol].c_str(), kI, kI3);
        pump_probe(nm.pp2[pol].c_str(), kII, kII3);
      } else {
        for (int i = 1; i <= p1; i++)
          for (int j = 1; j <= p3; j++) {
            IR[i][j] = kI[p1 - i + 1][j] + kII[i][j];
            IRi[i][j] = kIi[p1 - i + 1][j] + kIIi[i][j];
          }
        write2d(nm.twoD[pol].c_str(), kII1, kII3, IRi, IR);
        pump_probe(nm.pp[pol].c_str(), IR, IR3);
      }
    }
    if (!wfile) break;
  }
  if (WTime) fclose(WTime);
  return rc;
}
