/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Thin driver compiled together with the reference's own, unmodified
 * sources (propagate.c, NISE_subs.c, polar.c, read_trajectory.c,
 * hamshift.c, randomlib.c from the reference's src/) into
 * oracle/_ref/libnise_ref.so by oracle/Makefile.  It exposes the reference
 * primitives to the tests so that the restatement in nise_oracle.c can be
 * pinned bit for bit.  The reference objects are built with hidden
 * visibility and linked with --gc-sections, so functions that would need
 * LAPACK (ssyev_/sgeev_) are dropped; nothing here calls them.
 *
 * The loop driver ref_absorption_files restates the loop structure of
 * src/absorption.c:92-146 (that translation unit calls do_1DFFT, which needs
 * FFTW, absent from this image), but every file read, propagation and
 * response operation inside it is the reference's own compiled function.
 * The 2D drivers calc_2DIR / calc_2DES are built whole, with MPI, by
 * ref2d_main.c (oracle/Makefile, target ref2d).
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "types.h"
#include "NISE_subs.h"
#include "propagate.h"
#include "polar.h"
#include "read_trajectory.h"
#include "hamshift.h"

#define EXPORT __attribute__((visibility("default")))

static void init_non(t_non *non, int N, float deltat, float cut) {
  memset(non, 0, sizeof(*non));
  non->singles = N;
  non->deltat = deltat;
  non->couplingcut = cut;
  non->ts = 1;
  non->propagation = 1;
  non->cluster = -1;
  strcpy(non->hamiltonian, "Full");
}

EXPORT void ref_coupling_S(int N, float *H, float *cr, float *ci, int m,
                           int sign, float deltat, float cut) {
  t_non non;
  init_non(&non, N, deltat, cut);
  propagate_vec_coupling_S(&non, H, cr, ci, m, sign);
}

EXPORT void ref_coupling_S_doubles(int N, float *H, float *cr, float *ci,
                                   int m, float *Anh, float anharmonicity,
                                   float deltat, float cut) {
  t_non non;
  init_non(&non, N, deltat, cut);
  non.anharmonicity = anharmonicity;
  propagate_vec_coupling_S_doubles(&non, H, cr, ci, m, Anh);
}

EXPORT void ref_dipole_double(int N, float *dipole, float *cr, float *ci,
                              float *fr, float *fi, float *over,
                              float anharmonicity) {
  t_non non;
  init_non(&non, N, 1, 0);
  non.anharmonicity = anharmonicity;
  dipole_double(&non, dipole, cr, ci, fr, fi, over);
}

EXPORT void ref_dipole_double_last(int N, float *dipole, float *cr, float *ci,
                                   float *fr, float *fi, float *over,
                                   float anharmonicity) {
  t_non non;
  init_non(&non, N, 1, 0);
  non.anharmonicity = anharmonicity;
  dipole_double_last(&non, dipole, cr, ci, fr, fi, over);
}

EXPORT void ref_coupling_S_doubles_ES(int N, float *H, float *cr, float *ci,
                                      int m, float deltat, float cut) {
  t_non non;
  init_non(&non, N, deltat, cut);
  propagate_vec_coupling_S_doubles_ES(&non, H, cr, ci, m);
}

EXPORT void ref_dipole_double_ES(int N, float *dipole, float *cr, float *ci,
                                 float *fr, float *fi) {
  t_non non;
  init_non(&non, N, 1, 0);
  dipole_double_ES(&non, dipole, cr, ci, fr, fi);
}

EXPORT void ref_dipole_double_last_ES(int N, float *dipole, float *cr, float *ci,
                                      float *fr, float *fi) {
  t_non non;
  init_non(&non, N, 1, 0);
  dipole_double_last_ES(&non, dipole, cr, ci, fr, fi);
}

/* dipole_double_ground (src/NISE_subs.c:512-530), with its write-back of */
/* over for a fixed anharmonicity */
EXPORT void ref_dipole_double_ground(int N, float *dipole, float *fr, float *fi, float *over,
                                     float anharmonicity) {
  t_non non;
  init_non(&non, N, 1, 0);
  non.anharmonicity = anharmonicity;
  dipole_double_ground(&non, dipole, fr, fi, over);
}

/* find_raman_pol (src/polar.c:107-116) as its return value */
EXPORT int ref_find_raman_pol(int p2, int p3) { return find_raman_pol(p2, p3, -1); }

/* read_alpha (src/read_trajectory.c:380-399): component x of frame pos of */
/* a Raman tensor file; returns its control word */
EXPORT int ref_read_alpha(int N, const char *afile, int pos, int x, float *alpha) {
  t_non non;
  init_non(&non, N, 1, 0);
  FILE *f = fopen(afile, "rb");
  if (!f) return -1;
  int c = read_alpha(&non, alpha, f, pos, x);
  fclose(f);
  return c;
}

EXPORT float ref_polarweight(int pol, int x) { return polarweight(pol, x); }
EXPORT void ref_polar(int *xp, int x) { polar(xp, x); }
EXPORT int ref_sindex(int a, int b, int N) { return Sindex(a, b, N); }

/* Layout of the reference's t_non, to pin include/nise_abi.h. */
EXPORT int ref_tnon_layout(long long *out, int n) {
  long long v[] = {
      sizeof(t_non), offsetof(t_non, tmax1), offsetof(t_non, length),
      offsetof(t_non, deltat), offsetof(t_non, singles),
      offsetof(t_non, energyFName), offsetof(t_non, hamiltonian),
      offsetof(t_non, pbcFName), offsetof(t_non, tmax),
      offsetof(t_non, shifte), offsetof(t_non, thres),
      offsetof(t_non, couplingcut), offsetof(t_non, anharmonicity),
      offsetof(t_non, psites), offsetof(t_non, window),
      offsetof(t_non, SingleShiftSite), offsetof(t_non, SingleShift),
      offsetof(t_non, printLevel)};
  int k = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < k && i < n; i++) out[i] = v[i];
  return k;
}

static void copy_name(char *dst, const char *src) {
  strncpy(dst, src, 255);
  dst[255] = 0;
}

/* One frame through the reference's read_Hamiltonian / read_coupling. */
EXPORT int ref_read_frame(int N, const char *hfile, const char *dfile,
                          const char *cfile, const char *shiftfile,
                          const char *htype, float shifte, float cut, int pos,
                          float *He, float *mu3) {
  t_non non;
  init_non(&non, N, 1, cut);
  copy_name(non.energyFName, hfile);
  copy_name(non.dipoleFName, dfile);
  copy_name(non.couplingFName, cfile ? cfile : "");
  copy_name(non.singleShiftFName, shiftfile ? shiftfile : "");
  copy_name(non.hamiltonian, htype);
  non.shifte = shifte;
  non.length = 1 << 29;
  FILE *H_traj, *mu_traj, *Cfile = NULL, *C_traj = NULL;
  open_files(&non, &H_traj, &mu_traj, &Cfile);
  float *mu_xyz = calloc(3 * N, sizeof(float));
  read_coupling(&non, C_traj, mu_traj, He, mu_xyz);
  read_Hamiltonian(&non, He, H_traj, pos);
  for (int x = 0; x < 3; x++) read_dipole(&non, mu_traj, mu3 + x * N, mu_xyz, x, pos);
  free(mu_xyz);
  fclose(H_traj), fclose(mu_traj);
  return 0;
}

/* One frame of a TransitionDipole / ExtendedDipole Hamiltonian through the
 * reference's read_Hamiltonian (read_He builds the couplings from the
 * position, dipole and optional pbc files). */
EXPORT int ref_read_frame_onfly(int N, const char *hfile, const char *dfile,
                                const char *posfile, const char *pbcfile,
                                const char *htype, float shifte, float cut,
                                int pos, float *He) {
  t_non non;
  init_non(&non, N, 1, cut);
  copy_name(non.energyFName, hfile);
  copy_name(non.dipoleFName, dfile);
  copy_name(non.positionFName, posfile);
  copy_name(non.pbcFName, pbcfile ? pbcfile : "");
  copy_name(non.hamiltonian, htype);
  non.shifte = shifte;
  non.length = 1 << 29;
  FILE *H_traj, *mu_traj, *Cfile = NULL;
  open_files(&non, &H_traj, &mu_traj, &Cfile);
  read_Hamiltonian(&non, He, H_traj, pos);
  fclose(H_traj), fclose(mu_traj);
  return 0;
}

static int absorption_loop(t_non *non, float *re_S, float *im_S);

/* Loop of src/absorption.c:92-146 over the reference's primitives.  calc_S1
 * (src/absorption.c:205-212) is the two-line running sum written inline.
 * posfile / pbcfile: geometry for the on-the-fly Hamiltonian types. */
EXPORT int ref_absorption_files_geo(int N, const char *hfile, const char *dfile,
                                    const char *cfile, const char *shiftfile,
                                    const char *posfile, const char *pbcfile,
                                    const char *htype, float min1, float max1,
                                    float deltat, float cut, int ts, int tmax,
                                    int sample, int begin, int end, int length,
                                    float *re_S, float *im_S) {
  t_non non;
  init_non(&non, N, deltat, cut);
  copy_name(non.energyFName, hfile);
  copy_name(non.dipoleFName, dfile);
  copy_name(non.couplingFName, cfile ? cfile : "");
  copy_name(non.singleShiftFName, shiftfile ? shiftfile : "");
  copy_name(non.positionFName, posfile ? posfile : "");
  copy_name(non.pbcFName, pbcfile ? pbcfile : "");
  copy_name(non.hamiltonian, htype);
  non.min1 = min1, non.max1 = max1;
  non.ts = ts;
  non.tmax = non.tmax1 = tmax;
  non.sample = sample, non.begin = begin, non.end = end;
  non.length = length;
  return absorption_loop(&non, re_S, im_S);
}

EXPORT int ref_absorption_files(int N, const char *hfile, const char *dfile,
                                const char *cfile, const char *shiftfile,
                                const char *htype, float min1, float max1,
                                float deltat, float cut, int ts, int tmax,
                                int sample, int begin, int end, int length,
                                float *re_S, float *im_S) {
  return ref_absorption_files_geo(N, hfile, dfile, cfile, shiftfile, "", "", htype, min1, max1, deltat, cut,
                                  ts, tmax, sample, begin, end, length, re_S, im_S);
}

static int absorption_loop_streams(t_non *non_, FILE *H_traj, FILE *mu_traj, float *re_S, float *im_S);

static int absorption_loop(t_non *non, float *re_S, float *im_S) {
  FILE *H_traj, *mu_traj, *Cfile = NULL;
  open_files(non, &H_traj, &mu_traj, &Cfile);
  return absorption_loop_streams(non, H_traj, mu_traj, re_S, im_S);
}

/* The same loop over a trajectory held in memory: the reference's read_He
 * and read_mue read it through stdio streams over the buffers (fmemopen),
 * so no file system or page cache is involved.  Full-type records only. */
EXPORT int ref_absorption_mem(int N, const void *H_buf, long long H_bytes, const void *mu_buf,
                              long long mu_bytes, float min1, float max1, float deltat, float cut,
                              int ts, int tmax, int sample, int begin, int end, int length,
                              float *re_S, float *im_S) {
  t_non non;
  init_non(&non, N, deltat, cut);
  non.min1 = min1, non.max1 = max1;
  non.ts = ts;
  non.tmax = non.tmax1 = tmax;
  non.sample = sample, non.begin = begin, non.end = end;
  non.length = length;
  FILE *H_traj = fmemopen((void *)H_buf, (size_t)H_bytes, "rb");
  FILE *mu_traj = fmemopen((void *)mu_buf, (size_t)mu_bytes, "rb");
  if (!H_traj || !mu_traj) return 1;
  return absorption_loop_streams(&non, H_traj, mu_traj, re_S, im_S);
}

static int absorption_loop_streams(t_non *non_, FILE *H_traj, FILE *mu_traj, float *re_S, float *im_S) {
  t_non non = *non_;
  const int N = non.singles;
  non.shifte = (non.max1 + non.min1) / 2; /* src/absorption.c:57-59 */
  int nn2 = N * (N + 1) / 2;
  float *Hamil_i_e = calloc(nn2, sizeof(float));
  float *vecr = calloc(N, sizeof(float)), *veci = calloc(N, sizeof(float));
  float *mu_eg = calloc(N, sizeof(float)), *mu_xyz = calloc(3 * N, sizeof(float));
  FILE *C_traj = NULL;
  read_coupling(&non, C_traj, mu_traj, Hamil_i_e, mu_xyz);
  for (int samples = non.begin; samples < non.end; samples++) {
    int ti = samples * non.sample;
    for (int x = 0; x < 3; x++) {
      read_dipole(&non, mu_traj, vecr, mu_xyz, x, ti);
      clearvec(veci, non.singles);
      for (int t1 = 0; t1 < non.tmax; t1++) {
        int tj = ti + t1;
        read_Hamiltonian(&non, Hamil_i_e, H_traj, tj);
        read_dipole(&non, mu_traj, mu_eg, mu_xyz, x, tj);
        for (int i = 0; i < non.singles; i++) {
          re_S[t1] += mu_eg[i] * vecr[i];
          im_S[t1] += mu_eg[i] * veci[i];
        }
        propagate_vec_coupling_S(&non, Hamil_i_e, vecr, veci, non.ts, 1);
      }
    }
  }
  fclose(mu_traj), fclose(H_traj);
  free(Hamil_i_e), free(vecr), free(veci), free(mu_eg), free(mu_xyz);
  return 0;
}

#ifdef REF_WITH_LAPACK
/* Exact-exponential primitives that call LAPACK's ssyev_ through
 * diagonalizeLPD (src/propagate.c:1185-1221); built only into
 * _ref/libnise_ref_lapack.so, which links MKL (oracle/Makefile). */
EXPORT void ref_DIA(int N, float *H, float *cr, float *ci, int sign, float deltat) {
  t_non non;
  init_non(&non, N, deltat, 0);
  propagate_vec_DIA(&non, H, cr, ci, sign);
}

EXPORT int ref_DIA_S(int N, float *H, float *cr, float *ci, int sign, float deltat, float thres) {
  t_non non;
  init_non(&non, N, deltat, 0);
  non.thres = thres;
  return propagate_vec_DIA_S(&non, H, cr, ci, sign);
}

/* propagate_t2_DIA on one vector and T1 more (vr: T1 x N rows). */
EXPORT void ref_t2_DIA(int N, float *H, float *cr, float *ci, float *vr, float *vi, int T1,
                       float deltat) {
  t_non non;
  init_non(&non, N, deltat, 0);
  non.tmax1 = T1;
  float **pr = malloc(T1 * sizeof(float *)), **pi = malloc(T1 * sizeof(float *));
  for (int t = 0; t < T1; t++) pr[t] = vr + (size_t)t * N, pi[t] = vi + (size_t)t * N;
  propagate_t2_DIA(&non, H, cr, ci, pr, pi, 1);
  free(pr), free(pi);
}

EXPORT int ref_time_evolution_mat(int N, float *H, float deltat, int m, float thres, float *Ur,
                                  float *Ui, int *R, int *C) {
  t_non non;
  init_non(&non, N, deltat, 0);
  non.thres = thres;
  return time_evolution_mat(&non, H, Ur, Ui, R, C, m);
}
#endif

/* propagate_double_sparce(_ES) (src/propagate.c:1304-1473) with a given
 * element list: no LAPACK involved. */
EXPORT void ref_double_sparce(int N, float *Ur, float *Ui, int *R, int *C, float *fr, float *fi,
                              int elements, int m, float *Anh, float anharmonicity, float deltat,
                              int es) {
  t_non non;
  init_non(&non, N, deltat, 0);
  non.anharmonicity = anharmonicity;
  if (es)
    propagate_double_sparce_ES(&non, Ur, Ui, R, C, fr, fi, elements, m);
  else
    propagate_double_sparce(&non, Ur, Ui, R, C, fr, fi, elements, m, Anh);
}

/* The order of src/absorption.c:76-92 before its sample loop, over the
 * reference's own functions: open_files (missing files), control (its
 * diagnostics; the caller ignores the result, :80), determine_samples
 * (sample count and EndPoint checks, src/NISE_subs.c:196-221),
 * read_coupling, then the loop of absorption_loop.  Fatal conditions print
 * the reference's messages and exit with its codes; the tests run this in a
 * child process.  determine_samples appends to NISE.log in the working
 * directory. */
EXPORT int ref_absorption_checked(int N, const char *hfile, const char *dfile, float min1, float max1,
                                  float deltat, int tmax, int sample, int begin, int end, int length,
                                  float *re_S, float *im_S) {
  t_non non;
  init_non(&non, N, deltat, 0);
  copy_name(non.energyFName, hfile);
  copy_name(non.dipoleFName, dfile);
  copy_name(non.couplingFName, "");
  copy_name(non.singleShiftFName, "");
  non.min1 = min1, non.max1 = max1;
  non.shifte = (max1 + min1) / 2;
  non.tmax = non.tmax1 = tmax;
  non.sample = sample, non.begin = begin, non.end = end;
  non.length = length;
  FILE *H_traj, *mu_traj, *Cfile = NULL;
  open_files(&non, &H_traj, &mu_traj, &Cfile);
  control(&non);
  determine_samples(&non);
  return absorption_loop_streams(&non, H_traj, mu_traj, re_S, im_S);
}
