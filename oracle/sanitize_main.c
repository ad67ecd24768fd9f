/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Exercises every entry point of the oracle (nise_oracle.c) on small
 * deterministic inputs so that the whole restatement runs under
 * AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile, target
 * sanitize; run by tests/test_oracle_sanitize.py).  It checks nothing
 * numerically (the pinning tests do); it exits 0 when every call returns,
 * and the sanitizers abort on the first invalid access or undefined
 * operation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nise_oracle.h"

static unsigned long long g_state = 88172645463325252ULL;
static double urand(void) { /* xorshift64 */
  g_state ^= g_state << 13, g_state ^= g_state >> 7, g_state ^= g_state << 17;
  return (double)(g_state >> 11) / 9007199254740992.0;
}
static float nrand(float sd) { return (float)((urand() + urand() + urand() - 1.5) * 2.0 * sd); }

static void frame(int N, float *H, float esd, float jsd, double density) {
  for (int a = 0; a < N; a++)
    for (int b = a; b < N; b++)
      H[or_sindex(a, b, N)] = a == b ? nrand(esd) : (urand() < density ? nrand(jsd) : 0.f);
}

/* Full records (T frames) of an N-site system, dipole records, anharmonicity
 * and overtone records. */
static void records(int N, int T, float **H, float **mu, float **A, float **O) {
  const int nn2 = N * (N + 1) / 2;
  *H = malloc((size_t)T * (1 + nn2) * sizeof(float));
  *mu = malloc((size_t)T * (1 + 3 * N) * sizeof(float));
  *A = malloc((size_t)T * (1 + N) * sizeof(float));
  *O = malloc((size_t)T * (1 + 3 * N) * sizeof(float));
  for (int f = 0; f < T; f++) {
    float *h = *H + (size_t)f * (1 + nn2);
    h[0] = 0;
    frame(N, h + 1, 10, 4, 0.6);
    for (int a = 0; a < N; a++) h[1 + or_sindex(a, a, N)] += 1650;
    float *m = *mu + (size_t)f * (1 + 3 * N), *o = *O + (size_t)f * (1 + 3 * N), *an = *A + (size_t)f * (1 + N);
    m[0] = o[0] = an[0] = 0;
    for (int i = 0; i < 3 * N; i++) m[1 + i] = nrand(1), o[1 + i] = 1.41421356f * m[1 + i];
    for (int i = 0; i < N; i++) an[1 + i] = 16 + nrand(2);
  }
}

static t_non base_non(int N) {
  t_non non;
  memset(&non, 0, sizeof(non));
  non.singles = N, non.deltat = 2, non.ts = 1, non.propagation = 1, non.cluster = -1;
  non.min1 = 1550, non.max1 = 1750, non.shifte = 1650, non.fft = 64, non.lifetime = 300;
  non.dt1 = non.dt2 = non.dt3 = 1;
  strcpy(non.hamiltonian, "Full");
  strcpy(non.technique, "2DIR");
  return non;
}

int main(void) {
  /* singles, doubles and their FP64 forms */
  for (int N = 1; N <= 9; N += 4) {
    const int nn2 = N * (N + 1) / 2;
    float *H = malloc(nn2 * sizeof(float));
    frame(N, H, 50, 30, 0.7);
    float *cr = malloc(N * sizeof(float)), *ci = malloc(N * sizeof(float));
    double *dr = malloc(N * sizeof(double)), *di = malloc(N * sizeof(double));
    for (int i = 0; i < N; i++) cr[i] = nrand(1), ci[i] = nrand(1), dr[i] = cr[i], di[i] = ci[i];
    or_coupling_S_f32(N, H, cr, ci, 2, 1, 2.0f, 1.0f);
    or_coupling_S_f64(N, H, dr, di, 2, -1, 2.0f, 0.0f);
    or_coupling_S_inv_f64(N, H, dr, di, 2, -1, 2.0f, 0.0f);
    or_DIA_f64(N, H, dr, di, 1, 2.0f);
    or_DIA_S_f64(N, H, dr, di, -1, 2.0f, 1e-6f);
    float *fr = calloc(nn2, sizeof(float)), *fi = calloc(nn2, sizeof(float)), *Anh = malloc(N * sizeof(float));
    float *over = malloc(N * sizeof(float)), *dip = malloc(N * sizeof(float));
    double *gr = calloc(nn2, sizeof(double)), *gi = calloc(nn2, sizeof(double));
    for (int i = 0; i < N; i++) Anh[i] = 16 + nrand(2), over[i] = nrand(1), dip[i] = nrand(1);
    or_dipole_double(N, dip, cr, ci, fr, fi, over, 16.0f);
    or_dipole_double(N, dip, cr, ci, fr, fi, over, 0.0f);
    for (int i = 0; i < nn2; i++) gr[i] = fr[i], gi[i] = fi[i];
    or_coupling_S_doubles_f32(N, H, fr, fi, 2, Anh, 0.0f, 2.0f, 0.0f);
    or_coupling_S_doubles_f64(N, H, gr, gi, 1, Anh, 16.0f, 2.0f, 0.5f);
    or_coupling_S_doubles_ES_f32(N, H, fr, fi, 1, 2.0f, 0.0f);
    or_dipole_double_last(N, dip, fr, fi, cr, ci, over, 16.0f);
    or_dipole_double_ES(N, dip, cr, ci, fr, fi);
    or_dipole_double_last_ES(N, dip, fr, fi, cr, ci);
    float *Ur = malloc(N * N * sizeof(float)), *Ui = malloc(N * N * sizeof(float));
    int *R = malloc(N * N * sizeof(int)), *C = malloc(N * N * sizeof(int));
    for (int es = 0; es < 2; es++) {
      int el = or_time_evolution_mat(N, H, 2.0f, 1 + es, 1e-8f, Ur, Ui, C, R);
      or_double_sparce_f32(N, Ur, Ui, R, C, fr, fi, el, 1 + es, Anh, es ? 0.0f : 16.0f, 2.0f, es);
    }
    double *A = malloc(N * N * sizeof(double)), *V = malloc(N * N * sizeof(double)), *w = malloc(N * sizeof(double));
    for (int a = 0; a < N; a++)
      for (int b = 0; b < N; b++) A[a * N + b] = H[or_sindex(a, b, N)];
    or_jacobi_f64(N, A, w, V);
    free(H), free(cr), free(ci), free(dr), free(di), free(fr), free(fi), free(Anh), free(over), free(dip);
    free(gr), free(gi), free(Ur), free(Ui), free(R), free(C), free(A), free(V), free(w);
  }
  int xp[4];
  for (int m = 0; m < 21; m++) {
    or_polar(xp, m);
    for (int p = 0; p < 3; p++) (void)or_polarweight(p, m);
  }

  /* absorption (Full and Coupling types, 1 and 2 threads), FP64 gradient
     path, 1D FFT */
  {
    const int N = 5, T = 40;
    float *H, *mu, *A, *O;
    records(N, T, &H, &mu, &A, &O);
    const long long hs = 1 + N * (N + 1) / 2, ms = 1 + 3 * N;
    t_non non = base_non(N);
    non.tmax = non.tmax1 = 12, non.sample = 5, non.begin = 0, non.end = 4;
    float re[12], im[12];
    double re64[12], im64[12];
    for (int th = 1; th <= 2; th++)
      or_absorption(&non, H, hs, mu, ms, NULL, re, im, re64, im64, th);
    non.propagation = 0;
    or_absorption(&non, H, hs, mu, ms, NULL, re, im, NULL, NULL, 1);
    non.thres = 1e-7f;
    or_absorption(&non, H, hs, mu, ms, NULL, re, im, NULL, NULL, 1);
    non.propagation = 1, non.thres = 0;
    double delta[5] = {1, -1, 0.5, 0, 2};
    or_absorption_param_f64(&non, H, hs, mu, ms, NULL, delta, 1.1, re64, im64, 2);
    /* Coupling type: diagonal records and one coupling record */
    float *Hc = malloc((size_t)T * (1 + N) * sizeof(float)), *Cr = malloc(hs * sizeof(float));
    for (int f = 0; f < T; f++) {
      Hc[(size_t)f * (1 + N)] = 0;
      for (int a = 0; a < N; a++) Hc[(size_t)f * (1 + N) + 1 + a] = H[(size_t)f * hs + 1 + or_sindex(a, a, N)];
    }
    memcpy(Cr, H, hs * sizeof(float));
    strcpy(non.hamiltonian, "Coupling");
    or_absorption(&non, Hc, 1 + N, mu, ms, Cr, re, im, NULL, NULL, 1);
    strcpy(non.hamiltonian, "Full");
    double fq[64], sr[64], si[64];
    int nout = 0;
    or_1dfft(&non, re, im, 4, fq, sr, si, &nout);
    free(Hc), free(Cr);

    /* 2D items: Trotter and sparse schemes, waiting time, per-frame
       anharmonicities, two-level sites, every pathway technique */
    const char *techs[] = {"2DIR", "GBIR", "SEIR", "EAIR", "noEAIR", "2DUVvis", "EAUVvis", "noEAUVvis"};
    for (int k = 0; k < 8; k++)
      for (int prop = 0; prop < 2; prop++)
        for (int t2 = 0; t2 <= 2; t2 += 2) {
          t_non n2 = base_non(N);
          strcpy(n2.technique, techs[k]);
          n2.tmax1 = n2.tmax3 = 5, n2.tmax2 = t2, n2.sample = 4, n2.begin = 0, n2.end = 2;
          n2.propagation = prop, n2.thres = prop ? 0.0f : 1e-7f;
          n2.anharmonicity = (k % 2) ? 16.0f : 0.0f;
          n2.ts = 1 + (k % 2);
          float *R12 = calloc((size_t)12 * 25, sizeof(float));
          for (int s = 0; s < 2; s++)
            for (int m = 0; m < 21; m++) or_2dir_item(&n2, H, hs, mu, ms, NULL, A, O, s, m, R12);
          free(R12);
        }
    /* 2DIR-Raman items: every technique variant, Raman tensor records */
    {
      float *al = malloc((size_t)T * (1 + 6 * N) * sizeof(float));
      for (int f = 0; f < T; f++) {
        al[(size_t)f * (1 + 6 * N)] = 0;
        for (int i = 0; i < 6 * N; i++) al[(size_t)f * (1 + 6 * N) + 1 + i] = nrand(1);
      }
      const char *rt[] = {"2DIRraman", "2DIRraman1", "2DIRraman2", "2DIRraman3", "2DIRramanI", "2DIRramanII"};
      for (int k = 0; k < 6; k++) {
        t_non n2 = base_non(N);
        strcpy(n2.technique, rt[k]);
        n2.tmax1 = n2.tmax3 = 5, n2.sample = 4, n2.begin = 0, n2.end = 2, n2.anharmonicity = 16.0f;
        float *R12 = calloc((size_t)12 * 25, sizeof(float));
        for (int s = 0; s < 2; s++)
          for (int m = 0; m < 21; m++) or_2dirraman_item(&n2, H, hs, mu, ms, NULL, al, s, m, R12);
        free(R12);
      }
      float dg[16], dgi[16], ov[5], dp[5] = {1, 2, 3, 4, 5};
      or_dipole_double_ground(5, dp, dg, dgi, ov, 16.0f);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) (void)or_raman_pol(a, b);
      free(al);
    }
    free(H), free(mu), free(A), free(O);
  }
  printf("sanitized oracle run completed\n");
  return 0;
}
