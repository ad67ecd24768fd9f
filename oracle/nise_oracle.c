/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see nise_oracle.h).
 *
 * CPU restatement of the reference's wavefunction-propagation hot path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or the timed CPU baseline.
 *
 * Build flags follow the reference build (src/CMakeLists.txt:70-77):
 *   gcc -O2 -ffast-math -std=gnu99 -fopenmp
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "nise_oracle.h"

/* Constants exactly as src/types.h:5-12 declares them (float). */
static const float OR_c_v = 2.99792458e-5;
static const float OR_twoPi = 2 * 3.14159265359;
static const float OR_icm2ifs = 2.99792458e-5;
static const float OR_sqrt2 = 1.41421356237;
static const float OR_ithirty = 1.0 / 30;

/* Packed upper-triangle index, src/NISE_subs.h:54-65. */
int or_sindex(int a, int b, int N) {
  if (a > b) return a + b * ((N << 1) - b - 1) / 2;
  return b + a * ((N << 1) - a - 1) / 2;
}

/* ---------------------------------------------------------------------- */
/* Trotter singles, src/propagate.c:821-899                                */
/* ---------------------------------------------------------------------- */
void or_coupling_S_f32(int N, const float *H, float *cr, float *ci, int m,
                       int sign, float deltat, float couplingcut) {
  int N2 = (N * (N - 1)) / 2;
  float f = deltat * OR_icm2ifs * OR_twoPi * sign / m; /* :836 */
  float *H0 = malloc(N * sizeof(float));
  float *H1 = malloc((N2 + 1) * sizeof(float));
  int *col = malloc((N2 + 1) * sizeof(int));
  int *row = malloc((N2 + 1) * sizeof(int));
  float *re_U = malloc(N * sizeof(float));
  float *im_U = malloc(N * sizeof(float));
  float *ocr = malloc(N * sizeof(float));
  float *oci = malloc(N * sizeof(float));
  int k = 0;
  for (int a = 0; a < N; a++) { /* :848-860 lexicographic a<b */
    H0[a] = H[or_sindex(a, a, N)];
    for (int b = a + 1; b < N; b++) {
      int index = or_sindex(a, b, N);
      if (fabs(H[index]) > couplingcut) {
        H1[k] = H[index];
        col[k] = a, row[k] = b;
        k++;
      }
    }
  }
  int kmax = k;
  for (int a = 0; a < N; a++) { /* :863-866, double libm */
    re_U[a] = cos(0.5 * H0[a] * f);
    im_U[a] = -sin(0.5 * H0[a] * f);
  }
  for (int i = 0; i < m; i++) {
    for (int a = 0; a < N; a++) { /* :870-873 */
      ocr[a] = cr[a] * re_U[a] - ci[a] * im_U[a];
      oci[a] = ci[a] * re_U[a] + cr[a] * im_U[a];
    }
    for (k = 0; k < kmax; k++) { /* :876-888 sequential */
      int a = col[k], b = row[k];
      float J = H1[k];
      J = J * f;
      float si = -sin(J);
      float co = sqrt(1 - si * si);
      float cr1 = co * ocr[a] - si * oci[b];
      float ci1 = co * oci[a] + si * ocr[b];
      float cr2 = co * ocr[b] - si * oci[a];
      float ci2 = co * oci[b] + si * ocr[a];
      ocr[a] = cr1, oci[a] = ci1, ocr[b] = cr2, oci[b] = ci2;
    }
    for (int a = 0; a < N; a++) { /* :891-894 */
      cr[a] = ocr[a] * re_U[a] - oci[a] * im_U[a];
      ci[a] = oci[a] * re_U[a] + ocr[a] * im_U[a];
    }
  }
  free(H0), free(H1), free(col), free(row);
  free(re_U), free(im_U), free(ocr), free(oci);
}

/* Same algorithm, double state and double trig (error reference). */
static void coupling_S_f64_impl(int N, const float *H, double *cr, double *ci,
                                int m, int sign, float deltat,
                                float couplingcut, int inverse,
                                const double *delta, double cscale) {
  float ff = deltat * OR_icm2ifs * OR_twoPi * sign / m;
  double f = ff;
  int N2 = (N * (N - 1)) / 2;
  double *co = malloc((N2 + 1) * sizeof(double));
  double *si = malloc((N2 + 1) * sizeof(double));
  int *col = malloc((N2 + 1) * sizeof(int));
  int *row = malloc((N2 + 1) * sizeof(int));
  double *reU = malloc(N * sizeof(double));
  double *imU = malloc(N * sizeof(double));
  int k = 0;
  for (int a = 0; a < N; a++) {
    double h0 = H[or_sindex(a, a, N)];
    if (delta) h0 += delta[a];
    reU[a] = cos(0.5 * h0 * f);
    imU[a] = -sin(0.5 * h0 * f);
    if (inverse) imU[a] = -imU[a];
    for (int b = a + 1; b < N; b++) {
      float J = H[or_sindex(a, b, N)];
      if (fabs(J) > couplingcut) {
        double s = -sin((double)J * cscale * f);
        si[k] = inverse ? -s : s;
        co[k] = cos((double)J * cscale * f);
        col[k] = a, row[k] = b;
        k++;
      }
    }
  }
  int kmax = k;
  for (int i = 0; i < m; i++) {
    for (int a = 0; a < N; a++) {
      double r = cr[a] * reU[a] - ci[a] * imU[a];
      double q = ci[a] * reU[a] + cr[a] * imU[a];
      cr[a] = r, ci[a] = q;
    }
    for (int kk = 0; kk < kmax; kk++) {
      int k2 = inverse ? kmax - 1 - kk : kk;
      int a = col[k2], b = row[k2];
      double c = co[k2], s = si[k2];
      double cr1 = c * cr[a] - s * ci[b];
      double ci1 = c * ci[a] + s * cr[b];
      double cr2 = c * cr[b] - s * ci[a];
      double ci2 = c * ci[b] + s * cr[a];
      cr[a] = cr1, ci[a] = ci1, cr[b] = cr2, ci[b] = ci2;
    }
    for (int a = 0; a < N; a++) {
      double r = cr[a] * reU[a] - ci[a] * imU[a];
      double q = ci[a] * reU[a] + cr[a] * imU[a];
      cr[a] = r, ci[a] = q;
    }
  }
  free(co), free(si), free(col), free(row), free(reU), free(imU);
}

void or_coupling_S_f64(int N, const float *H, double *cr, double *ci, int m,
                       int sign, float deltat, float couplingcut) {
  coupling_S_f64_impl(N, H, cr, ci, m, sign, deltat, couplingcut, 0, NULL, 1.0);
}

void or_coupling_S_inv_f64(int N, const float *H, double *cr, double *ci,
                           int m, int sign, float deltat, float couplingcut) {
  coupling_S_f64_impl(N, H, cr, ci, m, sign, deltat, couplingcut, 1, NULL, 1.0);
}

/* ---------------------------------------------------------------------- */
/* Trotter doubles, src/propagate.c:902-1048                               */
/* ---------------------------------------------------------------------- */
/* Applies one coupling's pair rotations over c in the five ranges of
 * src/propagate.c:976-1036 (c<a, c==a, a<c<b, c==b, c>b). */
#define OR_ROT(T, co_, si_, i1, i2)                                        \
  do {                                                                     \
    T cr1 = co_ * ocr[i1] - si_ * oci[i2];                                 \
    T ci1 = co_ * oci[i1] + si_ * ocr[i2];                                 \
    T cr2 = co_ * ocr[i2] - si_ * oci[i1];                                 \
    T ci2 = co_ * oci[i2] + si_ * ocr[i1];                                 \
    ocr[i1] = cr1, oci[i1] = ci1, ocr[i2] = cr2, oci[i2] = ci2;            \
  } while (0)

void or_coupling_S_doubles_f32(int N, const float *H, float *cr, float *ci,
                               int m, const float *Anh, float anharmonicity,
                               float deltat, float couplingcut) {
  int N2 = N * (N + 1) / 2;
  const float f = deltat * OR_icm2ifs * OR_twoPi / m; /* :905 */
  float *H0 = malloc(N2 * sizeof(float));
  float *H1 = malloc((N * N / 2 + 1) * sizeof(float));
  int *col = malloc((N * N / 2 + 1) * sizeof(int));
  int *row = malloc((N * N / 2 + 1) * sizeof(int));
  float *re_U = malloc(N2 * sizeof(float));
  float *im_U = malloc(N2 * sizeof(float));
  float *ocr = malloc(N2 * sizeof(float));
  float *oci = malloc(N2 * sizeof(float));
  for (int a = 0; a < N; a++) { /* :918-933 */
    int indexa = or_sindex(a, a, N);
    for (int b = a; b < N; b++) {
      int index = or_sindex(a, b, N);
      H0[index] = H[indexa] + H[or_sindex(b, b, N)];
      if (a == b) {
        if (anharmonicity == 0) H0[index] -= Anh[a];
        else H0[index] -= anharmonicity;
      }
    }
  }
  int kmax = 0;
  for (int a = 0; a < N; a++) { /* :936-948 */
    for (int b = a + 1; b < N; b++) {
      int index = b + a * ((N << 1) - a - 1) / 2;
      if (fabsf(H[index]) > couplingcut) {
        H1[kmax] = H[index];
        col[kmax] = a, row[kmax] = b;
        kmax++;
      }
    }
  }
  for (int a = 0; a < N2; a++) { /* :951-954, float libm */
    re_U[a] = cosf(0.5f * H0[a] * f);
    im_U[a] = -sinf(0.5f * H0[a] * f);
  }
  for (int i = 0; i < m; i++) {
    for (int a = 0; a < N2; a++) {
      ocr[a] = cr[a] * re_U[a] - ci[a] * im_U[a];
      oci[a] = cr[a] * im_U[a] + ci[a] * re_U[a];
    }
    for (int k = 0; k < kmax; k++) {
      int a = col[k], b = row[k], c = 0;
      float J = H1[k] * f;
      int aNsum = a * ((N << 1) - a - 1) / 2;
      int bNsum = b * ((N << 1) - b - 1) / 2;
      float si = -sinf(J);
      float co = sqrtf(1 - si * si);
      for (; c < a; c++) {
        int i1 = a + c * ((N << 1) - c - 1) / 2;
        int i2 = b + c * ((N << 1) - c - 1) / 2;
        OR_ROT(float, co, si, i1, i2);
      }
      float si2 = -sinf(J * OR_sqrt2);
      float co2 = sqrtf(1 - si2 * si2);
      for (; c == a; c++) {
        int i1 = a + c * ((N << 1) - c - 1) / 2;
        int i2 = b + c * ((N << 1) - c - 1) / 2;
        OR_ROT(float, co2, si2, i1, i2);
      }
      for (; c < b; c++) {
        int i1 = c + aNsum;
        int i2 = b + c * ((N << 1) - c - 1) / 2;
        OR_ROT(float, co, si, i1, i2);
      }
      for (; c == b; c++) {
        int i1 = c + aNsum;
        int i2 = b + c * ((N << 1) - c - 1) / 2;
        OR_ROT(float, co2, si2, i1, i2);
      }
      for (; c < N; c++) {
        int i1 = c + aNsum;
        int i2 = c + bNsum;
        OR_ROT(float, co, si, i1, i2);
      }
    }
    for (int a = 0; a < N2; a++) {
      cr[a] = ocr[a] * re_U[a] - oci[a] * im_U[a];
      ci[a] = ocr[a] * im_U[a] + oci[a] * re_U[a];
    }
  }
  free(H0), free(H1), free(col), free(row);
  free(re_U), free(im_U), free(ocr), free(oci);
}

/* Two-level (fermionic) doubles, src/propagate.c:1051-1181: H0_ab = e_a +
   e_b with no anharmonicity; for each kept coupling (a, b) only the
   rotations of (|ca>, |cb>) with c != a, b, all with the angle of J.       */
void or_coupling_S_doubles_ES_f32(int N, const float *H, float *cr, float *ci,
                                  int m, float deltat, float couplingcut) {
  int N2 = N * (N + 1) / 2;
  const float f = deltat * OR_icm2ifs * OR_twoPi / m; /* :1055 */
  float *re_U = malloc(N2 * sizeof(float)), *im_U = malloc(N2 * sizeof(float));
  float *ocr = malloc(N2 * sizeof(float)), *oci = malloc(N2 * sizeof(float));
  for (int a = 0; a < N; a++) /* :1066-1074, :1091-1094 (float libm) */
    for (int b = a; b < N; b++) {
      float h0 = H[or_sindex(a, a, N)] + H[or_sindex(b, b, N)];
      re_U[or_sindex(a, b, N)] = cosf(0.5f * h0 * f);
      im_U[or_sindex(a, b, N)] = -sinf(0.5f * h0 * f);
    }
  for (int i = 0; i < m; i++) {
    for (int a = 0; a < N2; a++) {
      ocr[a] = cr[a] * re_U[a] - ci[a] * im_U[a];
      oci[a] = cr[a] * im_U[a] + ci[a] * re_U[a];
    }
    for (int a = 0; a < N; a++) /* kept couplings in lexicographic order */
      for (int b = a + 1; b < N; b++) {
        float Hab = H[b + a * ((N << 1) - a - 1) / 2];
        if (!(fabsf(Hab) > couplingcut)) continue;
        float J = Hab * f;
        float si = -sinf(J);
        float co = sqrtf(1 - si * si);
        for (int c = 0; c < N; c++) {
          if (c == a || c == b) continue;
          int i1 = or_sindex(c, a, N), i2 = or_sindex(c, b, N);
          OR_ROT(float, co, si, i1, i2);
        }
      }
    for (int a = 0; a < N2; a++) {
      cr[a] = ocr[a] * re_U[a] - oci[a] * im_U[a];
      ci[a] = ocr[a] * im_U[a] + oci[a] * re_U[a];
    }
  }
  free(re_U), free(im_U), free(ocr), free(oci);
}

void or_coupling_S_doubles_f64(int N, const float *H, double *cr, double *ci,
                               int m, const float *Anh, float anharmonicity,
                               float deltat, float couplingcut) {
  int N2 = N * (N + 1) / 2;
  const float ff = deltat * OR_icm2ifs * OR_twoPi / m;
  double f = ff;
  double *reU = malloc(N2 * sizeof(double));
  double *imU = malloc(N2 * sizeof(double));
  double *ocr = cr, *oci = ci;
  for (int a = 0; a < N; a++) {
    for (int b = a; b < N; b++) {
      int index = or_sindex(a, b, N);
      double h0 = (double)H[or_sindex(a, a, N)] + (double)H[or_sindex(b, b, N)];
      if (a == b) h0 -= (anharmonicity == 0) ? (double)Anh[a] : (double)anharmonicity;
      reU[index] = cos(0.5 * h0 * f);
      imU[index] = -sin(0.5 * h0 * f);
    }
  }
  for (int i = 0; i < m; i++) {
    for (int a = 0; a < N2; a++) {
      double r = cr[a] * reU[a] - ci[a] * imU[a];
      double q = cr[a] * imU[a] + ci[a] * reU[a];
      cr[a] = r, ci[a] = q;
    }
    for (int a = 0; a < N; a++) {
      for (int b = a + 1; b < N; b++) {
        float Jf = H[b + a * ((N << 1) - a - 1) / 2];
        if (!(fabsf(Jf) > couplingcut)) continue;
        double J = (double)Jf * f;
        double si = -sin(J), co = cos(J);
        double si2 = -sin(J * sqrt(2.0)), co2 = cos(J * sqrt(2.0));
        for (int c = 0; c < N; c++) {
          int i1 = or_sindex(c, a, N), i2 = or_sindex(c, b, N);
          if (c == a || c == b) OR_ROT(double, co2, si2, i1, i2);
          else OR_ROT(double, co, si, i1, i2);
        }
      }
    }
    for (int a = 0; a < N2; a++) {
      double r = ocr[a] * reU[a] - oci[a] * imU[a];
      double q = ocr[a] * imU[a] + oci[a] * reU[a];
      cr[a] = r, ci[a] = q;
    }
  }
  free(reU), free(imU);
}

/* ---------------------------------------------------------------------- */
/* Exact exponential (DIA), src/propagate.c:60-125, in double precision.   */
/* The reference diagonalises with LAPACK ssyev_ (src/propagate.c:1185-    */
/* 1221); here a cyclic Jacobi method in double does it.                   */
/* ---------------------------------------------------------------------- */
int or_jacobi_f64(int N, double *A, double *w, double *V) {
  /* A: row-major symmetric, destroyed.  V[i*N+j] = component i of
   * eigenvector j.  Returns the number of sweeps, or -1 without
   * convergence. */
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) V[i * N + j] = (i == j);
  int sweep;
  for (sweep = 0; sweep < 100; sweep++) {
    double off = 0, diag = 0;
    for (int i = 0; i < N; i++) {
      diag += A[i * N + i] * A[i * N + i];
      for (int j = i + 1; j < N; j++) off += A[i * N + j] * A[i * N + j];
    }
    if (off <= 1e-30 * (diag + 1e-300) || off == 0) break;
    for (int p = 0; p < N - 1; p++) {
      for (int q = p + 1; q < N; q++) {
        double apq = A[p * N + q];
        if (apq == 0) continue;
        double theta = (A[q * N + q] - A[p * N + p]) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        double c = 1 / sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < N; k++) { /* A <- A J (columns p,q) */
          double akp = A[k * N + p], akq = A[k * N + q];
          A[k * N + p] = c * akp - s * akq;
          A[k * N + q] = s * akp + c * akq;
        }
        for (int k = 0; k < N; k++) { /* A <- J^T A (rows p,q) */
          double apk = A[p * N + k], aqk = A[q * N + k];
          A[p * N + k] = c * apk - s * aqk;
          A[q * N + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < N; k++) {
          double vkp = V[k * N + p], vkq = V[k * N + q];
          V[k * N + p] = c * vkp - s * vkq;
          V[k * N + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < N; i++) w[i] = A[i * N + i];
  /* ascending order, as ssyev_ returns */
  for (int i = 0; i < N; i++) {
    int mn = i;
    for (int j = i + 1; j < N; j++)
      if (w[j] < w[mn]) mn = j;
    if (mn != i) {
      double tw = w[i];
      w[i] = w[mn], w[mn] = tw;
      for (int k = 0; k < N; k++) {
        double tv = V[k * N + i];
        V[k * N + i] = V[k * N + mn], V[k * N + mn] = tv;
      }
    }
  }
  return sweep < 100 ? sweep : -1;
}

void or_DIA_f64(int N, const float *H, double *cr, double *ci, int sign,
                float deltat) {
  float ff = deltat * OR_icm2ifs * OR_twoPi * sign; /* :69, no /m */
  double f = ff;
  double *A = malloc(N * N * sizeof(double));
  double *V = malloc(N * N * sizeof(double));
  double *w = malloc(N * sizeof(double));
  double *yr = malloc(N * sizeof(double)), *yi = malloc(N * sizeof(double));
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++) A[a * N + b] = H[or_sindex(a, b, N)];
  or_jacobi_f64(N, A, w, V);
  /* y = V^T psi ; y_j *= exp(-i E_j f) ; psi = V y */
  for (int j = 0; j < N; j++) {
    double sr = 0, si = 0;
    for (int a = 0; a < N; a++) sr += V[a * N + j] * cr[a], si += V[a * N + j] * ci[a];
    double c = cos(w[j] * f), s = -sin(w[j] * f);
    yr[j] = sr * c - si * s;
    yi[j] = si * c + sr * s;
  }
  for (int a = 0; a < N; a++) {
    double sr = 0, si = 0;
    for (int j = 0; j < N; j++) sr += V[a * N + j] * yr[j], si += V[a * N + j] * yi[j];
    cr[a] = sr, ci[a] = si;
  }
  free(A), free(V), free(w), free(yr), free(yi);
}

/* Sparse scheme, src/propagate.c:185-258: U = V exp(-iEf) V^T in the site
   basis (:217-234), then psi_a <- sum over the kept elements U_ab psi_b,
   an element kept when |U_ab|^2 > thres (:242); the reference tests the FP32
   elements, so the test here is on the FP32-rounded element. */
int or_DIA_S_f64(int N, const float *H, double *cr, double *ci, int sign,
                 float deltat, float thres) {
  float ff = deltat * OR_icm2ifs * OR_twoPi * sign; /* :197, no /m */
  double f = ff;
  double *A = malloc(N * N * sizeof(double));
  double *V = malloc(N * N * sizeof(double));
  double *w = malloc(N * sizeof(double));
  double *ur = malloc(N * N * sizeof(double)), *ui = malloc(N * N * sizeof(double));
  double *yr = malloc(N * sizeof(double)), *yi = malloc(N * sizeof(double));
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++) A[a * N + b] = H[or_sindex(a, b, N)];
  or_jacobi_f64(N, A, w, V);
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++) {
      double sr = 0, si = 0;
      for (int c = 0; c < N; c++) {
        double p = V[a * N + c] * V[b * N + c];
        sr += p * cos(w[c] * f);
        si -= p * sin(w[c] * f);
      }
      ur[a * N + b] = sr, ui[a * N + b] = si;
    }
  int kept = 0;
  for (int a = 0; a < N; a++) {
    double sr = 0, si = 0;
    for (int b = 0; b < N; b++) {
      float r = (float)ur[a * N + b], i = (float)ui[a * N + b];
      if (r * r + i * i > thres) {
        kept++;
        sr += ur[a * N + b] * cr[b] - ui[a * N + b] * ci[b];
        si += ur[a * N + b] * ci[b] + ui[a * N + b] * cr[b];
      }
    }
    yr[a] = sr, yi[a] = si;
  }
  for (int a = 0; a < N; a++) cr[a] = yr[a], ci[a] = yi[a];
  free(A), free(V), free(w), free(ur), free(ui), free(yr), free(yi);
  return kept;
}

/* ---------------------------------------------------------------------- */
/* mu_ef application, src/NISE_subs.c:485-509 and :552-575                 */
/* ---------------------------------------------------------------------- */
void or_dipole_double(int N, const float *dipole, const float *cr,
                      const float *ci, float *fr, float *fi, float *over,
                      float anharmonicity) {
  int N2 = N * (N + 1) / 2;
  for (int i = 0; i < N2; i++) fr[i] = 0, fi[i] = 0;
  if (anharmonicity != 0)
    for (int i = 0; i < N; i++) over[i] = OR_sqrt2 * dipole[i];
  for (int i = 0; i < N; i++) {
    int index = or_sindex(i, i, N);
    fr[index] += over[i] * cr[i];
    fi[index] += over[i] * ci[i];
    for (int j = i + 1; j < N; j++) {
      index = or_sindex(i, j, N);
      fr[index] += dipole[i] * cr[j];
      fi[index] += dipole[i] * ci[j];
      fr[index] += dipole[j] * cr[i];
      fi[index] += dipole[j] * ci[i];
    }
  }
}

void or_dipole_double_last(int N, const float *dipole, const float *cr,
                           const float *ci, float *fr, float *fi, float *over,
                           float anharmonicity) {
  for (int i = 0; i < N; i++) fr[i] = 0, fi[i] = 0;
  if (anharmonicity != 0)
    for (int i = 0; i < N; i++) over[i] = OR_sqrt2 * dipole[i];
  for (int i = 0; i < N; i++) {
    int index = or_sindex(i, i, N);
    fr[i] += over[i] * cr[index];
    fi[i] += over[i] * ci[index];
    for (int j = i + 1; j < N; j++) {
      index = or_sindex(i, j, N);
      fr[j] += dipole[i] * cr[index];
      fi[j] += dipole[i] * ci[index];
      fr[i] += dipole[j] * cr[index];
      fi[i] += dipole[j] * ci[index];
    }
  }
}

/* Two-level mu_ef, src/NISE_subs.c:533-549 and :578-593: no |aa> states. */
void or_dipole_double_ES(int N, const float *dipole, const float *cr,
                         const float *ci, float *fr, float *fi) {
  int N2 = N * (N + 1) / 2;
  for (int i = 0; i < N2; i++) fr[i] = 0, fi[i] = 0;
  for (int i = 0; i < N; i++)
    for (int j = i + 1; j < N; j++) {
      int index = or_sindex(i, j, N);
      fr[index] += dipole[i] * cr[j];
      fi[index] += dipole[i] * ci[j];
      fr[index] += dipole[j] * cr[i];
      fi[index] += dipole[j] * ci[i];
    }
}

void or_dipole_double_last_ES(int N, const float *dipole, const float *cr,
                              const float *ci, float *fr, float *fi) {
  for (int i = 0; i < N; i++) fr[i] = 0, fi[i] = 0;
  for (int i = 0; i < N; i++)
    for (int j = i + 1; j < N; j++) {
      int index = or_sindex(i, j, N);
      fr[j] += dipole[i] * cr[index];
      fi[j] += dipole[i] * ci[index];
      fr[i] += dipole[j] * cr[index];
      fi[i] += dipole[j] * ci[index];
    }
}

/* Isotropic weights and molecular-frame indices, src/polar.c:12-105. */
float or_polarweight(int pol, int x) {
  float weight = 0;
  if (pol == 0) weight = (x <= 2) ? 6 * OR_ithirty : 2 * OR_ithirty;
  if (pol == 1) {
    if (x <= 2) weight = 2 * OR_ithirty;
    else if (x <= 8) weight = 4 * OR_ithirty;
    else weight = -1 * OR_ithirty;
  }
  if (pol == 2) {
    if (x <= 8) weight = 0.0;
    else if (x <= 14) weight = 2.5 * OR_ithirty;
    else weight = -2.5 * OR_ithirty;
  }
  return weight;
}

void or_polar(int *xp, int x) {
  /* XXXX (0-2), XXYY (3-8), XYXY (9-14), XYYX (15-20) */
  static const int tab[21][4] = {
      {0, 0, 0, 0}, {1, 1, 1, 1}, {2, 2, 2, 2}, {0, 0, 1, 1}, {0, 0, 2, 2},
      {1, 1, 0, 0}, {1, 1, 2, 2}, {2, 2, 0, 0}, {2, 2, 1, 1}, {0, 1, 0, 1},
      {0, 2, 0, 2}, {1, 0, 1, 0}, {1, 2, 1, 2}, {2, 0, 2, 0}, {2, 1, 2, 1},
      {0, 1, 1, 0}, {0, 2, 2, 0}, {1, 0, 0, 1}, {1, 2, 2, 1}, {2, 0, 0, 2},
      {2, 1, 1, 2}};
  for (int i = 0; i < 4; i++) xp[i] = tab[x][i];
}

/* ---------------------------------------------------------------------- */
/* Frame ingest: read_He / read_Dia / apply_singleshift semantics          */
/* (src/read_trajectory.c:47-87,106-178,305-341, src/hamshift.c:59-66).    */
/* ---------------------------------------------------------------------- */
static int is_coupling(const t_non *non) {
  return strcmp(non->hamiltonian, "Coupling") == 0;
}

/* On-the-fly couplings (src/read_trajectory.c:181-303): geometry records set
   by or_set_geometry; 1 TransitionDipole, 2 ExtendedDipole.  Evaluated in
   IEEE arithmetic in source order, as the reference's ACCURATE_MATHS build
   does (under -ffast-math GCC turns 1.0/sqrt(x) into an approximate
   reciprocal square root plus a Newton step, whose low bits depend on the
   CPU). */
#pragma GCC push_options
#pragma GCC optimize("no-fast-math")
static const float *g_pos, *g_dip, *g_box;
static long long g_pos_stride, g_dip_stride;

void or_set_geometry(const float *pos_rec, long long pos_stride, const float *dip_rec,
                     long long dip_stride, const float *box_rec) {
  g_pos = pos_rec, g_pos_stride = pos_stride, g_dip = dip_rec, g_dip_stride = dip_stride;
  g_box = box_rec;
}

static int onfly_kind(const t_non *non) {
  if (strcmp(non->hamiltonian, "TransitionDipole") == 0) return 1;
  if (strcmp(non->hamiltonian, "ExtendedDipole") == 0) return 2;
  return 0;
}

/* minimum image along axis x (pbc1, src/NISE_subs.c:641-648) */
static float image(float r, int x, const float *box) {
  if (box && box[0] > 0.0f) {
    if (r > box[x] / 2) r = r - box[x];
    if (r < -box[x] / 2) r = r + box[x];
  }
  return r;
}

/* |p_i - q_j| with the minimum image, float as the reference */
static float sep(const float *p, const float *q, int i, int j, int N, const float *box) {
  float x = image(p[i] - q[j], 0, box);
  float y = image(p[N + i] - q[N + j], 1, box);
  float z = image(p[2 * N + i] - q[2 * N + j], 2, box);
  return sqrt(x * x + y * y + z * z);
}

/* Couplings of frame pos into the packed triangle He (i < j). */
static void onfly_couplings(const t_non *non, long long pos, float *He) {
  const int N = non->singles, kind = onfly_kind(non);
  const float A = 5034.11861687f;
  const float *box = g_box ? g_box + pos * 4 + 1 : NULL;
  const float *m = g_dip + pos * g_dip_stride + 1;
  if (kind == 1) { /* :198-216 */
    const float *R = g_pos + pos * g_pos_stride + 1;
    for (int i = 0; i < N; i++)
      for (int j = i + 1; j < N; j++) {
        float rx = image(R[i] - R[j], 0, box);
        float ry = image(R[N + i] - R[N + j], 1, box);
        float rz = image(R[2 * N + i] - R[2 * N + j], 2, box);
        float d = sqrt(rx * rx + ry * ry + rz * rz);
        float id = 1.0 / d;
        float id3 = id * id * id;
        float dot = m[i] * m[j] + m[N + i] * m[N + j] + m[2 * N + i] * m[2 * N + j];
        float a1 = m[i] * rx + m[N + i] * ry + m[2 * N + i] * rz;
        float a2 = m[j] * rx + m[N + j] * ry + m[2 * N + j] * rz;
        float orient = -3 * a1 * a2 * id * id;
        He[or_sindex(i, j, N)] = A * (dot + orient) * id3;
      }
  } else { /* :244-297: charges on the two position records of the frame */
    const float *R = g_pos + 2 * pos * g_pos_stride + 1, *R2 = R + g_pos_stride;
    for (int i = 0; i < N; i++) {
      float mi2 = m[i] * m[i] + m[N + i] * m[N + i] + m[2 * N + i] * m[2 * N + i];
      float qi = sqrt(mi2) / sep(R2, R, i, i, N, box);
      for (int j = i + 1; j < N; j++) {
        float qj = sqrt(mi2) / sep(R2, R, j, j, N, box); /* site i's dipole, as :266 */
        /* each 1/r is rounded to float before it enters the float sum */
        float J = 1.0 / sep(R, R, i, j, N, box);
        float id = 1.0 / sep(R, R2, i, j, N, box);
        J = J - id;
        id = 1.0 / sep(R2, R, i, j, N, box);
        J = J - id;
        id = 1.0 / sep(R2, R2, i, j, N, box);
        J = J + id;
        He[or_sindex(i, j, N)] = A * qi * qj * J;
      }
    }
  }
}
#pragma GCC pop_options

void or_read_H(const t_non *non, const float *H_rec, long long H_stride,
               const float *C_rec, long long pos, float *He, int apply_shift) {
  int N = non->singles;
  if (is_coupling(non)) {
    /* couplings from the one Full record of the coupling file, cut once */
    const float *c = C_rec + 1;
    for (int i = 0; i < N * (N + 1) / 2; i++) He[i] = c[i];
    for (int i = 0; i < N; i++) He[or_sindex(i, i, N)] -= non->shifte;
    for (int i = 0; i < N; i++)
      for (int j = i + 1; j < N; j++)
        if (fabs(He[or_sindex(i, j, N)]) < non->couplingcut) He[or_sindex(i, j, N)] = 0;
    const float *d = H_rec + pos * H_stride + 1;
    for (int i = 0; i < N; i++) He[i * N + i - (i * (i + 1)) / 2] = d[i] - non->shifte;
  } else if (onfly_kind(non)) {
    /* site energies from the diagonal records (:135-154), couplings from
       the geometry (:181-303), then the cut (:305-311) */
    const float *d = H_rec + pos * H_stride + 1;
    for (int i = 0; i < N; i++) He[i * N + i - (i * (i + 1)) / 2] = d[i] - non->shifte;
    onfly_couplings(non, pos, He);
    for (int i = 0; i < N; i++)
      for (int j = i + 1; j < N; j++)
        if (fabs(He[or_sindex(i, j, N)]) < non->couplingcut) He[or_sindex(i, j, N)] = 0;
  } else {
    const float *h = H_rec + pos * H_stride + 1;
    for (int i = 0; i < N * (N + 1) / 2; i++) He[i] = h[i];
    for (int i = 0; i < N; i++) He[i * N + i - (i * (i + 1)) / 2] -= non->shifte;
    for (int i = 0; i < N; i++)
      for (int j = i + 1; j < N; j++)
        if (fabs(He[or_sindex(i, j, N)]) < non->couplingcut) He[or_sindex(i, j, N)] = 0;
  }
  if (apply_shift) {
    for (int i = 0; i < non->SingleShiftSites; i++) {
      int index = or_sindex(non->SingleShiftSite[i], non->SingleShiftSite[i], N);
      He[index] = He[index] * non->SingleShift[2 * i] + non->SingleShift[2 * i + 1];
    }
  }
}

static void read_mu(const t_non *non, const float *mu_rec, long long mu_stride,
                    long long pos, int x, float *mu) {
  int N = non->singles;
  if (is_coupling(non)) pos = 0; /* static dipoles, read_coupling :60-67 */
  const float *p = mu_rec + pos * mu_stride + 1 + (long long)x * N;
  for (int i = 0; i < N; i++) mu[i] = p[i];
}

/* ---------------------------------------------------------------------- */
/* Linear absorption, src/absorption.c:92-146 and calc_S1 :205-212.        */
/* ---------------------------------------------------------------------- */
int or_absorption(const t_non *non, const float *H_rec, long long H_stride,
                  const float *mu_rec, long long mu_stride,
                  const float *C_rec, float *re_S, float *im_S,
                  double *re_S64, double *im_S64, int nthreads) {
  int N = non->singles, tmax = non->tmax;
  int nn2 = N * (N + 1) / 2;
  if (non->propagation != 1 && non->propagation != 0) return NISE_ERR_ARG;
  /* src/propagate.c:25-28: a threshold in (0, 1] selects the sparse scheme */
  const int sparse = non->propagation == 0 && non->thres > 0 && non->thres <= 1;
  if (re_S) memset(re_S, 0, tmax * sizeof(float)), memset(im_S, 0, tmax * sizeof(float));
  if (re_S64) memset(re_S64, 0, tmax * sizeof(double)), memset(im_S64, 0, tmax * sizeof(double));
  int nsamp = non->end - non->begin;
  if (nthreads < 1) nthreads = 1;
  /* nthreads==1 keeps the reference's single running sum; more threads
   * give each thread a private partial sum over a contiguous sample block
   * (the reference's BeginPoint/EndPoint split done in one process). */
  float *pr = calloc((size_t)nthreads * tmax, sizeof(float));
  float *pi = calloc((size_t)nthreads * tmax, sizeof(float));
  double *pr64 = calloc((size_t)nthreads * tmax, sizeof(double));
  double *pi64 = calloc((size_t)nthreads * tmax, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0, nt = 1;
#ifdef _OPENMP
    tid = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int s0 = non->begin + (int)((long long)nsamp * tid / nt);
    int s1 = non->begin + (int)((long long)nsamp * (tid + 1) / nt);
    float *He = malloc(nn2 * sizeof(float));
    float *vecr = malloc(N * sizeof(float)), *veci = malloc(N * sizeof(float));
    float *mu = malloc(N * sizeof(float));
    double *dr = malloc(N * sizeof(double)), *di = malloc(N * sizeof(double));
    float *myr = pr + (size_t)tid * tmax, *myi = pi + (size_t)tid * tmax;
    double *myr64 = pr64 + (size_t)tid * tmax, *myi64 = pi64 + (size_t)tid * tmax;
    for (int samples = s0; samples < s1; samples++) {
      int ti = samples * non->sample;
      for (int x = 0; x < 3; x++) {
        read_mu(non, mu_rec, mu_stride, ti, x, vecr);
        for (int i = 0; i < N; i++) veci[i] = 0, dr[i] = vecr[i], di[i] = 0;
        for (int t1 = 0; t1 < tmax; t1++) {
          int tj = ti + t1;
          or_read_H(non, H_rec, H_stride, C_rec, tj, He, 1);
          read_mu(non, mu_rec, mu_stride, tj, x, mu);
          if (re_S) {
            for (int i = 0; i < N; i++) { /* calc_S1 */
              myr[t1] += mu[i] * vecr[i];
              myi[t1] += mu[i] * veci[i];
            }
            if (non->propagation == 1) {
              or_coupling_S_f32(N, He, vecr, veci, non->ts, 1, non->deltat, non->couplingcut);
            } else {
              for (int i = 0; i < N; i++) dr[i] = vecr[i], di[i] = veci[i];
              if (sparse) or_DIA_S_f64(N, He, dr, di, 1, non->deltat, non->thres);
              else or_DIA_f64(N, He, dr, di, 1, non->deltat);
              for (int i = 0; i < N; i++) vecr[i] = (float)dr[i], veci[i] = (float)di[i];
            }
          }
          if (re_S64) {
            for (int i = 0; i < N; i++) {
              myr64[t1] += (double)mu[i] * dr[i];
              myi64[t1] += (double)mu[i] * di[i];
            }
            if (non->propagation == 1)
              or_coupling_S_f64(N, He, dr, di, non->ts, 1, non->deltat, non->couplingcut);
            else if (sparse)
              or_DIA_S_f64(N, He, dr, di, 1, non->deltat, non->thres);
            else
              or_DIA_f64(N, He, dr, di, 1, non->deltat);
          }
        }
      }
    }
    free(He), free(vecr), free(veci), free(mu), free(dr), free(di);
  }
  for (int t = 0; t < nthreads; t++)
    for (int t1 = 0; t1 < tmax; t1++) {
      if (re_S) re_S[t1] += pr[(size_t)t * tmax + t1], im_S[t1] += pi[(size_t)t * tmax + t1];
      if (re_S64) re_S64[t1] += pr64[(size_t)t * tmax + t1], im_S64[t1] += pi64[(size_t)t * tmax + t1];
    }
  free(pr), free(pi), free(pr64), free(pi64);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* FP64 absorption with gradient-run parameters: delta[a] added to every    */
/* frame's shifted site energy, kept couplings scaled by cscale (the        */
/* parameterisation of nise_abs_plan_set_params).  Reference for central   */
/* finite differences of the adjoint; propagation 1 only.                  */
/* ---------------------------------------------------------------------- */
int or_absorption_param_f64(const t_non *non, const float *H_rec, long long H_stride,
                            const float *mu_rec, long long mu_stride,
                            const float *C_rec, const double *delta, double cscale,
                            double *re_S, double *im_S, int nthreads) {
  int N = non->singles, tmax = non->tmax;
  int nn2 = N * (N + 1) / 2;
  if (non->propagation != 1) return NISE_ERR_ARG;
  memset(re_S, 0, tmax * sizeof(double)), memset(im_S, 0, tmax * sizeof(double));
  int nsamp = non->end - non->begin;
  if (nthreads < 1) nthreads = 1;
  double *pr = calloc((size_t)nthreads * tmax, sizeof(double));
  double *pi = calloc((size_t)nthreads * tmax, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0, nt = 1;
#ifdef _OPENMP
    tid = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int s0 = non->begin + (int)((long long)nsamp * tid / nt);
    int s1 = non->begin + (int)((long long)nsamp * (tid + 1) / nt);
    float *He = malloc(nn2 * sizeof(float));
    float *mu = malloc(N * sizeof(float));
    double *dr = malloc(N * sizeof(double)), *di = malloc(N * sizeof(double));
    double *myr = pr + (size_t)tid * tmax, *myi = pi + (size_t)tid * tmax;
    for (int samples = s0; samples < s1; samples++) {
      int ti = samples * non->sample;
      for (int x = 0; x < 3; x++) {
        read_mu(non, mu_rec, mu_stride, ti, x, mu);
        for (int i = 0; i < N; i++) dr[i] = mu[i], di[i] = 0;
        for (int t1 = 0; t1 < tmax; t1++) {
          int tj = ti + t1;
          or_read_H(non, H_rec, H_stride, C_rec, tj, He, 1);
          read_mu(non, mu_rec, mu_stride, tj, x, mu);
          for (int i = 0; i < N; i++) {
            myr[t1] += (double)mu[i] * dr[i];
            myi[t1] += (double)mu[i] * di[i];
          }
          coupling_S_f64_impl(N, He, dr, di, non->ts, 1, non->deltat, non->couplingcut, 0,
                              delta, cscale);
        }
      }
    }
    free(He), free(mu), free(dr), free(di);
  }
  for (int t = 0; t < nthreads; t++)
    for (int t1 = 0; t1 < tmax; t1++)
      re_S[t1] += pr[(size_t)t * tmax + t1], im_S[t1] += pi[(size_t)t * tmax + t1];
  free(pr), free(pi);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* 1D FFT and output grid, src/1DFFT.c:144-228 (FFTW forward, double).     */
/* The transform is a direct DFT in double with exact index reduction.     */
/* Returns the rows Absorption.dat would hold, in file order.              */
/* ---------------------------------------------------------------------- */
int or_1dfft(const t_non *non, const float *re_S, const float *im_S,
             int samples, double *freq, double *spec_r, double *spec_i,
             int *nout) {
  int fft = non->fft;
  if (fft < non->tmax1) fft = non->tmax1;
  float shift1 = non->shifte;
  double *inr = calloc(fft, sizeof(double)), *ini = calloc(fft, sizeof(double));
  for (int i = 0; i < non->tmax1; i++) {
    inr[i] = im_S[i] / samples;
    ini[i] = re_S[i] / samples;
    if (non->lifetime > 0.0) {
      inr[i] *= exp(-i * non->deltat / (2 * non->lifetime));
      ini[i] *= exp(-i * non->deltat / (2 * non->lifetime));
    }
    if (non->homogen > 0.0) {
      inr[i] *= exp(-i * non->deltat / (2 * non->homogen));
      ini[i] *= exp(-i * non->deltat / (2 * non->homogen));
    }
    if (non->inhomogen > 0.0) {
      inr[i] *= exp(-i * non->deltat * i * non->deltat / (2 * non->inhomogen * non->inhomogen));
      ini[i] *= exp(-i * non->deltat * i * non->deltat / (2 * non->inhomogen * non->inhomogen));
    }
    if (non->window == 1) {
      double wv = cos(i * OR_twoPi / (4 * non->tmax - 4)) * cos(i * OR_twoPi / (4 * non->tmax - 4));
      inr[i] *= wv;
      ini[i] *= wv;
    }
  }
  inr[0] *= 0.5, ini[0] *= 0.5;
  double *ctab = malloc(fft * sizeof(double)), *stab = malloc(fft * sizeof(double));
  for (int j = 0; j < fft; j++) {
    ctab[j] = cos(2.0 * M_PI * j / fft);
    stab[j] = sin(2.0 * M_PI * j / fft);
  }
  float *sr = malloc(fft * sizeof(float)), *si = malloc(fft * sizeof(float));
  for (int k = 0; k < fft; k++) {
    double xr = 0, xi = 0;
    for (int n = 0; n < non->tmax1; n++) {
      int j = (int)(((long long)n * k) % fft);
      /* forward: e^{-2 pi i n k / fft} */
      xr += inr[n] * ctab[j] + ini[n] * stab[j];
      xi += ini[n] * ctab[j] - inr[n] * stab[j];
    }
    sr[k] = xi * 2 * non->deltat * OR_c_v; /* :203-206 re/im swap */
    si[k] = xr * 2 * non->deltat * OR_c_v;
  }
  int n = 0;
  for (int i = fft / 2; i <= fft - 1; i++) {
    float w = -((fft - i) / non->deltat / OR_c_v / fft - shift1);
    if (w > non->min1 && w < non->max1) {
      freq[n] = w, spec_r[n] = sr[i], spec_i[n] = si[i];
      n++;
    }
  }
  for (int i = 0; i <= fft / 2 - 1; i++) {
    float w = -((-i) / non->deltat / OR_c_v / fft - shift1);
    if (w > non->min1 && w < non->max1) {
      freq[n] = w, spec_r[n] = sr[i], spec_i[n] = si[i];
      n++;
    }
  }
  *nout = n;
  free(inr), free(ini), free(ctab), free(stab), free(sr), free(si);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Truncated one-exciton propagator, src/propagate.c:1241-1302:            */
/* U = V exp(-i E f/m) V^T (f = dt icm2ifs 2 pi / m, sign +1) in the site  */
/* basis, elements listed a-major, b-minor where FP32 |U_ab|^2 > thres:    */
/* Ur/Ui = U_ab, R = a, C = b.  The eigensystem is the FP64 Jacobi one     */
/* (the reference's is LAPACK ssyev_ in FP32); each element is rounded to  */
/* FP32 once, so U is exactly symmetric.  Returns the element count.      */
/* ---------------------------------------------------------------------- */
int or_time_evolution_mat(int N, const float *H, float deltat, int m, float thres,
                          float *Ur, float *Ui, int *R, int *C) {
  float ff = deltat * OR_icm2ifs * OR_twoPi / m; /* :1253 */
  double f = ff;
  double *A = malloc(N * N * sizeof(double));
  double *V = malloc(N * N * sizeof(double));
  double *w = malloc(N * sizeof(double));
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++) A[a * N + b] = H[or_sindex(a, b, N)];
  or_jacobi_f64(N, A, w, V);
  int elements = 0;
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++) {
      double sr = 0, si = 0;
      for (int c = 0; c < N; c++) {
        double p = V[a * N + c] * V[b * N + c];
        sr += p * cos(w[c] * f);
        si -= p * sin(w[c] * f);
      }
      float r = (float)sr, i = (float)si;
      if (r * r + i * i > thres) { /* :1291 */
        Ur[elements] = r, Ui[elements] = i;
        R[elements] = a, C[elements] = b;
        elements++;
      }
    }
  free(A), free(V), free(w);
  return elements;
}

/* ---------------------------------------------------------------------- */
/* Sparse two-exciton step, src/propagate.c:1304-1398 (three-level sites)  */
/* and :1400-1473 (two-level sites), in the reference's FP32 order: per    */
/* substep the anharmonic half phase on |aa> (a copy for two-level sites), */
/* the sum over element pairs b <= a of U_a U_b g f[Sindex(t,v)] into      */
/* v[Sindex(s,u)] with (s,t) = (R[a],C[a]), (u,v) = (R[b],C[b]), and the   */
/* second half phase.  g: sqrt2 when exactly one of s!=u, t!=v holds, else */
/* 1 (two-level sites: 0 there, else 1).  The statements keep the          */
/* reference's shapes: under -ffast-math the compiler's reassociation of   */
/* each update depends on them, and the results are bit-identical only so. */
/* ---------------------------------------------------------------------- */
static void sparce3(int N, const float *Ur, const float *Ui, const int *R, const int *C,
                    float *fr, float *fi, int elements, int m, const float *Anh,
                    float anharmonicity, float deltat) {
  int a, b, indexA, indexB;
  int Nf, s, t, u, v;
  float g;
  float *vr, *vi;
  float Uar, Ubr, Uai, Ubi;
  float *si, *co;
  float f, fm;
  int i;
  f = deltat * OR_icm2ifs * OR_twoPi;
  fm = f * 0.5 / m;
  Nf = N * (N + 1) / 2;
  vr = (float *)calloc(Nf, sizeof(float));
  vi = (float *)calloc(Nf, sizeof(float));
  co = (float *)calloc(N, sizeof(float));
  si = (float *)calloc(N, sizeof(float));
  if (anharmonicity != 0) {
    for (i = 0; i < N; i++) co[i] = cos(fm * anharmonicity), si[i] = sin(fm * anharmonicity);
  } else {
    for (i = 0; i < N; i++) co[i] = cos(fm * Anh[i]), si[i] = sin(fm * Anh[i]);
  }
  for (i = 0; i < m; i++) {
    for (a = 0; a < N; a++)
      for (b = 0; b <= a; b++) {
        indexA = or_sindex(a, b, N);
        vr[indexA] = fr[indexA];
        vi[indexA] = fi[indexA];
      }
    for (a = 0; a < N; a++) {
      indexA = or_sindex(a, a, N);
      fr[indexA] = co[a] * vr[indexA] - si[a] * vi[indexA];
      fi[indexA] = co[a] * vi[indexA] + si[a] * vr[indexA];
    }
    for (a = 0; a < N; a++)
      for (b = 0; b <= a; b++) {
        indexA = or_sindex(a, b, N);
        vr[indexA] = 0;
        vi[indexA] = 0;
      }
    for (a = 0; a < elements; a++) {
      s = R[a], t = C[a], Uar = Ur[a], Uai = Ui[a];
      for (b = 0; b <= a; b++) {
        u = R[b], v = C[b], Ubr = Ur[b], Ubi = Ui[b];
        g = 1;
        indexA = or_sindex(s, u, N), indexB = or_sindex(t, v, N);
        if (s != u) g = OR_sqrt2;
        if (v != t) g = OR_sqrt2;
        if (s != u && v != t) g = 1;
        vr[indexA] += (Uar * Ubr - Uai * Ubi) * fr[indexB] * g;
        vr[indexA] -= (Uai * Ubr + Uar * Ubi) * fi[indexB] * g;
        vi[indexA] += (Uar * Ubr - Uai * Ubi) * fi[indexB] * g;
        vi[indexA] += (Uai * Ubr + Uar * Ubi) * fr[indexB] * g;
      }
    }
    for (a = 0; a < N; a++)
      for (b = 0; b <= a; b++) {
        indexA = or_sindex(a, b, N);
        fr[indexA] = vr[indexA];
        fi[indexA] = vi[indexA];
      }
    for (a = 0; a < N; a++) {
      indexA = or_sindex(a, a, N);
      fr[indexA] = co[a] * vr[indexA] - si[a] * vi[indexA];
      fi[indexA] = co[a] * vi[indexA] + si[a] * vr[indexA];
    }
  }
  free(si), free(co), free(vr), free(vi);
}

static void sparce2(int N, const float *Ur, const float *Ui, const int *R, const int *C,
                    float *fr, float *fi, int elements, int m) {
  int a, b, indexA, indexB;
  int Nf, s, t, u, v;
  float g;
  float *vr, *vi;
  float Uar, Ubr, Uai, Ubi;
  int i;
  Nf = N * (N + 1) / 2;
  vr = (float *)calloc(Nf, sizeof(float));
  vi = (float *)calloc(Nf, sizeof(float));
  for (i = 0; i < m; i++) {
    for (a = 0; a < N; a++)
      for (b = 0; b <= a; b++) {
        indexA = or_sindex(a, b, N);
        vr[indexA] = fr[indexA];
        vi[indexA] = fi[indexA];
      }
    for (a = 0; a < N; a++) {
      indexA = or_sindex(a, a, N);
      fr[indexA] = vr[indexA];
      fi[indexA] = vi[indexA];
    }
    for (a = 0; a < N; a++)
      for (b = 0; b <= a; b++) {
        indexA = or_sindex(a, b, N);
        vr[indexA] = 0;
        vi[indexA] = 0;
      }
    for (a = 0; a < elements; a++) {
      s = R[a], t = C[a], Uar = Ur[a], Uai = Ui[a];
      for (b = 0; b <= a; b++) {
        u = R[b], v = C[b], Ubr = Ur[b], Ubi = Ui[b];
        g = 1;
        indexA = or_sindex(s, u, N), indexB = or_sindex(t, v, N);
        if (s != u) g = 0;
        if (v != t) g = 0;
        if (s != u && v != t) g = 1;
        vr[indexA] += (Uar * Ubr - Uai * Ubi) * fr[indexB] * g;
        vr[indexA] -= (Uai * Ubr + Uar * Ubi) * fi[indexB] * g;
        vi[indexA] += (Uar * Ubr - Uai * Ubi) * fi[indexB] * g;
        vi[indexA] += (Uai * Ubr + Uar * Ubi) * fr[indexB] * g;
      }
    }
    for (a = 0; a < N; a++)
      for (b = 0; b <= a; b++) {
        indexA = or_sindex(a, b, N);
        fr[indexA] = vr[indexA];
        fi[indexA] = vi[indexA];
      }
    for (a = 0; a < N; a++) {
      indexA = or_sindex(a, a, N);
      fr[indexA] = vr[indexA];
      fi[indexA] = vi[indexA];
    }
  }
  free(vr), free(vi);
}

void or_double_sparce_f32(int N, const float *Ur, const float *Ui, const int *R,
                          const int *C, float *fr, float *fi, int elements, int m,
                          const float *Anh, float anharmonicity, float deltat,
                          int es) {
  if (es)
    sparce2(N, Ur, Ui, R, C, fr, fi, elements, m);
  else
    sparce3(N, Ur, Ui, R, C, fr, fi, elements, m, Anh, anharmonicity, deltat);
}

/* src/NISE_subs.c:512-530: the ground-state double excitation, |ii> =    */
/* over_i (over = sqrt(2) dipole when the anharmonicity is fixed, written  */
/* back to over as the reference does), every other state 0.               */
void or_dipole_double_ground(int N, const float *dipole, float *fr, float *fi, float *over,
                             float anharmonicity) {
  int N2 = N * (N + 1) / 2;
  for (int i = 0; i < N2; i++) fr[i] = 0, fi[i] = 0;
  if (anharmonicity != 0)
    for (int i = 0; i < N; i++) over[i] = OR_sqrt2 * dipole[i];
  for (int i = 0; i < N; i++) fr[or_sindex(i, i, N)] = over[i];
}

/* src/polar.c:107-116: the Raman tensor component (xx, xy, xz, yy, yz, zz) */
/* of the polarisation pair (p2, p3), as its return value.                 */
int or_raman_pol(int p2, int p3) {
  if (p2 == 0 && p3 == 0) return 0;
  if ((p2 == 0 && p3 == 1) || (p2 == 1 && p3 == 0)) return 1;
  if ((p2 == 0 && p3 == 2) || (p2 == 2 && p3 == 0)) return 2;
  if (p2 == 1 && p3 == 1) return 3;
  if ((p2 == 1 && p3 == 2) || (p2 == 2 && p3 == 1)) return 4;
  return 5;
}

/* One singles step of propagate_vector (src/propagate.c:20-39): the        */
/* Trotter scheme for propagation 1; for propagation 0 the exact            */
/* exponential when thres is 0 or above 1, else the sparse scheme (DIA_S,   */
/* always when force_sparse, as src/propagate.c:1519-1525 calls it); the    */
/* exponential schemes in FP64, the vector rounded to FP32 after the step.  */
static void or_singles_step(const t_non *non, const float *He, float *cr, float *ci,
                            int sign, int force_sparse) {
  int N = non->singles;
  if (non->propagation == 1) {
    or_coupling_S_f32(N, He, cr, ci, non->ts, sign, non->deltat, non->couplingcut);
    return;
  }
  double *dr = malloc(N * sizeof(double)), *di = malloc(N * sizeof(double));
  for (int i = 0; i < N; i++) dr[i] = cr[i], di[i] = ci[i];
  if (force_sparse || (non->thres > 0 && non->thres <= 1))
    or_DIA_S_f64(N, He, dr, di, sign, non->deltat, non->thres);
  else
    or_DIA_f64(N, He, dr, di, sign, non->deltat);
  for (int i = 0; i < N; i++) cr[i] = (float)dr[i], ci[i] = (float)di[i];
  free(dr), free(di);
}

/* One waiting-time step, src/propagate.c:130-180: the exact exponential    */
/* (no Trotter split, sign +1) for whatever propagation scheme is chosen.   */
static void or_t2_step(int N, const float *He, float *cr, float *ci, float deltat) {
  double *dr = malloc(N * sizeof(double)), *di = malloc(N * sizeof(double));
  for (int i = 0; i < N; i++) dr[i] = cr[i], di[i] = ci[i];
  or_DIA_f64(N, He, dr, di, 1, deltat);
  for (int i = 0; i < N; i++) cr[i] = (float)dr[i], ci[i] = (float)di[i];
  free(dr), free(di);
}

/* ---------------------------------------------------------------------- */
/* One 2DIR work item, src/calc_2DIR.c:184-506 (propagation 1, the Trotter */
/* scheme, or 0, the exact exponential / sparse scheme of                   */
/* src/propagate.c:1476-1533; waiting time tmax2 >= 0 propagated as        */
/* src/calc_2DIR.c:303-326 does).  Adds into R12 = 12 arrays [tmax3][tmax1] */
/* in the order of src/calc_2DIR.c:540-542.                                */
/* ---------------------------------------------------------------------- */
int or_2dir_item(const t_non *non, const float *H_rec, long long H_stride,
                 const float *mu_rec, long long mu_stride, const float *C_rec,
                 const float *A_rec, const float *over_rec, int sample,
                 int molPol, float *R12) {
  int N = non->singles, T1 = non->tmax1, T3 = non->tmax3;
  int nn2 = N * (N + 1) / 2;
  if (non->tmax2 < 0 || (non->propagation != 1 && non->propagation != 0)) return NISE_ERR_ARG;
  const char *tech = non->technique;
  /* calc_2DES (src/calc_2DES.c:22-586) is calc_2DIR for two-level sites:
     its techniques end in UVvis and its doubles are the _ES forms */
  int es = !strcmp(tech, "GBUVvis") || !strcmp(tech, "SEUVvis") || !strcmp(tech, "EAUVvis") ||
           !strcmp(tech, "2DUVvis") || !strcmp(tech, "noEAUVvis");
  int do_gb = !strcmp(tech, "GBIR") || !strcmp(tech, "2DIR") || !strcmp(tech, "noEAIR") ||
              !strcmp(tech, "GBUVvis") || !strcmp(tech, "2DUVvis") || !strcmp(tech, "noEAUVvis");
  int do_se = !strcmp(tech, "SEIR") || !strcmp(tech, "2DIR") || !strcmp(tech, "noEAIR") ||
              !strcmp(tech, "SEUVvis") || !strcmp(tech, "2DUVvis") || !strcmp(tech, "noEAUVvis");
  int do_ea = !strcmp(tech, "EAIR") || !strcmp(tech, "2DIR") || !strcmp(tech, "EAUVvis") ||
              !strcmp(tech, "2DUVvis");
  size_t plane = (size_t)T3 * T1;
#define RA(q, t3, t1) R12[(size_t)(q) * plane + (size_t)(t3) * T1 + (t1)]
  float *He = malloc(nn2 * sizeof(float));
  float *Anh = calloc(N, sizeof(float)), *over = calloc(N, sizeof(float));
  float *leftrr = calloc(N, sizeof(float)), *leftri = calloc(N, sizeof(float));
  float *leftnr = calloc((size_t)T1 * N, sizeof(float)), *leftni = calloc((size_t)T1 * N, sizeof(float));
  float *rightrr = calloc((size_t)T1 * N, sizeof(float)), *rightri = calloc((size_t)T1 * N, sizeof(float));
  float *rightnr = calloc(N, sizeof(float)), *rightni = calloc(N, sizeof(float));
  float *mut2 = calloc(N, sizeof(float)), *mut3r = calloc(N, sizeof(float));
  float *mut3i = calloc(N, sizeof(float)), *mut4 = calloc(N, sizeof(float));
  float *fr = calloc(nn2, sizeof(float)), *fi = calloc(nn2, sizeof(float));
  float *ft1r = calloc((size_t)T1 * nn2, sizeof(float)), *ft1i = calloc((size_t)T1 * nn2, sizeof(float));
  float *t1nr = calloc(T1, sizeof(float)), *t1ni = calloc(T1, sizeof(float));
  float *t1rr = calloc(T1, sizeof(float)), *t1ri = calloc(T1, sizeof(float));
  float *lt = malloc(plane * sizeof(float));
  float *Urs = malloc((size_t)N * N * sizeof(float)), *Uis = malloc((size_t)N * N * sizeof(float));
  int *Rs = malloc((size_t)N * N * sizeof(int)), *Cs = malloc((size_t)N * N * sizeof(int));
  for (int t1 = 0; t1 < T1; t1++)
    for (int t3 = 0; t3 < T3; t3++)
      lt[t3 * T1 + t1] = (non->lifetime > 0)
          ? (float)exp(-(double)(t1 + t3) * non->deltat / (2 * non->lifetime)) : 1.0f;
  float wpol[3];
  for (int p = 0; p < 3; p++) wpol[p] = or_polarweight(p, molPol);
  int tj = sample * non->sample + T1, tk = tj + non->tmax2;
  int px[4];
  or_polar(px, molPol);
  float dt = non->deltat, cut = non->couplingcut;
  int ts = non->ts;

  read_mu(non, mu_rec, mu_stride, tj, px[1], mut2);
  /* GB: t1 seeds propagated t1 steps (:215-239) */
  for (int t1 = 0; t1 < T1; t1++) {
    read_mu(non, mu_rec, mu_stride, tj - t1, px[0], leftnr + (size_t)t1 * N);
    for (int tm = 0; tm < t1; tm++) {
      or_read_H(non, H_rec, H_stride, C_rec, tj - t1 + tm, He, 0);
      or_singles_step(non, He, leftnr + (size_t)t1 * N, leftni + (size_t)t1 * N, 1, 0);
    }
  }
  for (int t1 = 0; t1 < T1; t1++) { /* :244-250 */
    t1nr[t1] = 0, t1ni[t1] = 0;
    for (int i = 0; i < N; i++) {
      t1nr[t1] += mut2[i] * leftnr[(size_t)t1 * N + i];
      t1ni[t1] += mut2[i] * leftni[(size_t)t1 * N + i];
    }
  }
  read_mu(non, mu_rec, mu_stride, tk, px[2], mut3r);
  for (int t3 = 0; t3 < T3; t3++) { /* :257-299 */
    int tl = tk + t3;
    read_mu(non, mu_rec, mu_stride, tl, px[3], mut4);
    if (do_gb) {
      float t3nr = 0, t3ni = 0;
      for (int i = 0; i < N; i++) t3nr += mut4[i] * mut3r[i], t3ni += mut4[i] * mut3i[i];
      for (int t1 = 0; t1 < T1; t1++) {
        for (int p = 0; p < 3; p++) {
          float pw = wpol[p] * lt[t3 * T1 + t1];
          RA(4 * p + 0, t3, t1) -= (t3ni * t1nr[t1] - t3nr * t1ni[t1]) * pw;
          RA(4 * p + 1, t3, t1) -= (t3nr * t1nr[t1] + t3ni * t1ni[t1]) * pw;
          RA(4 * p + 2, t3, t1) -= (t3ni * t1nr[t1] + t3nr * t1ni[t1]) * pw;
          RA(4 * p + 3, t3, t1) -= (t3nr * t1nr[t1] - t3ni * t1ni[t1]) * pw;
        }
      }
    }
    or_read_H(non, H_rec, H_stride, C_rec, tl, He, 0);
    or_singles_step(non, He, mut3r, mut3i, 1, 0);
  }
  /* SE (:303-443), after the waiting time (:305-326) */
  read_mu(non, mu_rec, mu_stride, tj, px[1], leftrr);
  for (int i = 0; i < N; i++) leftri[i] = 0;
  for (int t2 = 0; t2 < non->tmax2; t2++) {
    or_read_H(non, H_rec, H_stride, C_rec, tj + t2, He, 0);
    or_t2_step(N, He, leftrr, leftri, dt);
    for (int t1 = 0; t1 < T1; t1++)
      or_t2_step(N, He, leftnr + (size_t)t1 * N, leftni + (size_t)t1 * N, dt);
  }
  read_mu(non, mu_rec, mu_stride, tk, px[2], mut3r);
  if (do_ea && es) { /* src/calc_2DES.c:318-323 */
    or_dipole_double_ES(N, mut3r, leftrr, leftri, fr, fi);
    for (int t1 = 0; t1 < T1; t1++)
      or_dipole_double_ES(N, mut3r, leftnr + (size_t)t1 * N, leftni + (size_t)t1 * N,
                          ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2);
  } else if (do_ea) { /* :331-348 */
    if (non->anharmonicity == 0) {
      const float *o = over_rec + (long long)tk * (1 + 3 * N) + 1 + px[2] * N;
      for (int i = 0; i < N; i++) over[i] = o[i];
    }
    or_dipole_double(N, mut3r, leftrr, leftri, fr, fi, over, non->anharmonicity);
    for (int t1 = 0; t1 < T1; t1++)
      or_dipole_double(N, mut3r, leftnr + (size_t)t1 * N, leftni + (size_t)t1 * N,
                       ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2, over, non->anharmonicity);
  }
  if (do_ea) {
    memcpy(rightrr, leftnr, (size_t)T1 * N * sizeof(float));
    memcpy(rightri, leftni, (size_t)T1 * N * sizeof(float));
    for (size_t i = 0; i < (size_t)T1 * N; i++) rightri[i] = -rightri[i];
    memcpy(rightnr, leftrr, N * sizeof(float));
    memcpy(rightni, leftri, N * sizeof(float));
    for (int i = 0; i < N; i++) rightni[i] = -rightni[i];
  }
  for (int i = 0; i < N; i++) mut3i[i] = 0;
  float t3nr = 0, t3ni = 0;
  for (int i = 0; i < N; i++) t3nr += leftrr[i] * mut3r[i], t3ni -= leftri[i] * mut3r[i];
  for (int t1 = 0; t1 < T1; t1++) {
    t1rr[t1] = 0, t1ri[t1] = 0;
    for (int i = 0; i < N; i++) {
      t1rr[t1] += leftnr[(size_t)t1 * N + i] * mut3r[i];
      t1ri[t1] -= leftni[(size_t)t1 * N + i] * mut3r[i];
    }
  }
  for (int t3 = 0; t3 < T3; t3++) {
    int tl = tk + t3;
    read_mu(non, mu_rec, mu_stride, tl, px[3], mut4);
    float t3rr = 0, t3ri = 0;
    for (int i = 0; i < N; i++) t3rr += mut4[i] * leftrr[i], t3ri += mut4[i] * leftri[i];
    for (int t1 = 0; t1 < T1; t1++) {
      t1nr[t1] = 0, t1ni[t1] = 0;
      for (int i = 0; i < N; i++) {
        t1nr[t1] += leftnr[(size_t)t1 * N + i] * mut4[i];
        t1ni[t1] += leftni[(size_t)t1 * N + i] * mut4[i];
      }
    }
    if (do_se) {
      for (int t1 = 0; t1 < T1; t1++) {
        for (int p = 0; p < 3; p++) {
          float pw = wpol[p] * lt[t3 * T1 + t1];
          RA(4 * p + 0, t3, t1) -= (t3rr * t1ri[t1] + t3ri * t1rr[t1]) * pw;
          RA(4 * p + 1, t3, t1) -= (t3rr * t1rr[t1] - t3ri * t1ri[t1]) * pw;
          RA(4 * p + 2, t3, t1) -= (t3ni * t1nr[t1] + t3nr * t1ni[t1]) * pw;
          RA(4 * p + 3, t3, t1) -= (t3nr * t1nr[t1] - t3ni * t1ni[t1]) * pw;
        }
      }
    }
    or_read_H(non, H_rec, H_stride, C_rec, tl, He, 0);
    or_singles_step(non, He, leftrr, leftri, 1, 0);
    for (int t1 = 0; t1 < T1; t1++)
      or_singles_step(non, He, leftnr + (size_t)t1 * N, leftni + (size_t)t1 * N, 1, 0);
  }
  if (do_ea) { /* :445-506 */
    for (int t3 = 0; t3 < T3; t3++) {
      int tl = tk + t3;
      read_mu(non, mu_rec, mu_stride, tl, px[3], mut4);
      if (es) { /* src/calc_2DES.c:428-432 */
        or_dipole_double_last_ES(N, mut4, fr, fi, leftrr, leftri);
        for (int t1 = 0; t1 < T1; t1++)
          or_dipole_double_last_ES(N, mut4, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2,
                                   leftnr + (size_t)t1 * N, leftni + (size_t)t1 * N);
      } else {
        if (non->anharmonicity == 0) {
          const float *o = over_rec + (long long)tl * (1 + 3 * N) + 1 + px[3] * N;
          for (int i = 0; i < N; i++) over[i] = o[i];
        }
        or_dipole_double_last(N, mut4, fr, fi, leftrr, leftri, over, non->anharmonicity);
        for (int t1 = 0; t1 < T1; t1++)
          or_dipole_double_last(N, mut4, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2,
                                leftnr + (size_t)t1 * N, leftni + (size_t)t1 * N, over,
                                non->anharmonicity);
      }
      for (int t1 = 0; t1 < T1; t1++) {
        float rrI = 0, riI = 0, rrII = 0, riII = 0;
        const float *Rr = rightrr + (size_t)t1 * N, *Ri = rightri + (size_t)t1 * N;
        const float *Lr = leftnr + (size_t)t1 * N, *Li = leftni + (size_t)t1 * N;
        for (int i = 0; i < N; i++) {
          rrI += leftri[i] * Rr[i] + leftrr[i] * Ri[i];
          riI += leftrr[i] * Rr[i] - Ri[i] * leftri[i];
          rrII += rightnr[i] * Li[i] + rightni[i] * Lr[i];
          riII += rightnr[i] * Lr[i] - rightni[i] * Li[i];
        }
        for (int p = 0; p < 3; p++) {
          float pw = wpol[p] * lt[t3 * T1 + t1];
          RA(4 * p + 0, t3, t1) += rrI * pw;
          RA(4 * p + 1, t3, t1) += riI * pw;
          RA(4 * p + 2, t3, t1) += rrII * pw;
          RA(4 * p + 3, t3, t1) += riII * pw;
        }
      }
      or_read_H(non, H_rec, H_stride, C_rec, tl, He, 0);
      if (non->propagation == 0) {
        /* src/propagate.c:1486-1525 (and :1606-1645 for two-level sites):
           the caller passes the row and column arrays swapped */
        if (non->anharmonicity == 0 && !es) {
          const float *a = A_rec + (long long)tl * (1 + N) + 1;
          for (int i = 0; i < N; i++) Anh[i] = a[i];
        }
        int el = or_time_evolution_mat(N, He, dt, ts, non->thres, Urs, Uis, Cs, Rs);
        or_double_sparce_f32(N, Urs, Uis, Rs, Cs, fr, fi, el, ts, Anh, non->anharmonicity, dt, es);
        for (int t1 = 0; t1 < T1; t1++)
          or_double_sparce_f32(N, Urs, Uis, Rs, Cs, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2,
                               el, ts, Anh, non->anharmonicity, dt, es);
        or_singles_step(non, He, rightnr, rightni, -1, 1);
        for (int t1 = 0; t1 < T1; t1++)
          or_singles_step(non, He, rightrr + (size_t)t1 * N, rightri + (size_t)t1 * N, -1, 1);
        continue;
      }
      if (es) { /* propagate_double_2DES_control, src/propagate.c:1664-1683 */
        or_coupling_S_doubles_ES_f32(N, He, fr, fi, ts, dt, cut);
        for (int t1 = 0; t1 < T1; t1++)
          or_coupling_S_doubles_ES_f32(N, He, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2, ts,
                                       dt, cut);
      } else {
        if (non->anharmonicity == 0) {
          const float *a = A_rec + (long long)tl * (1 + N) + 1;
          for (int i = 0; i < N; i++) Anh[i] = a[i];
        }
        or_coupling_S_doubles_f32(N, He, fr, fi, ts, Anh, non->anharmonicity, dt, cut);
        for (int t1 = 0; t1 < T1; t1++)
          or_coupling_S_doubles_f32(N, He, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2, ts,
                                    Anh, non->anharmonicity, dt, cut);
      }
      or_coupling_S_f32(N, He, rightnr, rightni, ts, -1, dt, cut);
      for (int t1 = 0; t1 < T1; t1++)
        or_coupling_S_f32(N, He, rightrr + (size_t)t1 * N, rightri + (size_t)t1 * N, ts, -1, dt, cut);
    }
  }
#undef RA
  free(He), free(Anh), free(over), free(leftrr), free(leftri), free(leftnr), free(leftni);
  free(rightrr), free(rightri), free(rightnr), free(rightni), free(mut2), free(mut3r);
  free(mut3i), free(mut4), free(fr), free(fi), free(ft1r), free(ft1i);
  free(t1nr), free(t1ni), free(t1rr), free(t1ri), free(lt);
  free(Urs), free(Uis), free(Rs), free(Cs);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* One 2DIR-Raman work item, src/calc_2DIRraman.c:176-441 with             */
/* propagate_1 (:591-633): the Trotter scheme, a fixed anharmonicity (the  */
/* reference exits for a per-frame one, :222-226, :330-334).  Two changes, */
/* both where the reference's result is not defined: the Raman component   */
/* is find_raman_pol's return value (the reference discards it, :319, and  */
/* reads its uninitialised alphapol), and lifetime 0 means no lifetime     */
/* decay (the reference divides by 2 lifetime, :106-107).  Everything else */
/* is the reference's, the technique variants' block conditions included   */
/* (2DIRramanI computes diagram 1 without propagating fr and rightrr in t3; */
/* 2DIRraman2 propagates them, :620-629).  Adds into R12 = 12 arrays        */
/* [tmax3][tmax1] in the order of :466-468 (rrIpar, riIpar, rrIIpar, ...).  */
/* alpha_rec: per frame 1 + 6N words, components xx xy xz yy yz zz         */
/* (read_alpha, src/read_trajectory.c:380-399).                            */
/* ---------------------------------------------------------------------- */
static int tech_in(const char *t, const char *const *set, int n) {
  for (int i = 0; i < n; i++)
    if (!strcmp(t, set[i])) return 1;
  return 0;
}

int or_2dirraman_item(const t_non *non, const float *H_rec, long long H_stride,
                      const float *mu_rec, long long mu_stride, const float *C_rec,
                      const float *alpha_rec, int sample, int molPol, float *R12) {
  int N = non->singles, T1 = non->tmax1, T3 = non->tmax3;
  int nn2 = N * (N + 1) / 2;
  if (non->propagation != 1 || non->anharmonicity == 0) return NISE_ERR_ARG;
  const char *tech = non->technique;
  static const char *const s_k1[] = {"2DIRraman", "2DIRraman1", "2DIRramanI"};
  static const char *const s_k2[] = {"2DIRraman", "2DIRraman2", "2DIRraman3", "2DIRramanII"};
  static const char *const s_d2[] = {"2DIRraman", "2DIRraman2", "2DIRramanII"};
  static const char *const s_d3[] = {"2DIRraman", "2DIRraman3", "2DIRramanII"};
  static const char *const s_p1[] = {"2DIRraman", "2DIRraman1", "2DIRraman2"};
  int k1 = tech_in(tech, s_k1, 3), k2 = tech_in(tech, s_k2, 4), d2 = tech_in(tech, s_d2, 3);
  int d3 = tech_in(tech, s_d3, 3), p1 = tech_in(tech, s_p1, 3);
  if (!k1 && !k2) return NISE_ERR_ARG;
  size_t plane = (size_t)T3 * T1;
#define RA(q, t3, t1) R12[(size_t)(q) * plane + (size_t)(t3) * T1 + (t1)]
  float *He = malloc(nn2 * sizeof(float));
  float *Anh = calloc(N, sizeof(float)), *over = calloc(N, sizeof(float));
  float *leftrr = calloc(N, sizeof(float)), *leftri = calloc(N, sizeof(float));
  float *leftnr = calloc((size_t)T1 * N, sizeof(float)), *leftni = calloc((size_t)T1 * N, sizeof(float));
  float *rightrr = calloc((size_t)T1 * N, sizeof(float)), *rightri = calloc((size_t)T1 * N, sizeof(float));
  float *rightnr = calloc(N, sizeof(float)), *rightni = calloc(N, sizeof(float));
  float *l3r = calloc(N, sizeof(float)), *l3i = calloc(N, sizeof(float));
  float *mut3r = calloc(N, sizeof(float)), *alpha4 = calloc(N, sizeof(float));
  float *fr = calloc(nn2, sizeof(float)), *fi = calloc(nn2, sizeof(float));
  float *ft1r = calloc((size_t)T1 * nn2, sizeof(float)), *ft1i = calloc((size_t)T1 * nn2, sizeof(float));
  float *t1nr = calloc((size_t)T1 * N, sizeof(float)), *t1ni = calloc((size_t)T1 * N, sizeof(float));
  float *lt = malloc(plane * sizeof(float));
  for (int t1 = 0; t1 < T1; t1++)
    for (int t3 = 0; t3 < T3; t3++)
      lt[t3 * T1 + t1] = (non->lifetime > 0)
          ? (float)exp(-(double)(t1 + t3) * non->deltat / (2 * non->lifetime)) : 1.0f;
  float wpol[3];
  for (int p = 0; p < 3; p++) wpol[p] = or_polarweight(p, molPol);
  int tj = sample * non->sample + T1;
  int px[4];
  or_polar(px, molPol);
  int apol = or_raman_pol(px[2], px[3]);
  float dt = non->deltat, cut = non->couplingcut, anh = non->anharmonicity;
  int ts = non->ts;

  /* t1: kI singles and kII doubles from mu_x0(tj - t1) (:197-246) */
  for (int t1 = 0; t1 < T1; t1++) {
    float *lr = leftnr + (size_t)t1 * N, *li = leftni + (size_t)t1 * N;
    float *gr = ft1r + (size_t)t1 * nn2, *gi = ft1i + (size_t)t1 * nn2;
    read_mu(non, mu_rec, mu_stride, tj - t1, px[0], lr);
    if (k2) or_dipole_double_ground(N, lr, gr, gi, over, anh);
    for (int tm = 0; tm < t1; tm++) {
      or_read_H(non, H_rec, H_stride, C_rec, tj - t1 + tm, He, 0);
      if (k1) or_coupling_S_f32(N, He, lr, li, ts, 1, dt, cut);
      if (k2) or_coupling_S_doubles_f32(N, He, gr, gi, ts, Anh, anh, dt, cut);
    }
  }
  read_mu(non, mu_rec, mu_stride, tj, px[1], mut3r); /* :249 */
  if (k1) { /* :255-264 */
    or_dipole_double_ground(N, mut3r, fr, fi, over, anh);
    memcpy(rightrr, leftnr, (size_t)T1 * N * sizeof(float));
    memcpy(rightri, leftni, (size_t)T1 * N * sizeof(float));
    for (size_t i = 0; i < (size_t)T1 * N; i++) rightri[i] = -rightri[i];
  }
  if (d2) { /* :267-274 */
    memcpy(rightnr, mut3r, N * sizeof(float));
    for (int i = 0; i < N; i++) rightni[i] = 0;
  }
  if (d3) /* :280-287 */
    for (int t1 = 0; t1 < T1; t1++)
      or_dipole_double_last(N, mut3r, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2,
                            t1nr + (size_t)t1 * N, t1ni + (size_t)t1 * N, over, anh);
  const float one = 1, zero = 0; /* the ground-state right side of diagram 3, :389 */
  for (int t3 = 0; t3 < T3; t3++) {
    int tl = tj + t3;
    const float *a = alpha_rec + (long long)tl * (1 + 6 * N) + 1 + (long long)apol * N;
    for (int i = 0; i < N; i++) alpha4[i] = a[i];
    if (k1) { /* diagram 1, :306-330 */
      or_dipole_double_last(N, alpha4, fr, fi, leftrr, leftri, over, anh);
      for (int t1 = 0; t1 < T1; t1++) {
        float rrI = 0, riI = 0;
        const float *Rr = rightrr + (size_t)t1 * N, *Ri = rightri + (size_t)t1 * N;
        for (int i = 0; i < N; i++) {
          rrI += leftri[i] * Rr[i] + leftrr[i] * Ri[i];
          riI += leftrr[i] * Rr[i] - Ri[i] * leftri[i];
        }
        for (int p = 0; p < 3; p++) {
          float pw = wpol[p] * lt[t3 * T1 + t1];
          RA(4 * p + 0, t3, t1) += rrI * pw;
          RA(4 * p + 1, t3, t1) += riI * pw;
        }
      }
    }
    if (d2) /* diagram 2, :333-356 */
      for (int t1 = 0; t1 < T1; t1++) {
        float *Lr = leftnr + (size_t)t1 * N, *Li = leftni + (size_t)t1 * N;
        or_dipole_double_last(N, alpha4, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2, Lr, Li,
                              over, anh);
        float rrII = 0, riII = 0;
        for (int i = 0; i < N; i++) {
          rrII += rightnr[i] * Li[i] + rightni[i] * Lr[i];
          riII += rightnr[i] * Lr[i] - rightni[i] * Li[i];
        }
        for (int p = 0; p < 3; p++) {
          float pw = wpol[p] * lt[t3 * T1 + t1];
          RA(4 * p + 2, t3, t1) += rrII * pw;
          RA(4 * p + 3, t3, t1) += riII * pw;
        }
      }
    if (d3) /* diagram 3, :359-402 */
      for (int t1 = 0; t1 < T1; t1++) {
        const float *nr = t1nr + (size_t)t1 * N, *ni = t1ni + (size_t)t1 * N;
        for (int i = 0; i < N; i++) l3r[i] = 0, l3i[i] = 0;
        for (int i = 0; i < N; i++) l3r[i] += alpha4[i] * nr[i], l3i[i] += alpha4[i] * ni[i];
        float rrII = 0, riII = 0;
        for (int i = 0; i < N; i++) {
          rrII += one * l3i[i] + zero * l3r[i];
          riII += one * l3r[i] - zero * l3i[i];
        }
        for (int p = 0; p < 3; p++) {
          float pw = wpol[p] * lt[t3 * T1 + t1];
          RA(4 * p + 2, t3, t1) -= rrII * pw;
          RA(4 * p + 3, t3, t1) -= riII * pw;
        }
      }
    or_read_H(non, H_rec, H_stride, C_rec, tl, He, 0); /* :421, then propagate_1 */
    if (d2) {
      for (int t1 = 0; t1 < T1; t1++)
        or_coupling_S_doubles_f32(N, He, ft1r + (size_t)t1 * nn2, ft1i + (size_t)t1 * nn2, ts, Anh,
                                  anh, dt, cut);
      or_coupling_S_f32(N, He, rightnr, rightni, ts, -1, dt, cut);
    }
    if (d3)
      for (int t1 = 0; t1 < T1; t1++)
        or_coupling_S_f32(N, He, t1nr + (size_t)t1 * N, t1ni + (size_t)t1 * N, ts, 1, dt, cut);
    if (p1) {
      or_coupling_S_doubles_f32(N, He, fr, fi, ts, Anh, anh, dt, cut);
      for (int t1 = 0; t1 < T1; t1++)
        or_coupling_S_f32(N, He, rightrr + (size_t)t1 * N, rightri + (size_t)t1 * N, ts, -1, dt, cut);
    }
  }
#undef RA
  free(He), free(Anh), free(over), free(leftrr), free(leftri), free(leftnr), free(leftni);
  free(rightrr), free(rightri), free(rightnr), free(rightni), free(l3r), free(l3i), free(mut3r);
  free(alpha4), free(fr), free(fi), free(ft1r), free(ft1i), free(t1nr), free(t1ni), free(lt);
  return 0;
}
