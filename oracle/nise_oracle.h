/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference NISE hot path (GHlacour/NISE_2017), used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in nise_2017_amd/ links, loads or calls it.
 *
 * Every function states the reference lines it follows.  *_f32 variants keep
 * the reference's FP32 arithmetic in the reference's order (compiled with the
 * reference's flags -O2 -ffast-math, so they are bit-identical to the
 * reference build where one can be made, see oracle/Makefile); *_f64
 * variants are the same algorithm in double precision and give the error
 * figure against which every FP32 result is reported.
 *
 * Pinning (what fixes each function independently of this file):
 *   or_coupling_S_f32 / or_coupling_S_doubles_f32 / _ES_f32 / or_dipole_double* /
 *   or_polar* / or_read_H* (including the TransitionDipole /
 *   ExtendedDipole couplings built from geometry) : bit-identical to the reference's own
 *     propagate.c / NISE_subs.c / polar.c / read_trajectory.c compiled from
 *     the reference sources (oracle/_ref), and to tests/golden fixtures.
 *   or_coupling_S_f64 : closed forms (uncoupled phases, dimer D.R.D,
 *     norm conservation), substep identity, FP64 expm limit.
 *   or_DIA_f64 / or_jacobi_f64 : eigen-decomposition residuals and FP64
 *     matrix exponential (numpy eigh / scipy expm) in tests.
 *   or_DIA_S_f64 : equals or_DIA_f64 when no element is dropped (tiny
 *     threshold); a dimer's analytic propagator (off-diagonal magnitude
 *     |sin(J f)|) with the off-diagonal pair dropped or kept on either side
 *     of the threshold; uncoupled sites (diagonal U) unchanged (tests).
 *   or_absorption : per-step composition of the pinned primitives; compared
 *     with the reference-primitive driver in oracle/_ref.
 *   or_1dfft : numpy.fft with the reference's re/im swap and scaling.
 *   or_absorption_param_f64 : or_absorption's FP64 path with the gradient
 *     parameters; at delta = 0, cscale = 1 it equals or_absorption's FP64
 *     result exactly (tests), and an energy offset on an uncoupled site is a
 *     pure phase e^{-i f delta t} (closed form in tests).
 *   or_time_evolution_mat / or_double_sparce_f32 : the sparse step is
 *     bit-identical to the reference's propagate_double_sparce(_ES) given
 *     the same element list; the element list agrees with the reference's
 *     time_evolution_mat built against LAPACK (oracle/_ref, MKL) to the
 *     eigensolver's rounding, with the same kept set (tests/golden).
 *   or_DIA_f64 / or_DIA_S_f64 / the t2 step : also checked against the
 *     reference's propagate_vec_DIA / _DIA_S / propagate_t2_DIA built
 *     against LAPACK, to the eigensolver's rounding.
 *   or_2dir_item : composition of pinned primitives, checked against the
 *     closed form of uncoupled anharmonic sites (each pathway alone, the
 *     two-level sum, per-frame anharmonicities and overtone dipoles), and
 *     against the R*.dat files of the reference's own calc_2DIR / calc_2DES
 *     (oracle/_ref/nise_ref2d, MPI singleton) for coupled systems, the
 *     waiting time and the sparse scheme (tests/golden/reference_2d.npz).
 */
#ifndef NISE_ORACLE_H
#define NISE_ORACLE_H

#include "../include/nise_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

int or_sindex(int a, int b, int N);

void or_coupling_S_f32(int N, const float *H, float *cr, float *ci, int m,
                       int sign, float deltat, float couplingcut);
void or_coupling_S_f64(int N, const float *H, double *cr, double *ci, int m,
                       int sign, float deltat, float couplingcut);
/* Exact inverse of one or_coupling_S_f64 step (reverse order, negated
 * angles): used to pin the adjoint kernel. */
void or_coupling_S_inv_f64(int N, const float *H, double *cr, double *ci,
                           int m, int sign, float deltat, float couplingcut);

void or_coupling_S_doubles_f32(int N, const float *H, float *cr, float *ci,
                               int m, const float *Anh, float anharmonicity,
                               float deltat, float couplingcut);
void or_coupling_S_doubles_ES_f32(int N, const float *H, float *cr, float *ci,
                                  int m, float deltat, float couplingcut);
void or_coupling_S_doubles_f64(int N, const float *H, double *cr, double *ci,
                               int m, const float *Anh, float anharmonicity,
                               float deltat, float couplingcut);

int or_jacobi_f64(int N, double *A, double *w, double *V);
void or_DIA_f64(int N, const float *H, double *cr, double *ci, int sign,
                float deltat);
/* the sparse scheme: U as in or_DIA_f64, elements with |U_ab|^2 <= thres
   (tested on the FP32-rounded element) dropped; returns the kept count */
int or_DIA_S_f64(int N, const float *H, double *cr, double *ci, int sign,
                 float deltat, float thres);

void or_dipole_double(int N, const float *dipole, const float *cr,
                      const float *ci, float *fr, float *fi, float *over,
                      float anharmonicity);
void or_dipole_double_ES(int N, const float *dipole, const float *cr,
                         const float *ci, float *fr, float *fi);
void or_dipole_double_last_ES(int N, const float *dipole, const float *cr,
                              const float *ci, float *fr, float *fi);
void or_dipole_double_last(int N, const float *dipole, const float *cr,
                           const float *ci, float *fr, float *fi, float *over,
                           float anharmonicity);
float or_polarweight(int pol, int x);
void or_polar(int *xp, int x);

/* One Hamiltonian frame as read_Hamiltonian()/read_He() leave it. */
/* geometry for TransitionDipole / ExtendedDipole frames read by or_read_H
   and or_absorption: position records (two per frame for ExtendedDipole),
   transition dipole records, optional box records (int + 3 floats) */
void or_set_geometry(const float *pos_rec, long long pos_stride, const float *dip_rec,
                     long long dip_stride, const float *box_rec);
void or_read_H(const t_non *non, const float *H_rec, long long H_stride,
               const float *C_rec, long long pos, float *He, int apply_shift);

int or_absorption(const t_non *non, const float *H_rec, long long H_stride,
                  const float *mu_rec, long long mu_stride,
                  const float *C_rec, float *re_S, float *im_S,
                  double *re_S64, double *im_S64, int nthreads);

int or_1dfft(const t_non *non, const float *re_S, const float *im_S,
             int samples, double *freq, double *spec_r, double *spec_i,
             int *nout);

int or_absorption_param_f64(const t_non *non, const float *H_rec,
                            long long H_stride, const float *mu_rec,
                            long long mu_stride, const float *C_rec,
                            const double *delta, double cscale,
                            double *re_S, double *im_S, int nthreads);

/* src/propagate.c:1241-1302: kept elements of the one-exciton propagator
   (Ur, Ui, R = row, C = column, a-major order); returns the count */
int or_time_evolution_mat(int N, const float *H, float deltat, int m, float thres,
                          float *Ur, float *Ui, int *R, int *C);
/* src/propagate.c:1304-1473: the sparse two-exciton step (es = 1: two-level
   sites) */
void or_double_sparce_f32(int N, const float *Ur, const float *Ui, const int *R,
                          const int *C, float *fr, float *fi, int elements, int m,
                          const float *Anh, float anharmonicity, float deltat,
                          int es);

int or_2dir_item(const t_non *non, const float *H_rec, long long H_stride,
                 const float *mu_rec, long long mu_stride, const float *C_rec,
                 const float *A_rec, const float *over_rec, int sample,
                 int molPol, float *R12);

/* src/NISE_subs.c:512-530 and src/polar.c:107-116 (as a return value) */
void or_dipole_double_ground(int N, const float *dipole, float *fr, float *fi, float *over,
                             float anharmonicity);
int or_raman_pol(int p2, int p3);
/* One 2DIR-Raman work item, src/calc_2DIRraman.c:176-441 (propagation 1,
   fixed anharmonicity); see nise_oracle.c for the two changes */
int or_2dirraman_item(const t_non *non, const float *H_rec, long long H_stride,
                      const float *mu_rec, long long mu_stride, const float *C_rec,
                      const float *alpha_rec, int sample, int molPol, float *R12);

#ifdef __cplusplus
}
#endif
#endif
