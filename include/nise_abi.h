/*
 * nise_abi.h -- C ABI of libnise_hip.so, the MI355X (gfx950) implementation of
 * NISE's wavefunction-propagation hot path.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in symbols with the exact C signatures of the reference
 *     (GHlacour/NISE_2017).  An unchanged NISE host program links these in
 *     place of the reference's definitions.  They keep the reference's
 *     conventions: caller-owned host arrays, split real/imaginary float
 *     vectors updated in place, the packed upper-triangle Hamiltonian with the
 *     RWA-shifted diagonal, void return, print-and-exit on fatal errors.
 *     Every one of them computes on the GPU; none has a CPU fallback.
 *
 *  2. New batched entry points (prefix nise_) that keep trajectories resident
 *     in HBM and run the frame-major kernels.  They return an int status
 *     (NISE_OK == 0) and never exit the process.
 *
 * If the reference's types.h is included before this header, its t_non is
 * used (the definition below is byte-identical to it, src/types.h:37-90).
 */
#ifndef NISE_ABI_H
#define NISE_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* t_non: run parameters.  Field order and sizes follow src/types.h:37-90 so  */
/* that a t_non built by the reference's readInput() can be passed as is.    */
/* ------------------------------------------------------------------------ */
#ifndef _TYPES_
typedef struct {
  int tmax1, tmax2, tmax3;
  int tmin1, tmin2, tmin3;
  int dt1, dt2, dt3;
  int d1, d2, d3;
  float min1, max1;
  float min2, max2;
  float min3, max3;
  int k[4];
  int length;
  int sample;
  int fft;
  int begin;
  int end;
  int is;
  int ts;
  int interpol;
  int propagation;
  float lifetime;
  float homogen;
  float inhomogen;
  float deltat;
  int buffer, singles, doubles;
  char energyFName[256];
  char dipoleFName[256];
  char alphaFName[256];
  char positionFName[256];
  char anharFName[256];
  char overdipFName[256];
  char singleShiftFName[256];
  char technique[256];
  char basis[256];
  char pdbFName[256];
  char hamiltonian[256];
  char couplingFName[256];
  char pbcFName[256];
  int tmax;
  int cluster;
  float shifte, shiftf;
  int labPol;
  float dephasing;
  float statstart, statend, statstep;
  int statsteps;
  float thres;
  float couplingcut;
  float temperature;
  float anharmonicity;
  int Npsites;
  int *psites;
  int window;
  int SingleShiftSites, *SingleShiftSite;
  float *SingleShift;
  int printLevel;
} t_non;
#endif

/* MPI communicator handles appear only in calc_2DIR's signature and are
 * never dereferenced here (one GPU, rank 0 does all work).  Without mpi.h
 * they are declared with MPICH's integer handle type; both an int and a
 * pointer handle occupy one argument slot in the SysV x86-64 ABI. */
#if defined(NISE_ABI_WITH_MPI)
#include <mpi.h>
#elif !defined(MPI_VERSION)
typedef int MPI_Comm;
#endif

/* ------------------------------------------------------------------------ */
/* 1. Drop-in symbols (reference signatures).                                */
/* ------------------------------------------------------------------------ */

/* Replaces src/propagate.c:20-39 (prototype src/propagate.h:4).  Dispatch on
 * non->propagation: 1 -> Trotter coupling scheme; 0 -> exact exponential
 * (DIA) when thres == 0 or thres > 1, else the sparse scheme (DIA_S: U with
 * elements |U_ab|^2 <= thres dropped, with the reference's sparsity report).
 * Only RK4 (propagation 3) is not built: it prints a message and exits, as
 * the reference does for an unknown scheme. */
void propagate_vector(t_non *non, float *Hamil_i_e, float *vecr, float *veci,
                      int sign, int samples, int display);

/* Replaces src/propagate.c:43-52 (prototype src/propagate.h:5): propagates
 * the non->singles column vectors of an N x N block in one launch. */
void propagate_matrix(t_non *non, float *Hamil_i_e, float *vecr, float *veci,
                      int sign, int samples, int display);

/* Replaces src/propagate.c:821-899 (prototype src/propagate.h:13). */
void propagate_vec_coupling_S(t_non *non, float *Hamiltonian_i, float *cr,
                              float *ci, int m, int sign);

/* Replaces src/propagate.c:902-1048 (prototype src/propagate.h:14-15). */
void propagate_vec_coupling_S_doubles(t_non *non, float *Hamiltonian_i,
                                      float *cr, float *ci, int m, float *Anh);

/* Replaces src/propagate.c:1051-1181 (prototype src/propagate.h:16), the
 * two-exciton step of two-level sites used by calc_2DES: no anharmonicity,
 * no rotations for c == a or c == b. */
void propagate_vec_coupling_S_doubles_ES(t_non *non, float *Hamiltonian_i,
                                         float *cr, float *ci, int m);

/* Replaces src/propagate.c:60-125 (prototype src/propagate.h:7). */
void propagate_vec_DIA(t_non *non, float *Hamiltonian_i, float *cr, float *ci,
                       int sign);

/* Replaces src/propagate.c:1185-1221 (prototype src/propagate.h:17):
 * H is a dense row-major symmetric N x N matrix.  On return
 * H[i*N + j] is component i of eigenvector j (the layout the reference
 * leaves after its transpose copy) and v holds the ascending eigenvalues.
 * Computed with the batched Jacobi kernel. */
void diagonalizeLPD(float *H, float *v, int N);

/* Replaces src/propagate.c:1476-1591 (prototype src/propagate.h:22-25) for
 * non->propagation == 1: doubles propagation of fr and ft1r[0..tmax1) and
 * sign -1 singles propagation of rightnr and rightrr[0..tmax1), one launch
 * each. */
void propagate_double_2DIR_control(t_non *non, float *Hamil_i_e, float **ft1r,
                                   float **ft1i, float *fr, float *fi,
                                   float **rightrr, float **rightri,
                                   float *rightnr, float *rightni,
                                   int currentSample, int molPol, int t3,
                                   float *Anh);

/* Replaces src/propagate.c:1594-1709 (prototype src/propagate.h:26-28) for
 * non->propagation == 1: as propagate_double_2DIR_control with the
 * two-level doubles step. */
void propagate_double_2DES_control(t_non *non, float *Hamil_i_e, float **ft1r,
                                   float **ft1i, float *fr, float *fi,
                                   float **rightrr, float **rightri,
                                   float *rightnr, float *rightni,
                                   int currentSample, int molPol, int t3);

/* Replaces src/absorption.c:205-212 (prototype src/absorption.h:4). */
void calc_S1(float *re_S_1, float *im_S_1, int t1, t_non *non, float *cr,
             float *ci, float *mu);

/* Replaces src/absorption.c:18-202 (prototype src/absorption.h:3): reads the
 * trajectory named in non, runs the batched frame-major kernel, writes
 * TD_Absorption.dat and Absorption.dat. */
void absorption(t_non *non);

/* Replaces src/1DFFT.c:144-274 (prototype src/1DFFT.h:5). */
void do_1DFFT(t_non *non, char fname[], float *re_S_1, float *im_S_1,
              int samples);

/* Replaces src/calc_2DIR.c:22-629 (prototype src/calc_2DIR.h:5).  Rank 0
 * computes every (sample, molPol) work item on its GPU and writes the six
 * R{par,per,cro}{I,II}.dat files; other ranks return at once. */
void calc_2DIR(t_non *non, int parentRank, int parentSize, int subRank,
               int subSize, MPI_Comm subComm, MPI_Comm rootComm);

/* Replaces src/calc_2DES.c:22-586 (prototype src/calc_2DES.h:5), the 2DUVvis
 * techniques (2DUVvis, GBUVvis, SEUVvis, EAUVvis, noEAUVvis): calc_2DIR for
 * two-level sites, with the same files and messages. */
void calc_2DES(t_non *non, int parentRank, int parentSize, int subRank,
               int subSize, MPI_Comm subComm, MPI_Comm rootComm);

/* Replaces src/calc_2DIRraman.c:22-663 (prototype src/calc_2DIRraman.h:5),
 * the techniques 2DIRraman, 2DIRraman1, 2DIRraman2, 2DIRraman3, 2DIRramanI
 * and 2DIRramanII: reads the Hamiltonian, dipole and Raman tensor (Alphafile)
 * trajectories and writes R{par,per,cro}{I,II}.IRraman.dat.  The Raman
 * component of each item is find_raman_pol's return value (the reference
 * discards it, src/calc_2DIRraman.c:319), lifetime 0 means no decay; the
 * Trotter scheme with a fixed anharmonicity only (the reference exits for a
 * per-frame one).  Rank 0 runs every work item; other ranks return. */
void calc_2DIRraman(t_non *non, int parentRank, int parentSize, int subRank,
                    int subSize, MPI_Comm subComm, MPI_Comm rootComm);

/* ------------------------------------------------------------------------ */
/* 2. Batched entry points (no counterpart in the reference).                */
/* ------------------------------------------------------------------------ */

enum {
  NISE_OK = 0,
  NISE_ERR_ARG = 1,       /* invalid argument or unsupported setting     */
  NISE_ERR_HIP = 2,       /* HIP runtime error                           */
  NISE_ERR_NOMEM = 3,     /* device allocation failed                    */
  NISE_ERR_RANGE = 4,     /* trajectory too short for the requested run  */
  NISE_ERR_NUMERIC = 5    /* non-finite input or eigensolver failure     */
};

/* Trajectory record layout (GROBIN, the reference's binary format): every
 * frame is one int32 header followed by float32 payload.  A pointer
 * argument named *_rec points at the header of frame 0; *_stride is the
 * record length in 4-byte words.
 *   Full Hamiltonian:   stride >= 1 + N(N+1)/2, packed upper triangle
 *   Coupling-type H:    stride >= 1 + N, diagonal only
 *   Dipole:             stride >= 1 + 3N, x block, y block, z block      */

/* Where the pointer arguments of a batched call live. */
enum { NISE_HOST = 0, NISE_DEVICE = 1 };

/* Batched linear absorption (src/absorption.c:92-146 over all samples in
 * [non->begin, non->end) and the three polarisations).  Uses non->singles,
 * tmax1, sample, begin, end, ts, deltat, couplingcut, shifte, hamiltonian
 * ("Full" or "Coupling"), SingleShift*.  For "Coupling" the Hamiltonian
 * records carry the diagonal and coupling_rec is one Full record with the
 * static couplings; the dipole of frame 0 is used at all times.
 * re_S/im_S (host, length tmax1) receive the sums over samples and
 * polarisations, un-normalised, exactly what the reference accumulates.
 * kernel: 0 = frame-major batched kernel (default), 1 = plain
 * one-sample-per-launch kernel (baseline).  propagation 0 (exact
 * exponential, N <= 64) runs the DIA conveyor of nise_absorption_multi. */
int nise_absorption_batched(const t_non *non, const float *H_rec,
                            int64_t H_stride, const float *mu_rec,
                            int64_t mu_stride, int64_t nframes,
                            const float *coupling_rec, int where,
                            int kernel, float *re_S, float *im_S);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): what nise_absorption_batched computes for device
 * records (where = NISE_DEVICE, kernel 0) with propagation 1, bit for bit,
 * using only a library-owned 32 MiB device array per GPU (held until the
 * process ends) for its mask, schedule, per-frame tables and partial sums.
 * Runs needing more return NISE_ERR_RANGE with a message.  No hipMalloc /
 * hipFree, no stream or event is created (null stream, synchronous copies;
 * re_S / im_S are host arrays).  Calls on one GPU share the array: they are
 * serialised inside this process, and overlapping use from elsewhere races. */
int nise_absorption_batched_noalloc(const t_non *non, const float *H_rec,
                                    int64_t H_stride, const float *mu_rec,
                                    int64_t mu_stride, int64_t nframes,
                                    const float *coupling_rec, float *re_S,
                                    float *im_S);

/* Absorption plan: the state of nise_absorption_batched kept on the device
 * so that a run can be repeated (timed, replayed as a hipGraph, re-shaped)
 * and differentiated.  Arguments as nise_absorption_batched; host records
 * are copied at creation, device records must outlive the plan. */
typedef struct nise_abs_plan nise_abs_plan;
int nise_abs_plan_create(nise_abs_plan **plan, const t_non *non,
                         const float *H_rec, int64_t H_stride,
                         const float *mu_rec, int64_t mu_stride,
                         int64_t nframes, const float *coupling_rec, int where,
                         int kernel);
/* As nise_abs_plan_create for HamiltonianType TransitionDipole or
 * ExtendedDipole: H_rec holds the N site energies per frame, the couplings
 * are built per frame on the device from positions (pos_rec: read_mue
 * records, int + 3N per frame, two per frame for ExtendedDipole) and the
 * transition dipoles mu_rec, with the minimum image in the box of box_rec
 * (int + 3 floats per frame) when it is given.  Replaces the couplings that
 * read_He computes on the fly, src/read_trajectory.c:181-303. */
int nise_abs_plan_create_onfly(nise_abs_plan **plan, const t_non *non,
                               const float *H_rec, int64_t H_stride,
                               const float *mu_rec, int64_t mu_stride,
                               int64_t nframes, const float *pos_rec,
                               int64_t pos_stride, const float *box_rec,
                               int where, int kernel);
/* Conveyor shape: vp vector slots per workgroup, samples_per_wg samples
 * (3 polarisations each, 3*samples_per_wg <= vp), block 256 or 512
 * (multi-wave), 64 (one wave) or -2 / -4 (one wave, 2 or 4 slots per lane). */
int nise_abs_plan_set_shape(nise_abs_plan *plan, int vp, int samples_per_wg,
                            int block);
/* Row conveyor (the default): one wave per workgroup holding 64/rows
 * vectors, rows rotations of a layer at a time (1, 2, 4, 8, 16, 32 or 64; 0
 * picks the count from the coupling structure). */
int nise_abs_plan_set_rows(nise_abs_plan *plan, int rows);
/* Autotuner: times each candidate configuration (row conveyor with 1..64
 * rows, slot-conveyor shapes) on the plan's own workload and keeps the
 * fastest.  ms / codes (host, n entries each, optional) receive each timed
 * candidate's conveyor time and code (rows R as R; slot shapes as
 * -(10000 vp + 1000 spw + block), block -2/-4 as 902/904).  Returns the
 * number of candidates timed, or minus an error code. */
int nise_abs_plan_autotune(nise_abs_plan *plan, float *ms, int *codes, int n);
/* Run the plan; re_S / im_S (host, tmax1 each) may be NULL to leave the
 * result on the device.  flags bit 0: time the stages with HIP events;
 * bit 1: launch the stages as one captured hipGraph. */
int nise_abs_plan_execute(nise_abs_plan *plan, float *re_S, float *im_S,
                          int flags);
int nise_abs_plan_sync(nise_abs_plan *plan);
/* info[0..11] = N, schedule entries per sweep, layers, workgroups, vp,
 * samples per workgroup, first frame, frames, block, steps, rows (0: slot
 * conveyor), rotations per sweep; returns the count. */
int nise_abs_plan_info(const nise_abs_plan *plan, int64_t *info, int n);
/* ms of the last timed execute (tables, conveyor or plain kernel, reduce)
 * and of the last adjoint run; returns the count (4). */
int nise_abs_plan_last_timing(const nise_abs_plan *plan, float *ms, int n);
void nise_abs_plan_destroy(nise_abs_plan *plan);

/* Parameters of gradient runs (no counterpart in the reference): delta
 * (host, N values or NULL for none) is added to every frame's shifted site
 * energy, cscale multiplies every kept coupling.  delta = NULL and
 * cscale = 1 give the reference's result bit for bit. */
int nise_abs_plan_set_params(nise_abs_plan *plan, const float *delta,
                             float cscale);
/* Record forward checkpoints psi(t) at t = 0, every, 2*every, ... and
 * tmax1-1 for every (sample, polarisation) vector on each execute. */
int nise_abs_plan_enable_adjoint(nise_abs_plan *plan, int every);
/* Adjoint of the last execute: given g_re[t] = dL/dRe S(t) and
 * g_im[t] = dL/dIm S(t) (host, tmax1 each), writes dL/d delta (host, N
 * doubles) and dL/d cscale.  Propagates the adjoint backwards with the
 * inverse Trotter steps, rebuilding the forward states from the
 * checkpoints. */
int nise_abs_plan_adjoint(nise_abs_plan *plan, const float *g_re,
                          const float *g_im, double *grad_delta,
                          double *grad_c);

/* One frame's propagation tables as the plan computes them (inspection):
 * diag_out = N (cos, -sin) half-step phases, cs_out = (cos, sin) of every
 * coupling of the strict upper triangle in lexicographic order, *K_out =
 * their count. */
int nise_frame_tables(const t_non *non, const float *H_rec, int64_t H_stride,
                      const float *coupling_rec, int64_t pos, int sign,
                      float *diag_out, float *cs_out, int *K_out);

/* The row schedule the row conveyor uses (inspection, host only): for the
 * kept-coupling mask over the strict upper triangle in lexicographic order
 * (N(N-1)/2 bytes), ab_out receives *nlayers layers of R entries
 * (a | b << 16); entry q of a padded layer is (N+q, N+q).  Returns the
 * entry count, or -1 if cap is too small. */
int nise_row_schedule(int N, const uint8_t *mask, int R, uint32_t *ab_out,
                      int cap, int *nlayers);

/* Batched 2DIR (src/calc_2DIR.c:184-506 over all samples in
 * [non->begin, non->end) and the 21 molecular polarisation items).  Uses
 * non->singles, tmax1, tmax3 (tmax2 must be 0), sample, begin, end (0: as
 * many samples as the trajectory holds), ts, deltat, couplingcut, min1 and
 * max1 (shifte = (min1+max1)/2), lifetime, anharmonicity, technique
 * ("2DIR", "GBIR", "SEIR", "EAIR", "noEAIR"; or the two-level-site forms of
 * calc_2DES: "2DUVvis", "GBUVvis", "SEUVvis", "EAUVvis", "noEAUVvis", which
 * take no anharmonicity), hamiltonian ("Full" or
 * "Coupling"); propagation must be 1.  anh_rec (1 + N words per frame) and
 * over_rec (1 + 3N words per frame) are read when anharmonicity == 0 and
 * the technique has the EA pathway.  R12 (host, 12 x tmax3 x tmax1, t1
 * fastest) receives the sums over samples and items of rrIpar, riIpar,
 * rrIIpar, riIIpar, then the per and cro planes in the same order,
 * un-normalised: what the reference reduces before print2D. */
int nise_2dir_batched(const t_non *non, const float *H_rec, int64_t H_stride,
                      const float *mu_rec, int64_t mu_stride, int64_t nframes,
                      const float *coupling_rec, const float *anh_rec,
                      const float *over_rec, int where, float *R12);

/* Absorption streamed from the trajectory files named in non (Full or
 * Coupling type, propagation 1 or 0), for frame ranges beyond device or host
 * memory: the samples begin..end-1 (end resolved by the caller) run in
 * windows of at most window_samples start points (0: sized to a quarter of
 * the free device memory); a reader thread reads the next window's records
 * into pinned host memory and uploads them on a copy stream while the
 * current window runs as nise_absorption_batched; the windows' responses
 * are summed in FP64.  No counterpart in the reference, which reads one
 * frame per step (src/read_trajectory.c:138-173); absorption() uses it when
 * the frame range exceeds a third of the free device memory.  re_S/im_S:
 * tmax1 floats (host), un-normalised as src/absorption.c:112-146 leaves
 * them. */
int nise_absorption_stream(const t_non *non, int window_samples, float *re_S, float *im_S);

/* 2DIR-Raman in one call, the computation of calc_2DIRraman
 * (src/calc_2DIRraman.c:176-441 per work item) for every sample and item:
 * alpha_rec holds the Raman tensor records (1 + 6N words per frame:
 * header, then the xx, xy, xz, yy, yz, zz components), the other records
 * as for nise_2dir_batched.  R12 receives the sums over samples and items
 * of rrIpar, riIpar, rrIIpar, riIIpar, then the per and cro planes,
 * un-normalised.  The technique selects the variant (2DIRraman,
 * 2DIRraman1, 2DIRraman2, 2DIRraman3, 2DIRramanI, 2DIRramanII); the
 * Raman component of an item is find_raman_pol's return value; propagation
 * 1 with a fixed anharmonicity only (NISE_ERR_ARG otherwise). */
int nise_2dirraman_batched(const t_non *non, const float *H_rec, int64_t H_stride,
                           const float *mu_rec, int64_t mu_stride, int64_t nframes,
                           const float *coupling_rec, const float *alpha_rec, int where,
                           float *R12);

/* One system of a multi-system batch. */
typedef struct {
  const t_non *non;          /* per-system run parameters */
  const float *H_rec;        /* Hamiltonian records (host or device) */
  int64_t H_stride;
  const float *mu_rec;       /* dipole records */
  int64_t mu_stride;
  int64_t nframes;
  const float *coupling_rec; /* "Coupling" type: one Full record, else NULL */
} nise_system;

/* Absorption of nsys independent systems of uneven size (no counterpart in
 * the reference; each system is what nise_absorption_batched computes).
 * Systems with propagation 1 share one row-conveyor launch; systems with
 * propagation 0 (exact exponential) and N <= 64 share one DIA-conveyor
 * launch fed by per-frame propagators from the batched Jacobi solver.
 * re_S / im_S (host) receive each system's tmax1 values, systems one after
 * another.  done_ms (host, nsys, optional) receives each system's
 * completion time in ms after the first workgroup of the run started
 * (device wall clock). */
int nise_absorption_multi(const nise_system *systems, int nsys, int where,
                          float *re_S, float *im_S, double *done_ms);

/* Autotuner of the multi-system batch: runs the batch once, then times it
 * with each system's own row count (the power of two nearest its mean layer
 * width, up to 32: code 0; up to 16 or 64: codes -16, -64) and with every
 * common row count of the row conveyor (1, 2, 4, ..., 64; best of two runs,
 * counts the batch cannot take skipped) and keeps the fastest for later
 * nise_absorption_multi calls of this thread on a batch of the same shape
 * (per system N, tmax, samples, spacing, scheme, Trotter steps); without it
 * each system takes its own count, up to 32.  R_best receives the code kept; ms / codes (n entries, optional)
 * each candidate's host wall time per call and its code.  Returns the number
 * of candidates timed, or minus an error code.  No counterpart in the
 * reference (its static work split, src/MPI_subs.c:36-110). */
int nise_absorption_multi_tune(const nise_system *systems, int nsys, int where, int *R_best,
                               float *ms, int *codes, int n);

/* Replaces the 2DFFT post-processing program (src/2DFFT.c:20-830, a
 * separate executable in the reference, run on the NISE input file) in
 * process.  Reads the response files calc_2DIR writes in the current
 * directory (R{par,per,cro}{I,II}.dat; R*_<t>fs.dat for each waiting time
 * listed in Waitingtime.dat; R*.IRraman.dat for the 2DIRraman techniques),
 * applies the apodisation (homogen, inhomogen, window), the edge halving and
 * the zero padding to fft x fft, transforms every response on the GPU
 * (forward 2D DFT in double precision) and writes the same files as the
 * program: Rw*.I.dat, Rw*.II.dat, 2D.*.dat and PP.*.dat.  The program's
 * input keywords are the t_non fields fft, deltat, homogen, inhomogen,
 * window, tmax1, tmax3, min1..max3 and technique; format is its Format
 * keyword: 0 Dislin / Python, 1 Matlab (kI / kII files and waxis.dat only),
 * 2 Gnuplot.  Returns NISE_OK, or NISE_ERR_ARG where the program prints a
 * message and exits (negative or inverted frequency ranges, fft == 0). */
int nise_2dfft(const t_non *non, int format);

/* The forward 2D transform of nise_2dfft on its own (no counterpart in the
 * reference): P grids of fft x fft complex doubles (interleaved re, im; row
 * major; host memory), out[p][k0][k1] = sum in[p][j0][j1]
 * exp(-2 pi i (j0 k0 + j1 k1) / fft), the sign and layout of FFTW_FORWARD.
 * Only rows < nrows and columns < ncols of the input may be nonzero. */
int nise_dft2d(const double *in, int P, int fft, int nrows, int ncols,
               double *out);

/* Last error message of the calling thread (static storage). */
const char *nise_last_error(void);

/* Library version string. */
const char *nise_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NISE_ABI_H */
