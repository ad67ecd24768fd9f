/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Program entry for the reference's 2D response drivers, built by
 * oracle/Makefile (target ref2d) together with the reference's own,
 * unmodified sources (readinput.c, calc_2DIR.c, calc_2DES.c, MPI_subs.c,
 * project.c, propagate.c, NISE_subs.c, polar.c, read_trajectory.c,
 * hamshift.c, randomlib.c, types_MPI.c) against MPICH and MKL's LAPACK.
 * It repeats the set-up of the reference's main (src/NISE.c:49-125: MPI
 * with funnelled threads, the shared-memory sub-communicator and the
 * communicator of its roots, readInput, control, autodetect_singles, the
 * log file) and dispatches the 2DIR and 2DUVvis techniques the way
 * src/NISE.c:278-300 does.  It is run as an MPI singleton in a directory
 * holding an input deck and its trajectory files, and writes the R*.dat
 * files of print2D there (tests/golden/make_golden_2d.py).
 */
#include <stdio.h>
#include <stdlib.h>
#include "mpi.h"
#include "types.h"
#include "NISE_subs.h"
#include "readinput.h"
#include "calc_2DIR.h"
#include "calc_2DES.h"

int main(int argc, char *argv[]) {
  int parentRank, subRank, parentSize, subSize, provided = 0;
  MPI_Comm subComm, rootComm;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
  if (provided < MPI_THREAD_FUNNELED) {
    printf("Error: MPI library does not support threading!\n");
    MPI_Abort(MPI_COMM_WORLD, 1);
  }
  MPI_Comm_rank(MPI_COMM_WORLD, &parentRank);
  MPI_Comm_size(MPI_COMM_WORLD, &parentSize);
  MPI_Comm_split_type(MPI_COMM_WORLD, MPI_COMM_TYPE_SHARED, 0, MPI_INFO_NULL, &subComm);
  MPI_Comm_rank(subComm, &subRank);
  MPI_Comm_size(subComm, &subSize);
  MPI_Comm_split(MPI_COMM_WORLD, subRank == 0 ? 0 : MPI_UNDEFINED, 0, &rootComm);
  if (parentSize != 1) {
    printf("run this driver as an MPI singleton\n");
    MPI_Abort(MPI_COMM_WORLD, 1);
  }
  t_non *non = calloc(1, sizeof(t_non));
  readInput(argc, argv, non);
  int initResult = control(non); /* overwritten by the next call, as in src/NISE.c:104-105 */
  initResult = autodetect_singles(non);
  if (initResult != 0) {
    MPI_Finalize();
    return 1;
  }
  FILE *logFile = fopen("NISE.log", "w");
  if (logFile == NULL) {
    MPI_Finalize();
    return 1;
  }
  fprintf(logFile, "Log\n");
  fclose(logFile);
  int rc = 2;
  if (string_in_array(non->technique, (char *[]){"2DIR", "GBIR", "SEIR", "EAIR", "noEAIR"}, 5)) {
    calc_2DIR(non, parentRank, parentSize, subRank, subSize, subComm, rootComm);
    rc = 0;
  }
  if (string_in_array(non->technique, (char *[]){"2DUVvis", "GBUVvis", "SEUVvis", "EAUVvis", "noEAUVvis"}, 5)) {
    calc_2DES(non, parentRank, parentSize, subRank, subSize, subComm, rootComm);
    rc = 0;
  }
  free(non->psites);
  free(non);
  MPI_Finalize();
  return rc;
}
