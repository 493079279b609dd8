/*
 * solver_main.c -- the `solver` program on the GPU potential solve.
 *
 * Same sequence as the reference tool (solver.c:18-130): options, the density
 * from --input (LoadVlasovQuantities, fileio.c:6-70: "density-kji", fluxes
 * made from it with the flux scales) or, with no input, zero moments and the
 * built-in sinusoidal forcing (fileio.c:21-26), one potential solve with the
 * KSP options given, then the fields to <output>.hdf (OutputHDF5 with an empty
 * insert, solver.c:70) and the elapsed-time report.  The LHS eigenvalue
 * report (--lhs-eigenvalues, SLEPc) is not provided.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "chybrid.h"
#include "chy_output.h"

#define CHECK(call)                                                                          \
  do {                                                                                       \
    int rc_ = (call);                                                                        \
    if (rc_) {                                                                               \
      fprintf(stderr, "[0] %s failed with error %d: %s\n", #call, rc_, chy_last_error(ctx)); \
      free(vl);                                                                              \
      chy_destroy(ctx);                                                                      \
      return -rc_;                                                                           \
    }                                                                                        \
  } while (0)

int main(int argc, char **argv)
{
  chy_ctx *ctx = NULL;
  double *vl = NULL;
  time_t start = time(NULL);
  printf("\n**************** START *****************\n\n");
  chy_params prm;
  char msg[8192] = "";
  int rc = chy_params_from_options(argc - 1, (const char *const *)(argv + 1), &prm, msg, sizeof msg);
  if (msg[0]) printf("%s", msg);
  if (rc) {
    fprintf(stderr, "[0] option error %d\n", rc);
    return -rc;
  }
  const int64_t nc = prm.N[0] * prm.N[1] * prm.N[2];
  vl = calloc((size_t)(4 * nc), sizeof(double));
  if (!vl) {
    fprintf(stderr, "[0] out of host memory for %lld nodes\n", (long long)nc);
    return -CHY_ERR_MEM;
  }
  if (prm.input[0] == 0) {
    printf("Warning: Got empty path to density file; using built-in sinusoidal form of RHS.\n");
    prm.rhs_type = CHY_RHS_SINUSOIDAL;
  } else {
    printf("Attempting to load density from %s\n", prm.input);
    rc = chy_read_vlasov(&prm, prm.input, vl, msg, sizeof msg);
    if (rc) {
      fprintf(stderr, "[0] chy_read_vlasov failed with error %d: %s\n", rc, msg);
      free(vl);
      return -rc;
    }
    printf("Loaded density\n");
    printf("Created x flux from density\nCreated y flux from density\nCreated z flux from density\n");
  }
  /* the solver holds no ions: capacity for one */
  rc = chy_create(&prm, 1, &ctx);
  if (rc) {
    fprintf(stderr, "[0] chy_create failed with error %d (is a GPU visible?)\n", rc);
    free(vl);
    return -rc;
  }
  printf("Grid %lld x %lld x %lld, precision %s\n", (long long)prm.N[0], (long long)prm.N[1], (long long)prm.N[2],
         prm.precision == CHY_PREC_MIXED ? "mixed" : "fp64");
  CHECK(chy_set_moments(ctx, vl));
  chy_solve_stats ss;
  CHECK(chy_solve_potential(ctx, NULL, &ss));
  printf("Potential: %lld iterations, residual %.3e, %s, %.3f ms\n", (long long)ss.iterations, ss.residual,
         ss.converged ? "converged" : "not converged", ss.ms);
  rc = chy_output_write(ctx, prm.output, "");
  if (rc == CHY_ERR_SUP) {
    printf("No HDF5 library: fields written as %s.<field>.f64 raw files\n", prm.output);
  } else {
    CHECK(rc);
  }
  free(vl);
  chy_destroy(ctx);
  time_t end = time(NULL);
  printf("\n----------------------------------------\n");
  printf("Start time: %s", asctime(localtime(&start)));
  printf("End time:   %s", asctime(localtime(&end)));
  printf("Elapsed time: %f s\n", (float)(end - start));
  printf("----------------------------------------\n");
  printf("\n***************** END ******************\n");
  return 0;
}
