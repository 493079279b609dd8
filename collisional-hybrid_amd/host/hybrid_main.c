/*
 * hybrid_main.c -- the `hybrid` driver on the GPU hot path.
 *
 * Same loop as the reference driver (hybrid.c:18-144): options, set-up,
 * initial ions, options.txt, initial moments and potential, then per time
 * step velocities -> positions -> moments -> potential -> output, and the
 * elapsed-time report.  Flags follow parameters.c; output files are
 * <output>-initial and <output>-<it> (hybrid.c:84-121), written every
 * -output_interval steps (0 = none).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "chybrid.h"
#include "chy_output.h"

#define CHECK(call)                                                                   \
  do {                                                                                \
    int rc_ = (call);                                                                 \
    if (rc_) {                                                                        \
      fprintf(stderr, "[0] %s failed with error %d: %s\n", #call, rc_, chy_last_error(ctx)); \
      chy_destroy(ctx);                                                               \
      return -rc_;                                                                    \
    }                                                                                 \
  } while (0)

/* An output call's result: CHY_ERR_SUP means the fields went to raw .f64
 * files because no HDF5 library could be loaded, which is not a failure of the
 * run (the solver program treats it the same way).  Reported once. */
static int output_rc(int rc, const char *stem)
{
  static int told = 0;
  if (rc != CHY_ERR_SUP) return rc;
  if (!told) printf("No HDF5 library: fields written as %s<insert>.<field>.f64 raw files\n", stem);
  told = 1;
  return CHY_OK;
}

static int rank_from_env(void)
{
  const char *v = getenv("PMI_RANK");
  if (!v) v = getenv("OMPI_COMM_WORLD_RANK");
  if (!v) v = getenv("MPI_LOCALRANKID");
  return v ? atoi(v) : 0;
}

int main(int argc, char **argv)
{
  chy_ctx *ctx = NULL;
  const int rank = rank_from_env();
  if (rank != 0) {
    /* one GPU runs the whole problem; further MPI ranks have nothing to do */
    printf("[%d] hybrid: the GPU hot path runs on rank 0; rank %d exits\n", rank, rank);
    return 0;
  }
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
  if (prm.dt == 1.0 / prm.nui) printf("Warning: Setting dt = 1 / nui\n");
  rc = chy_create(&prm, 0, &ctx);
  if (rc) {
    fprintf(stderr, "[0] chy_create failed with error %d (is a GPU visible?)\n", rc);
    return -rc;
  }
  chy_get_params(ctx, &prm);
  printf("Grid %lld x %lld x %lld, Np = %lld, precision %s\n", (long long)prm.N[0], (long long)prm.N[1],
         (long long)prm.N[2], (long long)prm.Np, prm.precision == CHY_PREC_MIXED ? "mixed" : "fp64");
  CHECK(chy_init_particles(ctx));
  CHECK(chy_write_options_txt(&prm, argc - 1, (const char *const *)(argv + 1), "options.txt"));
  CHECK(chy_collect_moments(ctx, NULL));
  chy_solve_stats ss;
  CHECK(chy_solve_potential(ctx, NULL, &ss));
  int width = 1 + (int)log10((double)(prm.Nt > 0 ? prm.Nt : 1));
  if (width > 19) width = 19;
  /* output runs behind the following steps (snapshot on the device, copy and
   * file write overlapped); each call first finishes the previous one */
  if (prm.output_interval > 0) CHECK(output_rc(chy_output_write_async(ctx, prm.output, "-initial"), prm.output));
  printf("\n*** Main time-step loop ***\n\n");
  double total_ms = 0.0, solve_ms = 0.0, part_ms = 0.0;
  for (int64_t it = 0; it < prm.Nt; it++) {
    printf("Time step %0*lld\n", width, (long long)it);
    chy_step_stats st;
    CHECK(chy_step(ctx, &st));
    if (prm.mn > 0.0)
      printf("Colliding %lld particles out of %lld ...\nCollision efficiency: %6.4f.\n",
             (long long)st.collisions.Nc, (long long)prm.Np,
             (st.collisions.accepted + st.collisions.refused) > 0
                 ? (double)st.collisions.accepted / (double)(st.collisions.accepted + st.collisions.refused)
                 : 1.0);
    printf("  particles %.3f ms, potential %.3f ms (%lld its, residual %.2e)\n", st.ms_particles, st.solve.ms,
           (long long)st.solve.iterations, st.solve.residual);
    total_ms += st.ms_total;
    solve_ms += st.solve.ms;
    part_ms += st.ms_particles;
    if (prm.output_interval > 0 && (it + 1) % prm.output_interval == 0) {
      char insert[96];
      snprintf(insert, sizeof insert, "-%0*lld", width, (long long)it);
      CHECK(output_rc(chy_output_write_async(ctx, prm.output, insert), prm.output));
    }
  }
  CHECK(output_rc(chy_output_wait(ctx), prm.output));
  if (prm.Nt > 0)
    printf("\nparticle-steps/s %.4e, mean step %.3f ms (particles %.3f, potential %.3f)\n",
           (double)prm.Np * (double)prm.Nt / (total_ms * 1e-3), total_ms / prm.Nt, part_ms / prm.Nt,
           solve_ms / prm.Nt);
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
