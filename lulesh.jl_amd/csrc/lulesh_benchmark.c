/*
 * The reference's benchmark driver (examples/benchmark.jl:22-122 with
 * parse_cmd, src/utils.jl:3-27) in plain C over the C ABI of
 * liblulesh_hip.so: no Python, no torch.  Same command line,
 *
 *   lulesh_benchmark [-s] [--mpi] [--enzyme] [N] [num_iters]
 *
 * (N = elements per edge, default 45; num_iters default 10), the same loop
 *
 *   while time < stoptime: timeIncrement!; lagrangeLeapFrog (or its
 *   Enzyme call); break if cycle >= num_iters
 *
 * and the same final report (verifyAndWriteFinalOutput, src/utils.jl:124-164)
 * on stdout.  The reference's @info lines go to stderr, as Julia's logger
 * sends them.  By default the loop runs on the device (lulesh_run, one
 * replayed graph per cycle, the loop test evaluated on the device), so no
 * per-cycle "Completed" line is printed; --per-cycle steps it from the host
 * and prints them.  One rank: --mpi is refused.  Exit status 0, or 1 with
 * the library's (the reference's) error text.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "lulesh_hip.h"

static double wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int check(lulesh_ctx *ctx, int rc, const char *what) {
  if (rc == LULESH_OK) return 0;
  fprintf(stderr, "ERROR: %s: %s\n", what,
          ctx ? lulesh_last_error(ctx) : lulesh_error_string(rc));
  return 1;
}

static double scalar(lulesh_ctx *ctx, const char *name) {
  double v = 0.0;
  lulesh_get_scalar(ctx, name, &v);
  return v;
}

/* shadowDomain = Domain(prob) (examples/benchmark.jl:49): the shadow's state
 * fields start as the fresh domain's */
static const char *kStateFields[] = {"x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss"};

static int init_shadow(lulesh_ctx *ctx) {
  const int64_t E = lulesh_num_elem(ctx), N = lulesh_num_node(ctx);
  double *buf = (double *)malloc(sizeof(double) * (size_t)(N > E ? N : E));
  if (!buf) return 1;
  int err = 0;
  for (int k = 0; k < 11 && !err; ++k) {
    const int64_t n = k < 6 ? N : E;
    err = check(ctx, lulesh_download(ctx, kStateFields[k], buf, n), "download") ||
          check(ctx, lulesh_shadow_upload(ctx, kStateFields[k], buf, n), "shadow upload");
  }
  free(buf);
  return err;
}

/* a number as Julia's @printf writes it: Inf, -Inf and NaN by those names */
static const char *jl(char *buf, size_t n, const char *fmt, double x) {
  if (isnan(x)) return "NaN";
  if (isinf(x)) return x > 0 ? "Inf" : "-Inf";
  snprintf(buf, n, fmt, x);
  return buf;
}

/* src/utils.jl:124-164 */
static int report(lulesh_ctx *ctx, double elapsed, int nx, int ranks) {
  double v[4];
  if (check(ctx, lulesh_verify(ctx, v), "verify")) return 1;
  const double cycle = scalar(ctx, "cycle");
  const double grind1 = ((elapsed * 1e6) / cycle) / ((double)nx * nx * nx);
  const double grind2 = ((elapsed * 1e6) / cycle) / ((double)nx * nx * nx * ranks);
  printf("Run completed:\n");
  printf("   Problem size        =  %d\n", nx);
  printf("   MPI tasks           =  %d\n", ranks);
  printf("   Iteration count     =  %lld\n", (long long)cycle);
  char a[64], b[64];
  printf("   Final Origin Energy =  %s\n", jl(a, sizeof a, "%.6e", v[0]));
  printf("   Testing Plane 0 of Energy Array on rank 0:\n");
  printf("        MaxAbsDiff   = %s\n", jl(a, sizeof a, "%.6e", v[1]));
  printf("        TotalAbsDiff = %s\n", jl(a, sizeof a, "%.6e", v[2]));
  printf("        MaxRelDiff   = %s\n", jl(a, sizeof a, "%.6e", v[3]));
  printf("\nElapsed time         = %.2f (s)\n", elapsed);
  printf("Grind time (us/z/c)  =  %s  (per dom)  (%s overall)\n", jl(a, sizeof a, "%.8f", grind1),
         jl(b, sizeof b, "%.8f", elapsed));
  printf("FOM                  = %s (z/s)\n\n", jl(a, sizeof a, "%.8f", 1000.0 / grind2));
  return 0;
}

static int usage(const char *prog) {
  fprintf(stderr, "usage: %s [-s] [--mpi] [--enzyme] [--per-cycle] [N] [num_iters]\n", prog);
  return 2;
}

int main(int argc, char **argv) {
  int structured = 0, mpi = 0, enzyme = 0, per_cycle = 0, npos = 0;
  long nx = 45, num_iters = 10;
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    if (!strcmp(a, "-s")) structured = 1;
    else if (!strcmp(a, "--mpi")) mpi = 1;
    else if (!strcmp(a, "--enzyme")) enzyme = 1;
    else if (!strcmp(a, "--per-cycle")) per_cycle = 1;
    else if (!strcmp(a, "-h") || !strcmp(a, "--help")) return usage(argv[0]) - 2;
    else {
      char *end = NULL;
      const long v = strtol(a, &end, 10);
      if (a[0] == '-' || !end || *end || npos > 1) return usage(argv[0]);
      if (npos++ == 0) nx = v;
      else num_iters = v;
    }
  }
  if (mpi) {
    fprintf(stderr, "ERROR: --mpi: this driver runs one rank on one GPU\n");
    return 2;
  }
  fprintf(stderr, "[ Info: Constructing LuleshProblem num_iters=%ld structured=%s nx=%ld nr=1 "
          "balance=1 cost=1 ranks=1\n", num_iters, structured ? "true" : "false", nx);
  if (!structured) {
    /* Domain(prob) without -s, src/domain.jl:612 */
    fprintf(stderr, "ERROR: Reading unstructured mesh is currently missing in the Julia "
            "version of LULESH.\n");
    return 1;
  }
  if (nx < 1 || nx > 800) {
    fprintf(stderr, "ERROR: N must lie in 1..800\n");
    return 2;
  }
  lulesh_ctx *ctx = NULL;
  if (check(NULL, lulesh_create((int32_t)nx, 1, 1, 1, LULESH_QUIRKS_REFERENCE, 0, &ctx),
            "Domain(prob)"))
    return 1;
  int err = 0;
  if (enzyme) {
    /* the first Enzyme call on a fresh domain, then a fresh domain again
     * (examples/benchmark.jl:47-51); the shadow carries over */
    err = init_shadow(ctx);
    if (!err) {
      const double t0 = wtime();
      err = check(ctx, lulesh_vjp_cycle(ctx, NULL), "Enzyme.autodiff");
      fprintf(stderr, "  %.6f seconds\n", wtime() - t0);
    }
    if (!err) err = check(ctx, lulesh_reset(ctx), "Domain(prob)");
  }
  if (!err)
    fprintf(stderr, "[ Info: Running until=%g domain=(%ld, %ld, %ld)\n", scalar(ctx, "stoptime"),
            nx, nx, nx);
  const double start = wtime();
  if (!err && !enzyme && !per_cycle) {
    int64_t ran = 0;
    err = check(ctx, lulesh_run(ctx, (int64_t)num_iters, &ran), "lagrangeLeapFrog");
  } else {
    while (!err && scalar(ctx, "time") < scalar(ctx, "stoptime")) {
      err = check(ctx, lulesh_time_increment(ctx), "timeIncrement!");
      if (err) break;
      if (enzyme) err = check(ctx, lulesh_vjp_cycle(ctx, NULL), "Enzyme.autodiff");
      else err = check(ctx, lulesh_lagrange_leapfrog(ctx), "lagrangeLeapFrog");
      if (err) break;
      fprintf(stderr, "[ Info: Completed cycle=%lld time=%.17g dt=%.17g\n",
              (long long)scalar(ctx, "cycle"), scalar(ctx, "time"), scalar(ctx, "deltatime"));
      if (scalar(ctx, "cycle") >= (double)num_iters) break;
    }
  }
  const double elapsed = wtime() - start;
  if (!err) err = report(ctx, elapsed, (int)nx, 1);
  lulesh_destroy(ctx);
  return err ? 1 : 0;
}
