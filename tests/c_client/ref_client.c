/* A C host of the rigid-fiber timestep through the C ABI alone
 * (include/rigidfibers.h), as the reference's src/main.cu:83-99 loop runs
 * Simulation::step: reads a fibers file (src/parameters.cc:75-116 layout),
 * runs `steps` timesteps with the default configuration (GMRES, numerical
 * integrals, the reference's .ini values) and compares the final positions
 * and orientations with reference vectors (one value per line).
 *
 *   ref_client <fibers file> <steps> <positions ref> <orientations ref>
 *
 * Exit status 0 when both agree to 1e-5 (max |a - b| / max |b|). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "rigidfibers.h"

static int read_fibers(const char* path, int* M, double** pos, double** ori) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  if (fscanf(f, "%d", M) != 1 || *M < 1) {
    fclose(f);
    return -1;
  }
  *pos = malloc(sizeof(double) * 3 * (size_t)*M);
  *ori = malloc(sizeof(double) * 3 * (size_t)*M);
  for (int i = 0; i < *M; ++i) {
    double* p = *pos + 3 * i;
    double* o = *ori + 3 * i;
    if (fscanf(f, "%lf %lf %lf %lf %lf %lf", p, p + 1, p + 2, o, o + 1, o + 2) != 6) {
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return 0;
}

static int read_column(const char* path, double* v, int n) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  for (int i = 0; i < n; ++i)
    if (fscanf(f, "%lf", v + i) != 1) {
      fclose(f);
      return -1;
    }
  fclose(f);
  return 0;
}

static double rel_err(const double* a, const double* b, int n) {
  double d = 0.0, m = 0.0;
  for (int i = 0; i < n; ++i) {
    d = fmax(d, fabs(a[i] - b[i]));
    m = fmax(m, fabs(b[i]));
  }
  return d / m;
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s <fibers file> <steps> <positions ref> <orientations ref>\n", argv[0]);
    return 2;
  }
  int M = 0;
  double *pos = NULL, *ori = NULL;
  if (read_fibers(argv[1], &M, &pos, &ori) != 0) {
    fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  const long steps = strtol(argv[2], NULL, 10);
  rf_config cfg;
  rf_config_init(&cfg);
  cfg.number_of_fibers = M;
  rf_simulation* sim = NULL;
  if (rf_create(&cfg, pos, ori, &sim) != RF_OK) {
    fprintf(stderr, "rf_create: %s\n", rf_last_error());
    return 2;
  }
  for (long t = 0; t < steps; ++t) {
    const rf_status s = rf_step(sim, t);
    if (s != RF_OK && s != RF_WARN_NOT_CONVERGED) {
      fprintf(stderr, "rf_step %ld: %s\n", t, rf_last_error());
      rf_destroy(sim);
      return 2;
    }
  }
  rf_solver_stats st;
  rf_get_solver_stats(sim, &st);
  rf_get_state(sim, pos, ori);
  rf_destroy(sim);
  double* ref = malloc(sizeof(double) * 3 * (size_t)M);
  if (read_column(argv[3], ref, 3 * M) != 0) return 2;
  const double ep = rel_err(pos, ref, 3 * M);
  if (read_column(argv[4], ref, 3 * M) != 0) return 2;
  const double eo = rel_err(ori, ref, 3 * M);
  printf("M=%d steps=%ld gmres iterations %d: positions %.2e orientations %.2e\n", M, steps, st.iterations, ep, eo);
  free(ref);
  free(pos);
  free(ori);
  return (ep < 1e-5 && eo < 1e-5) ? 0 : 1;
}
