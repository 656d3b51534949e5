/* TEST INFRASTRUCTURE: a driver of the CPU oracle for the AddressSanitizer test (tests/test_host.py).
 *
 *   oracle_driver <control.bin> <scene.f64> <index.u32> <out prefix>
 *
 * reads a tcv_ba_control (its bytes as the C struct lays them out), a packed scene and index,
 * runs the oracle's solve (tcv_oracle_solve), chart gradient (tcv_oracle_gradient) and dense solve
 * (tcv_oracle_dense_solve), and writes <prefix>.solve.f64, <prefix>.gradient.f64 and
 * <prefix>.dense.f64. Built with -fsanitize=address together with oracle/ba_oracle.c.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "typecv_ba.h"

int tcv_oracle_set_threads(int threads);
int tcv_oracle_solve(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                     double* out, double* stats, uint32_t max_lm);
int tcv_oracle_gradient(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                        double log_focal, double* grad, double* half_cost);
int tcv_oracle_dense_solve(const tcv_ba_control* ctl, const double* scene, const uint32_t* index,
                           double* out);

/* The packed sizes (include/typecv_ba.h), restated so that the driver links the oracle only. */
static size_t scene_words(const tcv_ba_control* c) {
  return (size_t)c->cameras * 18 + (size_t)c->points * 3 + (size_t)c->observations * 2;
}
static size_t index_words(const tcv_ba_control* c) {
  return (size_t)c->observations * 3 + (size_t)c->points + (size_t)c->cameras * 2 + 2;
}
static size_t output_words(const tcv_ba_control* c) {
  return (size_t)c->cameras * 12 + (size_t)c->points * 3 + 26 + (c->shared_focal ? 1 : 0);
}

static void* slurp(const char* path, size_t* bytes) {
  FILE* f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  void* data = malloc(n > 0 ? (size_t)n : 1);
  if (data && n > 0 && fread(data, 1, (size_t)n, f) != (size_t)n) {
    free(data);
    data = NULL;
  }
  fclose(f);
  *bytes = (size_t)n;
  return data;
}

static int spill(const char* prefix, const char* name, const double* v, size_t n) {
  char path[4096];
  snprintf(path, sizeof path, "%s.%s.f64", prefix, name);
  FILE* f = fopen(path, "wb");
  if (!f) return 0;
  const size_t w = fwrite(v, sizeof *v, n, f);
  fclose(f);
  return w == n;
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: oracle_driver <control.bin> <scene.f64> <index.u32> <out prefix>\n");
    return 2;
  }
  size_t cb = 0, sb = 0, ib = 0;
  tcv_ba_control* ctl = slurp(argv[1], &cb);
  double* scene = slurp(argv[2], &sb);
  uint32_t* index = slurp(argv[3], &ib);
  if (!ctl || !scene || !index || cb != sizeof *ctl || sb != scene_words(ctl) * 8 ||
      ib != index_words(ctl) * 4) {
    fprintf(stderr, "inputs do not match the control sizes\n");
    return 2;
  }
  tcv_oracle_set_threads(2);
  const size_t n_out = output_words(ctl);
  const size_t n_grad = (size_t)ctl->cameras * 6 + (size_t)ctl->points * 3 + (ctl->shared_focal ? 1 : 0);
  double* out = calloc(n_out, sizeof *out);
  double* dense = calloc(n_out, sizeof *dense);
  double* grad = calloc(n_grad, sizeof *grad);
  double stats[4] = {0, 0, 0, 0}, half = 0;
  int status = tcv_oracle_solve(ctl, scene, index, out, stats, 0);
  if (status == 0) status = tcv_oracle_gradient(ctl, scene, index, 0.0, grad, &half) < 0 ? 1 : 0;
  if (status == 0) status = tcv_oracle_dense_solve(ctl, scene, index, dense);
  if (status == 0 && !(spill(argv[4], "solve", out, n_out) && spill(argv[4], "gradient", grad, n_grad) &&
                       spill(argv[4], "dense", dense, n_out)))
    status = 3;
  printf("{\"status\": %d, \"lm_iterations\": %.0f, \"cg_iterations\": %.0f}\n", status, stats[0],
         stats[1]);
  free(out);
  free(dense);
  free(grad);
  free(ctl);
  free(scene);
  free(index);
  return status == 0 ? 0 : 1;
}
