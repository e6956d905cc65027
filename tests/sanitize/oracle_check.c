/* Sanitizer check driver for the oracle (test infrastructure).  Runs
 * ora_create and N steps of oracle/dgbgk_oracle.c on a case read from raw
 * files, serially and with T OpenMP threads, and writes the per-step
 * records and f of both runs back for comparison with the oracle library.
 * Built with -fsanitize=address,undefined by tests/sanitize/Makefile.
 *   oracle_check <case_dir> <out_dir> <steps> <threads> */
#include "dgbgk_oracle.h"
#include "io.h"

static int run(const case_input *in, int steps, int threads, const char *out, const char *tag) {
  ora_ctx *c = NULL;
  int st = ora_create(in->nelem, in->npoin, in->nboun, in->intma, in->coord, in->bsido,
                      in->rorder, in->torder, in->eta, in->w, (const ora_params *)in->params,
                      &c);
  if (st) { fprintf(stderr, "ora_create: status %d\n", st); return 1; }
  if (threads > 1 && ora_set_threads(c, threads)) { ora_destroy(c); return 1; }
  ora_step_info *info = (ora_step_info *)calloc((size_t)steps, sizeof *info);
  st = ora_step(c, steps, info);
  if (st) { fprintf(stderr, "ora_step: status %d\n", st); free(info); ora_destroy(c); return 1; }
  size_t nf = 3 * (size_t)ora_nvel(c) * (size_t)in->nelem;
  double *f = (double *)malloc(nf * sizeof *f);
  ora_get_f(c, f);
  char name[64];
  snprintf(name, sizeof name, "info_%s.f64", tag);
  write_file(out, name, info, (size_t)steps * sizeof *info);
  snprintf(name, sizeof name, "f_%s.f64", tag);
  write_file(out, name, f, nf * sizeof *f);
  free(f);
  free(info);
  ora_destroy(c);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: oracle_check <case_dir> <out_dir> <steps> <threads>\n");
    return 2;
  }
  case_input in = read_case(argv[1]);
  check_size("params.bin", in.params_bytes, sizeof(ora_params));
  int steps = atoi(argv[3]), threads = atoi(argv[4]);
  int rc = run(&in, steps, 1, argv[2], "serial");
  if (!rc && threads > 1) rc = run(&in, steps, threads, argv[2], "threads");
  free_case(&in);
  return rc;
}
