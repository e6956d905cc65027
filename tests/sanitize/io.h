/* Raw-file reading and writing for the sanitizer check drivers (test
 * infrastructure).  A case directory holds sizes.i32 (nelem npoin nboun
 * rorder torder), intma.i32, coord.f64, bsido.i32, eta.f64, w.f64 and a
 * parameter block params.bin, written by tests/test_host_sanitizers.py. */
#pragma once
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *read_file(const char *dir, const char *name, size_t *bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *fp = fopen(path, "rb");
  if (!fp) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  fseek(fp, 0, SEEK_END);
  long n = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  void *buf = malloc(n > 0 ? (size_t)n : 1);
  if (!buf || (n > 0 && fread(buf, 1, (size_t)n, fp) != (size_t)n)) {
    fprintf(stderr, "cannot read %s\n", path); exit(2);
  }
  fclose(fp);
  *bytes = (size_t)n;
  return buf;
}

static void write_file(const char *dir, const char *name, const void *data, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *fp = fopen(path, "wb");
  if (!fp || (bytes && fwrite(data, 1, bytes, fp) != bytes)) {
    fprintf(stderr, "cannot write %s\n", path); exit(2);
  }
  fclose(fp);
}

typedef struct {
  int nelem, npoin, nboun, rorder, torder;
  int *intma, *bsido;
  double *coord, *eta, *w;
  void *params;
  size_t params_bytes;
} case_input;

static void check_size(const char *name, size_t got, size_t want) {
  if (got != want) {
    fprintf(stderr, "%s: %zu bytes, expected %zu\n", name, got, want);
    exit(2);
  }
}

static case_input read_case(const char *dir) {
  case_input c;
  size_t n;
  int *sz = (int *)read_file(dir, "sizes.i32", &n);
  check_size("sizes.i32", n, 5 * sizeof(int));
  c.nelem = sz[0]; c.npoin = sz[1]; c.nboun = sz[2]; c.rorder = sz[3]; c.torder = sz[4];
  free(sz);
  c.intma = (int *)read_file(dir, "intma.i32", &n);
  check_size("intma.i32", n, 3 * sizeof(int) * (size_t)c.nelem);
  c.coord = (double *)read_file(dir, "coord.f64", &n);
  check_size("coord.f64", n, 2 * sizeof(double) * (size_t)c.npoin);
  c.bsido = (int *)read_file(dir, "bsido.i32", &n);
  check_size("bsido.i32", n, 6 * sizeof(int) * (size_t)c.nboun);
  c.eta = (double *)read_file(dir, "eta.f64", &n);
  check_size("eta.f64", n, sizeof(double) * (size_t)c.rorder);
  c.w = (double *)read_file(dir, "w.f64", &n);
  check_size("w.f64", n, sizeof(double) * (size_t)c.rorder);
  c.params = read_file(dir, "params.bin", &c.params_bytes);
  return c;
}

static void free_case(case_input *c) {
  free(c->intma); free(c->bsido); free(c->coord); free(c->eta); free(c->w); free(c->params);
}
