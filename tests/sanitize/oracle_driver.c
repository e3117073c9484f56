/* Drives the oracle (oracle/fh_oracle.c) through its entry points on small
 * problems -- uniform and graded busbar pairs, a skew random cloud, the LU,
 * the mesh system and its sweep, the pFFT operator, GMRES and the iterative
 * system -- for an AddressSanitizer + UndefinedBehaviorSanitizer build
 * (oracle/Makefile target `sanitize`).  Exits non-zero on a wrong result;
 * the sanitizers abort on a memory or UB error. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fh_oracle.h"

static int failures = 0;
#define CHECK(cond, what)                          \
  do {                                             \
    if (!(cond)) {                                 \
      fprintf(stderr, "FAILED: %s\n", what);       \
      ++failures;                                  \
    }                                              \
  } while (0)

static uint64_t state = 7;
static double rnd(void) {
  state = state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(state >> 11) / 9007199254740992.0;
}

/* Two parallel bars of `segs` segments shorted at the far end, one port. */
static void busbar(int segs, uint32_t nw, uint32_t nh, double ratio, double **nodes, size_t *nn,
                   fho_segment **sg, fho_grid **gr, size_t *ns, fho_port *port) {
  *nn = 2 * (size_t)(segs + 1);
  *ns = 2 * (size_t)segs + 1;
  *nodes = calloc(3 * *nn, sizeof(double));
  *sg = calloc(*ns, sizeof(fho_segment));
  *gr = calloc(*ns, sizeof(fho_grid));
  for (int k = 0; k <= segs; ++k) {
    (*nodes)[3 * k] = 1e-3 * k;
    (*nodes)[3 * (segs + 1 + k)] = 1e-3 * k;
    (*nodes)[3 * (segs + 1 + k) + 1] = 1.5e-3;
  }
  size_t s = 0;
  for (int k = 0; k < segs; ++k) {
    (*sg)[s].a = k;
    (*sg)[s++].b = k + 1;
  }
  for (int k = 0; k < segs; ++k) {
    (*sg)[s].a = segs + 2 + k;
    (*sg)[s++].b = segs + 1 + k;
  }
  (*sg)[s].a = segs;
  (*sg)[s++].b = 2 * segs + 1;
  for (size_t q = 0; q < *ns; ++q) {
    (*sg)[q].width = 1e-3;
    (*sg)[q].height = 1e-4;
    (*sg)[q].sigma = 5.8e7;
    (*gr)[q].nw = nw;
    (*gr)[q].nh = nh;
    (*gr)[q].width_ratio = ratio;
    (*gr)[q].height_ratio = ratio;
    (*gr)[q].graded = ratio != 1.0;
  }
  port->positive = 0;
  port->negative = segs + 1;
}

static void mesh_case(uint32_t nw, uint32_t nh, double ratio) {
  double *nodes;
  size_t nn, ns;
  fho_segment *sg;
  fho_grid *gr;
  fho_port port;
  busbar(5, nw, nh, ratio, &nodes, &nn, &sg, &gr, &ns, &port);
  fho_system *sys = NULL;
  CHECK(fho_system_create(nodes, nn, sg, ns, gr, &port, 1, 2, &sys) == FHO_OK, "system");
  if (sys) {
    size_t c[6];
    fho_system_counts(sys, c);
    const size_t m = c[4], p = c[5];
    double *blk = calloc(2 * (m * m + m * p + p * p), sizeof(double));
    CHECK(fho_system_blocks(sys, blk, blk + m * m, blk + 2 * m * m, blk + 2 * m * m + m * p,
                            blk + 2 * m * m + 2 * m * p, blk + 2 * m * m + 2 * m * p + p * p) == FHO_OK,
          "blocks");
    const double freqs[4] = {0.0, 1e3, 1e6, 1e9};
    double z[8];
    CHECK(fho_system_sweep(sys, freqs, 4, z, 2) == FHO_OK, "sweep");
    for (int k = 0; k < 8; ++k) CHECK(isfinite(z[k]), "finite Z");
    CHECK(z[1] == 0.0, "DC is real");
    const double bad[2] = {1e3, -1.0};
    CHECK(fho_system_sweep(sys, bad, 2, z, 1) == FHO_E_INVALID_FREQUENCY, "invalid frequency");
    free(blk);
    fho_system_destroy(sys);
  }
  /* the iterative system on the same problem */
  fho_pfft_params pp = {1, 8.0, 0.0, 8.0, 5};
  fho_iter *it = NULL;
  CHECK(fho_iter_create(nodes, nn, sg, ns, gr, &port, 1, &pp, 1e-10, 30, 500, &it) == FHO_OK, "iterative");
  if (it) {
    double z[2];
    size_t its = 0;
    CHECK(fho_iter_impedance(it, 1e6, z, &its) == FHO_OK && isfinite(z[0]), "iterative Z");
    fho_iter_destroy(it);
  }
  free(nodes);
  free(sg);
  free(gr);
}

static void cloud(void) {
  enum { N = 60 };
  double nodes[6 * N];
  fho_segment sg[N];
  fho_grid gr[N];
  memset(sg, 0, sizeof sg);
  memset(gr, 0, sizeof gr);
  for (int k = 0; k < N; ++k) {
    for (int a = 0; a < 3; ++a) nodes[6 * k + a] = 1e-2 * rnd();
    for (int a = 0; a < 3; ++a) nodes[6 * k + 3 + a] = nodes[6 * k + a] + 2e-3 * (rnd() - 0.5);
    sg[k].a = 2 * k;
    sg[k].b = 2 * k + 1;
    sg[k].width = 2e-4 * (0.5 + rnd());
    sg[k].height = 1e-4 * (0.5 + rnd());
    sg[k].sigma = 5.8e7;
    gr[k].nw = 1 + k % 3;
    gr[k].nh = 1 + k % 2;
    gr[k].width_ratio = gr[k].height_ratio = 1.0 + (k % 4);
    gr[k].graded = k % 4 != 0;
  }
  size_t n = 0;
  CHECK(fho_discretize_geometry(nodes, 2 * N, sg, N, gr, NULL, 0, &n) == FHO_OK, "cloud count");
  fho_filament *f = calloc(n, sizeof(fho_filament));
  CHECK(fho_discretize_geometry(nodes, 2 * N, sg, N, gr, f, n, &n) == FHO_OK, "cloud filaments");
  double *L = calloc(n * n, sizeof(double));
  CHECK(fho_fill(f, n, L, 2) == FHO_OK, "fill");
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) CHECK(L[i * n + j] == L[j * n + i], "symmetric");
  fho_pfft_params pp = {1, 8.0, 0.0, 8.0, 5};
  fho_pfft *op = NULL;
  CHECK(fho_pfft_create(f, n, &pp, &op) == FHO_OK, "pfft");
  if (op) {
    double *x = calloc(n, sizeof(double)), *y = calloc(n, sizeof(double));
    for (size_t i = 0; i < n; ++i) x[i] = rnd() - 0.5;
    fho_pfft_apply(op, x, y);
    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
      double d = 0;
      for (size_t j = 0; j < n; ++j) d += L[i * n + j] * x[j];
      num += (y[i] - d) * (y[i] - d);
      den += d * d;
    }
    CHECK(sqrt(num / den) < 1e-6, "pfft product against the dense fill");
    free(x);
    free(y);
    fho_pfft_destroy(op);
  }
  /* invalid inputs (the count query alone does not discretize) */
  fho_filament few[64];
  fho_grid bad = gr[1];
  bad.width_ratio = 0.5;
  bad.graded = 1;
  CHECK(fho_discretize_geometry(nodes, 2, sg, 1, &bad, few, 64, &n) != FHO_OK, "bad ratio");
  bad = gr[1];
  bad.nw = 0xFFFFFFFFu;
  bad.nh = 0xFFFFFFFFu;
  CHECK(fho_discretize_geometry(nodes, 2, sg, 1, &bad, few, 64, &n) != FHO_OK, "overflow");
  free(L);
  free(f);
}

static void dense(void) {
  for (size_t order = 1; order < 70; order += 17) {
    double *a = calloc(2 * order * order, sizeof(double)), *b = calloc(2 * order, sizeof(double));
    for (size_t k = 0; k < 2 * order * order; ++k) a[k] = rnd() - 0.5;
    for (size_t k = 0; k < order; ++k) a[2 * (k * order + k)] += (double)order;
    for (size_t k = 0; k < 2 * order; ++k) b[k] = rnd();
    double *a2 = malloc(2 * order * order * sizeof(double)), *x = calloc(2 * order, sizeof(double));
    memcpy(a2, a, 2 * order * order * sizeof(double));
    size_t its = 0;
    double res = 0;
    int conv = 0;
    CHECK(fho_gmres_dense(order, a2, b, x, 1e-10, 8, 400, 1, &its, &res, &conv) == FHO_OK && conv, "gmres");
    CHECK(fho_lu_solve_complex(a, order, b, 1, 2) == FHO_OK, "lu");
    free(a);
    free(a2);
    free(b);
    free(x);
  }
  double z[4] = {0, 0, 0, 0}, r[2] = {1, 1};
  CHECK(fho_lu_solve_real(z, 2, r, 1, 1) == FHO_E_SINGULAR, "singular");
}

int main(void) {
  for (int portable = 1; portable >= 0; --portable) {
    fho_set_math(portable);
    mesh_case(3, 2, 1.0);
    mesh_case(4, 3, 2.0);
    cloud();
    dense();
  }
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle sanitizer driver: all checks passed\n");
  return 0;
}
