/*
 * cac_oracle.c -- CPU FP64 oracle for the CAC force-and-neighbor path.
 *
 * TEST INFRASTRUCTURE ONLY (see cac_oracle.h).  Plain C restatement of the
 * reference arithmetic, compiled with -ffp-contract=off -fno-fast-math so
 * every multiply and add rounds on its own, in the order the reference
 * writes it.  The reference itself cannot be built here (it needs MPI and
 * its hot path is never wired, run.cpp:93-99, verlet.cpp:91-97), so these
 * functions are pinned by closed forms, brute force, finite differences and
 * the finest-mesh invariant in tests/test_oracle_*.py.
 *
 * Where the reference has a defect on this path the oracle fixes it and the
 * fix is named at the spot:
 *   - element neighbor search is complete (the reference bins elements with
 *     the atom stencil, neighbor.cpp:963-975, which cannot reach cut_all);
 *   - elepbc subtracts the period at x-hi (domain.cpp:682 adds it);
 *   - the element / sub-element pruning radii are true bounds (see
 *     orc_ip_neighbors), so the integration-point lists are complete.
 */
#include "cac_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXD(a, b) ((a) > (b) ? (a) : (b))
#define MIND(a, b) ((a) < (b) ? (a) : (b))

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static void img_shift(int img, int s[3]) {
  s[0] = img % 3 - 1;
  s[1] = (img / 3) % 3 - 1;
  s[2] = img / 9 - 1;
}

/* Image copy of a coordinate: x + s*prd, one add per component, the value a
 * reference ghost carries (atom_vec_atomic.cpp:202-221). */
static inline double shifted(double x, int s, double prd) {
  return s ? x + (double)s * prd : x;
}

/* ------------------------------------------------------------------------
 * shape functions: element.cpp:1361-1386 (sites), :1326-1354 (integration
 * points).  N_k = 0.125*(1+-px)*(1+-py)*(1+-pz), evaluated left to right.
 * ---------------------------------------------------------------------- */
void orc_shape(double px, double py, double pz, double N[8]) {
  N[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
  N[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
  N[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
  N[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
  N[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
  N[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
  N[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
  N[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
}

/* Element::inter_ele, element.cpp:1394-1423: parent coordinates of all
 * nae0*nae1*nae2 sites, x index outermost, z innermost. */
void orc_site_pcoords(const int nae[3], double *pc) {
  double inx = 2 / ((double)nae[0] - 1);
  double iny = 2 / ((double)nae[1] - 1);
  double inz = 2 / ((double)nae[2] - 1);
  int n = 0;
  for (int ii = 0; ii < nae[0]; ii++)
    for (int jj = 0; jj < nae[1]; jj++)
      for (int kk = 0; kk < nae[2]; kk++) {
        pc[3 * n + 0] = inx * (double)ii - 1;
        pc[3 * n + 1] = iny * (double)jj - 1;
        pc[3 * n + 2] = inz * (double)kk - 1;
        n++;
      }
}

/* Element::shape_array_inter_ele, element.cpp:1361-1386. */
void orc_site_shapes(const int nae[3], double *spa) {
  int ns = nae[0] * nae[1] * nae[2];
  double *pc = (double *)malloc(sizeof(double) * 3 * ns);
  orc_site_pcoords(nae, pc);
  for (int s = 0; s < ns; s++) orc_shape(pc[3 * s], pc[3 * s + 1], pc[3 * s + 2], spa + 8 * s);
  free(pc);
}

/* Element::intp_pcoord (element.cpp:1170-1191) then shape_array_int
 * (element.cpp:1326-1354): lattice indices -> parent coords -> N_k. */
void orc_intp_shapes(const int nae[3], int nint, const int *ijk, double *spa_int) {
  double dx = 2 / ((double)nae[0] - 1);
  double dy = 2 / ((double)nae[1] - 1);
  double dz = 2 / ((double)nae[2] - 1);
  for (int g = 0; g < nint; g++) {
    double px = dx * ((double)ijk[3 * g + 0]) - 1;
    double py = dy * ((double)ijk[3 * g + 1]) - 1;
    double pz = dz * ((double)ijk[3 * g + 2]) - 1;
    orc_shape(px, py, pz, spa_int + 8 * g);
  }
}

/* Element::find_node_list, element.cpp:1296-1319 (TOL 0.01, element.cpp:21):
 * node k is carried by the first integration point with N_k within TOL of 1. */
int orc_node_list(int nint, const double *spa_int, int node_list[8]) {
  for (int k = 0; k < 8; k++) {
    int g;
    for (g = 0; g < nint; g++)
      if (fabs(spa_int[8 * g + k] - 1.0) < 0.01) break;
    if (g == nint) return -1;
    node_list[k] = g;
  }
  return 0;
}

/* Element::element_size, element.cpp:567-625: running max over elements of
 * the body diagonals 0-6, 1-7, 2-4, 3-5. */
double orc_element_max_size(int ne, const double *nodex) {
  double length = 0.0;
  for (int i = 0; i < ne; i++) {
    const double *X = nodex + 24 * i;
    for (int j = 0; j < 4; j++) {
      int a = j, b = (j < 2) ? j + 6 : j + 2;
      double lx = X[3 * a + 0] - X[3 * b + 0];
      double ly = X[3 * a + 1] - X[3 * b + 1];
      double lz = X[3 * a + 2] - X[3 * b + 2];
      length = MAXD(length, sqrt(lx * lx + ly * ly + lz * lz));
    }
  }
  return length;
}

/* Element center = node mean, fix_nve.cpp:139-153 (sum from zero in node
 * order, then divide by npe). */
void orc_element_centers(int ne, const double *nodex, double *xc) {
  for (int i = 0; i < ne; i++) {
    double c[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) c[d] += nodex[24 * i + 3 * k + d];
    for (int d = 0; d < 3; d++) xc[3 * i + d] = c[d] / 8;
  }
}

/* ------------------------------------------------------------------------
 * EAM tables
 * ---------------------------------------------------------------------- */

/* 4-point Lagrange regrid of one funcfl array, pair_eam.cpp:357-372
 * (frho), :396-411 (rhor), :451-463 (zr).  f1 is 1-based (f1[1..nf]). */
static void regrid(int n, double delta, int nf, double delta_f, const double *f1, double *out1) {
  const double sixth = 1.0 / 6.0;
  for (int m = 1; m <= n; m++) {
    double r = (m - 1) * delta;
    double p = r / delta_f + 1.0;
    int k = (int)p;
    k = k < nf - 2 ? k : nf - 2;
    k = k > 2 ? k : 2;
    p -= k;
    p = MIND(p, 2.0);
    double cof1 = -sixth * p * (p - 1.0) * (p - 2.0);
    double cof2 = 0.5 * (p * p - 1.0) * (p - 2.0);
    double cof3 = -0.5 * p * (p + 1.0) * (p - 2.0);
    double cof4 = sixth * p * (p * p - 1.0);
    out1[m] = cof1 * f1[k - 1] + cof2 * f1[k] + cof3 * f1[k + 1] + cof4 * f1[k + 2];
  }
}

/* PairEAM::file2array for one funcfl file, pair_eam.cpp:311-523.
 * Inputs are the file's arrays, 0-based (value at r=(m-1)*dr is f[m-1]).
 * Outputs are 1-based arrays of length n+1 (index 0 unused).  The z2r array
 * is the file's third block used as-is (pair_eam.cpp:486-487). */
int orc_funcfl_regrid(int nrho_f, double drho_f, int nr_f, double dr_f,
                      const double *frho_f, const double *rhor_f, const double *zr_f,
                      int *nrho, double *drho, int *nr, double *dr,
                      double *frho, double *rhor, double *z2r) {
  if (nrho_f < 5 || nr_f < 5) return -1;
  double rmax = (nr_f - 1) * dr_f;
  double rhomax = (nrho_f - 1) * drho_f;
  *dr = dr_f;
  *drho = drho_f;
  *nr = (int)(rmax / dr_f + 0.5);
  *nrho = (int)(rhomax / drho_f + 0.5);
  double *f1 = (double *)malloc(sizeof(double) * (MAXD(nrho_f, nr_f) + 1));
  memcpy(f1 + 1, frho_f, sizeof(double) * nrho_f);
  regrid(*nrho, *drho, nrho_f, drho_f, f1, frho);
  memcpy(f1 + 1, rhor_f, sizeof(double) * nr_f);
  regrid(*nr, *dr, nr_f, dr_f, f1, rhor);
  memcpy(f1 + 1, zr_f, sizeof(double) * nr_f);
  regrid(*nr, *dr, nr_f, dr_f, f1, z2r);
  free(f1);
  return 0;
}

/* PairEAM::interpolate, pair_eam.cpp:552-580.  f is 1-based [n+1]; spline is
 * [(n+1)][7], row 0 unused. */
void orc_spline(int n, double delta, const double *f, double *sp) {
#define S(m, c) sp[7 * (m) + (c)]
  for (int m = 1; m <= n; m++) S(m, 6) = f[m];
  S(1, 5) = S(2, 6) - S(1, 6);
  S(2, 5) = 0.5 * (S(3, 6) - S(1, 6));
  S(n - 1, 5) = 0.5 * (S(n, 6) - S(n - 2, 6));
  S(n, 5) = S(n, 6) - S(n - 1, 6);
  for (int m = 3; m <= n - 2; m++)
    S(m, 5) = ((S(m - 2, 6) - S(m + 2, 6)) + 8.0 * (S(m + 1, 6) - S(m - 1, 6))) / 12.0;
  for (int m = 1; m <= n - 1; m++) {
    S(m, 4) = 3.0 * (S(m + 1, 6) - S(m, 6)) - 2.0 * S(m, 5) - S(m + 1, 5);
    S(m, 3) = S(m, 5) + S(m + 1, 5) - 2.0 * (S(m + 1, 6) - S(m, 6));
  }
  S(n, 4) = 0.0;
  S(n, 3) = 0.0;
  for (int m = 1; m <= n; m++) {
    S(m, 2) = S(m, 5) / delta;
    S(m, 1) = 2.0 * S(m, 4) / delta;
    S(m, 0) = 3.0 * S(m, 3) / delta;
  }
  for (int c = 0; c < 7; c++) S(0, c) = 0.0;
#undef S
}

/* r-table lookup, pair_eam.cpp:649-655. */
static inline const double *rtab(const double *sp, double r, double rdr, int nr, double *pout) {
  double p = r * rdr + 1.0;
  int m = (int)p;
  m = m < nr - 1 ? m : nr - 1;
  p -= m;
  p = MIND(p, 1.0);
  *pout = p;
  return sp + 7 * m;
}

/* rho-table lookup, pair_eam.cpp:672-678. */
static inline const double *rhotab(const double *sp, double rho, double rdrho, int nrho, double *pout) {
  double p = rho * rdrho + 1.0;
  int m = (int)p;
  m = MAXD(1, MIND(m, nrho - 1));
  p -= m;
  p = MIND(p, 1.0);
  *pout = p;
  return sp + 7 * m;
}

/* Embedding energy F(rho).  The reference computes no energy (eflag forced
 * to 0, pair_eam.cpp:594); the value polynomial is the one its table
 * defines, plus the linear extension past rhomax that keeps dF/drho = fp. */
static inline double embed_energy(const orc_eam *t, double rho, double fp) {
  double p;
  const double *c = rhotab(t->frho_spline, rho, t->rdrho, t->nrho, &p);
  double phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
  double rhomax = (t->nrho - 1) / t->rdrho;
  if (rho > rhomax) phi += fp * (rho - rhomax);
  return phi;
}

static inline double embed_fp(const orc_eam *t, double rho) {
  double p;
  const double *c = rhotab(t->frho_spline, rho, t->rdrho, t->nrho, &p);
  return (c[0] * p + c[1]) * p + c[2];
}

/* ------------------------------------------------------------------------
 * cell helper for periodic neighbor searches (cells >= cut, >= 3 per dim,
 * else brute force over the 27 images)
 * ---------------------------------------------------------------------- */
typedef struct {
  int n[3];
  double inv[3];
  int *head, *next;
  int ok;
} cells_t;

static void cells_build(cells_t *c, int np, const double *p, const orc_box *box, double cut) {
  c->ok = 1;
  for (int d = 0; d < 3; d++) {
    double prd = box->hi[d] - box->lo[d];
    c->n[d] = (int)(prd / (cut * 1.0001));
    if (c->n[d] < 3) c->ok = 0;
    c->inv[d] = c->n[d] > 0 ? c->n[d] / prd : 0.0;
  }
  c->head = c->next = NULL;
  if (!c->ok) return;
  int nc = c->n[0] * c->n[1] * c->n[2];
  c->head = (int *)malloc(sizeof(int) * nc);
  c->next = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  for (int i = 0; i < nc; i++) c->head[i] = -1;
  for (int i = np - 1; i >= 0; i--) {
    int ci[3];
    for (int d = 0; d < 3; d++) {
      int k = (int)floor((p[3 * i + d] - box->lo[d]) * c->inv[d]);
      ci[d] = k < 0 ? 0 : (k >= c->n[d] ? c->n[d] - 1 : k);
    }
    int id = (ci[2] * c->n[1] + ci[1]) * c->n[0] + ci[0];
    c->next[i] = c->head[id];
    c->head[id] = i;
  }
}

static void cells_free(cells_t *c) {
  free(c->head);
  free(c->next);
}

static int cell_of(const cells_t *c, const double *pt, const orc_box *box, int ci[3]) {
  for (int d = 0; d < 3; d++) {
    int k = (int)floor((pt[d] - box->lo[d]) * c->inv[d]);
    ci[d] = k < 0 ? 0 : (k >= c->n[d] ? c->n[d] - 1 : k);
  }
  return 0;
}

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : x > y;
}
static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

/* ------------------------------------------------------------------------
 * Atom neighbor list: Neighbor::half_bin_no_newton, neighbor.cpp:546-625 --
 * a FULL list of every j != i (local or ghost) with rsq <= cutneighmaxsq,
 * i over local atoms.  Ghosts are periodic images (comm_brick.cpp:641-914,
 * self-swap branch).  Entries per atom are returned sorted.
 * ---------------------------------------------------------------------- */
static int atom_pair_test(const double *x, int i, int j, int img, const double prd[3], double cutsq) {
  int s[3];
  img_shift(img, s);
  if (j == i && img == ORC_IMG_SELF) return 0;
  double xj = shifted(x[3 * j + 0], s[0], prd[0]);
  double yj = shifted(x[3 * j + 1], s[1], prd[1]);
  double zj = shifted(x[3 * j + 2], s[2], prd[2]);
  double delx = x[3 * i + 0] - xj;
  double dely = x[3 * i + 1] - yj;
  double delz = x[3 * i + 2] - zj;
  double rsq = delx * delx + dely * dely + delz * delz;
  return rsq <= cutsq;
}

int orc_atom_neighbors_brute(int n, const double *x, const orc_box *box, double cutneigh,
                             int maxneigh, int *numneigh, uint32_t *neigh) {
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  double cutsq = cutneigh * cutneigh;
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow)
  for (int i = 0; i < n; i++) {
    int cnt = 0;
    for (int img = 0; img < 27; img++)
      for (int j = 0; j < n; j++)
        if (atom_pair_test(x, i, j, img, prd, cutsq)) {
          if (cnt < maxneigh) neigh[(size_t)i * maxneigh + cnt] = (uint32_t)j | ((uint32_t)img << 27);
          cnt++;
        }
    if (cnt > maxneigh) overflow = 1;
    numneigh[i] = cnt;
    if (cnt <= maxneigh) qsort(neigh + (size_t)i * maxneigh, cnt, sizeof(uint32_t), cmp_u32);
  }
  return overflow ? -1 : 0;
}

int orc_atom_neighbors(int n, const double *x, const orc_box *box, double cutneigh,
                       int maxneigh, int *numneigh, uint32_t *neigh) {
  cells_t c;
  cells_build(&c, n, x, box, cutneigh);
  if (!c.ok) return orc_atom_neighbors_brute(n, x, box, cutneigh, maxneigh, numneigh, neigh);
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  double cutsq = cutneigh * cutneigh;
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
  for (int i = 0; i < n; i++) {
    int ci[3], cnt = 0;
    cell_of(&c, x + 3 * i, box, ci);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz}, s[3];
          for (int d = 0; d < 3; d++) {
            s[d] = 0;
            if (cc[d] < 0) { cc[d] += c.n[d]; s[d] = -1; }
            else if (cc[d] >= c.n[d]) { cc[d] -= c.n[d]; s[d] = 1; }
          }
          int img = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
          for (int j = c.head[(cc[2] * c.n[1] + cc[1]) * c.n[0] + cc[0]]; j >= 0; j = c.next[j])
            if (atom_pair_test(x, i, j, img, prd, cutsq)) {
              if (cnt < maxneigh) neigh[(size_t)i * maxneigh + cnt] = (uint32_t)j | ((uint32_t)img << 27);
              cnt++;
            }
        }
    if (cnt > maxneigh) overflow = 1;
    numneigh[i] = cnt;
    if (cnt <= maxneigh) qsort(neigh + (size_t)i * maxneigh, cnt, sizeof(uint32_t), cmp_u32);
  }
  cells_free(&c);
  return overflow ? -1 : 0;
}

/* ------------------------------------------------------------------------
 * Integration-point neighbor list: Neighbor::node_neighbor,
 * full_node_neighbor.cpp:19-239, fed by Neighbor::element_neighbor,
 * element_neighbor.cpp:11-75, with cutoffs from Neighbor::setup_cut,
 * neighbor.cpp:284-293 and sub-element tables from element.cpp:725-906.
 * ---------------------------------------------------------------------- */
typedef struct {
  int nsub;
  int *nsite;          /* sites per sub-element */
  int *site;           /* [nsub][nsite_total] site ids, list_inter_subele */
  double *spa_center;  /* [nsub][8] Spa_center_subele */
  double *spa_corner;  /* [nsub][8][8] N at the sub-element's 8 corners */
} subele_t;

/* Element::specify_sub_ele_pc / count_sub_element / shape_array_inter_subele,
 * element.cpp:725-906. */
static int subele_build(subele_t *sb, const int nae[3], int nsplit) {
  int ns = nae[0] * nae[1] * nae[2];
  int nsub = nsplit * nsplit * nsplit;
  double ls = 2.0 / nsplit;
  double pc0[8][3] = {{-1.0, -1.0, -1.0}, {ls - 1.0, -1.0, -1.0}, {ls - 1.0, ls - 1.0, -1.0},
                      {-1.0, ls - 1.0, -1.0}, {-1.0, -1.0, ls - 1.0}, {ls - 1.0, -1.0, ls - 1.0},
                      {ls - 1.0, ls - 1.0, ls - 1.0}, {-1.0, ls - 1.0, ls - 1.0}};
  double *pc = (double *)malloc(sizeof(double) * 3 * ns);
  orc_site_pcoords(nae, pc);
  sb->nsub = nsub;
  sb->nsite = (int *)calloc(nsub, sizeof(int));
  sb->site = (int *)malloc(sizeof(int) * (size_t)nsub * ns);
  sb->spa_center = (double *)malloc(sizeof(double) * 8 * nsub);
  sb->spa_corner = (double *)malloc(sizeof(double) * 64 * nsub);
  int in = 0, total = 0;
  for (int ii = 0; ii < nsplit; ii++)
    for (int jj = 0; jj < nsplit; jj++)
      for (int kk = 0; kk < nsplit; kk++, in++) {
        double sp[8][3];
        for (int l = 0; l < 8; l++) {
          sp[l][0] = in == 0 ? pc0[l][0] : pc0[l][0] + ii * ls;
          sp[l][1] = in == 0 ? pc0[l][1] : pc0[l][1] + jj * ls;
          sp[l][2] = in == 0 ? pc0[l][2] : pc0[l][2] + kk * ls;
        }
        double cpc[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < 8; l++)
          for (int d = 0; d < 3; d++) cpc[d] += sp[l][d];
        for (int d = 0; d < 3; d++) cpc[d] = cpc[d] / 8;
        orc_shape(cpc[0], cpc[1], cpc[2], sb->spa_center + 8 * in);
        for (int l = 0; l < 8; l++) orc_shape(sp[l][0], sp[l][1], sp[l][2], sb->spa_corner + 64 * in + 8 * l);
        double xlo = sp[0][0], xhi = sp[1][0], ylo = sp[0][1], yhi = sp[3][1], zlo = sp[0][2], zhi = sp[4][2];
        if (xlo < -0.99) xlo = -1.01;
        if (xhi > 0.99) xhi = 1.01;
        if (ylo < -0.99) ylo = -1.01;
        if (yhi > 0.99) yhi = 1.01;
        if (zlo < -0.99) zlo = -1.01;
        if (zhi > 0.99) zhi = 1.01;
        for (int s = 0; s < ns; s++) {
          double px = pc[3 * s], py = pc[3 * s + 1], pz = pc[3 * s + 2];
          if (px > xlo && px <= xhi && py > ylo && py <= yhi && pz > zlo && pz <= zhi)
            sb->site[(size_t)in * ns + sb->nsite[in]++] = s;
        }
        total += sb->nsite[in];
      }
  free(pc);
  return total == ns ? 0 : -1; /* element.cpp:819 */
}

static void subele_free(subele_t *sb) {
  free(sb->spa_corner);
  free(sb->nsite);
  free(sb->site);
  free(sb->spa_center);
}

/* interpolate x = sum_k N_k * X'_k in node order (full_node_neighbor.cpp:141-159) */
static inline void interp3(const double *N, const double X[8][3], double out[3]) {
  double a = 0, b = 0, c = 0;
  for (int k = 0; k < 8; k++) {
    a += N[k] * X[k][0];
    b += N[k] * X[k][1];
    c += N[k] * X[k][2];
  }
  out[0] = a;
  out[1] = b;
  out[2] = c;
}

static inline void shifted_nodes(const double *nodex, int j, int img, const double prd[3], double X[8][3]) {
  int s[3];
  img_shift(img, s);
  for (int k = 0; k < 8; k++)
    for (int d = 0; d < 3; d++) X[k][d] = shifted(nodex[24 * j + 3 * k + d], s[d], prd[d]);
}

static inline void ip_position(const orc_elems *el, int i, int g, double out[3]) {
  double X[8][3];
  for (int k = 0; k < 8; k++)
    for (int d = 0; d < 3; d++) X[k][d] = el->nodex[24 * i + 3 * k + d];
  interp3(el->spa_int + 8 * g, X, out);
}

typedef struct {
  int j, img;
} eref_t;

/* Pruning radii.  The reference prunes with cut_ele = cn + max_size/2 and
 * cut_subele = cn + max_size/(2*nsplit) (neighbor.cpp:284-293), which are not
 * bounds once elements are distorted: with jittered nodes it misses sites at
 * the cutoff (tests/test_oracle_neighbors.py).  Fixed here: each element
 * (sub-element) prunes with cn + the largest distance from its center to its
 * own corners, a true bound since every site is a convex combination of the
 * corners (trilinear weights are >= 0 and sum to 1).  Centers are the
 * reference's (node mean; Spa_center_subele image of the sub-cube center,
 * which equals the mean of the sub-cube corner images). */
static double corner_radius(const double *Nc /* [8][8] */, const double X[8][3], const double c[3]) {
  double r2 = 0.0;
  for (int l = 0; l < 8; l++) {
    double p[3];
    interp3(Nc + 8 * l, X, p);
    double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
    double q = dx * dx + dy * dy + dz * dz;
    r2 = MAXD(r2, q);
  }
  return sqrt(r2) * (1.0 + 1e-12) + 1e-9;
}

static double node_radius(const double X[8][3], const double c[3]) {
  double r2 = 0.0;
  for (int k = 0; k < 8; k++) {
    double dx = X[k][0] - c[0], dy = X[k][1] - c[1], dz = X[k][2] - c[2];
    double q = dx * dx + dy * dy + dz * dz;
    r2 = MAXD(r2, q);
  }
  return sqrt(r2) * (1.0 + 1e-12) + 1e-9;
}

/* Sites of element image (j, img) within cutneigh of ip, reference order:
 * sub-elements, then their sites (full_node_neighbor.cpp:127-170,176-224). */
static int scan_element(const orc_elems *el, const subele_t *sb, int j, int img, int self, const double ip[3],
                        const double prd[3], double cutneigh, double Rj, const double cj[3], int n, int maxnip,
                        uint64_t *out) {
  double cutsq = cutneigh * cutneigh;
  double npdis = (ip[0] - cj[0]) * (ip[0] - cj[0]) + (ip[1] - cj[1]) * (ip[1] - cj[1]) +
                 (ip[2] - cj[2]) * (ip[2] - cj[2]);
  double ce = cutneigh + Rj;
  if (npdis > ce * ce) return n;
  double X[8][3];
  shifted_nodes(el->nodex, j, img, prd, X);
  for (int isub = 0; isub < sb->nsub; isub++) {
    double sc[3];
    interp3(sb->spa_center + 8 * isub, X, sc);
    double nps = (ip[0] - sc[0]) * (ip[0] - sc[0]) + (ip[1] - sc[1]) * (ip[1] - sc[1]) +
                 (ip[2] - sc[2]) * (ip[2] - sc[2]);
    double cs = cutneigh + corner_radius(sb->spa_corner + 64 * isub, X, sc);
    if (nps > cs * cs) continue;
    for (int mm = 0; mm < sb->nsite[isub]; mm++) {
      int site = sb->site[(size_t)isub * el->nsite + mm];
      double sx[3];
      interp3(el->spa_site + 8 * site, X, sx);
      double delx = sx[0] - ip[0], dely = sx[1] - ip[1], delz = sx[2] - ip[2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq <= cutsq && (!self || rsq > 1.0e-4)) {
        if (n < maxnip) out[n] = (uint64_t)j | ((uint64_t)img << 27) | ((uint64_t)site << 32);
        n++;
      }
    }
  }
  return n;
}

int orc_ip_neighbors(const orc_elems *el, const orc_box *box, double cutneigh, int nsplit,
                     int maxnip, int *numneigh, uint64_t *neigh) {
  int ne = el->ne, nint = el->nint;
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  subele_t sb;
  if (subele_build(&sb, el->nae, nsplit)) { subele_free(&sb); return -2; }
  double *xc = (double *)malloc(sizeof(double) * 3 * (ne > 0 ? ne : 1));
  double *R = (double *)malloc(sizeof(double) * (ne > 0 ? ne : 1));
  orc_element_centers(ne, el->nodex, xc);
  double Rmax = 0.0;
  for (int i = 0; i < ne; i++) {
    double X[8][3];
    shifted_nodes(el->nodex, i, ORC_IMG_SELF, prd, X);
    R[i] = node_radius(X, xc + 3 * i);
    Rmax = MAXD(Rmax, R[i]);
  }
  /* element list reach (reference: cut_all = cn + max_size): two elements can
   * share an in-range (integration point, site) pair only if their centers
   * are within cn + R_i + R_j */
  double cut_all = cutneigh + 2.0 * Rmax;
  cells_t c;
  cells_build(&c, ne, xc, box, cut_all);
  int overflow = 0;

#pragma omp parallel for schedule(dynamic, 4) reduction(| : overflow)
  for (int i = 0; i < ne; i++) {
    /* element_neighbor.cpp:37-66: every other element (local or ghost image)
     * whose center is in reach of i's center.  Complete search (the reference
     * bins elements with the atom stencil, neighbor.cpp:963-975, which cannot
     * reach cut_all). */
    int cap = 256, nel = 0;
    eref_t *elist = (eref_t *)malloc(sizeof(eref_t) * cap);
#define ELEM_TEST(j, img)                                                                  \
  do {                                                                                     \
    int s_[3];                                                                             \
    img_shift(img, s_);                                                                    \
    if ((j) == i && (img) == ORC_IMG_SELF) break;                                          \
    double dx = xc[3 * i + 0] - shifted(xc[3 * (j) + 0], s_[0], prd[0]);                   \
    double dy = xc[3 * i + 1] - shifted(xc[3 * (j) + 1], s_[1], prd[1]);                   \
    double dz = xc[3 * i + 2] - shifted(xc[3 * (j) + 2], s_[2], prd[2]);                   \
    double rsq = dx * dx + dy * dy + dz * dz;                                              \
    double reach = cutneigh + R[i] + R[(j)];                                               \
    if (rsq <= reach * reach) {                                                            \
      if (nel == cap) { cap *= 2; elist = (eref_t *)realloc(elist, sizeof(eref_t) * cap); } \
      elist[nel].j = (j);                                                                  \
      elist[nel].img = (img);                                                              \
      nel++;                                                                               \
    }                                                                                      \
  } while (0)
    if (c.ok) {
      int ci[3];
      cell_of(&c, xc + 3 * i, box, ci);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz}, s[3];
            for (int d = 0; d < 3; d++) {
              s[d] = 0;
              if (cc[d] < 0) { cc[d] += c.n[d]; s[d] = -1; }
              else if (cc[d] >= c.n[d]) { cc[d] -= c.n[d]; s[d] = 1; }
            }
            int img = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
            for (int j = c.head[(cc[2] * c.n[1] + cc[1]) * c.n[0] + cc[0]]; j >= 0; j = c.next[j])
              ELEM_TEST(j, img);
          }
    } else {
      for (int img = 0; img < 27; img++)
        for (int j = 0; j < ne; j++) ELEM_TEST(j, img);
    }
#undef ELEM_TEST
    for (int g = 0; g < nint; g++) {
      size_t p = (size_t)i * nint + g;
      uint64_t *out = neigh + p * maxnip;
      int n = 0;
      double ip[3];
      ip_position(el, i, g, ip);
      for (int l = 0; l < nel; l++) {
        int j = elist[l].j, img = elist[l].img, s[3];
        img_shift(img, s);
        double cj[3] = {shifted(xc[3 * j + 0], s[0], prd[0]), shifted(xc[3 * j + 1], s[1], prd[1]),
                        shifted(xc[3 * j + 2], s[2], prd[2])};
        n = scan_element(el, &sb, j, img, 0, ip, prd, cutneigh, R[j], cj, n, maxnip, out);
      }
      /* self element last, full_node_neighbor.cpp:176-224: rsq > 1e-4 drops
       * the integration point's own site */
      n = scan_element(el, &sb, i, ORC_IMG_SELF, 1, ip, prd, cutneigh, R[i], xc + 3 * i, n, maxnip, out);
      if (n > maxnip) overflow = 1;
      numneigh[p] = n;
      if (n <= maxnip) qsort(out, n, sizeof(uint64_t), cmp_u64);
    }
    free(elist);
  }
  cells_free(&c);
  free(xc);
  free(R);
  subele_free(&sb);
  return overflow ? -1 : 0;
}

/* Brute force: every site of every element image within cutneigh of the
 * integration point, the own site of the own element excluded by rsq>1e-4.
 * Independent of any pruning; pins orc_ip_neighbors. */
int orc_ip_neighbors_brute(const orc_elems *el, const orc_box *box, double cutneigh,
                           int maxnip, int *numneigh, uint64_t *neigh) {
  int ne = el->ne, nint = el->nint;
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  double cutsq = cutneigh * cutneigh;
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : overflow)
  for (int p = 0; p < ne * nint; p++) {
    int i = p / nint, g = p % nint, n = 0;
    uint64_t *out = neigh + (size_t)p * maxnip;
    double ip[3];
    ip_position(el, i, g, ip);
    for (int img = 0; img < 27; img++)
      for (int j = 0; j < ne; j++) {
        double X[8][3];
        shifted_nodes(el->nodex, j, img, prd, X);
        for (int site = 0; site < el->nsite; site++) {
          double sx[3];
          interp3(el->spa_site + 8 * site, X, sx);
          double delx = sx[0] - ip[0], dely = sx[1] - ip[1], delz = sx[2] - ip[2];
          double rsq = delx * delx + dely * dely + delz * delz;
          if (rsq > cutsq) continue;
          if (j == i && img == ORC_IMG_SELF && !(rsq > 1.0e-4)) continue;
          if (n < maxnip) out[n] = (uint64_t)j | ((uint64_t)img << 27) | ((uint64_t)site << 32);
          n++;
        }
      }
    if (n > maxnip) overflow = 1;
    numneigh[p] = n;
    if (n <= maxnip) qsort(out, n, sizeof(uint64_t), cmp_u64);
  }
  return overflow ? -1 : 0;
}

/* ------------------------------------------------------------------------
 * EAM, atoms: PairEAM::compute, pair_eam.cpp:584-746.  rho pass
 * (:629-658), fp pass (:670-680), fp "forward comm" (ghost fp = owner fp,
 * :683), force pass without Newton's third law (:688-745).
 * Energy and virial are not computed by the reference (eflag forced 0,
 * pair_eam.cpp:594); they are defined here as
 *   e_i = F(rho_i) + 1/2 sum_j phi(r_ij),  v_i = 1/2 sum_j del (x) del * fpair
 * with v in the order xx yy zz xy xz yz.
 * ---------------------------------------------------------------------- */
void orc_eam_atoms(int n, const double *x, const orc_box *box,
                   const int *numneigh, const uint32_t *neigh, int maxneigh,
                   const orc_eam *t, double *rho, double *fp, double *f,
                   double *eatom, double *vatom) {
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
    double r_i = 0.0;
    for (int jj = 0; jj < numneigh[i]; jj++) {
      uint32_t e = neigh[(size_t)i * maxneigh + jj];
      int j = e & 0x7FFFFFF, s[3];
      img_shift(e >> 27, s);
      double delx = xtmp - shifted(x[3 * j], s[0], prd[0]);
      double dely = ytmp - shifted(x[3 * j + 1], s[1], prd[1]);
      double delz = ztmp - shifted(x[3 * j + 2], s[2], prd[2]);
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < t->cutforcesq) {
        double p;
        const double *c = rtab(t->rhor_spline, sqrt(rsq), t->rdr, t->nr, &p);
        r_i += ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
      }
    }
    rho[i] = r_i;
    fp[i] = embed_fp(t, r_i);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
    double fx = 0.0, fy = 0.0, fz = 0.0, epair = 0.0;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = 0; jj < numneigh[i]; jj++) {
      uint32_t e = neigh[(size_t)i * maxneigh + jj];
      int j = e & 0x7FFFFFF, s[3];
      img_shift(e >> 27, s);
      double delx = xtmp - shifted(x[3 * j], s[0], prd[0]);
      double dely = ytmp - shifted(x[3 * j + 1], s[1], prd[1]);
      double delz = ztmp - shifted(x[3 * j + 2], s[2], prd[2]);
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < t->cutforcesq) {
        double r = sqrt(rsq), p;
        const double *c = rtab(t->rhor_spline, r, t->rdr, t->nr, &p);
        double rhoip = (c[0] * p + c[1]) * p + c[2];
        double rhojp = rhoip; /* one element type: type2rhor[i][j] == type2rhor[j][i] */
        const double *z = t->z2r_spline + (c - t->rhor_spline);
        double z2p = (z[0] * p + z[1]) * p + z[2];
        double z2 = ((z[3] * p + z[4]) * p + z[5]) * p + z[6];
        double recip = 1.0 / r;
        double phi = z2 * recip;
        double phip = z2p * recip - phi * recip;
        double psip = fp[i] * rhojp + fp[j] * rhoip + phip;
        double fpair = -psip * recip;
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        epair += 0.5 * phi;
        v[0] += 0.5 * delx * delx * fpair;
        v[1] += 0.5 * dely * dely * fpair;
        v[2] += 0.5 * delz * delz * fpair;
        v[3] += 0.5 * delx * dely * fpair;
        v[4] += 0.5 * delx * delz * fpair;
        v[5] += 0.5 * dely * delz * fpair;
      }
    }
    f[3 * i] = fx;
    f[3 * i + 1] = fy;
    f[3 * i + 2] = fz;
    if (eatom) eatom[i] = embed_energy(t, rho[i], fp[i]) + epair;
    if (vatom)
      for (int a = 0; a < 6; a++) vatom[6 * i + a] = v[a];
  }
}

/* ------------------------------------------------------------------------
 * EAM, elements: PairEAM::ecompute, pair_eam.cpp:927-1137 (the debug block
 * :826-925 is not reproduced).
 *   1. erho[e][k] = sum over the list of the integration point on node k
 *      of rho(|X_{e,k} - x_site|), rsq < cutforcesq (:927-986), using the
 *      node coordinate itself (:955-957).
 *   2. ghost erho = owner erho (eforward_comm_pair, :991).
 *   3. per integration point: x_g, rho_g interpolated, F'(rho_g); per
 *      neighbor site x_s, rho_s interpolated, F'(rho_s), rho', z2, z2';
 *      psip = (F'_g + F'_s) rho' + phi'; f_g += del * (-psip/r) (:1013-1129).
 *   4. nodef[e][k] += w_g*N_k(g)*f_g in integration point order (:1130-1134).
 * ipf receives f_g; eip/vip receive the integration-point energy
 * w_g*(F(rho_g) + 1/2 sum phi) and virial w_g*1/2 sum del(x)del*fpair
 * (build-defined, not in the reference).
 * ---------------------------------------------------------------------- */
static inline void site_of(const orc_elems *el, uint64_t e, const double prd[3], int *j, int *site,
                           double X[8][3]) {
  *j = (int)(e & 0x7FFFFFF);
  int img = (int)((e >> 27) & 31);
  *site = (int)(e >> 32);
  shifted_nodes(el->nodex, *j, img, prd, X);
}

void orc_eam_elements(const orc_elems *el, const orc_box *box,
                      const int *numneigh, const uint64_t *neigh, int maxnip,
                      const orc_eam *t, double *erho, double *nodef, double *ipf,
                      double *eip, double *vip) {
  int ne = el->ne, nint = el->nint;
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ne; i++) {
    for (int k = 0; k < 8; k++) {
      double ixtmp = el->nodex[24 * i + 3 * k], iytmp = el->nodex[24 * i + 3 * k + 1],
             iztmp = el->nodex[24 * i + 3 * k + 2];
      size_t nt = (size_t)i * nint + el->node_list[k];
      double acc = 0.0;
      for (int ll = 0; ll < numneigh[nt]; ll++) {
        int j, site;
        double X[8][3], s[3];
        site_of(el, neigh[nt * maxnip + ll], prd, &j, &site, X);
        interp3(el->spa_site + 8 * site, X, s);
        double delx = ixtmp - s[0], dely = iytmp - s[1], delz = iztmp - s[2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t->cutforcesq) {
          double p;
          const double *c = rtab(t->rhor_spline, sqrt(rsq), t->rdr, t->nr, &p);
          acc += ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
        }
      }
      erho[8 * i + k] = acc;
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ne; i++) {
    double nf[8][3];
    for (int k = 0; k < 8; k++) nf[k][0] = nf[k][1] = nf[k][2] = 0.0;
    for (int g = 0; g < nint; g++) {
      const double *Ng = el->spa_int + 8 * g;
      double ixtmp = 0, iytmp = 0, iztmp = 0, rhoi = 0;
      for (int kk = 0; kk < 8; kk++) {
        ixtmp += Ng[kk] * el->nodex[24 * i + 3 * kk];
        iytmp += Ng[kk] * el->nodex[24 * i + 3 * kk + 1];
        iztmp += Ng[kk] * el->nodex[24 * i + 3 * kk + 2];
        rhoi += Ng[kk] * erho[8 * i + kk];
      }
      double fpi = embed_fp(t, rhoi);
      double fx = 0, fy = 0, fz = 0, epair = 0;
      double v[6] = {0, 0, 0, 0, 0, 0};
      size_t p = (size_t)i * nint + g;
      for (int ll = 0; ll < numneigh[p]; ll++) {
        int j, site;
        double X[8][3];
        site_of(el, neigh[p * maxnip + ll], prd, &j, &site, X);
        const double *Ns = el->spa_site + 8 * site;
        double jx = 0, jy = 0, jz = 0, rhoj = 0;
        for (int kk = 0; kk < 8; kk++) {
          jx += Ns[kk] * X[kk][0];
          jy += Ns[kk] * X[kk][1];
          jz += Ns[kk] * X[kk][2];
          rhoj += Ns[kk] * erho[8 * j + kk];
        }
        double delx = ixtmp - jx, dely = iytmp - jy, delz = iztmp - jz;
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t->cutforcesq) {
          double fpj = embed_fp(t, rhoj);
          double r = sqrt(rsq), pp;
          const double *c = rtab(t->rhor_spline, r, t->rdr, t->nr, &pp);
          double rhoip = (c[0] * pp + c[1]) * pp + c[2];
          const double *z = t->z2r_spline + (c - t->rhor_spline);
          double z2p = (z[0] * pp + z[1]) * pp + z[2];
          double z2 = ((z[3] * pp + z[4]) * pp + z[5]) * pp + z[6];
          double recip = 1.0 / r;
          double phi = z2 * recip;
          double phip = z2p * recip - phi * recip;
          double psip = (fpi + fpj) * rhoip + phip;
          double fpair = -psip * recip;
          fx += delx * fpair;
          fy += dely * fpair;
          fz += delz * fpair;
          epair += 0.5 * phi;
          v[0] += 0.5 * delx * delx * fpair;
          v[1] += 0.5 * dely * dely * fpair;
          v[2] += 0.5 * delz * delz * fpair;
          v[3] += 0.5 * delx * dely * fpair;
          v[4] += 0.5 * delx * delz * fpair;
          v[5] += 0.5 * dely * delz * fpair;
        }
      }
      double w = el->weight[g];
      for (int kk = 0; kk < 8; kk++) {
        nf[kk][0] += w * Ng[kk] * fx;
        nf[kk][1] += w * Ng[kk] * fy;
        nf[kk][2] += w * Ng[kk] * fz;
      }
      if (ipf) { ipf[3 * p] = fx; ipf[3 * p + 1] = fy; ipf[3 * p + 2] = fz; }
      if (eip) eip[p] = w * (embed_energy(t, rhoi, fpi) + epair);
      if (vip)
        for (int a = 0; a < 6; a++) vip[6 * p + a] = w * v[a];
    }
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) nodef[24 * i + 3 * k + d] = nf[k][d];
  }
}

/* ------------------------------------------------------------------------
 * LJ: PairLJCut::compute, pair_lj_cut.cpp:227-288, and ::ecompute,
 * pair_lj_cut.cpp:295-442 (the element loop indexes elements by ii, equal to
 * enlist[ii] since element_neighbor.cpp:67 stores them in order).
 * Energy e = 1/2 sum r6inv*(lj3*r6inv - lj4) (no offset), build-defined.
 * ---------------------------------------------------------------------- */
void orc_lj_atoms(int n, const double *x, const orc_box *box,
                  const int *numneigh, const uint32_t *neigh, int maxneigh,
                  const orc_lj *pp, double *f, double *eatom, double *vatom) {
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
    double fx = 0, fy = 0, fz = 0, e = 0;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = 0; jj < numneigh[i]; jj++) {
      uint32_t en = neigh[(size_t)i * maxneigh + jj];
      int j = en & 0x7FFFFFF, s[3];
      img_shift(en >> 27, s);
      double delx = xtmp - shifted(x[3 * j], s[0], prd[0]);
      double dely = ytmp - shifted(x[3 * j + 1], s[1], prd[1]);
      double delz = ztmp - shifted(x[3 * j + 2], s[2], prd[2]);
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < pp->cutsq) {
        double r2inv = 1.0 / rsq;
        double r6inv = r2inv * r2inv * r2inv;
        double forcelj = r6inv * (pp->lj1 * r6inv - pp->lj2);
        double fpair = 1.0 * forcelj * r2inv; /* factor_lj = special_lj[0] = 1 */
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        e += 0.5 * (r6inv * (pp->lj3 * r6inv - pp->lj4));
        v[0] += 0.5 * delx * delx * fpair;
        v[1] += 0.5 * dely * dely * fpair;
        v[2] += 0.5 * delz * delz * fpair;
        v[3] += 0.5 * delx * dely * fpair;
        v[4] += 0.5 * delx * delz * fpair;
        v[5] += 0.5 * dely * delz * fpair;
      }
    }
    f[3 * i] = fx;
    f[3 * i + 1] = fy;
    f[3 * i + 2] = fz;
    if (eatom) eatom[i] = e;
    if (vatom)
      for (int a = 0; a < 6; a++) vatom[6 * i + a] = v[a];
  }
}

void orc_lj_elements(const orc_elems *el, const orc_box *box,
                     const int *numneigh, const uint64_t *neigh, int maxnip,
                     const orc_lj *pp, double *nodef, double *ipf, double *eip, double *vip) {
  int ne = el->ne, nint = el->nint;
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ne; i++) {
    double nf[8][3];
    for (int k = 0; k < 8; k++) nf[k][0] = nf[k][1] = nf[k][2] = 0.0;
    for (int g = 0; g < nint; g++) {
      const double *Ng = el->spa_int + 8 * g;
      double ixtmp = 0, iytmp = 0, iztmp = 0;
      for (int kk = 0; kk < 8; kk++) {
        ixtmp += Ng[kk] * el->nodex[24 * i + 3 * kk];
        iytmp += Ng[kk] * el->nodex[24 * i + 3 * kk + 1];
        iztmp += Ng[kk] * el->nodex[24 * i + 3 * kk + 2];
      }
      double fx = 0, fy = 0, fz = 0, e = 0;
      double v[6] = {0, 0, 0, 0, 0, 0};
      size_t p = (size_t)i * nint + g;
      for (int ll = 0; ll < numneigh[p]; ll++) {
        int j, site;
        double X[8][3], s[3];
        site_of(el, neigh[p * maxnip + ll], prd, &j, &site, X);
        interp3(el->spa_site + 8 * site, X, s);
        double delx = ixtmp - s[0], dely = iytmp - s[1], delz = iztmp - s[2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < pp->cutsq) {
          double r2inv = 1.0 / rsq;
          double r6inv = r2inv * r2inv * r2inv;
          double forcelj = r6inv * (pp->lj1 * r6inv - pp->lj2);
          double fpair = forcelj * r2inv;
          fx += delx * fpair;
          fy += dely * fpair;
          fz += delz * fpair;
          e += 0.5 * (r6inv * (pp->lj3 * r6inv - pp->lj4));
          v[0] += 0.5 * delx * delx * fpair;
          v[1] += 0.5 * dely * dely * fpair;
          v[2] += 0.5 * delz * delz * fpair;
          v[3] += 0.5 * delx * dely * fpair;
          v[4] += 0.5 * delx * delz * fpair;
          v[5] += 0.5 * dely * delz * fpair;
        }
      }
      double w = el->weight[g];
      for (int kk = 0; kk < 8; kk++) {
        nf[kk][0] += w * Ng[kk] * fx;
        nf[kk][1] += w * Ng[kk] * fy;
        nf[kk][2] += w * Ng[kk] * fz;
      }
      if (ipf) { ipf[3 * p] = fx; ipf[3 * p + 1] = fy; ipf[3 * p + 2] = fz; }
      if (eip) eip[p] = w * e;
      if (vip)
        for (int a = 0; a < 6; a++) vip[6 * p + a] = w * v[a];
    }
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) nodef[24 * i + 3 * k + d] = nf[k][d];
  }
}

/* ------------------------------------------------------------------------
 * Atom-element coupling (not in the reference, where atoms see only atoms
 * and integration points only element sites: pair_eam.cpp:584-746 vs
 * 927-1137).  Build-defined extension in the same force-based
 * scheme: an atom's list also holds the interpolated element sites within
 * cutneigh, an integration point's list also holds the atoms within
 * cutneigh, and every pair term is the reference's own (atom pass formula
 * for atoms, element pass formula for integration points).  With nae = 2
 * and the 8 corners as integration points of weight 1 the coupled system is
 * molecular dynamics of all sites (the invariant stated at
 * pair_eam.cpp:996), which is how tests/test_oracle_coupled.py pins it.
 *
 * Entry encoding (uint64): (j | img << 27) | ((uint64)site << 32); site ==
 * ORC_ATOM_SITE marks an atom j, otherwise j is an element and site one of
 * its nae^3 sites.
 * ---------------------------------------------------------------------- */
static inline void entry_pos(const double *x, const orc_elems *el, uint64_t e, const double prd[3], double out[3]) {
  int j = (int)(e & 0x7FFFFFF), img = (int)((e >> 27) & 31), s[3];
  uint32_t site = (uint32_t)(e >> 32);
  if (site == ORC_ATOM_SITE) {
    img_shift(img, s);
    for (int d = 0; d < 3; d++) out[d] = shifted(x[3 * j + d], s[d], prd[d]);
  } else {
    double X[8][3];
    shifted_nodes(el->nodex, j, img, prd, X);
    interp3(el->spa_site + 8 * site, X, out);
  }
}

/* ------------------------------------------------------------------------
 * Centro-symmetry parameter, Output::centro_atom (output.cpp:352-441): the
 * list entries with rsq < cutneighmaxsq (:374, :398-410), 0 with fewer than
 * nnn (:414-417), the nnn nearest (select2, :421), the nnn(nnn-1)/2 values
 * |x_jj + x_kk - 2 x_i|^2 (:426-434) and the sum of the nnn/2 smallest
 * (select, :438-442).  The reference's quickselect leaves the order of the
 * selected values, and the choice among entries tied at the nnn-th distance,
 * to its partitioning; here both are fixed (full sorts, ties by key), so the
 * value depends on the neighbor set alone.
 * ---------------------------------------------------------------------- */
typedef struct {
  double rsq;
  uint64_t key;
  uint64_t e;
} csp_nb_t;

static int cmp_csp_nb(const void *a, const void *b) {
  const csp_nb_t *p = (const csp_nb_t *)a, *q = (const csp_nb_t *)b;
  if (p->rsq != q->rsq) return p->rsq < q->rsq ? -1 : 1;
  return p->key < q->key ? -1 : p->key > q->key;
}

static int cmp_f64(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : x > y;
}

void orc_centro_atoms(int n, const double *x, const orc_elems *el, const orc_box *box, const int *anum,
                      const uint64_t *anb, int maxneigh, double cutneigh, int nnn, double *centro) {
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  double cutsq = cutneigh * cutneigh;
  int nhalf = nnn / 2, npairs = nnn * (nnn - 1) / 2;
#pragma omp parallel
  {
    csp_nb_t *nb = (csp_nb_t *)malloc(sizeof(csp_nb_t) * (size_t)(maxneigh > 0 ? maxneigh : 1));
    double *pairs = (double *)malloc(sizeof(double) * (size_t)(npairs > 0 ? npairs : 1));
#pragma omp for schedule(static)
    for (int i = 0; i < n; i++) {
      double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
      int cnt = 0;
      for (int jj = 0; jj < anum[i]; jj++) {
        uint64_t e = anb[(size_t)i * maxneigh + jj];
        double s[3];
        entry_pos(x, el, e, prd, s);
        double delx = xtmp - s[0], dely = ytmp - s[1], delz = ztmp - s[2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < cutsq) {
          nb[cnt].rsq = rsq;
          nb[cnt].key = (uint32_t)(e >> 32) == ORC_ATOM_SITE ? (e & 0xFFFFFFFFu) : ((1ull << 63) | e);
          nb[cnt].e = e;
          cnt++;
        }
      }
      if (cnt < nnn) {
        centro[i] = 0.0;
        continue;
      }
      qsort(nb, cnt, sizeof(csp_nb_t), cmp_csp_nb);
      double p[16][3];
      for (int a = 0; a < nnn; a++) entry_pos(x, el, nb[a].e, prd, p[a]);
      int m = 0;
      for (int j = 0; j < nnn; j++)
        for (int k = j + 1; k < nnn; k++) {
          double delx = p[j][0] + p[k][0] - 2.0 * xtmp;
          double dely = p[j][1] + p[k][1] - 2.0 * ytmp;
          double delz = p[j][2] + p[k][2] - 2.0 * ztmp;
          pairs[m++] = delx * delx + dely * dely + delz * delz;
        }
      qsort(pairs, npairs, sizeof(double), cmp_f64);
      double value = 0.0;
      for (int h = 0; h < nhalf; h++) value += pairs[h];
      centro[i] = value;
    }
    free(nb);
    free(pairs);
  }
}

/* F'(rho) of an entry: the atom's fp, or F' of the interpolated site
 * density sum_k N_k erho_k (pair_eam.cpp:1077, 1089-1096). */
static inline double entry_fp(const orc_elems *el, const orc_eam *t, const double *fp, const double *erho,
                              uint64_t e) {
  int j = (int)(e & 0x7FFFFFF);
  uint32_t site = (uint32_t)(e >> 32);
  if (site == ORC_ATOM_SITE) return fp[j];
  const double *Ns = el->spa_site + 8 * site;
  double r = 0.0;
  for (int kk = 0; kk < 8; kk++) r += Ns[kk] * erho[8 * j + kk];
  return embed_fp(t, r);
}

int orc_coupled_neighbors(int n, const double *x, const orc_elems *el, const orc_box *box, double cutneigh,
                          int nsplit, int brute, int maxneigh, int *anum, uint64_t *anb, int maxnip, int *inum,
                          uint64_t *inb) {
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  double cutsq = cutneigh * cutneigh;
  int ne = el->ne, nint = el->nint, nip = ne * nint, rc = 0;
  /* atom-atom and site-site parts: the reference lists */
  uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1) * maxneigh);
  if (n > 0) {
    rc = brute ? orc_atom_neighbors_brute(n, x, box, cutneigh, maxneigh, anum, tmp)
               : orc_atom_neighbors(n, x, box, cutneigh, maxneigh, anum, tmp);
    if (rc) { free(tmp); return rc; }
  }
  for (int i = 0; i < n; i++)
    for (int k = 0; k < anum[i]; k++)
      anb[(size_t)i * maxneigh + k] = (uint64_t)tmp[(size_t)i * maxneigh + k] | ((uint64_t)ORC_ATOM_SITE << 32);
  free(tmp);
  if (nip > 0) {
    rc = brute ? orc_ip_neighbors_brute(el, box, cutneigh, maxnip, inum, inb)
               : orc_ip_neighbors(el, box, cutneigh, nsplit, maxnip, inum, inb);
    if (rc) return rc;
  }
  if (n == 0 || ne == 0) return 0;

  /* element geometry: centers and corner radii (as orc_ip_neighbors) */
  subele_t sb;
  if (subele_build(&sb, el->nae, nsplit)) { subele_free(&sb); return -2; }
  double *xc = (double *)malloc(sizeof(double) * 3 * ne);
  double *R = (double *)malloc(sizeof(double) * ne);
  orc_element_centers(ne, el->nodex, xc);
  double Rmax = 0.0;
  for (int i = 0; i < ne; i++) {
    double X[8][3];
    shifted_nodes(el->nodex, i, ORC_IMG_SELF, prd, X);
    R[i] = node_radius(X, xc + 3 * i);
    Rmax = MAXD(Rmax, R[i]);
  }
  cells_t ec, ac;
  cells_build(&ec, ne, xc, box, cutneigh + 2.0 * Rmax);
  cells_build(&ac, n, x, box, cutneigh);
  int overflow = 0;

  /* sites for every atom: element images whose center is within
   * cutneigh + R_j, then their sub-elements and sites (scan_element) */
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow)
  for (int i = 0; i < n; i++) {
    uint64_t *out = anb + (size_t)i * maxneigh;
    int cnt = anum[i];
    const double *xi = x + 3 * i;
    if (brute) {
      for (int img = 0; img < 27; img++)
        for (int j = 0; j < ne; j++) {
          double X[8][3];
          shifted_nodes(el->nodex, j, img, prd, X);
          for (int site = 0; site < el->nsite; site++) {
            double sx[3];
            interp3(el->spa_site + 8 * site, X, sx);
            double delx = sx[0] - xi[0], dely = sx[1] - xi[1], delz = sx[2] - xi[2];
            double rsq = delx * delx + dely * dely + delz * delz;
            if (rsq <= cutsq) {
              if (cnt < maxneigh) out[cnt] = (uint64_t)j | ((uint64_t)img << 27) | ((uint64_t)site << 32);
              cnt++;
            }
          }
        }
    } else {
#define SCAN_ATOM(j, img)                                                                             \
  do {                                                                                                \
    int s_[3];                                                                                        \
    img_shift(img, s_);                                                                               \
    double cj[3] = {shifted(xc[3 * (j)], s_[0], prd[0]), shifted(xc[3 * (j) + 1], s_[1], prd[1]),    \
                    shifted(xc[3 * (j) + 2], s_[2], prd[2])};                                         \
    cnt = scan_element(el, &sb, (j), (img), 0, xi, prd, cutneigh, R[(j)], cj, cnt, maxneigh, out);    \
  } while (0)
      if (ec.ok) {
        int ci[3];
        cell_of(&ec, xi, box, ci);
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz}, s[3];
              for (int d = 0; d < 3; d++) {
                s[d] = 0;
                if (cc[d] < 0) { cc[d] += ec.n[d]; s[d] = -1; }
                else if (cc[d] >= ec.n[d]) { cc[d] -= ec.n[d]; s[d] = 1; }
              }
              int img = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
              for (int j = ec.head[(cc[2] * ec.n[1] + cc[1]) * ec.n[0] + cc[0]]; j >= 0; j = ec.next[j])
                SCAN_ATOM(j, img);
            }
      } else {
        for (int img = 0; img < 27; img++)
          for (int j = 0; j < ne; j++) SCAN_ATOM(j, img);
      }
#undef SCAN_ATOM
    }
    if (cnt > maxneigh) overflow = 1;
    anum[i] = cnt;
    if (cnt <= maxneigh) qsort(out, cnt, sizeof(uint64_t), cmp_u64);
  }

  /* atoms for every integration point */
#pragma omp parallel for schedule(dynamic, 16) reduction(| : overflow)
  for (int p = 0; p < nip; p++) {
    uint64_t *out = inb + (size_t)p * maxnip;
    int cnt = inum[p];
    double ip[3], q[3];
    int t[3];
    ip_position(el, p / nint, p % nint, ip);
    /* an integration point can lie outside the box (element nodes follow
     * the element center, domain.cpp:647-745): search the cells around its
     * wrapped position q = ip - t*prd; an atom found with cell image s is
     * the image s + t of the integration point */
    for (int d = 0; d < 3; d++) {
      t[d] = ip[d] < box->lo[d] ? -1 : (ip[d] >= box->hi[d] ? 1 : 0);
      q[d] = ip[d] - (double)t[d] * prd[d];
    }
#define TEST_ATOM(j, img)                                                                          \
  do {                                                                                             \
    int s_[3];                                                                                     \
    img_shift(img, s_);                                                                            \
    double delx = shifted(x[3 * (j)], s_[0], prd[0]) - ip[0];                                      \
    double dely = shifted(x[3 * (j) + 1], s_[1], prd[1]) - ip[1];                                  \
    double delz = shifted(x[3 * (j) + 2], s_[2], prd[2]) - ip[2];                                  \
    double rsq = delx * delx + dely * dely + delz * delz;                                          \
    if (rsq <= cutsq) {                                                                            \
      if (cnt < maxnip) out[cnt] = (uint64_t)(j) | ((uint64_t)(img) << 27) | ((uint64_t)ORC_ATOM_SITE << 32); \
      cnt++;                                                                                       \
    }                                                                                              \
  } while (0)
    if (ac.ok && !brute) {
      int ci[3];
      cell_of(&ac, q, box, ci);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz}, s[3], far = 0;
            for (int d = 0; d < 3; d++) {
              s[d] = t[d];
              if (cc[d] < 0) { cc[d] += ac.n[d]; s[d] -= 1; }
              else if (cc[d] >= ac.n[d]) { cc[d] -= ac.n[d]; s[d] += 1; }
              far |= s[d] < -1 || s[d] > 1;
            }
            if (far) continue; /* beyond one period: out of reach when prd > 2*(cutneigh + R) */
            int img = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
            for (int j = ac.head[(cc[2] * ac.n[1] + cc[1]) * ac.n[0] + cc[0]]; j >= 0; j = ac.next[j])
              TEST_ATOM(j, img);
          }
    } else {
      for (int img = 0; img < 27; img++)
        for (int j = 0; j < n; j++) TEST_ATOM(j, img);
    }
#undef TEST_ATOM
    if (cnt > maxnip) overflow = 1;
    inum[p] = cnt;
    if (cnt <= maxnip) qsort(out, cnt, sizeof(uint64_t), cmp_u64);
  }
  cells_free(&ec);
  cells_free(&ac);
  free(xc);
  free(R);
  subele_free(&sb);
  return overflow ? -1 : 0;
}

/* Reactions of the atom-site pair forces (build-defined coupling).  Every
 * (atom i, site s of element j at image img) pair that exerts a force f_is
 * on the atom puts -f_is on the site, and the site passes it to its nodes
 * through its shape functions: nodef[j][k] += N_k(xi_s) * R_s, with
 * R_s = sum of -f_is over the atoms in ascending entry order (atom | img<<27),
 * sites of an element in ascending index.  Since sum_k N_k = 1 the atoms
 * and the nodes receive equal and opposite totals for every such pair.
 * fsite holds f_is of atom list slot (i, jj) where has[i*maxneigh + jj]. */
typedef struct {
  uint64_t skey;  /* element << 32 | site */
  uint32_t akey;  /* atom | img << 27 */
  size_t slot;    /* i * maxneigh + jj */
} react_t;

static int cmp_react(const void *a, const void *b) {
  const react_t *p = (const react_t *)a, *q = (const react_t *)b;
  if (p->skey != q->skey) return p->skey < q->skey ? -1 : 1;
  return p->akey < q->akey ? -1 : p->akey > q->akey;
}

static void add_site_reactions(int n, const orc_elems *el, const int *anum, const uint64_t *anb, int maxneigh,
                               const double *fsite, const unsigned char *has, double *nodef) {
  size_t np = 0;
  for (int i = 0; i < n; i++)
    for (int jj = 0; jj < anum[i]; jj++) np += has[(size_t)i * maxneigh + jj];
  if (np == 0) return;
  react_t *r = (react_t *)malloc(sizeof(react_t) * np);
  size_t q = 0;
  for (int i = 0; i < n; i++)
    for (int jj = 0; jj < anum[i]; jj++) {
      size_t slot = (size_t)i * maxneigh + jj;
      if (!has[slot]) continue;
      uint64_t e = anb[slot];
      r[q].skey = ((e & 0x7FFFFFFull) << 32) | (e >> 32);
      r[q].akey = (uint32_t)i | (uint32_t)(((e >> 27) & 31) << 27);
      r[q].slot = slot;
      q++;
    }
  qsort(r, np, sizeof(react_t), cmp_react);
  for (size_t a = 0; a < np;) {
    size_t b = a;
    double R[3] = {0.0, 0.0, 0.0};
    while (b < np && r[b].skey == r[a].skey) {
      for (int d = 0; d < 3; d++) R[d] -= fsite[3 * r[b].slot + d];
      b++;
    }
    int j = (int)(r[a].skey >> 32), site = (int)(r[a].skey & 0xFFFFFFFFu);
    const double *Ns = el->spa_site + 8 * site;
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) nodef[24 * j + 3 * k + d] += Ns[k] * R[d];
    a = b;
  }
  free(r);
}

/* EAM over atoms and elements together (the passes of pair_eam.cpp:584-746
 * and 927-1137 on the coupled lists):
 *   1. erho[e][k]: the list of the integration point at node k, from the
 *      node coordinate (sites interpolated, atoms as they are);
 *   2. rho_i, fp_i of atoms over their lists (atoms and sites);
 *   3. atom force: psip = fp_i*rho' + fp_j*rho' + phi', fp_j of a site
 *      F'(sum_k N_k erho_k);
 *   4. integration-point force and node scatter as orc_eam_elements over the
 *      site entries of the lists (atoms count in the density of step 1 only);
 *   5. the reactions of step 3's atom-site forces on the sites' nodes
 *      (add_site_reactions).
 * Energy / virial tallies as in orc_eam_atoms / orc_eam_elements, except that
 * an atom-site pair puts its whole pair energy phi and virial on the atom
 * (the site's half is not in the integration-point sums). */
void orc_eam_coupled(int n, const double *x, const orc_elems *el, const orc_box *box, const int *anum,
                     const uint64_t *anb, int maxneigh, const int *inum, const uint64_t *inb, int maxnip,
                     const orc_eam *t, double *rho, double *fp, double *f, double *eatom, double *vatom,
                     double *erho, double *nodef, double *ipf, double *eip, double *vip) {
  int ne = el->ne, nint = el->nint;
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  double *fsite = NULL;
  unsigned char *has = NULL;
  if (n > 0 && maxneigh > 0) {
    fsite = (double *)malloc(sizeof(double) * 3 * (size_t)n * maxneigh);
    has = (unsigned char *)calloc((size_t)n * maxneigh, 1);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ne; i++)
    for (int k = 0; k < 8; k++) {
      const double *nk = el->nodex + 24 * i + 3 * k;
      size_t nt = (size_t)i * nint + el->node_list[k];
      double acc = 0.0;
      for (int ll = 0; ll < inum[nt]; ll++) {
        double s[3];
        entry_pos(x, el, inb[nt * maxnip + ll], prd, s);
        double delx = nk[0] - s[0], dely = nk[1] - s[1], delz = nk[2] - s[2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t->cutforcesq) {
          double p;
          const double *c = rtab(t->rhor_spline, sqrt(rsq), t->rdr, t->nr, &p);
          acc += ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
        }
      }
      erho[8 * i + k] = acc;
    }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const double *xi = x + 3 * i;
    double r_i = 0.0;
    for (int jj = 0; jj < anum[i]; jj++) {
      double s[3];
      entry_pos(x, el, anb[(size_t)i * maxneigh + jj], prd, s);
      double delx = xi[0] - s[0], dely = xi[1] - s[1], delz = xi[2] - s[2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < t->cutforcesq) {
        double p;
        const double *c = rtab(t->rhor_spline, sqrt(rsq), t->rdr, t->nr, &p);
        r_i += ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
      }
    }
    rho[i] = r_i;
    fp[i] = embed_fp(t, r_i);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const double *xi = x + 3 * i;
    double fx = 0.0, fy = 0.0, fz = 0.0, epair = 0.0;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = 0; jj < anum[i]; jj++) {
      uint64_t e = anb[(size_t)i * maxneigh + jj];
      double s[3];
      entry_pos(x, el, e, prd, s);
      double delx = xi[0] - s[0], dely = xi[1] - s[1], delz = xi[2] - s[2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < t->cutforcesq) {
        int site_e = (uint32_t)(e >> 32) != ORC_ATOM_SITE;
        double share = site_e ? 1.0 : 0.5;
        double fpj = entry_fp(el, t, fp, erho, e);
        double r = sqrt(rsq), p;
        const double *c = rtab(t->rhor_spline, r, t->rdr, t->nr, &p);
        double rhoip = (c[0] * p + c[1]) * p + c[2];
        const double *z = t->z2r_spline + (c - t->rhor_spline);
        double z2p = (z[0] * p + z[1]) * p + z[2];
        double z2 = ((z[3] * p + z[4]) * p + z[5]) * p + z[6];
        double recip = 1.0 / r;
        double phi = z2 * recip;
        double phip = z2p * recip - phi * recip;
        double psip = fp[i] * rhoip + fpj * rhoip + phip;
        double fpair = -psip * recip;
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        if (fsite && site_e) {
          size_t slot = (size_t)i * maxneigh + jj;
          fsite[3 * slot] = delx * fpair;
          fsite[3 * slot + 1] = dely * fpair;
          fsite[3 * slot + 2] = delz * fpair;
          has[slot] = 1;
        }
        epair += share * phi;
        v[0] += share * delx * delx * fpair;
        v[1] += share * dely * dely * fpair;
        v[2] += share * delz * delz * fpair;
        v[3] += share * delx * dely * fpair;
        v[4] += share * delx * delz * fpair;
        v[5] += share * dely * delz * fpair;
      }
    }
    f[3 * i] = fx;
    f[3 * i + 1] = fy;
    f[3 * i + 2] = fz;
    if (eatom) eatom[i] = embed_energy(t, rho[i], fp[i]) + epair;
    if (vatom)
      for (int a = 0; a < 6; a++) vatom[6 * i + a] = v[a];
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ne; i++) {
    double nf[8][3];
    for (int k = 0; k < 8; k++) nf[k][0] = nf[k][1] = nf[k][2] = 0.0;
    for (int g = 0; g < nint; g++) {
      const double *Ng = el->spa_int + 8 * g;
      double ixtmp = 0, iytmp = 0, iztmp = 0, rhoi = 0;
      for (int kk = 0; kk < 8; kk++) {
        ixtmp += Ng[kk] * el->nodex[24 * i + 3 * kk];
        iytmp += Ng[kk] * el->nodex[24 * i + 3 * kk + 1];
        iztmp += Ng[kk] * el->nodex[24 * i + 3 * kk + 2];
        rhoi += Ng[kk] * erho[8 * i + kk];
      }
      double fpi = embed_fp(t, rhoi);
      double fx = 0, fy = 0, fz = 0, epair = 0;
      double v[6] = {0, 0, 0, 0, 0, 0};
      size_t p = (size_t)i * nint + g;
      for (int ll = 0; ll < inum[p]; ll++) {
        uint64_t e = inb[p * maxnip + ll];
        if ((uint32_t)(e >> 32) == ORC_ATOM_SITE) continue; /* density only */
        double s[3];
        entry_pos(x, el, e, prd, s);
        double delx = ixtmp - s[0], dely = iytmp - s[1], delz = iztmp - s[2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < t->cutforcesq) {
          double fpj = entry_fp(el, t, fp, erho, e);
          double r = sqrt(rsq), pp;
          const double *c = rtab(t->rhor_spline, r, t->rdr, t->nr, &pp);
          double rhoip = (c[0] * pp + c[1]) * pp + c[2];
          const double *z = t->z2r_spline + (c - t->rhor_spline);
          double z2p = (z[0] * pp + z[1]) * pp + z[2];
          double z2 = ((z[3] * pp + z[4]) * pp + z[5]) * pp + z[6];
          double recip = 1.0 / r;
          double phi = z2 * recip;
          double phip = z2p * recip - phi * recip;
          double psip = (fpi + fpj) * rhoip + phip;
          double fpair = -psip * recip;
          fx += delx * fpair;
          fy += dely * fpair;
          fz += delz * fpair;
          epair += 0.5 * phi;
          v[0] += 0.5 * delx * delx * fpair;
          v[1] += 0.5 * dely * dely * fpair;
          v[2] += 0.5 * delz * delz * fpair;
          v[3] += 0.5 * delx * dely * fpair;
          v[4] += 0.5 * delx * delz * fpair;
          v[5] += 0.5 * dely * delz * fpair;
        }
      }
      double w = el->weight[g];
      for (int kk = 0; kk < 8; kk++) {
        nf[kk][0] += w * Ng[kk] * fx;
        nf[kk][1] += w * Ng[kk] * fy;
        nf[kk][2] += w * Ng[kk] * fz;
      }
      if (ipf) { ipf[3 * p] = fx; ipf[3 * p + 1] = fy; ipf[3 * p + 2] = fz; }
      if (eip) eip[p] = w * (embed_energy(t, rhoi, fpi) + epair);
      if (vip)
        for (int a = 0; a < 6; a++) vip[6 * p + a] = w * v[a];
    }
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) nodef[24 * i + 3 * k + d] = nf[k][d];
  }
  if (fsite) {
    add_site_reactions(n, el, anum, anb, maxneigh, fsite, has, nodef);
    free(fsite);
    free(has);
  }
}

/* LJ over the coupled lists: PairLJCut::compute / ::ecompute pair terms
 * (pair_lj_cut.cpp:255-287, 376-428) with every neighbor position from
 * entry_pos.  Integration points sum their site entries only; the atom-site
 * forces reach the sites' nodes as reactions (add_site_reactions), and an
 * atom-site pair puts its whole energy and virial on the atom. */
void orc_lj_coupled(int n, const double *x, const orc_elems *el, const orc_box *box, const int *anum,
                    const uint64_t *anb, int maxneigh, const int *inum, const uint64_t *inb, int maxnip,
                    const orc_lj *pp, double *f, double *eatom, double *vatom, double *nodef, double *ipf,
                    double *eip, double *vip) {
  int ne = el->ne, nint = el->nint;
  double prd[3] = {box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]};
  double *fsite = NULL;
  unsigned char *has = NULL;
  if (n > 0 && maxneigh > 0) {
    fsite = (double *)malloc(sizeof(double) * 3 * (size_t)n * maxneigh);
    has = (unsigned char *)calloc((size_t)n * maxneigh, 1);
  }
#pragma omp parallel for schedule(static)
  for (int q = 0; q < n + ne * nint; q++) {
    int is_atom = q < n;
    double c[3];
    const int *num = is_atom ? anum : inum;
    const uint64_t *nb = is_atom ? anb : inb;
    int row = is_atom ? q : q - n, cap = is_atom ? maxneigh : maxnip;
    if (is_atom) {
      for (int d = 0; d < 3; d++) c[d] = x[3 * q + d];
    } else {
      ip_position(el, row / nint, row % nint, c);
    }
    double fx = 0, fy = 0, fz = 0, e = 0;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = 0; jj < num[row]; jj++) {
      uint64_t ent = nb[(size_t)row * cap + jj];
      int site_e = (uint32_t)(ent >> 32) != ORC_ATOM_SITE;
      if (!is_atom && !site_e) continue; /* atoms: reactions below */
      double s[3];
      entry_pos(x, el, ent, prd, s);
      double delx = c[0] - s[0], dely = c[1] - s[1], delz = c[2] - s[2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < pp->cutsq) {
        double share = (is_atom && site_e) ? 1.0 : 0.5;
        double r2inv = 1.0 / rsq;
        double r6inv = r2inv * r2inv * r2inv;
        double forcelj = r6inv * (pp->lj1 * r6inv - pp->lj2);
        double fpair = forcelj * r2inv;
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        if (fsite && is_atom && site_e) {
          size_t slot = (size_t)row * maxneigh + jj;
          fsite[3 * slot] = delx * fpair;
          fsite[3 * slot + 1] = dely * fpair;
          fsite[3 * slot + 2] = delz * fpair;
          has[slot] = 1;
        }
        e += share * (r6inv * (pp->lj3 * r6inv - pp->lj4));
        v[0] += share * delx * delx * fpair;
        v[1] += share * dely * dely * fpair;
        v[2] += share * delz * delz * fpair;
        v[3] += share * delx * dely * fpair;
        v[4] += share * delx * delz * fpair;
        v[5] += share * dely * delz * fpair;
      }
    }
    if (is_atom) {
      f[3 * q] = fx;
      f[3 * q + 1] = fy;
      f[3 * q + 2] = fz;
      if (eatom) eatom[q] = e;
      if (vatom)
        for (int a = 0; a < 6; a++) vatom[6 * q + a] = v[a];
    } else {
      double w = el->weight[row % nint];
      if (ipf) { ipf[3 * row] = fx; ipf[3 * row + 1] = fy; ipf[3 * row + 2] = fz; }
      if (eip) eip[row] = w * e;
      if (vip)
        for (int a = 0; a < 6; a++) vip[6 * row + a] = w * v[a];
    }
  }
  /* node scatter in integration point order, pair_lj_cut.cpp:424-428
   * (ipf is required here: it carries f_g to the scatter) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ne; i++) {
    double nf[8][3];
    for (int k = 0; k < 8; k++) nf[k][0] = nf[k][1] = nf[k][2] = 0.0;
    for (int g = 0; g < nint; g++) {
      const double *Ng = el->spa_int + 8 * g;
      size_t p = (size_t)i * nint + g;
      double w = el->weight[g];
      for (int kk = 0; kk < 8; kk++) {
        nf[kk][0] += w * Ng[kk] * ipf[3 * p];
        nf[kk][1] += w * Ng[kk] * ipf[3 * p + 1];
        nf[kk][2] += w * Ng[kk] * ipf[3 * p + 2];
      }
    }
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) nodef[24 * i + 3 * k + d] = nf[k][d];
  }
  if (fsite) {
    add_site_reactions(n, el, anum, anb, maxneigh, fsite, has, nodef);
    free(fsite);
    free(has);
  }
}

/* ------------------------------------------------------------------------
 * FixNVE, fix_nve.cpp:40-242: dtf = 0.5*dt*ftm2v (:43); atoms use
 * dtfm = dtf/mass (:81); nodes use dtfm = dtf/(inv_npe*mass_e) (:124,:233).
 * ---------------------------------------------------------------------- */
void orc_nve_atoms_initial(int n, double *x, double *v, const double *f,
                           double dtv, double dtf, double mass) {
  double dtfm = dtf / mass;
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      v[3 * i + d] += dtfm * f[3 * i + d];
      x[3 * i + d] += dtv * v[3 * i + d];
    }
}

void orc_nve_atoms_final(int n, double *v, const double *f, double dtf, double mass) {
  double dtfm = dtf / mass;
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) v[3 * i + d] += dtfm * f[3 * i + d];
}

void orc_nve_nodes_initial(int ne, double *nodex, double *nodev, const double *nodef,
                           double *xc, double dtv, double dtf, double mass_e) {
  double inv_npe = 1 / (double)8;
  double dtfm = dtf / (inv_npe * mass_e);
  for (int i = 0; i < ne; i++)
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) {
        size_t a = 24 * (size_t)i + 3 * k + d;
        nodev[a] += dtfm * nodef[a];
        nodex[a] += dtv * nodev[a];
      }
  if (xc) orc_element_centers(ne, nodex, xc);
}

void orc_nve_nodes_final(int ne, double *nodev, const double *nodef, double dtf, double mass_e) {
  double inv_npe = 1 / (double)8;
  double dtfm = dtf / (inv_npe * mass_e);
  for (size_t a = 0; a < 24 * (size_t)ne; a++) nodev[a] += dtfm * nodef[a];
}

/* Domain::pbc, domain.cpp:535-633 (no deform remap). */
void orc_pbc_atoms(int n, double *x, int *image, const orc_box *box) {
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double prd = box->hi[d] - box->lo[d];
      if (x[3 * i + d] < box->lo[d]) {
        x[3 * i + d] += prd;
        if (image) image[3 * i + d]--;
      }
      if (x[3 * i + d] >= box->hi[d]) {
        x[3 * i + d] -= prd;
        x[3 * i + d] = MAXD(x[3 * i + d], box->lo[d]);
        if (image) image[3 * i + d]++;
      }
    }
}

/* Domain::elepbc, domain.cpp:647-745: wrap by center, nodes move with it.
 * Fixed: at x-hi the reference adds the period to the nodes (domain.cpp:682);
 * here the nodes move with the center. */
void orc_pbc_elements(int ne, double *xc, double *nodex, int *image, const orc_box *box) {
  for (int i = 0; i < ne; i++)
    for (int d = 0; d < 3; d++) {
      double prd = box->hi[d] - box->lo[d];
      if (xc[3 * i + d] < box->lo[d]) {
        xc[3 * i + d] += prd;
        for (int k = 0; k < 8; k++) nodex[24 * i + 3 * k + d] += prd;
        if (image) image[3 * i + d]--;
      }
      if (xc[3 * i + d] >= box->hi[d]) {
        xc[3 * i + d] -= prd;
        for (int k = 0; k < 8; k++) nodex[24 * i + 3 * k + d] -= prd;
        xc[3 * i + d] = MAXD(xc[3 * i + d], box->lo[d]);
        if (image) image[3 * i + d]++;
      }
    }
}

/* Neighbor::check_distance, neighbor.cpp:1176-1217: any point moved more
 * than sqrt(triggersq) since the last build. */
int orc_max_displacement_exceeds(int n, const double *x, const double *xhold, double triggersq) {
  int flag = 0;
  for (int i = 0; i < n; i++) {
    double delx = x[3 * i] - xhold[3 * i];
    double dely = x[3 * i + 1] - xhold[3 * i + 1];
    double delz = x[3 * i + 2] - xhold[3 * i + 2];
    double rsq = delx * delx + dely * dely + delz * delz;
    if (rsq > triggersq) flag = 1;
  }
  return flag;
}
