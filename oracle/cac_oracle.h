/*
 * cac_oracle.h -- CPU FP64 oracle for the CAC force-and-neighbor path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the
 * arithmetic of the reference CAC code (fangyuanxiao/CAC_Code) on the
 * force-and-neighbor path.  It exists to check the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker / CPU baseline.  The product
 * library (libcac_hip.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Build:
 * oracle/Makefile (gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp).
 *
 * Neighbor entry encodings (shared with the HIP library so sets can be
 * compared bit for bit after sorting):
 *   atom list entry  (uint32): j | (img << 27)
 *   IP list entry    (uint64): j | (img << 27) | ((uint64)site << 32)
 * where img = (sx+1) + 3*(sy+1) + 9*(sz+1), s in {-1,0,1}^3, is the
 * periodic image of j; img 13 is the unshifted copy.  The image
 * position is x_j + s*prd per component, the value the reference's
 * ghost copy carries (atom_vec_atomic.cpp:202-221, element.cpp:933-953).
 */
#ifndef CAC_ORACLE_H
#define CAC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_IMG_SELF 13

typedef struct {
  double lo[3];
  double hi[3];
} orc_box;

/* EAM spline tables in the reference layout: rows 1..n of 7 coefficients
 * (pair_eam.cpp:527-580); row 0 unused. */
typedef struct {
  int nrho;
  int nr;
  double rdrho;
  double rdr;
  double cutforcesq;
  const double *frho_spline; /* [(nrho+1)*7] */
  const double *rhor_spline; /* [(nr+1)*7]   */
  const double *z2r_spline;  /* [(nr+1)*7]   */
} orc_eam;

typedef struct {
  double lj1, lj2, lj3, lj4, cutsq;
} orc_lj;

/* One element type (the reference's element path hard-codes type 1,
 * pair_eam.cpp:980,1047,1091). */
typedef struct {
  int ne;
  const double *nodex;   /* [ne][8][3] */
  int nae[3];
  int nsite;             /* nae0*nae1*nae2 */
  const double *spa_site;/* [nsite][8]  Spa_inter_ele */
  int nint;
  const double *spa_int; /* [nint][8]   Spa_int */
  const double *weight;  /* [nint] */
  int node_list[8];      /* integration point at node k */
} orc_elems;

/* ---- shape functions and parent coordinates ---- */
void orc_shape(double px, double py, double pz, double N[8]);
void orc_site_pcoords(const int nae[3], double *pc);
void orc_site_shapes(const int nae[3], double *spa);
void orc_intp_shapes(const int nae[3], int nint, const int *ijk, double *spa_int);
int  orc_node_list(int nint, const double *spa_int, int node_list[8]);
double orc_element_max_size(int ne, const double *nodex);
void orc_element_centers(int ne, const double *nodex, double *xc);

/* ---- EAM tables ---- */
int  orc_funcfl_regrid(int nrho_f, double drho_f, int nr_f, double dr_f,
                       const double *frho_f, const double *rhor_f, const double *zr_f,
                       int *nrho, double *drho, int *nr, double *dr,
                       double *frho, double *rhor, double *z2r);
void orc_spline(int n, double delta, const double *f, double *spline);

/* ---- neighbor builds ---- */
int orc_atom_neighbors(int n, const double *x, const orc_box *box, double cutneigh,
                       int maxneigh, int *numneigh, uint32_t *neigh);
int orc_atom_neighbors_brute(int n, const double *x, const orc_box *box, double cutneigh,
                             int maxneigh, int *numneigh, uint32_t *neigh);
int orc_ip_neighbors(const orc_elems *el, const orc_box *box, double cutneigh, int nsplit,
                     int maxnip, int *numneigh, uint64_t *neigh);
int orc_ip_neighbors_brute(const orc_elems *el, const orc_box *box, double cutneigh,
                           int maxnip, int *numneigh, uint64_t *neigh);

/* ---- forces ---- */
void orc_eam_atoms(int n, const double *x, const orc_box *box,
                   const int *numneigh, const uint32_t *neigh, int maxneigh,
                   const orc_eam *t, double *rho, double *fp, double *f,
                   double *eatom, double *vatom);
void orc_eam_elements(const orc_elems *el, const orc_box *box,
                      const int *numneigh, const uint64_t *neigh, int maxnip,
                      const orc_eam *t, double *erho, double *nodef, double *ipf,
                      double *eip, double *vip);
void orc_lj_atoms(int n, const double *x, const orc_box *box,
                  const int *numneigh, const uint32_t *neigh, int maxneigh,
                  const orc_lj *p, double *f, double *eatom, double *vatom);
void orc_lj_elements(const orc_elems *el, const orc_box *box,
                     const int *numneigh, const uint64_t *neigh, int maxnip,
                     const orc_lj *p, double *nodef, double *ipf, double *eip, double *vip);

/* ---- atom-element coupling (build-defined; not in the reference) ----
 * Lists hold uint64 entries (j | img << 27) | ((uint64)site << 32) with
 * site == ORC_ATOM_SITE for an atom j.  Atom lists: atoms and element
 * sites within cutneigh; integration-point lists: sites (reference rules)
 * and atoms within cutneigh.  brute != 0 searches every image of every
 * atom and site.  Rows are sorted.
 * Forces (orc_eam_coupled / orc_lj_coupled): an atom takes the pair force
 * f_is of every site s in its list; the integration points count atoms in
 * their densities only, and each site s passes the reaction -f_is to its
 * element's nodes through its shape functions (nodef[e][k] += N_k(s) R_s,
 * R_s summed over the atoms in ascending (atom | img << 27) order, sites
 * in ascending index, after the integration-point scatter), so atom-element
 * forces are equal and opposite.  An atom-site pair tallies its whole
 * energy and virial on the atom. */
#define ORC_ATOM_SITE 0xFFFFFFFFu
int orc_coupled_neighbors(int n, const double *x, const orc_elems *el, const orc_box *box, double cutneigh,
                          int nsplit, int brute, int maxneigh, int *anum, uint64_t *anb, int maxnip, int *inum,
                          uint64_t *inb);
void orc_eam_coupled(int n, const double *x, const orc_elems *el, const orc_box *box, const int *anum,
                     const uint64_t *anb, int maxneigh, const int *inum, const uint64_t *inb, int maxnip,
                     const orc_eam *t, double *rho, double *fp, double *f, double *eatom, double *vatom,
                     double *erho, double *nodef, double *ipf, double *eip, double *vip);
void orc_lj_coupled(int n, const double *x, const orc_elems *el, const orc_box *box, const int *anum,
                    const uint64_t *anb, int maxneigh, const int *inum, const uint64_t *inb, int maxnip,
                    const orc_lj *pp, double *f, double *eatom, double *vatom, double *nodef, double *ipf,
                    double *eip, double *vip);

/* ---- centro-symmetry (Output::centro_atom, output.cpp:352-441) ----
 * Over the entries of each atom's list (uint64 entries as above; el may be
 * NULL when no entry is a site) with r^2 < cutneigh^2: fewer than nnn give
 * 0; else the nnn nearest by (r^2, tie key), and the sum, in ascending
 * order, of the nnn/2 smallest |x_j + x_k - 2 x_i|^2 over their pairs.  The
 * tie key orders atoms by their 32-bit entry (j | img << 27) and puts sites
 * after atoms.  nnn even, 2 <= nnn <= 16. */
void orc_centro_atoms(int n, const double *x, const orc_elems *el, const orc_box *box, const int *anum,
                      const uint64_t *anb, int maxneigh, double cutneigh, int nnn, double *centro);

/* ---- integration and periodic wrap ---- */
void orc_nve_atoms_initial(int n, double *x, double *v, const double *f,
                           double dtv, double dtf, double mass);
void orc_nve_atoms_final(int n, double *v, const double *f, double dtf, double mass);
void orc_nve_nodes_initial(int ne, double *nodex, double *nodev, const double *nodef,
                           double *xc, double dtv, double dtf, double mass_e);
void orc_nve_nodes_final(int ne, double *nodev, const double *nodef, double dtf, double mass_e);
void orc_pbc_atoms(int n, double *x, int *image, const orc_box *box);
void orc_pbc_elements(int ne, double *xc, double *nodex, int *image, const orc_box *box);
int  orc_max_displacement_exceeds(int n, const double *x, const double *xhold, double triggersq);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
