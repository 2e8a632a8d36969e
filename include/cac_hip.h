/*
 * cac_hip.h -- C-ABI of the MI355X-native CAC force-and-neighbor path.
 *
 * One context per host thread, one HIP stream per context, all device
 * buffers owned by the context.  Plain pointers and sizes only; host arrays
 * stay the owner of record for I/O, the device copy is authoritative while a
 * run is in progress.  Every entry point returns 0 (CAC_OK) or a CAC_ERR_*
 * status; cac_last_error() returns the message.  The reference has no return
 * codes: a host style built on this ABI maps a non-zero status to its
 * Error::all (error.cpp:72-91).
 *
 * Each declaration names the reference interface it replaces.  Reference
 * paths are relative to the fangyuanxiao/CAC_Code tree.
 *
 * One atom type and one element type per context: the reference's element
 * path hard-codes type 1 (pair_eam.cpp:980, 1047, 1091, 1109, 1111;
 * pair_lj_cut.cpp:327-331), so the uploads take one atom mass and one
 * element shape, and the host bindings refuse more than one type of each.
 *
 * Neighbor entry encoding (cac_get_atom_neighbors / cac_get_ip_neighbors),
 * uint64: j | (img << 27) | ((uint64)site << 32), where site ==
 * CAC_ATOM_SITE names atom j and any other value site `site` of element j.
 * img = (sx+1) + 3*(sy+1) + 9*(sz+1), s in {-1,0,1}^3 the periodic image of
 * j (position x_j + s*prd); img 13 is the unshifted copy.  Without coupling
 * (the reference) atom lists hold only atoms and integration-point lists
 * only element sites.
 */
#ifndef CAC_HIP_H
#define CAC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cac_ctx cac_ctx;

enum {
  CAC_OK = 0,
  CAC_ERR_ARG = 1,      /* bad argument or call order */
  CAC_ERR_HIP = 2,      /* HIP runtime error */
  CAC_ERR_OVERFLOW = 3, /* neighbor capacity exceeded (after regrow attempts) */
  CAC_ERR_BOX = 4,      /* box too small for the cutoffs */
  CAC_ERR_DATA = 5,     /* inconsistent element data, or a malformed file */
  CAC_ERR_IO = 6        /* a file could not be opened, read or written */
};

enum { CAC_PREC_FP64 = 0, CAC_PREC_MIXED = 1 };
#define CAC_ATOM_SITE 0xFFFFFFFFu
enum { CAC_PAIR_NONE = 0, CAC_PAIR_EAM = 1, CAC_PAIR_LJ = 2 };

/* ---- context ------------------------------------------------------------
 * Replaces the CAC god-object set-up (cac.cpp:157-213) for this path. */
int cac_ctx_create(int device, cac_ctx **out);
int cac_ctx_destroy(cac_ctx *ctx);
const char *cac_last_error(const cac_ctx *ctx);
int cac_synchronize(cac_ctx *ctx);

/* Domain::set_initial_box / set_global_box (domain.cpp), fully periodic only
 * (the C0-C4 decks use "boundary p p p"). */
int cac_set_box(cac_ctx *ctx, const double lo[3], const double hi[3], const int periodic[3]);

/* Atom container upload, replaces AtomVecAtomic::data_atom + grow
 * (atom_vec_atomic.cpp:41-71).  x, v: [n][3] AoS like Atom::x (atom.h:37);
 * v may be NULL (zero).  One atom type of mass `mass` (Atom::mass[1]). */
int cac_upload_atoms(cac_ctx *ctx, int n, const double *x, const double *v, double mass);

/* Element container upload, replaces Element::data_elements/data_nodes/
 * set_nae (element.cpp:195-294, 478-540, 1086-1110).  nodex, nodev:
 * [ne][8][3] like Element::nodex (element.h:25); nodev may be NULL.
 * mass_e = nae0*nae1*nae2*mass (element.cpp:1104). */
int cac_upload_elements(cac_ctx *ctx, int ne, const int nae[3], const double *nodex,
                        const double *nodev, double mass_e);

/* Integration points of the element type, replaces the Numint/Intp sections
 * (element.cpp:1113-1166) plus intp_pcoord / shape_array_int /
 * find_node_list (element.cpp:1170-1191, 1326-1354, 1296-1319).  ijk: [nint][3]
 * lattice indices, w: [nint] weights.  Fails with CAC_ERR_DATA if a node has
 * no coincident integration point. */
int cac_set_quadrature(cac_ctx *ctx, int nint, const int *ijk, const double *w);

/* Host-side table construction shared by every path: PairEAM::file2array +
 * array2spline for one funcfl file (pair_eam.cpp:311-580).  Inputs are the
 * file arrays (0-based, value at r = m*dr is f[m]); outputs are 1-based spline
 * arrays [(n+1)][7] with n = *nrho_out / *nr_out; pass NULL outputs to query
 * the sizes first. */
int cac_eam_tables_from_funcfl(int nrho_f, double drho_f, int nr_f, double dr_f,
                               const double *frho_f, const double *rhor_f, const double *zr_f,
                               int *nrho_out, double *drho_out, int *nr_out, double *dr_out,
                               double *frho_spline, double *rhor_spline, double *z2r_spline);

/* The same tables from a LAMMPS setfl file of one element (eam/alloy).  No
 * counterpart in the reference, which reads funcfl only
 * (pair_eam.cpp:197-268).  The setfl arrays F(rho) [nrho], rho(r) [nr] and
 * r*phi(r) [nr] are already on one grid: they are the knot values as they
 * are, with the knot coefficients of array2spline (pair_eam.cpp:527-580).
 * Outputs [(nrho+1)][7] and [(nr+1)][7] as above. */
int cac_eam_tables_from_setfl(int nrho, double drho, int nr, double dr, const double *frho,
                              const double *rhor, const double *z2r, double *frho_spline,
                              double *rhor_spline, double *z2r_spline);

/* pair_style eam + pair_coeff (pair_eam.cpp:124-160, init_style :274-283,
 * init_one :288-304): upload the spline tables (rows 1..n of 7) and
 * cutforcesq. */
int cac_set_eam(cac_ctx *ctx, int nrho, double drho, const double *frho_spline,
                int nr, double dr, const double *rhor_spline, const double *z2r_spline,
                double cutforcesq);

/* pair_style lj/cut + pair_coeff (pair_lj_cut.cpp:96-198): lj1..lj4 from
 * PairLJCut::init_one and cutsq = cut^2. */
int cac_set_lj(cac_ctx *ctx, double lj1, double lj2, double lj3, double lj4, double cutsq);

/* neighbor + neigh_modify (neighbor.cpp:194-262, init :297-481).
 * cutneigh is the full list cutoff (the reference's "skin" value, which it
 * uses as the whole cutoff, neighbor.cpp:357-358).  triggersq: squared
 * displacement that forces a rebuild (neighbor.cpp:334).  every/delay/check
 * as in Neighbor::decide (neighbor.cpp:1150-1163).  maxneigh / maxnip:
 * initial per-atom / per-integration-point capacities (grown on overflow). */
int cac_set_neighbor(cac_ctx *ctx, double cutneigh, double triggersq, int every, int delay,
                     int check, int maxneigh, int maxnip);

/* Order of the entries within each atom list (no counterpart in the
 * reference, whose lists are in bin order, neighbor.cpp:583-613): entries are
 * sorted by their displacement x_j' - x_i rounded to multiples of `spacing`
 * (Angstrom), k, ordered by |k|^2 and then k_z, k_y, k_x, then by value.  On a
 * lattice whose neighbor vectors are multiples of `spacing` (a/2 for fcc)
 * every atom then lists its neighbors in the same relative order, nearest
 * shells first, which makes the force passes' gathers coalesce and lets them
 * stop at the last shell within the cutoff.  0 (the default) sorts by value.
 * The neighbor sets do not change. */
int cac_set_list_order(cac_ctx *ctx, double spacing);

/* Atom-element coupling (no counterpart in the reference, where atoms see
 * only atoms, pair_eam.cpp:584-746, and integration points only element
 * sites, pair_eam.cpp:927-1137).  With on != 0 the next build adds to every
 * atom list the interpolated element sites within cutneigh and to every
 * integration-point list the atoms within cutneigh.  Atoms take the atom
 * pass's pair force from each site (the site's F' interpolated from its
 * element's node densities); integration points count atoms in their
 * densities but not in their forces; instead every site passes the
 * reaction of its atom pairs to its element's nodes through its shape
 * functions, so atom-element forces are equal and opposite (Newton's third
 * law across the interface).  An atom-site pair's energy and virial are
 * tallied on the atom.  Off by default. */
int cac_set_coupling(cac_ctx *ctx, int on);

/* Affine remap of the box and everything in it (no counterpart in the
 * reference, whose input has no deformation command; it is LAMMPS' fix
 * deform with "remap x").  Every coordinate maps to lo + (x - lo)*scale[d]:
 * atoms, element nodes and centers, and the positions held at the last
 * build (so the rebuild check measures motion, not the remap); box hi
 * follows, velocities do not change.  The neighbor lists stay in use until
 * the usual check asks for a rebuild. */
int cac_deform(cac_ctx *ctx, const double scale[3]);

/* Constant engineering strain rate (per time unit of dtv) that cac_run
 * applies before every step: the box edge at elapsed time t is
 * L0*(1 + rate*t), L0 the edge when the rate was set.  A zero rate stops
 * it. */
int cac_set_deform_rate(cac_ctx *ctx, const double rate[3]);

/* Vertex clusters of the element integration points (no counterpart in the
 * reference: a grouping of work, the forces do not change).  Cluster c holds
 * the node slots slots[start[c] .. start[c+1]), slot 8*e + k naming node k
 * of element e, of the elements meeting at one mesh vertex: at most 8 per
 * cluster, every slot of every element in exactly one cluster.  When set
 * and every integration point is a node point, the node-density and
 * integration-point force passes run one cluster per wavefront: the union of
 * the cluster's neighbor lists is interpolated once into LDS and the
 * integration-point sums read it there.  ncl = 0 clears it.  Takes effect at
 * the next build. */
int cac_set_ip_clusters(cac_ctx *ctx, int ncl, const int *start, const int *slots);

/* State of the vertex-cluster element passes after the last build: *active
 * = 1 when the force calls use them, *max_union = the largest union of a
 * cluster's lists (its LDS rows).  No counterpart in the reference. */
int cac_cluster_info(cac_ctx *ctx, int *ncl, int *active, int *max_union);

/* Neighbor rebuild: Domain::pbc/elepbc + Neighbor::build/ebuild/nbuild
 * (verlet.cpp:173-193; neighbor.cpp:546-625, element_neighbor.cpp:11-75,
 * full_node_neighbor.cpp:19-239).  Regrows capacities on overflow. */
int cac_build_neighbors(cac_ctx *ctx);

/* Neighbor::decide + check_distance (neighbor.cpp:1150-1217). */
int cac_decide(cac_ctx *ctx, int *rebuild);

/* Pair::compute (atoms) + Pair::ecompute (elements), pair.h:96-97:
 * PairEAM::compute/ecompute (pair_eam.cpp:584-746, 761-1137) or
 * PairLJCut::compute/ecompute (pair_lj_cut.cpp:227-288, 295-442), after
 * Verlet::force_clear/node_force_clear (verlet.cpp:105-144).  eflag/vflag
 * enable the energy/virial tallies (build-defined; the reference has none).
 * precision: CAC_PREC_FP64 or CAC_PREC_MIXED (FP32 pair/embedding math,
 * FP64 accumulation). */
int cac_force(cac_ctx *ctx, int eflag, int vflag, int precision);

/* FixNVE::initial_integrate + einitial_integrate (fix_nve.cpp:54-157) and
 * FixNVE::final_integrate + efinal_integrate (fix_nve.cpp:181-242);
 * dtf = 0.5*dt*ftm2v (fix_nve.cpp:43). */
int cac_nve_initial(cac_ctx *ctx, double dtv, double dtf);
int cac_nve_final(cac_ctx *ctx, double dtf);

/* Verlet::run (verlet.cpp:149-219) for nsteps with FixNVE, device-resident:
 * initial integrate, decide/rebuild, force, final integrate.  *nbuilds (may
 * be NULL) receives the number of rebuilds. */
int cac_run(cac_ctx *ctx, int nsteps, double dtv, double dtf, int precision, int *nbuilds);

/* Device -> host copies (Atom::x/v/f, Element::nodex/nodev/nodef layouts). */
int cac_download_atoms(cac_ctx *ctx, double *x, double *v, double *f);
int cac_download_elements(cac_ctx *ctx, double *nodex, double *nodev, double *nodef);

/* Energy and virial of the last cac_force with eflag/vflag set:
 * eng[0] atoms, eng[1] elements (integration-point weighted);
 * virial[6] = xx yy zz xy xz yz summed over both (build-defined, not in the
 * reference: integrate.cpp:97-114 forces eflag = vflag = 0). */
int cac_get_energy_virial(cac_ctx *ctx, double eng[2], double virial[6]);

/* Thermodynamic output of the current state (the reference's thermo prints
 * only the step, thermo.cpp:195-209, and its computes are empty,
 * compute_pe.cpp:16-49, compute_temp.cpp:15-32, compute_pressure.cpp:18-138;
 * these are LAMMPS' definitions).  Runs one FP64 force call with the energy
 * and virial tallies (forces are recomputed on the same positions, so they
 * do not change) and returns out = {pe, ke, temp, press, etotal}:
 *   pe    = atom + integration-point energies (cac_get_energy_virial),
 *   ke    = 0.5 * mvv2e * (sum m v^2 over atoms + sum (m_e/8) v^2 over nodes),
 *   temp  = 2 ke / (dof * boltz), dof = 3 (n + 8 ne) - 3,
 *   press = (dof*boltz*temp + W_xx + W_yy + W_zz) / (3 V) * nktv2p,
 *   etotal = pe + ke.
 * Unit constants from the caller (metal: mvv2e = 1.0364269e-4,
 * boltz = 8.617343e-5, nktv2p = 1.6021765e6, update.cpp:107-125). */
int cac_thermo(cac_ctx *ctx, double mvv2e, double boltz, double nktv2p, double out[5]);

/* Centro-symmetry parameter of every atom (replaces Output::centro_atom,
 * output.cpp:352-441, which verlet.cpp:93-94 calls after Neighbor::build):
 * over the list entries with r^2 < cutneigh^2 (cutneighmaxsq), the nnn
 * nearest (nnn = 12 in the reference, output.cpp:84; 8 for bcc), and
 * centro[i] = the sum of the nnn/2 smallest |x_j + x_k - 2 x_i|^2 over their
 * pairs, or 0 with fewer than nnn entries.  centro: [n] on the host.  Uses
 * the current lists; called right after cac_build_neighbors it is the
 * reference's call.  With coupling on, element sites in the atom lists
 * count as neighbors.  Ties at the nnn-th distance go to the smaller entry
 * word (atom index | image << 27); the reference leaves them to its
 * quickselect. */
int cac_centro_atoms(cac_ctx *ctx, int nnn, double *centro);

/* Every interpolated site of every element (replaces
 * Element::interpolate_atom_vec, element.cpp:1226-1265): x[(e*nsite + s)*3 + d]
 * = sum_k N_k(site s) nodex[e][k][d], summed in node order, sites in the
 * inter_ele order (element.cpp:1394-1423), nsite = nae[0]*nae[1]*nae[2].
 * The reference's only executed path (verlet.cpp:91-95) replaces its atoms
 * by these sites, builds the atom list over them and computes their
 * centro-symmetry: upload them as the atoms of a context and call
 * cac_build_neighbors and cac_centro_atoms there. */
int cac_interpolate_sites(cac_ctx *ctx, double *x);

/* Parity hooks: neighbor lists (row-major [n][maxneigh] / [ne*nint][maxnip],
 * device order, entries as above), per-atom rho/fp, per-node erho ([ne][8]),
 * per-integration-point force ([ne*nint][3]).  Return CAC_ERR_ARG if the
 * capacity given is below the device list length (see cac_get_capacities). */
int cac_get_capacities(cac_ctx *ctx, int *maxneigh, int *maxnip);
int cac_get_atom_neighbors(cac_ctx *ctx, int maxneigh, int *numneigh, uint64_t *neigh);
int cac_get_ip_neighbors(cac_ctx *ctx, int maxnip, int *numneigh, uint64_t *neigh);
int cac_get_rho(cac_ctx *ctx, double *rho, double *fp);
int cac_get_erho(cac_ctx *ctx, double *erho);
int cac_get_ip_forces(cac_ctx *ctx, double *ipf);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops with fixed memory):
 * the EAM atom passes of cac_force (PairEAM::compute's rho, fp and force
 * loops, pair_eam.cpp:620-745) over caller-owned DEVICE arrays, allocating
 * nothing.  x: [n][3] positions; numneigh: [n]; neigh: [n][maxneigh] full
 * lists of entries j | img << 27 (atom j < n at image img, as in
 * cac_get_atom_neighbors' low 32 bits; other entries are skipped); f: [n][3]
 * forces out; rho: [n] densities out, or NULL.  Tables, box and stream are
 * the context's (cac_set_eam, cac_set_box).  F'(rho) lives in a fixed
 * per-device scratch array of the library: n <= 1,048,576 (8 MB), larger n
 * fails with CAC_ERR_ARG.  Queued on the context stream (cac_synchronize
 * waits); calls that overlap on one device share the scratch and race. */
int cac_eam_atoms_noalloc(cac_ctx *ctx, int n, const double *x, int maxneigh, const int *numneigh,
                          const uint32_t *neigh, double *f, double *rho);

/* Timing hook: average duration (ms) of the named kernel family over the
 * launches recorded since the last reset, measured with HIP events on the
 * context stream ("atom_force", "ip_force", ...); returns -1 if unknown. */
int cac_profile_enable(cac_ctx *ctx, int on);
double cac_profile_avg_ms(cac_ctx *ctx, const char *name, int *count);

/* ---- learned element surrogate (no counterpart in the reference) --------
 * A BF16-MFMA multilayer perceptron in -> nhidden (x nlayers, SiLU) -> out
 * with FP32 accumulation and FP32 master weights, trained on the device with
 * Adam (lr given, beta 0.9/0.999, eps 1e-8) on the mean-squared error.
 * Weights are [out][in] row-major FP32 on the host side.  Inputs and targets
 * are FP32 [n][in] / [n][out]; the caller normalizes them. */
int cac_mlp_create(cac_ctx *ctx, int nin, int nhidden, int nlayers, int nout, uint64_t seed);
int cac_mlp_set_weights(cac_ctx *ctx, int layer, const float *w, const float *b);
int cac_mlp_get_weights(cac_ctx *ctx, int layer, float *w, float *b);
int cac_mlp_forward(cac_ctx *ctx, int n, const float *x, float *y);
/* loss and gradients of one batch, no update: grads[l] = dW_l [out][in]
 * then db_l [out]; NULL entries are skipped */
int cac_mlp_grads(cac_ctx *ctx, int n, const float *x, const float *t, float *loss, float **grads);
int cac_mlp_train_step(cac_ctx *ctx, int n, const float *x, const float *t, float lr, float *loss);
int cac_mlp_set_dataset(cac_ctx *ctx, int n, const float *x, const float *t);
int cac_mlp_train(cac_ctx *ctx, int epochs, int batch, float lr, uint64_t seed, float *loss);

/* Surrogate on the force path (no counterpart in the reference).  X0:
 * [ne][8][3] reference node positions; nbr: [ne][26] mesh neighbors of each
 * element (-1 where the neighbor block is not an element); in_scale
 * [CAC_SURR_NFEAT], out_scale [CAC_SURR_NOUT] normalize the features and the
 * nodal forces.  Features: deviations from the element's best-fit affine map
 * of its 8 nodes (24), of the 26 neighbor centroids (78) and of the 7
 * neighbor nodes meeting each corner (168), then the strain F - I (9).
 * Prediction of nodal force c = 3k + d (normalized):
 *   b[c] + A[c] . x + sum_{i<=j} Q[c][ij] z_i z_j + MLP(x)[c]
 * with z the 33 local inputs of corner k (own node k, its 7 corner
 * neighbors, the strain); A, Q, b are set by cac_surrogate_set_fit (FP32),
 * the MLP (CAC_SURR_NFEAT -> 24) by cac_mlp_set_weights.  While enabled,
 * cac_force (EAM, eflag = vflag = 0) takes the nodal forces of every element
 * from the surrogate except fallback elements (interface, or Green-Lagrange
 * strain above strain_limit), which keep the plain integration-point path;
 * classification is refreshed at each rebuild.  cac_surrogate_features
 * writes the raw features [ne][CAC_SURR_NFEAT] and the class per element
 * (0 surrogate, 1 interface, 2 strained).
 * With an MLP of CAC_SURR_NFEAT_CONFIG inputs instead (the configured
 * network: the first 102 features, own nodes and neighbor centroids), the
 * prediction is MLP(x)[c] alone and cac_surrogate_set_fit is not used. */
#define CAC_SURR_NFEAT 279
#define CAC_SURR_NFEAT_CONFIG 102
#define CAC_SURR_NOUT 24
#define CAC_SURR_NQUAD 561
int cac_surrogate_setup(cac_ctx *ctx, const double *X0, const int *nbr, const float *in_scale,
                        const float *out_scale, double strain_limit);
int cac_surrogate_set_fit(cac_ctx *ctx, const float *A, const float *Q, const float *b);
int cac_surrogate_enable(cac_ctx *ctx, int on);
int cac_surrogate_features(cac_ctx *ctx, float *feats, int *cls);
int cac_surrogate_info(cac_ctx *ctx, int *ninterface, int *nstrained, int *nsurrogate);

#ifdef __cplusplus
}
#endif
#endif
