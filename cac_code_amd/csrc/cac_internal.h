// cac_internal.h -- device context and launcher declarations shared by the
// translation units of libcac_hip.so.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/cac_hip.h"

namespace cac {

// Grow-only device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

constexpr int kMaxNint = 64;         // integration points per element
constexpr int kImgSelf = 13;
// site field of a raw list entry that names an atom (CAC_ATOM_SITE)
constexpr uint32_t kAtomSite = 0xFFFFFFFFu;
// Lanes cooperating on one atom / node / integration point in the pair
// passes, and the row-group size of the blocked neighbor layout: entry k of
// row r lives at ((r/8)*M + k)*8 + r%8, so a wavefront serving 8 rows with
// 8 lanes each reads 64 consecutive entries per step.
constexpr int kLanes = 8;
constexpr int kRowGroup = 8;
// Atom lists use groups of 64 rows and runs of 4 entries: entry k of atom i
// lives at ((i/64)*M + (k & ~3))*64 + (i%64)*4 + (k & 3), so the 64 threads
// of a wavefront (one atom each) read 4 entries each with one 16-byte load,
// 1 KB contiguous per step (M is a multiple of 8).
constexpr int kAtomGroup = 64;
constexpr int kAtomRun = 4;
// Vertex clusters of integration points: at most 8 node slots per cluster.
constexpr int kClMax = 8;
constexpr int kUSmall = 256;
__host__ __device__ __forceinline__ size_t atom_slot(int i, int k, int M) {
  return ((size_t)(i >> 6) * M + (k & ~(kAtomRun - 1))) * kAtomGroup + (size_t)(i & 63) * kAtomRun +
         (k & (kAtomRun - 1));
}
// Cluster-local integration-point lists (uint16 union indices): row p at
// stride M rounded to 32; entry k, taken by lane k % 8 at its step k / 8,
// lives in runs of 4 steps per lane, so a lane reads 4 consecutive entries
// of its own with one 8-byte load and the 8 lanes of a point read 64
// contiguous bytes.
__host__ __device__ __forceinline__ size_t local_stride(int M) { return (size_t)((M + 31) & ~31); }
__host__ __device__ __forceinline__ size_t local_slot(int p, int k, int M) {
  const int l = k & 7, u = k >> 3;
  return (size_t)p * local_stride(M) + (size_t)(u >> 2) * 32 + (size_t)l * 4 + (u & 3);
}
// offset of entry k from atom_slot(i, 0, M)
__host__ __device__ __forceinline__ size_t atom_off(int k) {
  return (size_t)(k & ~(kAtomRun - 1)) * kAtomGroup + (k & (kAtomRun - 1));
}

// Workgroups are handed to the 8 XCDs of an MI355X round-robin.  Kernels
// launched with xcd_grid() call xcd_block() to remap: XCD x then processes
// the contiguous logical blocks [x*q, (x+1)*q), so its L2 serves one spatial
// region of the (spatially ordered) particle arrays instead of all of them.
__device__ __forceinline__ int xcd_block() {
  const int q = (int)(gridDim.x >> 3);
  const int b = (int)blockIdx.x;
  return (b & 7) * q + (b >> 3);
}
inline unsigned xcd_grid(long nthreads, int block) {
  long nb = (nthreads + block - 1) / block;
  return (unsigned)(((nb + 7) / 8) * 8);
}
inline int round_up(int v, int m) { return ((v + m - 1) / m) * m; }

struct Timer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double total_ms = 0.0;
  int count = 0;
};

struct EamTables {
  int nrho = 0, nr = 0;
  double rdrho = 0, rdr = 0, cutforcesq = 0, rhomax = 0;
  // r-knot rows: rv [(nr+1)][4] = rhor value coefficients c3..c6 (density
  // pass); rz [(nr+1)][8] = rhor c3..c6, z2r c3..c6 (force pass, derivative
  // polynomial formed from the value coefficients).  rho-knot rows fr
  // [(nrho+1)][8] = frho c0..c6.
  double *rv = nullptr, *rz = nullptr, *fr = nullptr;
  float *rv_f = nullptr, *rz_f = nullptr, *fr_f = nullptr;
  DevBuf b_rv, b_rz, b_fr, b_rv_f, b_rz_f, b_fr_f;
  // knot form [(nr+1)][4] = rhor c6, c5, z2r c6, c5 (FP64 force passes);
  // knots: row c3, c4 recompute bit-exactly from it
  double *rk = nullptr;
  bool knots = false;
  DevBuf b_rk;
};

struct MlpState;

// Learned element surrogate (k_surrogate.hip): reference nodes, mesh
// adjacency, scales, and the element / active-site subsets of the plain
// path for fallback elements.
struct SurrState {
  bool setup = false, enabled = false, ready = false;
  double strain_limit = 0.02;
  double *ref = nullptr;      // [nref][279] reference geometry classes (r0, d0, c0, G0^-1)
  int *rcls = nullptr;        // [ne] reference class of each element
  int nref = 0;
  std::vector<int> hrcls, hnbr;  // host copies (classification builds the row metadata from them)
  int *meta = nullptr;        // [nS][28] surrogate rows: element, class, 26 neighbors
  float *fit_A = nullptr, *fit_Q = nullptr, *fit_b = nullptr;  // [24][279], [24][561], [24] least-squares part
  int *nbr = nullptr;         // [ne][26] mesh neighbors, -1 if absent
  float *inv_in = nullptr;    // [279] 1 / input scale
  float *out_scale = nullptr; // [24]
  double *ubar = nullptr;     // [ne][3] current element centroids (per step)
  int nF = 0, nR = 0, nR2 = 0, nS = 0, nsR = 0, nsR2 = 0, ninterface = 0, nstrained = 0;
  int *elemF = nullptr, *elemR = nullptr, *elemR2 = nullptr, *elemS = nullptr, *siteR = nullptr, *siteR2 = nullptr;
  DevBuf b_meta, b_ref, b_rcls, b_A, b_Q, b_b, b_xf, b_nbr, b_inv_in, b_out, b_cls, b_mark, b_F, b_S, b_R, b_R2, b_sR, b_sR2, b_tmp, b_ubar, b_feat;
};

struct LjParams {
  double lj1 = 0, lj2 = 0, lj3 = 0, lj4 = 0, cutsq = 0;
};

}  // namespace cac

struct cac_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t lds_bytes = 65536;  // shared memory per workgroup on this device
  int ncu = 256;             // compute units on this device
  std::string err;

  // box (fully periodic)
  bool box_set = false;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, prd[3] = {0, 0, 0};

  // ---- particles: x4 = (x, y, z, fp) rows, the n atoms first, then the
  // nactive active element sites (xs4 = x4 + 4n).  Every neighbor entry
  // indexes this array, so a pass gathers atoms and sites alike.
  // v, f, xhold of the atoms are SoA [3][n].
  int n = 0;
  double mass = 1.0;
  double *x4 = nullptr, *v = nullptr, *f = nullptr, *xhold = nullptr, *rho = nullptr;
  double *eatom = nullptr, *vatom = nullptr;  // [n], [6][n]
  cac::DevBuf b_x4, b_v, b_f, b_xhold, b_rho, b_eatom, b_vatom;
  // atom neighbor list, blocked: entry k of atom i at atom_slot(i, k, maxneigh).
  // Each row is ordered by its displacement vectors rounded to list_spacing
  // (shell by shell, then lattice order; by entry value when 0), so on a
  // lattice every atom lists its neighbors in the same relative order, the
  // nearest shells first.
  int maxneigh = 128;
  double list_spacing = 0.0;
  uint32_t *aneigh = nullptr;
  int *anum = nullptr;
  cac::DevBuf b_aneigh, b_anum;
  // rank of each rounded displacement of the build's code box in list-key
  // order (box half-width crank_qk; the table is made when it changes)
  int16_t *crank = nullptr;
  int crank_qk = 0;
  std::vector<int16_t> crank_host;
  cac::DevBuf b_crank;
  // in-range bits of each atom row (r^2 < cutforcesq), written by the
  // density pass of every EAM force call for its force pass
  uint32_t *amask = nullptr;
  cac::DevBuf b_amask;
  // inner mask: bits of the entries with r^2 < inner_cutsq, (cutforce +
  // margin)^2, at the positions xhold_in of the last full density pass; the
  // density pass gathers only those while no atom has moved more than
  // margin/2 since (device flag *ibad == 0, raised by cac_run's step kernel,
  // set by the host on any other change)
  uint32_t *amask_in = nullptr;
  double *xhold_in = nullptr;
  cac::DevBuf b_amask_in, b_xhold_in;
  double inner_cutsq = -1.0, inner_margin = 0.0, inner_cut = 0.0;
  double inner_prod[3] = {1.0, 1.0, 1.0};  // remap scales applied since the last invalidation
  // atom cells (kept from the last build: the coupled integration-point
  // search reads them)
  int acell[3] = {0, 0, 0};
  uint32_t *abin = nullptr;
  int *abin_s = nullptr, *abin_e = nullptr;
  cac::DevBuf b_akey, b_aval, b_akey2, b_aval2, b_acs, b_ace, b_sorttmp;
  cac::DevBuf b_ahead, b_acells;  // atom list build: cell-start flags and scan, non-empty cells
  int *acells = nullptr, nacells = 0;  // non-empty atom cells of the last build
  double rmax = 0.0;                   // largest element radius of the last build
  // atom-element coupling (build-defined, off by default = the reference):
  // atom lists also hold element sites, integration-point lists also atoms.
  // Interface atoms are those within cutneigh + R_j of some element center.
  bool coupled = false;
  int niface = 0;
  int *iflag = nullptr, *ipos = nullptr, *iface = nullptr, *asnum = nullptr;
  uint64_t *atmp = nullptr;  // [niface][maxneigh] raw (element | img<<27 | site<<32)
  int *etouch = nullptr;     // [ne] element has an atom in a list or a site in an atom list
  cac::DevBuf b_iflag, b_ipos, b_iface, b_asnum, b_atmp, b_etouch;
  // Every atom-site pair of the coupled atom lists once, as (active site
  // a) << 32 | (atom | img << 27): interface atom r's site entries at
  // sent[sbase[r] ..], nsent in all.  Sorted, they are the reaction lists.
  long nsent = 0;
  int *sbase = nullptr;
  uint64_t *sent = nullptr, *sent2 = nullptr;  // sent2: the sorted copy
  cac::DevBuf b_sbase, b_sent, b_sent2;
  // Reactions of the atom-site pair forces (coupling): per active site the
  // atoms whose lists hold it, as sorted entries (atom | img << 27) in
  // rkey[rstart[a] .. rstart[a+1]); the sites with at least one, ascending
  // (rsite[nrs]); the elements owning them (relem[nre]) with their range
  // rel_start[r] .. rel_start[r+1] in rsite; per step the summed reaction of
  // each such site (rforce [3][nrs]).  Built at every coupled rebuild.
  int nrs = 0, nre = 0;
  long nrev = 0;
  int *rcount = nullptr, *rstart = nullptr, *rsite = nullptr, *relem = nullptr, *rel_start = nullptr;
  int *rsl = nullptr;  // [nrs] site index of each reaction site in its element
  // [nrs] per reaction site: active site a, rstart[a], rstart[a+1] (one load
  // where the reaction pass would chain rsite and rstart)
  int4 *rmeta = nullptr;
  uint32_t *rkey = nullptr;
  double *rforce = nullptr;
  cac::DevBuf b_rcount, b_rstart, b_rkey, b_rsite, b_relem, b_rel_start, b_rforce, b_rtmp, b_rsl, b_rmeta;

  // ---- elements (one type)
  int ne = 0, nae[3] = {0, 0, 0}, nsite = 0;
  double mass_e = 1.0;
  double *nodex = nullptr, *nodev = nullptr, *nodef = nullptr, *nodexhold = nullptr, *xc = nullptr;
  cac::DevBuf b_nodex, b_nodev, b_nodef, b_nodexhold, b_xc;
  double *spa_site = nullptr;  // [nsite][8]
  cac::DevBuf b_spa_site;
  int nint = 0;
  double *spa_int = nullptr, *weight = nullptr;  // [nint][8], [nint]
  cac::DevBuf b_spa_int, b_weight;
  std::vector<double> h_spa_int, h_weight;
  int node_list[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool quad_set = false;
  // element geometry at build
  double *erad = nullptr, *ehull = nullptr;  // erad: [ne] center, radius; ehull: [ne][16] parallelepiped hull
  cac::DevBuf b_erad, b_ehull;
  // element neighbor list, column-major [maxelist][ne]
  int maxelist = 192;
  uint32_t *eneigh = nullptr;
  int *enumb = nullptr;
  cac::DevBuf b_eneigh, b_enumb;
  int ecell[3] = {0, 0, 0};  // 0: box too small for element cells (brute force)
  uint32_t *ebin = nullptr;
  int *ebin_s = nullptr, *ebin_e = nullptr;
  cac::DevBuf b_ekey, b_eval, b_ekey2, b_eval2, b_ecs, b_ece;
  // integration-point neighbor list: raw (j | img<<27 | site<<32) entries
  // from the search (site = kAtomSite for an atom j), then rewritten as
  // particle entries (row of x4 | img<<27: an atom j, or n + active-site id)
  // in the blocked layout; maxnip is a multiple of kLanes
  int maxnip = 128;
  uint64_t *iptmp = nullptr;
  uint32_t *ipneigh = nullptr;
  int *ipnum = nullptr;
  cac::DevBuf b_iptmp, b_ipneigh, b_ipnum;
  // in-range bits of the integration-point rows from the node-density pass
  // (only when every integration point is a node point, ip_on_nodes), per
  // lane of the 8 serving a row
  bool ip_on_nodes = false;
  uint32_t *ipmask = nullptr;
  cac::DevBuf b_ipmask;
  // vertex clusters (cac_set_ip_clusters): node slots 8e+k meeting at a mesh
  // vertex, CSR.  Per build: each cluster's union of list entries
  // (cl_union [ncl][cl_stride], count cl_nu, largest cl_umax) and every
  // integration-point list rewritten as indices into its cluster's union
  // (iplocal, the ipneigh layout); cl_ok when built and the largest union
  // fits the LDS budget of the cluster passes.  aref: the active sites atom
  // lists refer to (coupling), the only ones still interpolated into x4.
  int ncl = 0;
  int *cl_start = nullptr, *cl_slot = nullptr;
  uint32_t *cl_union = nullptr;
  int *cl_nu = nullptr;
  uint16_t *iplocal = nullptr;
  int cl_stride = 0, cl_umax = 0;
  // launch order of the cluster passes: the clusters with unions of more
  // than kUSmall entries (cl_nbig, rows in a cl_scratch slot) at the start of
  // every XCD's share, the rest (largest union cl_umax_small) after.
  // cl_order[o]: launch slot o's cluster, slot start, union size and
  // (scratch slot or -1) * 16 + slot count, one load for a wavefront's start
  int4 *cl_order = nullptr;
  int cl_nbig = 0, cl_umax_small = 0;
  double *cl_scratch = nullptr;
  cac::DevBuf b_cl_scratch;
  // centro-symmetry values of the last cac_centro_atoms call
  double *centro = nullptr;
  cac::DevBuf b_centro;
  // every interpolated site of every element (cac_interpolate_sites)
  double *sitex = nullptr;
  cac::DevBuf b_sitex;
  bool cl_ok = false;
  int naref = 0;
  int *aref = nullptr, *amark = nullptr;
  cac::DevBuf b_cl_start, b_cl_slot, b_cl_union, b_cl_nu, b_iplocal, b_aref, b_amark, b_cl_order;
  // active sites: the (element, site) pairs some list refers to; per step
  // xs4 = (x, y, z, fp) of each, interpolated once
  int nactive = 0, nwords = 0;
  uint32_t *bitmap = nullptr;   // [ne][nwords] referenced-site bits
  int *wprefix = nullptr;       // [ne][nwords] first active id of each word
  int *ecount = nullptr;        // [ne+1] active sites per element
  int *estart = nullptr;        // [ne+1] first active id per element
  int2 *asite = nullptr;        // [nactive] (element, site)
  double *xs4 = nullptr;        // x4 + 4n: [nactive][4]
  cac::DevBuf b_bitmap, b_wprefix, b_ecount, b_estart, b_asite, b_scantmp;
  double *erho = nullptr, *ipf = nullptr, *eip = nullptr, *vip = nullptr;  // [ne*8], [3][nip], [nip], [6][nip]
  cac::DevBuf b_erho, b_ipf, b_eip, b_vip;

  // ---- potential
  int pair = CAC_PAIR_NONE;
  cac::EamTables eam;
  cac::LjParams lj;

  // ---- deformation (cac_set_deform_rate): rate per time unit, edge L0 at
  // the time the rate was set, time elapsed since then
  bool deform_on = false;
  double deform_rate[3] = {0, 0, 0}, deform_l0[3] = {0, 0, 0}, deform_t = 0.0;

  // ---- neighbor settings / state
  double cutneigh = 0.0, triggersq = 0.0;
  int every = 1, delay = 0, check = 1;
  bool neigh_set = false, built = false;
  int ago = 0;
  double build_prod[3] = {1.0, 1.0, 1.0};  // remap scales applied since the last build
  long ncalls = 0;

  // ---- energy / virial of the last force call
  bool have_ev = false;
  double eng[2] = {0, 0}, virial[6] = {0, 0, 0, 0, 0, 0};
  double *red = nullptr;  // reduction scratch [64 blocks][16]
  cac::DevBuf b_red;

  // ---- flags: [0] overflow atoms, [1] max atom count, [2] overflow elist,
  // [3] max elist count, [4] overflow ip, [5] max ip count, [6] rebuild flag,
  // [7..10] first out-of-range list entry (set, row, slot, entry)
  int *dflags = nullptr;
  int *hflags = nullptr;  // pinned mirror
  // cac_run's rebuild check: pinned host flag (host and device addresses)
  // and the event recorded after the kernel that writes it
  volatile int *hcheck = nullptr;
  int *dcheck = nullptr;
  hipEvent_t ev_check = nullptr;
  cac::DevBuf b_dflags;

  // ---- learned element surrogate (k_mlp.hip, k_surrogate.hip)
  cac::MlpState *mlp = nullptr;
  cac::SurrState surr;

  // ---- profiling
  bool profiling = false;
  std::map<std::string, cac::Timer> timers;
};

namespace cac {

// error helpers (return status, store message in ctx)
int fail(cac_ctx *ctx, int code, const std::string &msg);
int hip_fail(cac_ctx *ctx, hipError_t e, const char *where);
#define CAC_HIP(ctx, call)                                     \
  do {                                                         \
    hipError_t e_ = (call);                                    \
    if (e_ != hipSuccess) return ::cac::hip_fail(ctx, e_, #call); \
  } while (0)

template <class T>
int ensure(cac_ctx *ctx, DevBuf &b, size_t count, T **out) {
  size_t need = count * sizeof(T);
  if (need == 0) need = sizeof(T);
  if (b.bytes < need) {
    if (b.p) {
      hipError_t e = hipFree(b.p);
      if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
      b.p = nullptr;
      b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.p, need);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    b.bytes = need;
  }
  *out = static_cast<T *>(b.p);
  return CAC_OK;
}

// Grow b to hold count T, keeping the first keep elements (stream-ordered
// copy into the new allocation).
template <class T>
int ensure_keep(cac_ctx *ctx, DevBuf &b, size_t count, size_t keep, T **out) {
  size_t need = count * sizeof(T);
  if (need == 0) need = sizeof(T);
  if (b.bytes < need) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    if (b.p && keep) {
      e = hipMemcpyAsync(p, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream);
      if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync");
      e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    }
    if (b.p) {
      e = hipFree(b.p);
      if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
    }
    b.p = p;
    b.bytes = need;
  }
  *out = static_cast<T *>(b.p);
  return CAC_OK;
}

// timing of kernel families on the context stream
void timer_begin(cac_ctx *ctx, const char *name, hipEvent_t *start);
void timer_end(cac_ctx *ctx, const char *name, hipEvent_t start);

// ---- launchers (k_neighbor.hip)
int launch_pbc_and_centers(cac_ctx *ctx);
int launch_atom_neighbors(cac_ctx *ctx);
int launch_element_neighbors(cac_ctx *ctx);
int launch_ip_neighbors(cac_ctx *ctx);
int launch_atom_sites(cac_ctx *ctx);
int launch_active_sites(cac_ctx *ctx);
int launch_clusters(cac_ctx *ctx);
int launch_reaction_lists(cac_ctx *ctx);
int validate_lists(cac_ctx *ctx);
int launch_check_displacement(cac_ctx *ctx, double triggersq);

// ---- launchers (k_pair.hip): all passes of one force call, atoms and
// elements interleaved in the order the coupled lists need
int launch_eam(cac_ctx *ctx, int eflag, int vflag, int precision);
int launch_lj(cac_ctx *ctx, int eflag, int vflag, int precision);
int atom_list_site_rows(cac_ctx *ctx);
int launch_centro(cac_ctx *ctx, int nnn, double *dout);
int launch_interp_sites(cac_ctx *ctx, double *dout);

// ---- launchers (cac_api.hip)
int launch_reduce_ev(cac_ctx *ctx, int eflag, int vflag);

// ---- the passes that allocate nothing (k_noalloc.hip): this device's
// address of their fixed scratch array, looked up at context creation
int noalloc_init(cac_ctx *ctx);

// ---- MLP and surrogate (k_mlp.hip, k_surrogate.hip)
void mlp_destroy(cac_ctx *ctx);
int surrogate_classify(cac_ctx *ctx);
int surrogate_forces(cac_ctx *ctx);
void surrogate_destroy(cac_ctx *ctx);

// host table construction (tables.cpp)
int eam_regrid_spline(int nrho_f, double drho_f, int nr_f, double dr_f, const double *frho_f,
                      const double *rhor_f, const double *zr_f, int *nrho, double *drho, int *nr,
                      double *dr, double *frho_spline, double *rhor_spline, double *z2r_spline);

}  // namespace cac
