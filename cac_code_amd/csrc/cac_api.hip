// cac_api.hip -- C-ABI entry points of libcac_hip.so: context, uploads,
// FixNVE kernels, energy/virial reduction and the device-resident Verlet
// loop.  The force and neighbor kernels live in k_pair.hip and
// k_neighbor.hip.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "cac_internal.h"

namespace cac {

int fail(cac_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(cac_ctx *ctx, hipError_t e, const char *where) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) at %s", (int)e, hipGetErrorString(e), where);
  if (ctx) ctx->err = buf;
  return CAC_ERR_HIP;
}

void timer_begin(cac_ctx *ctx, const char *name, hipEvent_t *start) {
  *start = nullptr;
  if (!ctx->profiling) return;
  if (hipEventCreate(start) != hipSuccess) { *start = nullptr; return; }
  hipEventRecord(*start, ctx->stream);
  (void)name;
}

void timer_end(cac_ctx *ctx, const char *name, hipEvent_t start) {
  if (!ctx->profiling || !start) return;
  hipEvent_t stop;
  if (hipEventCreate(&stop) != hipSuccess) return;
  hipEventRecord(stop, ctx->stream);
  ctx->timers[name].pending.emplace_back(start, stop);
}

namespace {

constexpr int kBlock = 256;
inline unsigned grid_for(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------------- FixNVE
// FixNVE::initial_integrate, fix_nve.cpp:54-98: v += dtfm*f; x += dtv*v.
__global__ void k_nve_atoms_initial(int n, double *__restrict__ x4, double *__restrict__ v,
                                    const double *__restrict__ f, double dtv, double dtfm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) {
    double vv = v[(size_t)d * n + i] + dtfm * f[(size_t)d * n + i];
    v[(size_t)d * n + i] = vv;
    x4[4 * i + d] += dtv * vv;
  }
}

// FixNVE::final_integrate, fix_nve.cpp:181-212.
__global__ void k_nve_atoms_final(int n, double *__restrict__ v, const double *__restrict__ f, double dtfm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) v[(size_t)d * n + i] += dtfm * f[(size_t)d * n + i];
}

// FixNVE::einitial_integrate, fix_nve.cpp:100-157: nodes with the lumped
// mass m_e/8, then the element center as the node mean.
__global__ void k_nve_nodes_initial(int ne, double *__restrict__ nodex, double *__restrict__ nodev,
                                    const double *__restrict__ nodef, double *__restrict__ xc, double dtv,
                                    double dtfm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const size_t b = 24 * (size_t)i;
  double c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) {
      double vv = nodev[b + 3 * k + d] + dtfm * nodef[b + 3 * k + d];
      nodev[b + 3 * k + d] = vv;
      double xx = nodex[b + 3 * k + d] + dtv * vv;
      nodex[b + 3 * k + d] = xx;
    }
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) c[d] += nodex[b + 3 * k + d];
  for (int d = 0; d < 3; ++d) xc[3 * i + d] = c[d] / 8;
}

// FixNVE::efinal_integrate, fix_nve.cpp:216-242.
__global__ void k_nve_nodes_final(long nq, double *__restrict__ nodev, const double *__restrict__ nodef,
                                  double dtfm) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nq) nodev[q] += dtfm * nodef[q];
}

// Step boundary of the device-resident loop: FixNVE::final_integrate of the
// previous step (when fin != 0) and initial_integrate of this one
// (fix_nve.cpp:181-212, then :54-98) in one pass over v, f and x, with the
// roundings of the separate kernels (v += dtfm*f twice, then x += dtv*v).
// When flag != null the same pass runs Neighbor::check_distance
// (neighbor.cpp:1176-1217) on the new positions and stores 1 to *flag (host
// memory the caller waits on) if anything moved more than sqrt(triggersq).
//
// With remap set, the step starts with the affine remap of k_remap (same
// operations) on the positions and the held positions; the velocities do not
// depend on it, so this equals k_remap followed by the step.
struct Scale3 {
  double lo[3], s[3];
};

// The flag goes to host memory; `abort` (device memory) tells the force
// passes the host has already queued for this step that their results will
// be discarded (a rebuild comes first), so they return at once.
__device__ __forceinline__ void raise_flag(int *flag, int *abort) {
  __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void nve_atom_step(int i, int n, double *__restrict__ x4, double *__restrict__ v,
                                              const double *__restrict__ f, double dtv, double dtfm, int fin,
                                              double *__restrict__ xhold, double triggersq, int *flag, int *abort,
                                              int remap, const Scale3 &m, double *__restrict__ xhold_in,
                                              double thr_in, int *__restrict__ ibad) {
  if (i >= n) return;
  double dsq = 0.0, din = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double fd = f[(size_t)d * n + i];
    double vv = v[(size_t)d * n + i];
    if (fin) vv += dtfm * fd;
    vv = vv + dtfm * fd;
    v[(size_t)d * n + i] = vv;
    double x0 = x4[4 * (size_t)i + d], xh = xhold[(size_t)d * n + i];
    double xq = xhold_in ? xhold_in[(size_t)d * n + i] : 0.0;
    if (remap) {
      x0 = m.lo[d] + (x0 - m.lo[d]) * m.s[d];
      xh = m.lo[d] + (xh - m.lo[d]) * m.s[d];
      xhold[(size_t)d * n + i] = xh;
      if (xhold_in) {
        xq = m.lo[d] + (xq - m.lo[d]) * m.s[d];
        xhold_in[(size_t)d * n + i] = xq;
      }
    }
    const double xx = x0 + dtv * vv;
    x4[4 * (size_t)i + d] = xx;
    const double dd = xx - xh;
    dsq = d == 0 ? dd * dd : dsq + dd * dd;
    const double dq = xx - xq;
    din += dq * dq;
  }
  if (flag && dsq > triggersq) raise_flag(flag, abort);
  // the density pass's inner mask (see k_eam_atom_rho) needs motion below thr_in
  if (xhold_in && din > thr_in) __hip_atomic_store(ibad, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Nodes: 8 lanes per element (one per node, 24 contiguous bytes each); the
// center is the node mean summed in node order as in k_nve_nodes_initial.
// Wave-uniform (every lane of the wavefront calls it).
__device__ __forceinline__ void nve_node_step(long t, int ne, double *__restrict__ nodex, double *__restrict__ nodev,
                                              const double *__restrict__ nodef, double *__restrict__ xc, double dtv,
                                              double dtfm, int fin, double *__restrict__ nodexhold, double triggersq,
                                              int *flag, int *abort, int remap, const Scale3 &m) {
  const int i = (int)(t >> 3), k = (int)(t & 7);
  const bool live = i < ne;
  const size_t b = 24 * (size_t)(live ? i : 0) + 3 * k;
  double xn[3] = {0.0, 0.0, 0.0};
  double dsq = 0.0;
  if (live) {
    for (int d = 0; d < 3; ++d) {
      const double fd = nodef[b + d];
      double vv = nodev[b + d];
      if (fin) vv += dtfm * fd;
      vv = vv + dtfm * fd;
      nodev[b + d] = vv;
      double x0 = nodex[b + d], xh = nodexhold[b + d];
      if (remap) {  // the center is recomputed below from the remapped nodes
        x0 = m.lo[d] + (x0 - m.lo[d]) * m.s[d];
        xh = m.lo[d] + (xh - m.lo[d]) * m.s[d];
        nodexhold[b + d] = xh;
      }
      xn[d] = x0 + dtv * vv;
      nodex[b + d] = xn[d];
      const double dd = xn[d] - xh;
      dsq = d == 0 ? dd * dd : dsq + dd * dd;
    }
  }
  for (int d = 0; d < 3; ++d) {
    double c = 0.0;
    for (int kk = 0; kk < 8; ++kk) c += __shfl(xn[d], kk, 8);
    if (live && k == 0) xc[3 * (size_t)i + d] = c / 8;
  }
  if (live && flag && dsq > triggersq) raise_flag(flag, abort);
}

// One launch for both: blocks [0, ablocks) take atoms, the rest take nodes.
struct NveStepArgs {
  int n, ne, ablocks, fin, remap;
  double *x4, *v, *xhold, *xhold_in, *nodex, *nodev, *xc, *nodexhold;
  const double *f, *nodef;
  double dtv, dtfm, dtfm_node, triggersq, thr_in;
  int *flag, *abort, *ibad;
  Scale3 m;
};

__global__ void k_nve_step(NveStepArgs a) {
  if ((int)blockIdx.x < a.ablocks) {
    nve_atom_step((int)(blockIdx.x * blockDim.x + threadIdx.x), a.n, a.x4, a.v, a.f, a.dtv, a.dtfm, a.fin, a.xhold,
                  a.triggersq, a.flag, a.abort, a.remap, a.m, a.xhold_in, a.thr_in, a.ibad);
  } else {
    nve_node_step((long)(blockIdx.x - a.ablocks) * blockDim.x + threadIdx.x, a.ne, a.nodex, a.nodev, a.nodef, a.xc,
                  a.dtv, a.dtfm_node, a.fin, a.nodexhold, a.triggersq, a.flag, a.abort, a.remap, a.m);
  }
}

// Affine remap x -> lo + (x - lo)*scale of atoms, their held positions,
// nodes, their held positions and element centers (one thread per atom or
// node).
__global__ void k_remap(int n, double *__restrict__ x4, double *__restrict__ xhold, int ne,
                        double *__restrict__ nodex, double *__restrict__ nodexhold, double *__restrict__ xc,
                        Scale3 m) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    for (int d = 0; d < 3; ++d) {
      x4[4 * (size_t)t + d] = m.lo[d] + (x4[4 * (size_t)t + d] - m.lo[d]) * m.s[d];
      if (xhold) xhold[(size_t)d * n + t] = m.lo[d] + (xhold[(size_t)d * n + t] - m.lo[d]) * m.s[d];
    }
  } else if (t < (long)n + 8L * ne) {
    const size_t q = (size_t)(t - n);
    for (int d = 0; d < 3; ++d) {
      nodex[3 * q + d] = m.lo[d] + (nodex[3 * q + d] - m.lo[d]) * m.s[d];
      if (nodexhold) nodexhold[3 * q + d] = m.lo[d] + (nodexhold[3 * q + d] - m.lo[d]) * m.s[d];
    }
    if ((q & 7) == 0)
      for (int d = 0; d < 3; ++d) xc[3 * (q >> 3) + d] = m.lo[d] + (xc[3 * (q >> 3) + d] - m.lo[d]) * m.s[d];
  }
}

// ---------------------------------------------------------------- reductions
// Deterministic sums: block b of the grid reduces array b with a fixed
// stride pattern and a fixed LDS tree.
struct RedArgs {
  const double *src[14];
  long len[14];
};

__global__ void k_reduce(RedArgs a, double *__restrict__ out) {
  __shared__ double s[1024];
  const int q = blockIdx.x;
  const double *src = a.src[q];
  const long len = a.len[q];
  double acc = 0.0;
  for (long i = threadIdx.x; i < len; i += blockDim.x) acc += src[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = s[0];
}

// Sum of m*v^2 over the atoms (block 0) and over the nodes (block 1, node
// mass m_e/8): a fixed stride pattern and a fixed LDS tree, so the same
// bits every call.
__global__ void k_mv2(int n, const double *__restrict__ v, double m, long nq, const double *__restrict__ nodev,
                      double mn, double *__restrict__ out) {
  __shared__ double s[1024];
  const bool atoms = blockIdx.x == 0;
  const long len = atoms ? 3L * n : nq;
  const double *src = atoms ? v : nodev;
  double acc = 0.0;
  for (long i = threadIdx.x; i < len; i += blockDim.x) acc += src[i] * src[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = (atoms ? m : mn) * s[0];
}

int upload_tables(cac_ctx *ctx, int nrho, double drho, const double *frho_spline, int nr, double dr,
                  const double *rhor_spline, const double *z2r_spline, double cutforcesq) {
  EamTables &t = ctx->eam;
  const size_t nrow = (size_t)nr + 1, nrrow = (size_t)nrho + 1;
  std::vector<double> rv(nrow * 4), rz(nrow * 8), fr(nrrow * 8, 0.0);
  std::vector<float> rvf(nrow * 4), rzf(nrow * 8), frf(nrrow * 8, 0.0f);
  for (size_t m = 0; m < nrow; ++m)
    for (int c = 0; c < 4; ++c) {
      rv[4 * m + c] = rhor_spline[7 * m + 3 + c];
      rz[8 * m + c] = rhor_spline[7 * m + 3 + c];
      rz[8 * m + 4 + c] = z2r_spline[7 * m + 3 + c];
      rvf[4 * m + c] = (float)rv[4 * m + c];
      rzf[8 * m + c] = (float)rz[8 * m + c];
      rzf[8 * m + 4 + c] = (float)rz[8 * m + 4 + c];
    }
  for (size_t m = 0; m < nrrow; ++m)
    for (int c = 0; c < 7; ++c) {
      fr[8 * m + c] = frho_spline[7 * m + c];
      frf[8 * m + c] = (float)frho_spline[7 * m + c];
    }
  // Knot form of rhor and z2r for the FP64 force passes: per knot m the
  // value and 5-point slope (c6, c5) of each, 32 bytes, from which row m's
  // c4 and c3 are recomputed with array2spline's own operations
  // (pair_eam.cpp:527-548).  Used only when that recomputation reproduces
  // every stored c3, c4 bit for bit (it does for tables built per the
  // reference, as cac_eam_tables_from_funcfl builds them).
  std::vector<double> rk(nrow * 4);
  bool knots = nr >= 3;
  for (size_t m = 0; m < nrow; ++m) {
    rk[4 * m] = rhor_spline[7 * m + 6];
    rk[4 * m + 1] = rhor_spline[7 * m + 5];
    rk[4 * m + 2] = z2r_spline[7 * m + 6];
    rk[4 * m + 3] = z2r_spline[7 * m + 5];
  }
  for (size_t m = 1; knots && m + 1 < nrow; ++m) {
    const double *sp[2] = {rhor_spline, z2r_spline};
    for (int q = 0; q < 2; ++q) {
      const double f0 = sp[q][7 * m + 6], f1 = sp[q][7 * (m + 1) + 6];
      const double s0 = sp[q][7 * m + 5], s1 = sp[q][7 * (m + 1) + 5];
      const double c4 = 3.0 * (f1 - f0) - 2.0 * s0 - s1;
      const double c3 = s0 + s1 - 2.0 * (f1 - f0);
      if (std::memcmp(&c4, &sp[q][7 * m + 4], sizeof(double)) || std::memcmp(&c3, &sp[q][7 * m + 3], sizeof(double)))
        knots = false;
    }
  }
  int rc;
  if ((rc = ensure(ctx, t.b_rk, rk.size(), &t.rk))) return rc;
  CAC_HIP(ctx, hipMemcpy(t.rk, rk.data(), rk.size() * sizeof(double), hipMemcpyHostToDevice));
  t.knots = knots;
  if ((rc = ensure(ctx, t.b_rv, rv.size(), &t.rv)) || (rc = ensure(ctx, t.b_rz, rz.size(), &t.rz)) ||
      (rc = ensure(ctx, t.b_fr, fr.size(), &t.fr)) || (rc = ensure(ctx, t.b_rv_f, rvf.size(), &t.rv_f)) ||
      (rc = ensure(ctx, t.b_rz_f, rzf.size(), &t.rz_f)) || (rc = ensure(ctx, t.b_fr_f, frf.size(), &t.fr_f)))
    return rc;
  CAC_HIP(ctx, hipMemcpy(t.rv, rv.data(), rv.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(t.rz, rz.data(), rz.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(t.fr, fr.data(), fr.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(t.rv_f, rvf.data(), rvf.size() * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(t.rz_f, rzf.data(), rzf.size() * sizeof(float), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(t.fr_f, frf.data(), frf.size() * sizeof(float), hipMemcpyHostToDevice));
  t.nrho = nrho;
  t.nr = nr;
  t.rdrho = 1.0 / drho;
  t.rdr = 1.0 / dr;
  t.cutforcesq = cutforcesq;
  t.rhomax = (nrho - 1) / t.rdrho;
  return CAC_OK;
}

// The density pass's inner mask no longer matches the positions (atoms moved
// by a path without its motion check, a new build or a remap): the next
// density pass is a full one.
int invalidate_inner(cac_ctx *ctx) {
  for (int d = 0; d < 3; ++d) ctx->inner_prod[d] = 1.0;
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 14, 0xFF, sizeof(int), ctx->stream));
  return CAC_OK;
}

int read_flags(cac_ctx *ctx) {
  CAC_HIP(ctx, hipMemcpyAsync(ctx->hflags, ctx->dflags, 8 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int clear_flags(cac_ctx *ctx) {
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags, 0, 8 * sizeof(int), ctx->stream));
  return CAC_OK;
}

int ready_for_force(cac_ctx *ctx) {
  if (!ctx->box_set) return fail(ctx, CAC_ERR_ARG, "cac_set_box not called");
  if (ctx->pair == CAC_PAIR_NONE) return fail(ctx, CAC_ERR_ARG, "no pair style set");
  if (ctx->ne > 0 && !ctx->quad_set) return fail(ctx, CAC_ERR_ARG, "cac_set_quadrature not called");
  if (!ctx->built) return fail(ctx, CAC_ERR_ARG, "neighbor lists not built");
  return CAC_OK;
}

// The tally kernels write energy and virial together when either is asked.
int alloc_ev(cac_ctx *ctx, int eflag, int vflag) {
  if (!eflag && !vflag) return CAC_OK;
  int rc;
  if (ctx->n && ((rc = ensure(ctx, ctx->b_eatom, (size_t)ctx->n, &ctx->eatom)) ||
                 (rc = ensure(ctx, ctx->b_vatom, 6 * (size_t)ctx->n, &ctx->vatom))))
    return rc;
  const size_t nip = (size_t)ctx->ne * ctx->nint;
  if (nip && ((rc = ensure(ctx, ctx->b_eip, nip, &ctx->eip)) || (rc = ensure(ctx, ctx->b_vip, 6 * nip, &ctx->vip))))
    return rc;
  return CAC_OK;
}

int force_impl(cac_ctx *ctx, int eflag, int vflag, int precision) {
  int rc;
  if ((rc = alloc_ev(ctx, eflag, vflag))) return rc;
  if (ctx->pair == CAC_PAIR_EAM) {
    if ((rc = launch_eam(ctx, eflag, vflag, precision))) return rc;
  } else {
    if ((rc = launch_lj(ctx, eflag, vflag, precision))) return rc;
  }
  ctx->have_ev = false;
  if (eflag || vflag) return launch_reduce_ev(ctx, eflag, vflag);
  return CAC_OK;
}

int build_impl(cac_ctx *ctx) {
  int rc;
  // a failed build leaves no usable lists (capacities may have grown)
  ctx->built = false;
  if (!ctx->neigh_set) return fail(ctx, CAC_ERR_ARG, "cac_set_neighbor not called");
  if (!ctx->box_set) return fail(ctx, CAC_ERR_ARG, "cac_set_box not called");
  if (ctx->ne > 0 && !ctx->quad_set) return fail(ctx, CAC_ERR_ARG, "cac_set_quadrature not called");
  if ((rc = launch_pbc_and_centers(ctx))) return rc;
  // regrow-and-retry on capacity overflow (the reference aborts instead,
  // neighbor.cpp:619-620, element_neighbor.cpp:71-72)
  for (int attempt = 0; attempt < 4; ++attempt) {
    if ((rc = clear_flags(ctx))) return rc;
    if ((rc = launch_atom_neighbors(ctx))) return rc;
    if ((rc = launch_element_neighbors(ctx))) return rc;  // synchronizes
    if ((rc = read_flags(ctx))) return rc;
    bool again = false;
    if (ctx->hflags[0]) { ctx->maxneigh = ctx->hflags[1] + 16; again = true; }
    if (ctx->hflags[2]) { ctx->maxelist = ctx->hflags[3] + 16; again = true; }
    if (again) continue;
    if ((rc = launch_ip_neighbors(ctx))) return rc;
    if ((rc = launch_atom_sites(ctx))) return rc;  // coupled: element sites of atoms
    if ((rc = read_flags(ctx))) return rc;
    if (ctx->hflags[0]) { ctx->maxneigh = ctx->hflags[1] + 16; again = true; }
    if (ctx->hflags[4]) { ctx->maxnip = ctx->hflags[5] + 16; again = true; }
    if (again) continue;
    if ((rc = launch_active_sites(ctx))) return rc;
    if ((rc = launch_clusters(ctx))) return rc;
    if ((rc = launch_reaction_lists(ctx))) return rc;
    if (ctx->surr.enabled && (rc = surrogate_classify(ctx))) return rc;
    ctx->built = true;
    ctx->ago = 0;
    ctx->ncalls++;
    for (int d = 0; d < 3; ++d) ctx->build_prod[d] = 1.0;
    // force passes queued after this run in full (see raise_flag)
    CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 15, 0, sizeof(int), ctx->stream));
    return invalidate_inner(ctx);
  }
  return fail(ctx, CAC_ERR_OVERFLOW, "neighbor capacity still exceeded after regrowth");
}

// Rebuild trigger under remaps.  The held positions are remapped with the
// positions, so the check measures motion alone; a compressive remap since
// the build (scales prod_d < 1) brings a pair that was just outside cutneigh
// to cutneigh * min_d prod_d, so each point may move only
// sqrt(triggersq) - eps*cutneigh/2 with eps = max_d (1 - prod_d) before a
// pair could enter the force cutoff unlisted.  Returns that squared
// allowance, or -1 when the remap alone has used it up (rebuild now).
double effective_triggersq(const cac_ctx *ctx) {
  double eps = 0.0;
  for (int d = 0; d < 3; ++d) eps = fmax(eps, 1.0 - ctx->build_prod[d]);
  if (eps <= 0.0) return ctx->triggersq;
  const double trig = sqrt(ctx->triggersq) - 0.5 * eps * ctx->cutneigh;
  return trig > 0.0 ? trig * trig : -1.0;
}

int decide_impl(cac_ctx *ctx, int *rebuild) {
  // Neighbor::decide, neighbor.cpp:1150-1163
  ctx->ago++;
  *rebuild = 0;
  const double trigsq = effective_triggersq(ctx);
  if (trigsq < 0.0) { *rebuild = 1; return CAC_OK; }
  if (ctx->ago >= ctx->delay && ctx->ago % ctx->every == 0) {
    if (!ctx->check) { *rebuild = 1; return CAC_OK; }
    int rc;
    if ((rc = clear_flags(ctx))) return rc;
    if ((rc = launch_check_displacement(ctx, trigsq))) return rc;
    if ((rc = read_flags(ctx))) return rc;
    *rebuild = ctx->hflags[6] ? 1 : 0;
  }
  return CAC_OK;
}

int nve_initial_impl(cac_ctx *ctx, double dtv, double dtf) {
  if (ctx->n) invalidate_inner(ctx);
  if (ctx->n)
    hipLaunchKernelGGL(k_nve_atoms_initial, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->x4,
                       ctx->v, ctx->f, dtv, dtf / ctx->mass);
  if (ctx->ne) {
    double inv_npe = 1 / (double)8;
    hipLaunchKernelGGL(k_nve_nodes_initial, dim3(grid_for(ctx->ne)), dim3(kBlock), 0, ctx->stream, ctx->ne,
                       ctx->nodex, ctx->nodev, ctx->nodef, ctx->xc, dtv, dtf / (inv_npe * ctx->mass_e));
  }
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

// Host-visible check flag (fine-grained pinned memory the check kernel
// stores to) and the event that orders it, created on the first cac_run.
int ensure_step_sync(cac_ctx *ctx) {
  if (ctx->hcheck) return CAC_OK;
  void *p = nullptr;
  CAC_HIP(ctx, hipHostMalloc(&p, 64, hipHostMallocMapped | hipHostMallocCoherent));
  ctx->hcheck = static_cast<volatile int *>(p);
  *ctx->hcheck = 0;
  void *dp = nullptr;
  CAC_HIP(ctx, hipHostGetDevicePointer(&dp, p, 0));
  ctx->dcheck = static_cast<int *>(dp);
  CAC_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_check, hipEventDisableTiming));
  return CAC_OK;
}

// One step boundary of cac_run (k_nve_step); fin: apply the
// previous step's final half-step first; remap: the step's affine remap.
int nve_step_impl(cac_ctx *ctx, double dtv, double dtf, bool fin, int *flag, double triggersq,
                  const Scale3 *remap = nullptr) {
  Scale3 m = {};
  if (remap) m = *remap;
  // inner-mask motion check (k_eam_atom_rho) with the held positions of
  // this atom set.  Remaps since the last invalidation scale pair distances
  // by at most eps = max |prod(s) - 1|: an atom pair outside the inner set
  // (r >= cin) is still beyond cutforce while 2*motion < margin - eps*cin.
  // Otherwise the next density pass is a full one.
  double thr = 0.0;
  if (remap)
    for (int d = 0; d < 3; ++d) ctx->inner_prod[d] *= remap->s[d];
  {
    double eps = 0.0;
    for (int d = 0; d < 3; ++d) eps = fmax(eps, fabs(ctx->inner_prod[d] - 1.0));
    thr = 0.5 * (ctx->inner_margin - eps * ctx->inner_cut);
  }
  const bool inner = thr > 0.0 && ctx->xhold_in && ctx->b_xhold_in.bytes >= 3 * (size_t)ctx->n * sizeof(double);
  if (ctx->n && !inner) invalidate_inner(ctx);
  NveStepArgs a;
  a.n = ctx->n;
  a.ne = ctx->ne;
  a.ablocks = (int)grid_for(ctx->n);
  a.fin = fin ? 1 : 0;
  a.remap = remap ? 1 : 0;
  a.m = m;
  a.x4 = ctx->x4;
  a.v = ctx->v;
  a.f = ctx->f;
  a.xhold = ctx->xhold;
  a.xhold_in = inner ? ctx->xhold_in : nullptr;
  a.nodex = ctx->nodex;
  a.nodev = ctx->nodev;
  a.nodef = ctx->nodef;
  a.xc = ctx->xc;
  a.nodexhold = ctx->nodexhold;
  a.dtv = dtv;
  a.dtfm = ctx->n ? dtf / ctx->mass : 0.0;
  double inv_npe = 1 / (double)8;
  a.dtfm_node = ctx->ne ? dtf / (inv_npe * ctx->mass_e) : 0.0;
  a.triggersq = triggersq;
  a.thr_in = thr * thr;
  a.flag = flag;
  a.abort = ctx->dflags + 15;
  a.ibad = ctx->dflags + 14;
  const unsigned blocks = (unsigned)a.ablocks + grid_for(8L * ctx->ne);
  if (blocks) hipLaunchKernelGGL(k_nve_step, dim3(blocks), dim3(kBlock), 0, ctx->stream, a);
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

int deform_impl(cac_ctx *ctx, const double scale[3]) {
  Scale3 m;
  for (int d = 0; d < 3; ++d) {
    if (!(scale[d] > 0)) return fail(ctx, CAC_ERR_ARG, "deformation scale must be positive");
    m.lo[d] = ctx->lo[d];
    m.s[d] = scale[d];
  }
  const long tot = (long)ctx->n + 8L * ctx->ne;
  if (ctx->n) invalidate_inner(ctx);
  if (tot > 0)
    hipLaunchKernelGGL(k_remap, dim3(grid_for(tot)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->x4, ctx->xhold,
                       ctx->ne, ctx->nodex, ctx->ne ? ctx->nodexhold : nullptr, ctx->xc, m);
  CAC_HIP(ctx, hipGetLastError());
  for (int d = 0; d < 3; ++d) {
    ctx->prd[d] *= scale[d];
    ctx->hi[d] = ctx->lo[d] + ctx->prd[d];
    ctx->build_prod[d] *= scale[d];
  }
  return CAC_OK;
}

// The remap of one cac_run step under a constant strain rate (engineering:
// the edge at elapsed time t is L0*(1 + rate*t)): the box changes here, the
// coordinates in the step kernels (Scale3 m).
int deform_step(cac_ctx *ctx, double dtv, Scale3 *m) {
  ctx->deform_t += dtv;
  double sc[3];
  for (int d = 0; d < 3; ++d) {
    sc[d] = ctx->deform_l0[d] * (1.0 + ctx->deform_rate[d] * ctx->deform_t) / ctx->prd[d];
    if (!(sc[d] > 0)) return fail(ctx, CAC_ERR_ARG, "deformation scale must be positive");
  }
  for (int d = 0; d < 3; ++d) {
    const double scale = sc[d];
    m->lo[d] = ctx->lo[d];
    m->s[d] = scale;
    ctx->prd[d] *= scale;
    ctx->hi[d] = ctx->lo[d] + ctx->prd[d];
    ctx->build_prod[d] *= scale;
  }
  return CAC_OK;
}

int nve_final_impl(cac_ctx *ctx, double dtf) {
  if (ctx->n)
    hipLaunchKernelGGL(k_nve_atoms_final, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->v,
                       ctx->f, dtf / ctx->mass);
  if (ctx->ne) {
    double inv_npe = 1 / (double)8;
    long nq = 24L * ctx->ne;
    hipLaunchKernelGGL(k_nve_nodes_final, dim3(grid_for(nq)), dim3(kBlock), 0, ctx->stream, nq, ctx->nodev,
                       ctx->nodef, dtf / (inv_npe * ctx->mass_e));
  }
  CAC_HIP(ctx, hipGetLastError());
  return CAC_OK;
}

}  // namespace

int launch_reduce_ev(cac_ctx *ctx, int eflag, int vflag) {
  RedArgs a;
  int nq = 0;
  const long n = ctx->n, nip = (long)ctx->ne * ctx->nint;
  // slots: 0 atom energy, 1 IP energy, 2..7 atom virial, 8..13 IP virial
  for (int q = 0; q < 14; ++q) {
    a.src[q] = ctx->x4;
    a.len[q] = 0;
  }
  if (eflag) {
    if (n) { a.src[0] = ctx->eatom; a.len[0] = n; }
    if (nip) { a.src[1] = ctx->eip; a.len[1] = nip; }
  }
  if (vflag)
    for (int c = 0; c < 6; ++c) {
      if (n) { a.src[2 + c] = ctx->vatom + c * n; a.len[2 + c] = n; }
      if (nip) { a.src[8 + c] = ctx->vip + c * nip; a.len[8 + c] = nip; }
    }
  nq = 14;
  int rc;
  if ((rc = ensure(ctx, ctx->b_red, 64 * 16, &ctx->red))) return rc;
  hipLaunchKernelGGL(k_reduce, dim3(nq), dim3(1024), 0, ctx->stream, a, ctx->red + 16);
  double h[14];
  CAC_HIP(ctx, hipMemcpyAsync(h, ctx->red + 16, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->eng[0] = eflag ? h[0] : 0.0;
  ctx->eng[1] = eflag ? h[1] : 0.0;
  for (int c = 0; c < 6; ++c) ctx->virial[c] = vflag ? h[2 + c] + h[8 + c] : 0.0;
  ctx->have_ev = true;
  return CAC_OK;
}

}  // namespace cac

using namespace cac;

// ============================================================== C-ABI
extern "C" {

int cac_ctx_create(int device, cac_ctx **out) {
  if (!out) return CAC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CAC_ERR_HIP;
  if (device < 0 || device >= ndev) return CAC_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return CAC_ERR_HIP;
  cac_ctx *ctx = new cac_ctx();
  ctx->device = device;
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
      ctx->lds_bytes = (size_t)v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0)
      ctx->ncu = v;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return CAC_ERR_HIP;
  }
  if (ensure(ctx, ctx->b_dflags, 16, &ctx->dflags) || hipMemset(ctx->dflags, 0, 16 * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void **)&ctx->hflags, 8 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    cac_ctx_destroy(ctx);
    return CAC_ERR_HIP;
  }
  if (noalloc_init(ctx) != CAC_OK) {
    cac_ctx_destroy(ctx);
    return CAC_ERR_HIP;
  }
  *out = ctx;
  return CAC_OK;
}

int cac_ctx_destroy(cac_ctx *ctx) {
  if (!ctx) return CAC_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  mlp_destroy(ctx);
  surrogate_destroy(ctx);
  DevBuf *bufs[] = {&ctx->b_x4, &ctx->b_v, &ctx->b_f, &ctx->b_xhold, &ctx->b_rho, &ctx->b_eatom, &ctx->b_vatom,
                    &ctx->b_aneigh, &ctx->b_anum, &ctx->b_crank, &ctx->b_akey, &ctx->b_aval, &ctx->b_akey2, &ctx->b_aval2,
                    &ctx->b_ahead, &ctx->b_acells, &ctx->b_cl_scratch, &ctx->b_centro, &ctx->b_sitex, &ctx->b_acs, &ctx->b_ace, &ctx->b_sorttmp, &ctx->b_nodex,
                    &ctx->b_nodev, &ctx->b_nodef, &ctx->b_amask_in, &ctx->b_xhold_in,
                    &ctx->b_nodexhold, &ctx->b_xc, &ctx->b_spa_site, &ctx->b_spa_int, &ctx->b_weight,
                    &ctx->b_erad, &ctx->b_ehull, &ctx->b_eneigh, &ctx->b_enumb, &ctx->b_ekey, &ctx->b_eval, &ctx->b_ekey2,
                    &ctx->b_eval2, &ctx->b_ecs, &ctx->b_ece, &ctx->b_ipneigh, &ctx->b_ipnum, &ctx->b_erho,
                    &ctx->b_iptmp, &ctx->b_bitmap, &ctx->b_wprefix, &ctx->b_ecount, &ctx->b_estart, &ctx->b_asite,
                    &ctx->b_scantmp, &ctx->b_iflag, &ctx->b_ipos, &ctx->b_iface, &ctx->b_asnum, &ctx->b_atmp,
                    &ctx->b_sbase, &ctx->b_sent, &ctx->b_sent2,
                    &ctx->b_etouch, &ctx->b_rcount, &ctx->b_rstart, &ctx->b_rkey, &ctx->b_rsite,
                    &ctx->b_relem, &ctx->b_rel_start, &ctx->b_rforce, &ctx->b_rtmp, &ctx->b_rsl, &ctx->b_rmeta, &ctx->b_amask, &ctx->b_ipmask, &ctx->b_cl_start, &ctx->b_cl_slot,
                    &ctx->b_cl_union, &ctx->b_cl_nu, &ctx->b_iplocal, &ctx->b_aref, &ctx->b_amark, &ctx->b_cl_order,
                    &ctx->b_ipf, &ctx->b_eip, &ctx->b_vip, &ctx->b_red, &ctx->b_dflags, &ctx->eam.b_rz,
                    &ctx->eam.b_fr, &ctx->eam.b_rz_f, &ctx->eam.b_fr_f, &ctx->eam.b_rv, &ctx->eam.b_rv_f, &ctx->eam.b_rk};
  for (DevBuf *b : bufs)
    if (b->p) hipFree(b->p);
  if (ctx->hflags) hipHostFree(ctx->hflags);
  if (ctx->hcheck) hipHostFree(const_cast<int *>(ctx->hcheck));
  if (ctx->ev_check) hipEventDestroy(ctx->ev_check);
  for (auto &kv : ctx->timers)
    for (auto &pr : kv.second.pending) {
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return CAC_OK;
}

const char *cac_last_error(const cac_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cac_synchronize(cac_ctx *ctx) {
  if (!ctx) return CAC_ERR_ARG;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_set_box(cac_ctx *ctx, const double lo[3], const double hi[3], const int periodic[3]) {
  if (!ctx || !lo || !hi) return CAC_ERR_ARG;
  for (int d = 0; d < 3; ++d) {
    if (periodic && !periodic[d]) return fail(ctx, CAC_ERR_ARG, "only fully periodic boxes are supported");
    if (!(hi[d] > lo[d])) return fail(ctx, CAC_ERR_ARG, "box hi must exceed lo");
    ctx->lo[d] = lo[d];
    ctx->hi[d] = hi[d];
    ctx->prd[d] = hi[d] - lo[d];
  }
  ctx->box_set = true;
  ctx->built = false;
  return CAC_OK;
}

int cac_upload_atoms(cac_ctx *ctx, int n, const double *x, const double *v, double mass) {
  if (!ctx || n < 0 || (n > 0 && !x) || !(mass > 0)) return CAC_ERR_ARG;
  if (n >= (1 << 27)) return fail(ctx, CAC_ERR_ARG, "at most 2^27-1 atoms");
  hipSetDevice(ctx->device);
  ctx->n = n;
  ctx->mass = mass;
  int rc;
  if ((rc = ensure(ctx, ctx->b_x4, 4 * (size_t)n, &ctx->x4)) || (rc = ensure(ctx, ctx->b_v, 3 * (size_t)n, &ctx->v)) ||
      (rc = ensure(ctx, ctx->b_f, 3 * (size_t)n, &ctx->f)) || (rc = ensure(ctx, ctx->b_rho, (size_t)n, &ctx->rho)) ||
      (rc = ensure(ctx, ctx->b_xhold, 3 * (size_t)n, &ctx->xhold)))
    return rc;
  std::vector<double> x4((size_t)4 * n), vs((size_t)3 * n, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      x4[4 * (size_t)i + d] = x[3 * (size_t)i + d];
      if (v) vs[(size_t)d * n + i] = v[3 * (size_t)i + d];
    }
    x4[4 * (size_t)i + 3] = 0.0;
  }
  if (n) {
    CAC_HIP(ctx, hipMemcpy(ctx->x4, x4.data(), x4.size() * sizeof(double), hipMemcpyHostToDevice));
    CAC_HIP(ctx, hipMemcpy(ctx->v, vs.data(), vs.size() * sizeof(double), hipMemcpyHostToDevice));
    CAC_HIP(ctx, hipMemset(ctx->f, 0, 3 * (size_t)n * sizeof(double)));
  }
  ctx->built = false;
  return CAC_OK;
}

int cac_upload_elements(cac_ctx *ctx, int ne, const int nae[3], const double *nodex, const double *nodev,
                        double mass_e) {
  if (!ctx || ne < 0 || !nae || (ne > 0 && !nodex) || !(mass_e > 0)) return CAC_ERR_ARG;
  if (ne >= (1 << 27)) return fail(ctx, CAC_ERR_ARG, "at most 2^27-1 elements");
  for (int d = 0; d < 3; ++d)
    if (nae[d] < 2 || nae[d] > 128)
      return fail(ctx, CAC_ERR_ARG, "nae must be in [2, 128] per edge");  // element.cpp:1104-1107
  if ((long)nae[0] * nae[1] * nae[2] >= 65536) return fail(ctx, CAC_ERR_ARG, "at most 65535 sites per element");
  hipSetDevice(ctx->device);
  const bool new_type = (nae[0] != ctx->nae[0] || nae[1] != ctx->nae[1] || nae[2] != ctx->nae[2]);
  if (ne != ctx->ne) {
    // vertex clusters and the surrogate's tables describe the old mesh
    ctx->ncl = 0;
    ctx->cl_ok = false;
    ctx->surr.setup = ctx->surr.enabled = ctx->surr.ready = false;
  }
  ctx->ne = ne;
  for (int d = 0; d < 3; ++d) ctx->nae[d] = nae[d];
  ctx->nsite = nae[0] * nae[1] * nae[2];
  ctx->mass_e = mass_e;
  int rc;
  const size_t nq = 24 * (size_t)ne;
  if ((rc = ensure(ctx, ctx->b_nodex, nq, &ctx->nodex)) || (rc = ensure(ctx, ctx->b_nodev, nq, &ctx->nodev)) ||
      (rc = ensure(ctx, ctx->b_nodef, nq, &ctx->nodef)) || (rc = ensure(ctx, ctx->b_nodexhold, nq, &ctx->nodexhold)) ||
      (rc = ensure(ctx, ctx->b_xc, 3 * (size_t)ne, &ctx->xc)) || (rc = ensure(ctx, ctx->b_erho, 8 * (size_t)ne, &ctx->erho)))
    return rc;
  if (ne) {
    CAC_HIP(ctx, hipMemcpy(ctx->nodex, nodex, nq * sizeof(double), hipMemcpyHostToDevice));
    if (nodev)
      CAC_HIP(ctx, hipMemcpy(ctx->nodev, nodev, nq * sizeof(double), hipMemcpyHostToDevice));
    else
      CAC_HIP(ctx, hipMemset(ctx->nodev, 0, nq * sizeof(double)));
    CAC_HIP(ctx, hipMemset(ctx->nodef, 0, nq * sizeof(double)));
  }
  if (ctx->quad_set && ctx->nint > 0 && (rc = ensure(ctx, ctx->b_ipf, 3 * (size_t)ne * ctx->nint, &ctx->ipf)))
    return rc;
  // site shape functions, Element::inter_ele + shape_array_inter_ele
  // (element.cpp:1361-1423): parent coordinate 2/(n-1)*i - 1, x outermost.
  {
    std::vector<double> spa((size_t)ctx->nsite * 8);
    const double inx = 2 / ((double)nae[0] - 1), iny = 2 / ((double)nae[1] - 1), inz = 2 / ((double)nae[2] - 1);
    size_t s = 0;
    for (int ii = 0; ii < nae[0]; ++ii)
      for (int jj = 0; jj < nae[1]; ++jj)
        for (int kk = 0; kk < nae[2]; ++kk, ++s) {
          const double px = inx * (double)ii - 1, py = iny * (double)jj - 1, pz = inz * (double)kk - 1;
          double *N = &spa[8 * s];
          N[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
          N[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
          N[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
          N[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
          N[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
          N[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
          N[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
          N[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
        }
    if ((rc = ensure(ctx, ctx->b_spa_site, spa.size(), &ctx->spa_site))) return rc;
    CAC_HIP(ctx, hipMemcpy(ctx->spa_site, spa.data(), spa.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (new_type) ctx->quad_set = false;
  ctx->built = false;
  return CAC_OK;
}

int cac_set_quadrature(cac_ctx *ctx, int nint, const int *ijk, const double *w) {
  if (!ctx || nint <= 0 || nint > kMaxNint || !ijk || !w) return CAC_ERR_ARG;
  if (ctx->nae[0] < 2) return fail(ctx, CAC_ERR_ARG, "upload elements before the quadrature");
  hipSetDevice(ctx->device);
  // Element::intp_pcoord (element.cpp:1170-1191) + shape_array_int
  // (element.cpp:1326-1354)
  std::vector<double> spa((size_t)nint * 8);
  const double dx = 2 / ((double)ctx->nae[0] - 1), dy = 2 / ((double)ctx->nae[1] - 1),
               dz = 2 / ((double)ctx->nae[2] - 1);
  for (int g = 0; g < nint; ++g) {
    for (int d = 0; d < 3; ++d)
      if (ijk[3 * g + d] < 0 || ijk[3 * g + d] >= ctx->nae[d])
        return fail(ctx, CAC_ERR_DATA, "integration point lattice index outside the element");
    const double px = dx * ((double)ijk[3 * g]) - 1, py = dy * ((double)ijk[3 * g + 1]) - 1,
                 pz = dz * ((double)ijk[3 * g + 2]) - 1;
    double *N = &spa[8 * (size_t)g];
    N[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
    N[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
    N[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
    N[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
    N[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
    N[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
    N[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
    N[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
  }
  // Element::find_node_list, element.cpp:1296-1319 (TOL 0.01)
  for (int k = 0; k < 8; ++k) {
    int g = 0;
    for (; g < nint; ++g)
      if (std::fabs(spa[8 * (size_t)g + k] - 1.0) < 0.01) break;
    if (g == nint) return fail(ctx, CAC_ERR_DATA, "didn't contain all node integration point");
    ctx->node_list[k] = g;
  }
  // Every integration point is a node point with shape values exactly 0
  // and 1: the density pass then visits every integration point's list from
  // the same position as the force pass, so it can compact the lists.
  {
    bool on_nodes = nint == 8;
    bool seen[8] = {false, false, false, false, false, false, false, false};
    for (int k = 0; on_nodes && k < 8; ++k) {
      const int g = ctx->node_list[k];
      if (seen[g]) on_nodes = false;
      seen[g] = true;
      for (int kk = 0; kk < 8; ++kk)
        if (spa[8 * (size_t)g + kk] != (kk == k ? 1.0 : 0.0)) on_nodes = false;
    }
    ctx->ip_on_nodes = on_nodes;
  }
  int rc;
  if ((rc = ensure(ctx, ctx->b_spa_int, spa.size(), &ctx->spa_int)) ||
      (rc = ensure(ctx, ctx->b_weight, (size_t)nint, &ctx->weight)))
    return rc;
  CAC_HIP(ctx, hipMemcpy(ctx->spa_int, spa.data(), spa.size() * sizeof(double), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(ctx->weight, w, nint * sizeof(double), hipMemcpyHostToDevice));
  ctx->h_spa_int = spa;
  ctx->h_weight.assign(w, w + nint);
  ctx->nint = nint;
  const size_t nip = (size_t)ctx->ne * nint;
  if ((rc = ensure(ctx, ctx->b_ipf, 3 * nip, &ctx->ipf))) return rc;
  ctx->quad_set = true;
  ctx->built = false;
  return CAC_OK;
}

int cac_set_eam(cac_ctx *ctx, int nrho, double drho, const double *frho_spline, int nr, double dr,
                const double *rhor_spline, const double *z2r_spline, double cutforcesq) {
  if (!ctx || nrho < 3 || nr < 3 || !(drho > 0) || !(dr > 0) || !frho_spline || !rhor_spline || !z2r_spline ||
      !(cutforcesq > 0))
    return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  int rc = upload_tables(ctx, nrho, drho, frho_spline, nr, dr, rhor_spline, z2r_spline, cutforcesq);
  if (rc) return rc;
  ctx->pair = CAC_PAIR_EAM;
  return CAC_OK;
}

int cac_set_lj(cac_ctx *ctx, double lj1, double lj2, double lj3, double lj4, double cutsq) {
  if (!ctx || !(cutsq > 0)) return CAC_ERR_ARG;
  ctx->lj.lj1 = lj1;
  ctx->lj.lj2 = lj2;
  ctx->lj.lj3 = lj3;
  ctx->lj.lj4 = lj4;
  ctx->lj.cutsq = cutsq;
  ctx->pair = CAC_PAIR_LJ;
  return CAC_OK;
}

int cac_set_neighbor(cac_ctx *ctx, double cutneigh, double triggersq, int every, int delay, int check, int maxneigh,
                     int maxnip) {
  if (!ctx || !(cutneigh > 0) || triggersq < 0 || every <= 0 || delay < 0) return CAC_ERR_ARG;
  if (delay > 0 && delay % every != 0)
    return fail(ctx, CAC_ERR_ARG, "Neighbor delay must be 0 or multiple of every setting");  // neighbor.cpp:309
  ctx->cutneigh = cutneigh;
  ctx->triggersq = triggersq;
  ctx->every = every;
  ctx->delay = delay;
  ctx->check = check ? 1 : 0;
  if (maxneigh > 0) ctx->maxneigh = maxneigh;
  if (maxnip > 0) ctx->maxnip = maxnip;
  ctx->neigh_set = true;
  ctx->built = false;
  return CAC_OK;
}

int cac_set_list_order(cac_ctx *ctx, double spacing) {
  if (!ctx || !(spacing >= 0.0)) return CAC_ERR_ARG;
  if (spacing != ctx->list_spacing) ctx->built = false;
  ctx->list_spacing = spacing;
  return CAC_OK;
}

int cac_set_coupling(cac_ctx *ctx, int on) {
  if (!ctx) return CAC_ERR_ARG;
  if (ctx->coupled != (on != 0)) ctx->built = false;
  ctx->coupled = on != 0;
  return CAC_OK;
}

int cac_set_ip_clusters(cac_ctx *ctx, int ncl, const int *start, const int *slots) {
  if (!ctx || ncl < 0 || (ncl > 0 && (!start || !slots))) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  ctx->built = false;
  ctx->cl_ok = false;
  ctx->ncl = 0;
  if (ncl == 0) return CAC_OK;
  const long nslot = 8L * ctx->ne;
  if (start[0] != 0 || start[ncl] != nslot)
    return fail(ctx, CAC_ERR_ARG, "clusters must cover the 8 node slots of every element");
  std::vector<char> seen((size_t)nslot, 0);
  for (int c = 0; c < ncl; ++c) {
    const int sz = start[c + 1] - start[c];
    if (sz < 1 || sz > kClMax) return fail(ctx, CAC_ERR_ARG, "a cluster holds 1 to 8 node slots");
    for (int q = start[c]; q < start[c + 1]; ++q) {
      if (slots[q] < 0 || slots[q] >= nslot || seen[slots[q]])
        return fail(ctx, CAC_ERR_ARG, "every node slot must appear in exactly one cluster");
      seen[slots[q]] = 1;
    }
  }
  int rc;
  if ((rc = ensure(ctx, ctx->b_cl_start, (size_t)ncl + 1, &ctx->cl_start)) ||
      (rc = ensure(ctx, ctx->b_cl_slot, (size_t)nslot, &ctx->cl_slot)))
    return rc;
  CAC_HIP(ctx, hipMemcpy(ctx->cl_start, start, ((size_t)ncl + 1) * sizeof(int), hipMemcpyHostToDevice));
  CAC_HIP(ctx, hipMemcpy(ctx->cl_slot, slots, (size_t)nslot * sizeof(int), hipMemcpyHostToDevice));
  ctx->ncl = ncl;
  return CAC_OK;
}

int cac_deform(cac_ctx *ctx, const double scale[3]) {
  if (!ctx || !scale) return CAC_ERR_ARG;
  if (!ctx->box_set) return fail(ctx, CAC_ERR_ARG, "cac_set_box not called");
  hipSetDevice(ctx->device);
  return deform_impl(ctx, scale);
}

int cac_set_deform_rate(cac_ctx *ctx, const double rate[3]) {
  if (!ctx || !rate) return CAC_ERR_ARG;
  if (!ctx->box_set) return fail(ctx, CAC_ERR_ARG, "cac_set_box not called");
  ctx->deform_on = false;
  for (int d = 0; d < 3; ++d) {
    ctx->deform_rate[d] = rate[d];
    ctx->deform_l0[d] = ctx->prd[d];
    if (rate[d] != 0.0) ctx->deform_on = true;
  }
  ctx->deform_t = 0.0;
  return CAC_OK;
}

int cac_cluster_info(cac_ctx *ctx, int *ncl, int *active, int *max_union) {
  if (!ctx) return CAC_ERR_ARG;
  if (ncl) *ncl = ctx->ncl;
  if (active) *active = ctx->cl_ok ? 1 : 0;
  if (max_union) *max_union = ctx->cl_umax;
  return CAC_OK;
}

int cac_build_neighbors(cac_ctx *ctx) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  return build_impl(ctx);
}

int cac_decide(cac_ctx *ctx, int *rebuild) {
  if (!ctx || !rebuild) return CAC_ERR_ARG;
  if (!ctx->built) return fail(ctx, CAC_ERR_ARG, "neighbor lists not built");
  hipSetDevice(ctx->device);
  return decide_impl(ctx, rebuild);
}

int cac_force(cac_ctx *ctx, int eflag, int vflag, int precision) {
  if (!ctx || (precision != CAC_PREC_FP64 && precision != CAC_PREC_MIXED)) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  int rc = ready_for_force(ctx);
  if (rc) return rc;
  // a cac_run that stopped on an error between a raised check flag and
  // its rebuild leaves the skip flag set
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 15, 0, sizeof(int), ctx->stream));
  return force_impl(ctx, eflag, vflag, precision);
}

int cac_nve_initial(cac_ctx *ctx, double dtv, double dtf) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  return nve_initial_impl(ctx, dtv, dtf);
}

int cac_nve_final(cac_ctx *ctx, double dtf) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  return nve_final_impl(ctx, dtf);
}

int cac_run(cac_ctx *ctx, int nsteps, double dtv, double dtf, int precision, int *nbuilds) {
  if (!ctx || nsteps < 0 || (precision != CAC_PREC_FP64 && precision != CAC_PREC_MIXED)) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  int rc = ready_for_force(ctx);
  if (rc) return rc;
  if ((rc = ensure_step_sync(ctx))) return rc;
  // The skip flag (dflags[15]) is zero here and after every rebuild; it is
  // raised with the host check flag and always followed by a rebuild.
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 15, 0, sizeof(int), ctx->stream));
  int builds = 0;
  // Verlet::run, verlet.cpp:149-219.  Same sequence as nve_initial, decide,
  // [build], force, nve_final per step, arranged so the device never waits
  // for the host:
  //  - the final half-step of step s-1 and the initial half-step of step s
  //    are one kernel (the force does not change between them), which also
  //    runs the displacement check when Neighbor::decide would;
  //  - the check's flag goes to host memory, and the host waits for it only
  //    after queueing the step's force on the current lists; when the flag
  //    asks for a rebuild that force is discarded: the lists are rebuilt
  //    and the force recomputed before the final half-step, as in the
  //    reference order.
  for (int step = 0; step < nsteps; ++step) {
    // Neighbor::decide, neighbor.cpp:1150-1163
    ctx->ago++;
    const bool due = ctx->ago >= ctx->delay && ctx->ago % ctx->every == 0;
    Scale3 m;
    if (ctx->deform_on && (rc = deform_step(ctx, dtv, &m))) return rc;
    // a compressive remap can use up the whole skin: then rebuild whatever
    // every/delay say (see effective_triggersq)
    const double trigsq = effective_triggersq(ctx);
    const bool forced = trigsq < 0.0;
    const bool probe = due && ctx->check && !forced;
    if (probe) *ctx->hcheck = 0;
    if ((rc = nve_step_impl(ctx, dtv, dtf, step > 0, probe ? ctx->dcheck : nullptr, trigsq,
                            ctx->deform_on ? &m : nullptr)))
      return rc;
    if (probe) CAC_HIP(ctx, hipEventRecord(ctx->ev_check, ctx->stream));
    if ((due && !ctx->check) || forced) {
      if ((rc = build_impl(ctx))) return rc;
      ++builds;
    }
    if ((rc = force_impl(ctx, 0, 0, precision))) return rc;
    if (probe) {
      CAC_HIP(ctx, hipEventSynchronize(ctx->ev_check));
      if (*ctx->hcheck) {
        if ((rc = build_impl(ctx))) return rc;
        ++builds;
        if ((rc = force_impl(ctx, 0, 0, precision))) return rc;
      }
    }
  }
  if (nsteps > 0 && (rc = nve_final_impl(ctx, dtf))) return rc;
  if (nbuilds) *nbuilds = builds;
  return CAC_OK;
}

int cac_download_atoms(cac_ctx *ctx, double *x, double *v, double *f) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  const int n = ctx->n;
  if (n == 0) return CAC_OK;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (x) {
    std::vector<double> x4((size_t)4 * n);
    CAC_HIP(ctx, hipMemcpy(x4.data(), ctx->x4, x4.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) x[3 * (size_t)i + d] = x4[4 * (size_t)i + d];
  }
  std::vector<double> soa((size_t)3 * n);
  const double *srcs[2] = {ctx->v, ctx->f};
  double *dsts[2] = {v, f};
  for (int q = 0; q < 2; ++q) {
    if (!dsts[q]) continue;
    CAC_HIP(ctx, hipMemcpy(soa.data(), srcs[q], soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) dsts[q][3 * (size_t)i + d] = soa[(size_t)d * n + i];
  }
  return CAC_OK;
}

int cac_download_elements(cac_ctx *ctx, double *nodex, double *nodev, double *nodef) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  const size_t nq = 24 * (size_t)ctx->ne;
  if (nq == 0) return CAC_OK;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nodex) CAC_HIP(ctx, hipMemcpy(nodex, ctx->nodex, nq * sizeof(double), hipMemcpyDeviceToHost));
  if (nodev) CAC_HIP(ctx, hipMemcpy(nodev, ctx->nodev, nq * sizeof(double), hipMemcpyDeviceToHost));
  if (nodef) CAC_HIP(ctx, hipMemcpy(nodef, ctx->nodef, nq * sizeof(double), hipMemcpyDeviceToHost));
  return CAC_OK;
}

int cac_get_energy_virial(cac_ctx *ctx, double eng[2], double virial[6]) {
  if (!ctx) return CAC_ERR_ARG;
  if (!ctx->have_ev) return fail(ctx, CAC_ERR_ARG, "no energy/virial tallied: call cac_force with eflag/vflag");
  if (eng) { eng[0] = ctx->eng[0]; eng[1] = ctx->eng[1]; }
  if (virial)
    for (int c = 0; c < 6; ++c) virial[c] = ctx->virial[c];
  return CAC_OK;
}

int cac_thermo(cac_ctx *ctx, double mvv2e, double boltz, double nktv2p, double out[5]) {
  if (!ctx || !out || !(boltz > 0)) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  int rc = ready_for_force(ctx);
  if (rc) return rc;
  CAC_HIP(ctx, hipMemsetAsync(ctx->dflags + 15, 0, sizeof(int), ctx->stream));
  if ((rc = force_impl(ctx, 1, 1, CAC_PREC_FP64))) return rc;
  if ((rc = ensure(ctx, ctx->b_red, 64 * 16, &ctx->red))) return rc;
  const double inv_npe = 1 / (double)8;
  hipLaunchKernelGGL(k_mv2, dim3(2), dim3(1024), 0, ctx->stream, ctx->n, ctx->v, ctx->mass, 24L * ctx->ne,
                     ctx->nodev, inv_npe * ctx->mass_e, ctx->red);
  CAC_HIP(ctx, hipGetLastError());
  double mv2[2];
  CAC_HIP(ctx, hipMemcpyAsync(mv2, ctx->red, sizeof(mv2), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // ComputeTemp / ComputePressure / thermo pe, ke, etotal as LAMMPS defines
  // them (the reference's computes are empty, compute_pe.cpp:16-49,
  // compute_temp.cpp:15-32, compute_pressure.cpp:18-138): dof = 3 N - 3 over
  // the integrated particles (atoms and nodes)
  const double pe = ctx->eng[0] + ctx->eng[1];
  const double ke = 0.5 * mvv2e * (mv2[0] + mv2[1]);
  const long np = (long)ctx->n + 8L * ctx->ne;
  const double dof = 3.0 * (double)np - 3.0;
  const double temp = dof > 0 ? 2.0 * ke / (dof * boltz) : 0.0;
  const double vol = ctx->prd[0] * ctx->prd[1] * ctx->prd[2];
  const double press = (dof * boltz * temp + ctx->virial[0] + ctx->virial[1] + ctx->virial[2]) / 3.0 / vol * nktv2p;
  out[0] = pe;
  out[1] = ke;
  out[2] = temp;
  out[3] = press;
  out[4] = pe + ke;
  return CAC_OK;
}

int cac_centro_atoms(cac_ctx *ctx, int nnn, double *centro) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  if (nnn != 8 && nnn != 12) return fail(ctx, CAC_ERR_ARG, "centro-symmetry: nnn must be 8 or 12");
  if (!ctx->box_set) return fail(ctx, CAC_ERR_ARG, "cac_set_box not called");
  if (!ctx->built) return fail(ctx, CAC_ERR_ARG, "neighbor lists not built");
  const int n = ctx->n;
  if (n == 0) return CAC_OK;
  if (!centro) return fail(ctx, CAC_ERR_ARG, "centro-symmetry: null output");
  int rc;
  if ((rc = ensure(ctx, ctx->b_centro, (size_t)n, &ctx->centro))) return rc;
  // site rows of coupled lists (the lists of the last build, whatever the
  // coupling switch says now)
  if ((rc = atom_list_site_rows(ctx))) return rc;
  if ((rc = launch_centro(ctx, nnn, ctx->centro))) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(centro, ctx->centro, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_interpolate_sites(cac_ctx *ctx, double *x) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  const size_t total = (size_t)ctx->ne * ctx->nsite;
  if (total == 0) return CAC_OK;
  if (!x) return fail(ctx, CAC_ERR_ARG, "interpolate sites: null output");
  int rc;
  if ((rc = ensure(ctx, ctx->b_sitex, 3 * total, &ctx->sitex))) return rc;
  if ((rc = launch_interp_sites(ctx, ctx->sitex))) return rc;
  CAC_HIP(ctx, hipMemcpyAsync(x, ctx->sitex, 3 * total * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_get_capacities(cac_ctx *ctx, int *maxneigh, int *maxnip) {
  if (!ctx) return CAC_ERR_ARG;
  if (maxneigh) *maxneigh = ctx->maxneigh;
  if (maxnip) *maxnip = ctx->maxnip;
  return CAC_OK;
}

}  // extern "C"

namespace {
// Blocked device rows of particle entries -> row-major host rows of the
// ABI's uint64 entries (atom j, or element/site of an active site).
int download_rows(cac_ctx *ctx, const uint32_t *dev, const int *dnum, size_t rows, int M, int group, int cap,
                  int *numneigh, uint64_t *neigh) {
  std::vector<uint32_t> cm((size_t)M * round_up((int)rows, group));
  std::vector<int2> as((size_t)ctx->nactive);
  CAC_HIP(ctx, hipMemcpy(numneigh, dnum, rows * sizeof(int), hipMemcpyDeviceToHost));
  CAC_HIP(ctx, hipMemcpy(cm.data(), dev, cm.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (ctx->nactive)
    CAC_HIP(ctx, hipMemcpy(as.data(), ctx->asite, as.size() * sizeof(int2), hipMemcpyDeviceToHost));
  const uint32_t n = (uint32_t)ctx->n;
  for (size_t p = 0; p < rows; ++p)
    for (int k = 0; k < numneigh[p] && k < M; ++k) {
      const uint32_t e = group == kAtomGroup ? cm[atom_slot((int)p, k, M)]
                                             : cm[((p / group) * M + k) * group + (p % group)];
      const uint32_t idx = e & 0x7FFFFFFu;
      const uint64_t img = (uint64_t)(e >> 27) << 27;
      if (idx < n) {
        neigh[p * cap + k] = (uint64_t)idx | img | ((uint64_t)CAC_ATOM_SITE << 32);
      } else {
        if (idx - n >= (uint32_t)as.size())
          return fail(ctx, CAC_ERR_DATA, "list entry " + std::to_string(e) + " beyond the active sites");
        const int2 js = as[idx - n];
        neigh[p * cap + k] = (uint64_t)js.x | img | ((uint64_t)js.y << 32);
      }
    }
  return CAC_OK;
}
}  // namespace

extern "C" {

int cac_get_atom_neighbors(cac_ctx *ctx, int maxneigh, int *numneigh, uint64_t *neigh) {
  if (!ctx || !numneigh || !neigh) return CAC_ERR_ARG;
  if (!ctx->built) return fail(ctx, CAC_ERR_ARG, "neighbor lists not built");
  if (maxneigh < ctx->maxneigh) return fail(ctx, CAC_ERR_ARG, "capacity below the device list stride");
  hipSetDevice(ctx->device);
  if (ctx->n == 0) return CAC_OK;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return download_rows(ctx, ctx->aneigh, ctx->anum, (size_t)ctx->n, ctx->maxneigh, kAtomGroup, maxneigh, numneigh,
                       neigh);
}

int cac_get_ip_neighbors(cac_ctx *ctx, int maxnip, int *numneigh, uint64_t *neigh) {
  if (!ctx || !numneigh || !neigh) return CAC_ERR_ARG;
  if (!ctx->built) return fail(ctx, CAC_ERR_ARG, "neighbor lists not built");
  if (maxnip < ctx->maxnip) return fail(ctx, CAC_ERR_ARG, "capacity below the device list stride");
  hipSetDevice(ctx->device);
  const size_t nip = (size_t)ctx->ne * ctx->nint;
  if (nip == 0) return CAC_OK;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return download_rows(ctx, ctx->ipneigh, ctx->ipnum, nip, ctx->maxnip, kRowGroup, maxnip, numneigh, neigh);
}

int cac_get_rho(cac_ctx *ctx, double *rho, double *fp) {
  if (!ctx) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  const int n = ctx->n;
  if (n == 0) return CAC_OK;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (rho) CAC_HIP(ctx, hipMemcpy(rho, ctx->rho, n * sizeof(double), hipMemcpyDeviceToHost));
  if (fp) {
    std::vector<double> x4((size_t)4 * n);
    CAC_HIP(ctx, hipMemcpy(x4.data(), ctx->x4, x4.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) fp[i] = x4[4 * (size_t)i + 3];
  }
  return CAC_OK;
}

int cac_get_erho(cac_ctx *ctx, double *erho) {
  if (!ctx || !erho) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  if (ctx->ne == 0) return CAC_OK;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  CAC_HIP(ctx, hipMemcpy(erho, ctx->erho, 8 * (size_t)ctx->ne * sizeof(double), hipMemcpyDeviceToHost));
  return CAC_OK;
}

int cac_get_ip_forces(cac_ctx *ctx, double *ipf) {
  if (!ctx || !ipf) return CAC_ERR_ARG;
  hipSetDevice(ctx->device);
  const size_t nip = (size_t)ctx->ne * ctx->nint;
  if (nip == 0) return CAC_OK;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<double> soa(3 * nip);
  CAC_HIP(ctx, hipMemcpy(soa.data(), ctx->ipf, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t p = 0; p < nip; ++p)
    for (int d = 0; d < 3; ++d) ipf[3 * p + d] = soa[d * nip + p];
  return CAC_OK;
}

int cac_profile_enable(cac_ctx *ctx, int on) {
  if (!ctx) return CAC_ERR_ARG;
  hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->timers) {
    for (auto &pr : kv.second.pending) {
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
  }
  ctx->timers.clear();
  ctx->profiling = on != 0;
  return CAC_OK;
}

double cac_profile_avg_ms(cac_ctx *ctx, const char *name, int *count) {
  if (count) *count = 0;
  if (!ctx || !name) return -1.0;
  auto it = ctx->timers.find(name);
  if (it == ctx->timers.end()) return -1.0;
  Timer &t = it->second;
  hipStreamSynchronize(ctx->stream);
  for (auto &pr : t.pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      t.total_ms += ms;
      t.count++;
    }
    hipEventDestroy(pr.first);
    hipEventDestroy(pr.second);
  }
  t.pending.clear();
  if (count) *count = t.count;
  return t.count ? t.total_ms / t.count : -1.0;
}

}  // extern "C"
