// pair_eam_hip.cpp -- PairEAM's tables and force calls through libcac_hip.so.
#include "pair_eam_hip.h"

#include <vector>

#include "cac.h"
#include "error.h"

using namespace CAC_NS;

PairEAMHip::PairEAMHip(CAC *cac) : PairEAM(cac), CacHipHost(), precision(CAC_PREC_FP64) {}

// PairEAM::init_style (pair_eam.cpp:274-283) builds the spline tables with
// file2array / array2spline; here the library builds the same tables bit for
// bit from the funcfl arrays (1-based, pair_eam.cpp:231-245) and uploads
// them with the reference's containers.
void PairEAMHip::init_style() {
  PairEAM::init_style();
  if (nfuncfl != 1) error->all(FLERR, "pair_style eam/hip takes one funcfl file");
  Funcfl *f = &funcfl[0];
  if (!ctx && cac_ctx_create(0, &ctx) != CAC_OK) error->all(FLERR, "cac_ctx_create failed: no HIP device");
  int nrho_t, nr_t;
  double drho_t, dr_t;
  check(error, ctx, cac_eam_tables_from_funcfl(f->nrho, f->drho, f->nr, f->dr, &f->frho[1], &f->rhor[1],
                                               &f->zr[1], &nrho_t, &drho_t, &nr_t, &dr_t, NULL, NULL, NULL));
  std::vector<double> fs(7 * ((size_t)nrho_t + 1)), rs(7 * ((size_t)nr_t + 1)), zs(7 * ((size_t)nr_t + 1));
  check(error, ctx, cac_eam_tables_from_funcfl(f->nrho, f->drho, f->nr, f->dr, &f->frho[1], &f->rhor[1],
                                               &f->zr[1], &nrho_t, &drho_t, &nr_t, &dr_t, fs.data(), rs.data(),
                                               zs.data()));
  check(error, ctx, cac_set_eam(ctx, nrho_t, drho_t, fs.data(), nr_t, dr_t, rs.data(), zs.data(),
                                f->cut * f->cut));
  uploaded = false;  // containers go up at the first force call, after read_data
}

void PairEAMHip::compute(int, int) {}

void PairEAMHip::ecompute() {
  if (!uploaded) upload(cac, funcfl[0].cut);
  check(error, ctx, cac_force(ctx, eflag_either, vflag_either, precision));
}
