// pair_lj_cut_hip.cpp -- PairLJCut's coefficients and force calls through
// libcac_hip.so.
#include "pair_lj_cut_hip.h"

#include "cac.h"
#include "error.h"

using namespace CAC_NS;

PairLJCutHip::PairLJCutHip(CAC *cac) : PairLJCut(cac), CacHipHost(), precision(CAC_PREC_FP64) {}

// PairLJCut::init_one (pair_lj_cut.cpp:140-198) computes lj1..lj4 for the
// pair; the 1-1 pair's go to the device (one atom type).
double PairLJCutHip::init_one(int i, int j) {
  const double c = PairLJCut::init_one(i, j);
  if (i == 1 && j == 1) {
    if (!ctx && cac_ctx_create(0, &ctx) != CAC_OK) error->all(FLERR, "cac_ctx_create failed: no HIP device");
    check(error, ctx, cac_set_lj(ctx, lj1[1][1], lj2[1][1], lj3[1][1], lj4[1][1], c * c));
    uploaded = false;  // containers go up at the first force call, after read_data
  }
  return c;
}

void PairLJCutHip::compute(int, int) {}

void PairLJCutHip::ecompute() {
  if (!uploaded) upload(cac, cut[1][1]);
  check(error, ctx, cac_force(ctx, eflag_either, vflag_either, precision));
}
