// fix_nve_hip.cpp -- FixNVE's half-steps on the device context of the HIP
// pair style.
#include "fix_nve_hip.h"

#include "cac.h"
#include "cac_hip_host.h"
#include "error.h"
#include "force.h"
#include "pair.h"

using namespace CAC_NS;

FixNVEHip::FixNVEHip(CAC *cac, int narg, char **arg) : FixNVE(cac, narg, arg), host(NULL) {}

void FixNVEHip::init() {
  FixNVE::init();  // dtv, dtf = 0.5*dt*ftm2v (fix_nve.cpp:40-48)
  host = dynamic_cast<CacHipHost *>(force->pair);
  if (!host) error->all(FLERR, "fix nve/hip needs pair_style eam/hip or lj/cut/hip");
}

void FixNVEHip::initial_integrate(int) {
  CacHipHost::check(error, host->ctx, cac_nve_initial(host->ctx, dtv, dtf));
}

void FixNVEHip::final_integrate() {
  CacHipHost::check(error, host->ctx, cac_nve_final(host->ctx, dtf));
}
