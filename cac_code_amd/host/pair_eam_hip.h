/* pair_style eam/hip: pair_style eam (PairEAM, pair_eam.h:1-3) with the
 * force-and-neighbor work on the MI355X.  Same settings/coeff arguments and
 * potential file (the DYNAMO-style funcfl of pair_eam.cpp:197-268). */
#ifdef PAIR_CLASS

PairStyle(eam/hip,PairEAMHip)

#else

#ifndef CAC_PAIR_EAM_HIP_H
#define CAC_PAIR_EAM_HIP_H

#include "cac_hip_host.h"
#include "pair_eam.h"

namespace CAC_NS {

class PairEAMHip : public PairEAM, public CacHipHost {
 public:
  PairEAMHip(class CAC *);
  void init_style();
  void compute(int, int);  // both halves run in ecompute's device call
  void ecompute();
  int precision;           // CAC_PREC_FP64 or CAC_PREC_MIXED
};

}  // namespace CAC_NS

#endif
#endif
