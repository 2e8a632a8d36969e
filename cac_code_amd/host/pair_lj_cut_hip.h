/* pair_style lj/cut/hip: pair_style lj/cut (PairLJCut, pair_lj_cut.h:1-3)
 * with the force-and-neighbor work on the MI355X.  Same settings (global
 * cutoff) and coeff (I J eps sigma [cut]) arguments, pair_lj_cut.cpp:74-121. */
#ifdef PAIR_CLASS

PairStyle(lj/cut/hip,PairLJCutHip)

#else

#ifndef CAC_PAIR_LJ_CUT_HIP_H
#define CAC_PAIR_LJ_CUT_HIP_H

#include "cac_hip_host.h"
#include "pair_lj_cut.h"

namespace CAC_NS {

class PairLJCutHip : public PairLJCut, public CacHipHost {
 public:
  PairLJCutHip(class CAC *);
  double init_one(int, int);
  void compute(int, int);  // both halves run in ecompute's device call
  void ecompute();
  int precision;           // CAC_PREC_FP64 or CAC_PREC_MIXED
};

}  // namespace CAC_NS

#endif
#endif
