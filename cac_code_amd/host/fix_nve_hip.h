/* fix nve/hip: fix nve (FixNVE, fix_nve.h:1-3) for atoms and element nodes
 * on the device, through the context of a HIP pair style (eam/hip or
 * lj/cut/hip). */
#ifdef FIX_CLASS

FixStyle(nve/hip,FixNVEHip)

#else

#ifndef CAC_FIX_NVE_HIP_H
#define CAC_FIX_NVE_HIP_H

#include "fix_nve.h"

namespace CAC_NS {

class FixNVEHip : public FixNVE {
 public:
  FixNVEHip(class CAC *, int, char **);
  void init();
  void initial_integrate(int);  // atoms and nodes (fix_nve.cpp:54-157)
  void einitial_integrate(int) {}
  void final_integrate();       // atoms and nodes (fix_nve.cpp:181-242)
  void efinal_integrate() {}

 protected:
  class CacHipHost *host;
};

}  // namespace CAC_NS

#endif
#endif
