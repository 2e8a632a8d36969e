/*
 * cac_hip_host.h -- what the reference's C++ host needs to drive
 * libcac_hip.so: one device context per pair style, filled from the
 * reference's own containers (Domain, Atom, Element, Neighbor), and the
 * mapping of a non-zero status to the reference's Error::all
 * (error.cpp:72-91).  Used by pair_eam_hip, pair_lj_cut_hip and
 * fix_nve_hip; compiled into the reference's binary by adding these files
 * to its source list (Makefile:11-30) and linking libcac_hip.so.
 *
 * One atom type and one element type, as the reference's element path
 * assumes (pair_eam.cpp:980, 1047, 1091; pair_lj_cut.cpp:327-331): set-up
 * fails with a clear error for more.
 */
#ifndef CAC_HIP_HOST_H
#define CAC_HIP_HOST_H

#include "cac_hip.h"
#include "pointers.h"

namespace CAC_NS {

class CacHipHost {
 public:
  cac_ctx *ctx;
  bool uploaded;

  CacHipHost();
  virtual ~CacHipHost();

  // status -> Error::all with the library's message
  static void check(class Error *error, cac_ctx *ctx, int rc);

  // box, atoms, elements, integration points and neighbor settings from the
  // reference's containers (domain.h:40, atom.h:40, element.h:25-67,
  // neighbor.h:12-24); rc is the force cutoff (the trigger is half the skin
  // cutneighmax - rc, neighbor.cpp:334, 357-358).  Then the first build.
  void upload(class CAC *cac, double rc);

  // device results back into the reference's containers
  void download(class CAC *cac);
};

}  // namespace CAC_NS

#endif
