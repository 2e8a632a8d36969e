// cac_hip_host.cpp -- the reference containers <-> one libcac_hip context.
#include "cac_hip_host.h"

#include <vector>

#include "atom.h"
#include "cac.h"
#include "domain.h"
#include "element.h"
#include "error.h"
#include "neighbor.h"

using namespace CAC_NS;

CacHipHost::CacHipHost() : ctx(NULL), uploaded(false) {}

CacHipHost::~CacHipHost() {
  if (ctx) cac_ctx_destroy(ctx);
}

void CacHipHost::check(Error *error, cac_ctx *c, int rc) {
  if (rc != CAC_OK) error->all(FLERR, cac_last_error(c));
}

void CacHipHost::upload(CAC *cac, double rc) {
  Error *error = cac->error;
  Atom *atom = cac->atom;
  Element *element = cac->element;
  if (!ctx && cac_ctx_create(0, &ctx) != CAC_OK) error->all(FLERR, "cac_ctx_create failed: no HIP device");
  if (atom->ntypes > 1 || element->ntypes > 1)
    error->all(FLERR, "The HIP path takes one atom type and one element type");
  int per[3] = {cac->domain->xperiodic, cac->domain->yperiodic, cac->domain->zperiodic};
  check(error, ctx, cac_set_box(ctx, cac->domain->boxlo, cac->domain->boxhi, per));
  const int n = atom->nlocal;
  check(error, ctx, cac_upload_atoms(ctx, n, n ? &atom->x[0][0] : NULL, n && atom->v ? &atom->v[0][0] : NULL,
                                     atom->mass[1]));
  const int ne = element->nlocal;
  int nae[3] = {element->nae[1][0], element->nae[1][1], element->nae[1][2]};
  check(error, ctx, cac_upload_elements(ctx, ne, nae, ne ? &element->nodex[0][0][0] : NULL,
                                        ne && element->nodev ? &element->nodev[0][0][0] : NULL, element->mass[1]));
  const int nint = element->numint[1];
  std::vector<int> ijk(3 * (size_t)nint);
  for (int g = 0; g < nint; g++)
    for (int d = 0; d < 3; d++) ijk[3 * g + d] = element->intp[0][g][d];   // Intp section, element.cpp:1144-1166
  if (ne) check(error, ctx, cac_set_quadrature(ctx, nint, ijk.data(), element->weight[0]));
  Neighbor *neighbor = cac->neighbor;
  const double cn = neighbor->cutneighmax, skin = cn - rc;
  if (skin <= 0.0) error->all(FLERR, "Neighbor cutoff must exceed the force cutoff");
  check(error, ctx, cac_set_neighbor(ctx, cn, 0.25 * skin * skin, neighbor->every, neighbor->delay,
                                     neighbor->dist_check, 0, 0));
  check(error, ctx, cac_build_neighbors(ctx));
  uploaded = true;
}

void CacHipHost::download(CAC *cac) {
  Atom *atom = cac->atom;
  Element *element = cac->element;
  check(cac->error, ctx, cac_download_atoms(ctx, atom->nlocal ? &atom->x[0][0] : NULL,
                                            atom->nlocal && atom->v ? &atom->v[0][0] : NULL,
                                            atom->nlocal && atom->f ? &atom->f[0][0] : NULL));
  check(cac->error, ctx, cac_download_elements(ctx, element->nlocal ? &element->nodex[0][0][0] : NULL,
                                               element->nlocal ? &element->nodev[0][0][0] : NULL,
                                               element->nlocal ? &element->nodef[0][0][0] : NULL));
}
