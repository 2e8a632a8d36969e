// tables.cpp -- host construction of the EAM spline tables, shared by every
// device path so all of them evaluate one set of coefficients.
//
// PairEAM::file2array (pair_eam.cpp:311-523) resamples each funcfl array onto
// the common grid with a 4-point Lagrange stencil; PairEAM::interpolate
// (pair_eam.cpp:552-580) turns the samples into 7 coefficients per knot:
// [6] value, [5] slope from a 5-point difference (one-sided at the ends),
// [4],[3] the cubic's p^2, p^3 terms, [2..0] the derivative polynomial
// divided by the spacing.  Built with -ffp-contract=off: each operation
// rounds where the formula writes it.
#include <cmath>
#include <vector>

#include "cac_internal.h"

namespace cac {
namespace {

// Lagrange resample of a 1-based source array src[1..nsrc] (spacing hsrc)
// onto dst[1..ndst] at r = (m-1)*hdst.
void lagrange_resample(const std::vector<double> &src, int nsrc, double hsrc, int ndst, double hdst,
                       std::vector<double> &dst) {
  const double sixth = 1.0 / 6.0;
  dst.assign(ndst + 1, 0.0);
  for (int m = 1; m <= ndst; ++m) {
    double r = (m - 1) * hdst;
    double p = r / hsrc + 1.0;
    int k = static_cast<int>(p);
    if (k > nsrc - 2) k = nsrc - 2;
    if (k < 2) k = 2;
    p -= k;
    if (p > 2.0) p = 2.0;
    double c1 = -sixth * p * (p - 1.0) * (p - 2.0);
    double c2 = 0.5 * (p * p - 1.0) * (p - 2.0);
    double c3 = -0.5 * p * (p + 1.0) * (p - 2.0);
    double c4 = sixth * p * (p * p - 1.0);
    dst[m] = c1 * src[k - 1] + c2 * src[k] + c3 * src[k + 1] + c4 * src[k + 2];
  }
}

}  // namespace

// Knot coefficients; out is [(n+1)][7], row 0 zero.
void knot_coefficients(const std::vector<double> &f, int n, double delta, double *out) {
  auto c = [out](int m, int j) -> double & { return out[7 * m + j]; };
  for (int j = 0; j < 7; ++j) c(0, j) = 0.0;
  for (int m = 1; m <= n; ++m) c(m, 6) = f[m];
  c(1, 5) = c(2, 6) - c(1, 6);
  c(2, 5) = 0.5 * (c(3, 6) - c(1, 6));
  c(n - 1, 5) = 0.5 * (c(n, 6) - c(n - 2, 6));
  c(n, 5) = c(n, 6) - c(n - 1, 6);
  for (int m = 3; m <= n - 2; ++m)
    c(m, 5) = ((c(m - 2, 6) - c(m + 2, 6)) + 8.0 * (c(m + 1, 6) - c(m - 1, 6))) / 12.0;
  for (int m = 1; m <= n - 1; ++m) {
    double d = c(m + 1, 6) - c(m, 6);
    c(m, 4) = 3.0 * d - 2.0 * c(m, 5) - c(m + 1, 5);
    c(m, 3) = c(m, 5) + c(m + 1, 5) - 2.0 * d;
  }
  c(n, 4) = 0.0;
  c(n, 3) = 0.0;
  for (int m = 1; m <= n; ++m) {
    c(m, 2) = c(m, 5) / delta;
    c(m, 1) = 2.0 * c(m, 4) / delta;
    c(m, 0) = 3.0 * c(m, 3) / delta;
  }
}

int eam_regrid_spline(int nrho_f, double drho_f, int nr_f, double dr_f, const double *frho_f,
                      const double *rhor_f, const double *zr_f, int *nrho, double *drho, int *nr,
                      double *dr, double *frho_spline, double *rhor_spline, double *z2r_spline) {
  if (nrho_f < 5 || nr_f < 5 || !(drho_f > 0) || !(dr_f > 0)) return CAC_ERR_ARG;
  // one funcfl file: grid = file grid, n = (int)(max/delta + 0.5)
  // (pair_eam.cpp:336-343)
  const double rmax = (nr_f - 1) * dr_f;
  const double rhomax = (nrho_f - 1) * drho_f;
  const int n_r = static_cast<int>(rmax / dr_f + 0.5);
  const int n_rho = static_cast<int>(rhomax / drho_f + 0.5);
  if (nrho) *nrho = n_rho;
  if (nr) *nr = n_r;
  if (drho) *drho = drho_f;
  if (dr) *dr = dr_f;
  if (!frho_spline && !rhor_spline && !z2r_spline) return CAC_OK;
  std::vector<double> src, dst;
  auto one = [&](const double *f0, int nsrc, double hsrc, int ndst, double hdst, double *out) {
    if (!out) return;
    src.assign(nsrc + 1, 0.0);
    for (int m = 0; m < nsrc; ++m) src[m + 1] = f0[m];
    lagrange_resample(src, nsrc, hsrc, ndst, hdst, dst);
    knot_coefficients(dst, ndst, hdst, out);
  };
  one(frho_f, nrho_f, drho_f, n_rho, drho_f, frho_spline);
  one(rhor_f, nr_f, dr_f, n_r, dr_f, rhor_spline);
  one(zr_f, nr_f, dr_f, n_r, dr_f, z2r_spline);
  return CAC_OK;
}

}  // namespace cac

extern "C" int cac_eam_tables_from_funcfl(int nrho_f, double drho_f, int nr_f, double dr_f,
                                          const double *frho_f, const double *rhor_f,
                                          const double *zr_f, int *nrho_out, double *drho_out,
                                          int *nr_out, double *dr_out, double *frho_spline,
                                          double *rhor_spline, double *z2r_spline) {
  if ((frho_spline || rhor_spline || z2r_spline) && (!frho_f || !rhor_f || !zr_f)) return CAC_ERR_ARG;
  return cac::eam_regrid_spline(nrho_f, drho_f, nr_f, dr_f, frho_f, rhor_f, zr_f, nrho_out, drho_out,
                                nr_out, dr_out, frho_spline, rhor_spline, z2r_spline);
}

// setfl (LAMMPS eam/alloy, one element): the file's arrays are already on
// the common grid, so they become the knot values as they are (no resample)
// and get the same knot coefficients as a funcfl's resampled arrays.
extern "C" int cac_eam_tables_from_setfl(int nrho, double drho, int nr, double dr, const double *frho,
                                         const double *rhor, const double *z2r, double *frho_spline,
                                         double *rhor_spline, double *z2r_spline) {
  if (nrho < 5 || nr < 5 || !(drho > 0) || !(dr > 0)) return CAC_ERR_ARG;
  if (!frho || !rhor || !z2r || !frho_spline || !rhor_spline || !z2r_spline) return CAC_ERR_ARG;
  std::vector<double> f;
  auto one = [&f](const double *src, int n, double h, double *out) {
    f.assign(n + 1, 0.0);
    for (int m = 0; m < n; ++m) f[m + 1] = src[m];
    cac::knot_coefficients(f, n, h, out);
  };
  one(frho, nrho, drho, frho_spline);
  one(rhor, nr, dr, rhor_spline);
  one(z2r, nr, dr, z2r_spline);
  return CAC_OK;
}
