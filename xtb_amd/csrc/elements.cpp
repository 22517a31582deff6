// Element tables of GFN2-xTB after the transforms xtb applies at load time.
//
// initGFN2 (src/xtb/gfn2.f90:731-967): shell hardness, third-order shell
// parameters, valence shells, reference occupations, number of primitives.
// slaterToGauss (src/slater.f90:373-483): STO-nG exponents and normalised
// contraction coefficients.  newD4Model (src/disp/dftd4.F90:106-203): reference
// polarisabilities alpha(iw) in the GFN2 charge mode and their C6 integrals.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "gfn2_common.h"
#include "gfn2_params.h"

namespace xtbamd {
namespace {

Element g_el[XTB_MAX_Z + 1];
// alpha(iw) per element and reference, filled for Z <= 86
double g_alpha[XTB_MAX_Z + 1][XTB_MAX_REF][XTB_NFREQ];
std::once_flag g_once;

double ipow(double x, int l) {
  double r = 1.0;
  for (int i = 0; i < l; ++i) r *= x;
  return r;
}

// valence charge (src/type/molecule.f90:560-600)
double valence_charge(int z) {
  int nc;
  if (z <= 2) nc = 0;
  else if (z <= 10) nc = 2;
  else if (z <= 18) nc = 10;
  else if (z <= 29) nc = 18;
  else if (z <= 36) nc = 28;
  else if (z <= 47) nc = 36;
  else if (z <= 54) nc = 46;
  else if (z <= 71) nc = 54;
  else if (z <= 79) nc = 68;
  else nc = 78;
  double v = double(z) - double(nc);
  if (z > 57 && z < 72) v = 3.0;
  return v;
}

int sto_ng(int ng, int n, int l, double zeta, double* alpha, double* coeff) {
  if ((n > 5 || n <= l) && !(n == 6 && ng == 6)) return 2;
  int ityp = (l == 0 ? n : l == 1 ? 4 + n : l == 2 ? 7 + n : l == 3 ? 9 + n : 10 + n) - 1;
  const double z2 = zeta * zeta;
  const double* pa = nullptr;
  const double* pc = nullptr;
  switch (ng) {
    case 1: alpha[0] = k_sto_alpha1[ityp] * z2; coeff[0] = 1.0; break;
    case 2: pa = k_sto_alpha2; pc = k_sto_coeff2; break;
    case 3: pa = k_sto_alpha3; pc = k_sto_coeff3; break;
    case 4: pa = k_sto_alpha4; pc = k_sto_coeff4; break;
    case 5: pa = k_sto_alpha5; pc = k_sto_coeff5; break;
    case 6:
      if (n == 6) {
        const double* a6 = l == 0 ? k_sto_alpha6s : k_sto_alpha6p;
        const double* c6 = l == 0 ? k_sto_coeff6s : k_sto_coeff6p;
        for (int k = 0; k < 6; ++k) { alpha[k] = a6[k] * z2; coeff[k] = c6[k]; }
      } else {
        pa = k_sto_alpha6; pc = k_sto_coeff6;
      }
      break;
    default: return 1;
  }
  if (pa)
    for (int k = 0; k < ng; ++k) { alpha[k] = pa[ityp * ng + k] * z2; coeff[k] = pc[ityp * ng + k]; }
  static const double dfact[4] = {1.0, 1.0, 3.0, 15.0};
  const double top = 2.0 / kPi;
  for (int k = 0; k < ng; ++k)
    coeff[k] = coeff[k] * std::pow(top * alpha[k], 0.75) * ipow(std::sqrt(4 * alpha[k]), l) / std::sqrt(dfact[l]);
  return 0;
}

double zeta_fn(double a, double c, double qref, double qmod) {
  if (qmod < 0.0) return std::exp(a);
  return std::exp(a * (1.0 - std::exp(c * (1.0 - qref / qmod))));
}

void build() {
  std::memset(g_el, 0, sizeof(g_el));
  for (int z = 1; z <= XTB_MAX_Z; ++z) {
    Element& e = g_el[z];
    const int zi = z - 1;
    e.nsh = k_n_shell[zi];
    bool seen[4] = {false, false, false, false};
    for (int m = 0; m < e.nsh; ++m) {
      const int l = k_ang_shell[m + 3 * zi];
      const int pqn = k_pqn_shell[m + 3 * zi];
      e.l[m] = l;
      const bool valence = !seen[l];
      seen[l] = true;
      e.refocc[m] = valence ? k_reference_occ_l[l + 3 * zi] : 0.0;
      e.selfe[m] = k_self_energy[m + 3 * zi];
      e.zeta[m] = k_slater_exponent[m + 3 * zi];
      e.kcn[m] = k_kcn_l[l + 4 * zi];
      e.hard[m] = k_chemical_hardness_gfn2[zi] * (1.0 + k_shell_hardness_l[l + 3 * zi]);
      const double g3 = (l == 0) ? 1.0 : (l == 1 ? 0.5 : 0.25);
      e.third[m] = (k_third_order_atom[zi] * K_THIRD_ORDER_ATOM_MULT) * g3;
      int np = 0;
      if (z <= 2) np = (l == 0) ? 3 : (l == 1 ? 4 : 0);
      else if (l <= 1) np = pqn > 5 ? 6 : 4;
      else if (l == 2) np = 3;
      else np = 4;
      e.nprim[m] = np;
      sto_ng(np, pqn, l, e.zeta[m], e.alp[m], e.cc[m]);
    }
    for (int l = 0; l < 4; ++l) e.shpoly[l] = k_shell_poly_l[l + 4 * zi];
    e.rad = kAAtoAU * k_atomic_rad_aa[zi];
    e.en = k_pauling_en[zi];
    e.dipk = k_dip_kernel_raw[zi] * K_DIP_KERNEL_MULT;
    e.quadk = k_quad_kernel_raw[zi] * K_QUAD_KERNEL_MULT;
    e.valcn = k_valence_cn[zi];
    e.mrad = k_multi_rad[zi];
    e.rep_alpha = k_rep_alpha[zi];
    e.rep_zeff = k_rep_zeff[zi];
    e.zval = valence_charge(z);
    e.rcov_d3 = ((k_covalent_rad_d3_aa[zi] * kAAtoAU) * 4.0) / 3.0;
    e.rcov_nc = ((4.0 / 3.0) * k_ncoord_rad_aa[zi]) / 0.52917726;
    e.en_nc = k_ncoord_en[zi];
    e.sqrtzr4r2 = std::sqrt(0.5 * (k_r4_over_r2[zi] * std::sqrt(double(z))));
    e.d4zeff = k_d4_zeff[zi];
    e.d4gam = k_chemical_hardness_d4[zi];
    e.eeq_en = k_eeq_en[zi];
    e.eeq_gam = k_eeq_gam[zi];
    e.eeq_kappa = k_eeq_kappa[zi];
    e.eeq_alpha = k_eeq_alpha[zi];
    // D4 references in the GFN2 charge mode (secq = 0)
    int cncount[19];
    std::memset(cncount, 0, sizeof(cncount));
    cncount[0] = 1;
    e.nref = k_d4_refn[zi];
    for (int j = 0; j < e.nref; ++j) {
      const int is = k_d4_refsys[zi * D4_MAX_REF + j];
      const double iz = k_d4_zeff[is - 1];
      const double zf = zeta_fn(kD4ga, k_chemical_hardness_d4[is - 1] * kD4gc, 0.0 + iz,
                                k_d4_refh[zi * D4_MAX_REF + j] + iz);
      e.refcn[j] = k_d4_refcovcn[zi * D4_MAX_REF + j];
      e.refq[j] = k_d4_refq[zi * D4_MAX_REF + j];
      cncount[(int)std::lround(k_d4_refcn[zi * D4_MAX_REF + j])] += 1;
      for (int f = 0; f < XTB_NFREQ; ++f) {
        const double sec = k_d4_sscale[is - 1] * k_d4_secaiw[(is - 1) * XTB_NFREQ + f] * zf;
        const double v = k_d4_ascale[zi * D4_MAX_REF + j] *
                         (k_d4_alphaiw[(zi * D4_MAX_REF + j) * XTB_NFREQ + f] - k_d4_hcount[zi * D4_MAX_REF + j] * sec);
        g_alpha[z][j][f] = v > 0.0 ? v : 0.0;
      }
    }
    for (int j = 0; j < e.nref; ++j) {
      const int icn = cncount[(int)std::lround(k_d4_refcn[zi * D4_MAX_REF + j])];
      e.ncount[j] = icn * (icn + 1) / 2;
    }
  }
}

// trapezoidal Casimir-Polder weights (trapzd: src/disp/dftd4.F90:395-446)
double trapzd(const double* pol) {
  static const double fr[23] = {0.000001, 0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0,
                                1.2, 1.4, 1.6, 1.8, 2.0, 2.5, 3.0, 4.0, 5.0, 7.5, 10.0};
  double w[23];
  w[0] = 0.5 * (fr[1] - fr[0]);
  for (int i = 1; i < 22; ++i) w[i] = 0.5 * ((fr[i] - fr[i - 1]) + (fr[i + 1] - fr[i]));
  w[22] = 0.5 * (fr[22] - fr[21]);
  double s = 0.0;
  for (int i = 0; i < 23; ++i) s += pol[i] * w[i];
  return s;
}

}  // namespace

const Element* host_elements() {
  std::call_once(g_once, build);
  return g_el;
}

double host_ref_c6(int za, int ra, int zb, int rb) {
  host_elements();
  double a[23];
  for (int f = 0; f < 23; ++f) a[f] = g_alpha[za][ra][f] * g_alpha[zb][rb][f];
  return (3.0 / kPi) * trapzd(a);
}

}  // namespace xtbamd
