// Shared host/device definitions of the MI355X GFN2-xTB engine.
//
// Element data are the GFN2-xTB tables of xtb 6.7.1 (src/xtb/gfn2.f90:73-967,
// src/xtb/data.f90:331-365) after the transforms of initGFN2; the D4 model is
// the one of newD4Model (src/disp/dftd4.F90:106-203) in the GFN2 charge mode.
#pragma once
#include <cstdint>

#define XTB_MAX_Z 86
#define XTB_MAX_REF 7
#define XTB_NFREQ 23

namespace xtbamd {

// physical constants: src/mctc/mctc_constants.f90:22-34, src/mctc/convert.f90:25-29
constexpr double kPi = 3.1415926535897932384626433832795029;
constexpr double kAutoEV = 27.21138505;
constexpr double kEVtoAU = 1.0 / 27.21138505;
constexpr double kAutoAA = 0.52917726;
constexpr double kAAtoAU = 1.0 / 0.52917726;
constexpr double kBoltzmann = 3.166808578545117e-06;

// D4 parameters of GFN2 (src/xtb/gfn2.f90:65-66, 762-774)
constexpr double kD4ga = 3.0, kD4gc = 2.0, kD4wf = 6.0;
constexpr double kD4s6 = 1.0, kD4s8 = 2.7, kD4s10 = 0.0, kD4a1 = 0.52, kD4a2 = 5.0, kD4s9 = 5.0;
constexpr int kD4alp = 16;

// Per-element data after the transforms of initGFN2.
struct Element {
  int nsh;
  int l[3];
  int nprim[3];
  double refocc[3];     // reference shell occupation (setzshell / iniqshell)
  double selfe[3];      // atomic level (eV)
  double kcn[3];        // CN dependence of the level
  double zeta[3];       // Slater exponent
  double hard[3];       // shell hardness (Coulomb)
  double third[3];      // shell third-order parameter
  double shpoly[4];     // shell polynomial by angular momentum
  double alp[3][6];     // STO-nG exponents
  double cc[3][6];      // normalised contraction coefficients
  double rad;           // atomic radius (bohr)
  double en;            // Pauling EN
  double dipk, quadk;   // AES XC kernels
  double valcn, mrad;   // AES radius parameters
  double rep_alpha, rep_zeff;
  double zval;          // valence charge
  double rcov_d3;       // D3 covalent radius (bohr)
  double rcov_nc;       // covalent radius of the ncoord module (bohr)
  double en_nc;         // EN of the ncoord module
  double sqrtzr4r2;
  // D4
  int nref;
  int ncount[XTB_MAX_REF];
  double refcn[XTB_MAX_REF];
  double refq[XTB_MAX_REF];
  double d4zeff, d4gam;  // effective nuclear charge, chemical hardness (zeta)
  // EEQ 2019 model (src/charge_model.f90:147-245)
  double eeq_en, eeq_gam, eeq_kappa, eeq_alpha;
};

// Host-side table of all elements (index Z, 0 unused) and reference C6 for
// element pairs; built once.
const Element* host_elements();
double host_ref_c6(int za, int ra, int zb, int rb);

}  // namespace xtbamd
