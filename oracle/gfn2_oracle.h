/*
 * FP64 CPU oracle for the GFN2-xTB SCC single point (energy + gradient).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C restatement of the algorithm of
 * grimme-lab/xtb 6.7.1 (scf: src/scf_module.F90:74-920, scc:
 * src/scc_core.f90:255-640) used as the parity checker for the HIP engine and
 * as the CPU baseline of bench.py.  Nothing in the product library links,
 * loads or calls it.
 *
 * Units: coordinates in bohr, energies in Hartree, orbital energies in eV
 * (as in the reference, H is kept in eV inside the SCC).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double accuracy;   /* xtb accuracy, default 1.0 (src/xtb/calculator.f90:138-142) */
  double etemp;      /* electronic temperature in K, default 300 */
  int    maxiter;    /* max SCC iterations, default 250 */
  int    guess;      /* 0 = EEQ charges (default API path), 1 = SAD (q = charge/n) */
  int    lapack;     /* 1: use host LAPACK (dlopen) for the eigensolve if present */
  const char *solvent; /* implicit solvent name, NULL: gas phase */
  int    solv_kernel;  /* Born kernel: 1 Still (GFN2/GBSA parameters), 2 P16 (GFN2/ALPB parameters) */
  int    solv_alpb;    /* 1: ALPB shape correction (alpha = 0.571412) */
  /* restart: shell charges [nsh], atomic dipoles [n][3] and quadrupoles [n][6]
   * of a previous result seed the SCC instead of the guess (NULL: guess) */
  const double *init_qsh, *init_dipm, *init_qp;
} om_options;

typedef struct {
  /* dimensions */
  int n, nsh, nao, ndim_d4, iterations, converged, ihomo;
  /* energies (Eh) */
  double e_total, e_elec, e_es, e_aes, e_axc, e_disp, e_rep, e_atm, hl_gap_ev, gnorm;
  double dipole[3];
  double e_solv;     /* solvation free energy (Eh), 0 in the gas phase */
  /* caller-allocated outputs, may be NULL */
  double *gradient;  /* [n][3] */
  double *q;         /* [n] */
  double *qsh;       /* [nsh] */
  double *dipm;      /* [n][3] */
  double *qp;        /* [n][6]  xx,xy,yy,xz,yz,zz */
  double *wbo;       /* [n][n] */
  double *emo;       /* [nao] eV */
  double *focc;      /* [nao] */
  double *density;   /* [nao][nao] */
  double *coeff;     /* [nao][nao]  row i = AO, column k = MO */
} om_result;

/* Dimensions of the basis for a molecule: returns 0 on success. */
int om_dims(int n, const int *at, int *nsh, int *nao, int *ndim_d4);

/* Full single point.  Returns 0 on success, >0 on input/solver error. */
int om_singlepoint(const om_options *opt, int n, const int *at, const double *xyz,
                   double charge, int uhf, om_result *res);

/* Component entry points used by the component parity tests. */
int om_integrals(int n, const int *at, const double *xyz, double accuracy,
                 double *S, double *D, double *Q, double *H0, double *H0_noovlp,
                 double *cn_out);
int om_basis_maps(int n, const int *at, int *sh_at, int *sh_l, int *sh_ao0,
                  int *ao_at, int *ao_sh, int *at_sh0, int *at_ao0);
int om_eeq_guess(int n, const int *at, const double *xyz, double charge, double *q);
int om_coulomb_matrix(int n, const int *at, const double *xyz, double *jmat);
int om_repulsion(int n, const int *at, const double *xyz, double *energy, double *grad);
int om_cn_gfn(int n, const int *at, const double *xyz, double *cn, double *dcndr);
int om_d4_setup(int n, const int *at, const double *xyz, double *gw, double *dispmat);
void om_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
