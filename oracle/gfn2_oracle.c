/*
 * FP64 CPU oracle for the GFN2-xTB SCC single point -- TEST INFRASTRUCTURE.
 *
 * A plain restatement of grimme-lab/xtb 6.7.1's GFN2 single point, used only
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Every function cites the reference routine it follows
 * (paths relative to the xtb source root).
 *
 * Parity pins (see tests/test_oracle_golden.py): H2O per-term energies
 * (test/unit/test_gfn2.f90:140-164), propyne C3H4 energy/gradient/charges/
 * dipole/WBO (test/unit/test_gfn2.f90:226-233, test/api/c_api_example.c:105-112),
 * water tetramer (test/unit/test_gfn2.f90:450-457), mindless01-10
 * (test/unit/test_gfn2.f90:698-712), plus finite differences of the energy,
 * closed-form integrals and invariants (Tr PS = nel, sum q = charge).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "gfn2_oracle.h"
#include "../xtb_amd/csrc/gfn2_params.h"
#include "../xtb_amd/csrc/solv_params.h"

/* ---- constants: src/mctc/mctc_constants.f90:22-34, src/mctc/convert.f90:25-29 */
static const double PI = 3.1415926535897932384626433832795029;
static const double AUTOEV = 27.21138505;
static const double AUTOAA = 0.52917726;
static const double KB = 3.166808578545117e-06;
#define EVTOAU (1.0 / AUTOEV)
#define AATOAU (1.0 / AUTOAA)

static int g_threads = 1;
void om_set_threads(int nthreads) { g_threads = nthreads > 0 ? nthreads : 1; }

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}
#define NEW(T, n) ((T *)xcalloc((size_t)(n), sizeof(T)))

static double ipow(double x, int l) { double r = 1.0; for (int i = 0; i < l; ++i) r *= x; return r; }

/* ========================================================================
 * Element data with the transforms of initGFN2 (src/xtb/gfn2.f90:731-967)
 * ====================================================================== */
typedef struct {
  int nsh, l[3], pqn[3], nprim[3], valence[3];
  double refocc[3], selfe[3], zeta[3], kcn[3], hard[3], third[3];
  double shpoly_l[4], rad, en, dipk, quadk, valcn, mrad, rep_alpha, rep_zeff, zval;
} elem_t;

static elem_t EL[GFN2_MAX_ELEM + 1];
static double KSCALE[4][4];
static int g_init = 0;

/* valence charge: src/type/molecule.f90:560-600 */
static double valence_charge(int z) {
  int nc;
  if (z <= 2) nc = 0; else if (z <= 10) nc = 2; else if (z <= 18) nc = 10;
  else if (z <= 29) nc = 18; else if (z <= 36) nc = 28; else if (z <= 47) nc = 36;
  else if (z <= 54) nc = 46; else if (z <= 71) nc = 54; else if (z <= 79) nc = 68;
  else if (z <= 86) nc = 78; else nc = 86;
  double v = (double)z - (double)nc;
  if (z > 57 && z < 72) v = 3.0;
  return v;
}

static void init_elements(void) {
  if (g_init) return;
  /* kScale: src/xtb/gfn2.f90:823-833 with kshell = [1.85, 2.23, 2.23, 2.23] */
  const double kshell[4] = {1.85, 2.23, 2.23, 2.23};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) KSCALE[j][i] = 0.5 * (kshell[i] + kshell[j]);
  KSCALE[0][2] = KSCALE[2][0] = 2.0; /* ksd */
  KSCALE[1][2] = KSCALE[2][1] = 2.0; /* kpd */
  for (int z = 1; z <= GFN2_MAX_ELEM; ++z) {
    elem_t *e = &EL[z];
    int zi = z - 1;
    e->nsh = k_n_shell[zi];
    int seen[4] = {0, 0, 0, 0};
    for (int m = 0; m < e->nsh; ++m) {
      int l = k_ang_shell[m + 3 * zi];
      e->l[m] = l;
      e->pqn[m] = k_pqn_shell[m + 3 * zi];
      /* generateValenceShellData: src/xtb/data.f90:469-495 */
      e->valence[m] = !seen[l];
      seen[l] = 1;
      /* setGFN2ReferenceOcc: src/xtb/gfn2.f90:869-890 */
      e->refocc[m] = e->valence[m] ? k_reference_occ_l[l + 3 * zi] : 0.0;
      e->selfe[m] = k_self_energy[m + 3 * zi];
      e->zeta[m] = k_slater_exponent[m + 3 * zi];
      /* angToShellData: src/xtb/data.f90:676-703 */
      e->kcn[m] = k_kcn_l[l + 4 * zi];
      /* setGFN1ShellHardness: src/xtb/gfn1.f90 (used by initCoulomb gfn2.f90:777-795) */
      e->hard[m] = k_chemical_hardness_gfn2[zi] * (1.0 + k_shell_hardness_l[l + 3 * zi]);
      /* setGFN2ThirdOrderShell: gam3shell(kind,l) = 1, 0.5, 0.25 for l = 0, 1, 2 */
      double g3 = (l == 0) ? 1.0 : (l == 1 ? 0.5 : 0.25);
      e->third[m] = (k_third_order_atom[zi] * K_THIRD_ORDER_ATOM_MULT) * g3;
      /* setGFN2NumberOfPrimitives: src/xtb/gfn2.f90:893-937 */
      int np = 0;
      if (z <= 2) { np = (l == 0) ? 3 : (l == 1 ? 4 : 0); }
      else if (l == 0 || l == 1) np = (e->pqn[m] > 5) ? 6 : 4;
      else if (l == 2) np = 3;
      else if (l == 3) np = 4;
      e->nprim[m] = np;
    }
    for (int l = 0; l < 4; ++l) e->shpoly_l[l] = k_shell_poly_l[l + 4 * zi];
    e->rad = AATOAU * k_atomic_rad_aa[zi];  /* src/param/atomicrad.f90 */
    e->en = k_pauling_en[zi];
    e->dipk = k_dip_kernel_raw[zi] * K_DIP_KERNEL_MULT;
    e->quadk = k_quad_kernel_raw[zi] * K_QUAD_KERNEL_MULT;
    e->valcn = k_valence_cn[zi];
    e->mrad = k_multi_rad[zi];
    e->rep_alpha = k_rep_alpha[zi];
    e->rep_zeff = k_rep_zeff[zi];
    e->zval = valence_charge(z);
  }
  g_init = 1;
}

/* covalentRadD3 in bohr: src/param/covalentradd3.f90 (values * aatoau * 4/3) */
static double rcov_d3(int z) { return ((k_covalent_rad_d3_aa[z - 1] * AATOAU) * 4.0) / 3.0; }
/* ncoord module radii: src/disp/ncoord.f90:65-88 (4/3*rad/0.52917726) */
static double rcov_nc(int z) { return ((4.0 / 3.0) * k_ncoord_rad_aa[z - 1]) / 0.52917726; }
/* sqrtZr4r2: src/param/sqrtzr4r2.f90:68-69 */
static double sqrt_zr4r2(int z) { return sqrt(0.5 * (k_r4_over_r2[z - 1] * sqrt((double)z))); }

/* ========================================================================
 * STO-nG expansion: slaterToGauss (src/slater.f90:373-483)
 * ====================================================================== */
static int slater_to_gauss(int ng, int n, int l, double zeta, double *alpha, double *coeff) {
  if ((n > 5 || n <= l) && !(n == 6 && ng == 6)) return 2;
  if (zeta <= 0.0) return 4;
  int ityp;
  switch (l) {
    case 0: ityp = n; break;
    case 1: ityp = 4 + n; break;
    case 2: ityp = 7 + n; break;
    case 3: ityp = 9 + n; break;
    case 4: ityp = 10 + n; break;
    default: return 3;
  }
  ityp -= 1; /* 0-based column */
  const double z2 = zeta * zeta;
  const double *pa = NULL, *pc = NULL;
  switch (ng) {
    case 1: alpha[0] = k_sto_alpha1[ityp] * z2; coeff[0] = 1.0; break;
    case 2: pa = k_sto_alpha2; pc = k_sto_coeff2; break;
    case 3: pa = k_sto_alpha3; pc = k_sto_coeff3; break;
    case 4: pa = k_sto_alpha4; pc = k_sto_coeff4; break;
    case 5: pa = k_sto_alpha5; pc = k_sto_coeff5; break;
    case 6:
      if (n == 6) {
        if (l == 0) { for (int k = 0; k < 6; ++k) { alpha[k] = k_sto_alpha6s[k] * z2; coeff[k] = k_sto_coeff6s[k]; } }
        else if (l == 1) { for (int k = 0; k < 6; ++k) { alpha[k] = k_sto_alpha6p[k] * z2; coeff[k] = k_sto_coeff6p[k]; } }
        else return 2;
      } else { pa = k_sto_alpha6; pc = k_sto_coeff6; }
      break;
    default: return 1;
  }
  if (pa) for (int k = 0; k < ng; ++k) { alpha[k] = pa[ityp * ng + k] * z2; coeff[k] = pc[ityp * ng + k]; }
  static const double dfact[8] = {1.0, 1.0, 3.0, 15.0, 105.0, 945.0, 10395.0, 135135.0};
  const double top = 2.0 / PI;
  for (int k = 0; k < ng; ++k)
    coeff[k] = coeff[k] * pow(top * alpha[k], 0.75) * ipow(sqrt(4 * alpha[k]), l) / sqrt(dfact[l]);
  return 0;
}

/* ========================================================================
 * Basis: newBasisset / dim_basis (src/xtb/basisset.f90:33-378)
 * Spherical AO order per shell: s; p = x,y,z; d = (x2-y2), z2, xy, xz, yz.
 * ====================================================================== */
typedef struct {
  int n, nsh, nao;
  int *at_sh0, *at_nsh, *at_ao0, *at_nao;
  int *sh_at, *sh_l, *sh_ao0, *sh_nao, *sh_np, *sh_m;
  double *sh_alp, *sh_cc; /* [nsh][6] */
  int *ao_at, *ao_sh;
} basis_t;

static void basis_free(basis_t *b) {
  free(b->at_sh0); free(b->at_nsh); free(b->at_ao0); free(b->at_nao);
  free(b->sh_at); free(b->sh_l); free(b->sh_ao0); free(b->sh_nao); free(b->sh_np); free(b->sh_m);
  free(b->sh_alp); free(b->sh_cc); free(b->ao_at); free(b->ao_sh);
  memset(b, 0, sizeof(*b));
}

static int basis_build(int n, const int *at, basis_t *b) {
  init_elements();
  memset(b, 0, sizeof(*b));
  b->n = n;
  int nsh = 0, nao = 0;
  for (int i = 0; i < n; ++i) {
    if (at[i] < 1 || at[i] > GFN2_MAX_ELEM) return 1;
    const elem_t *e = &EL[at[i]];
    nsh += e->nsh;
    for (int m = 0; m < e->nsh; ++m) nao += 2 * e->l[m] + 1;
  }
  b->nsh = nsh; b->nao = nao;
  b->at_sh0 = NEW(int, n); b->at_nsh = NEW(int, n); b->at_ao0 = NEW(int, n); b->at_nao = NEW(int, n);
  b->sh_at = NEW(int, nsh); b->sh_l = NEW(int, nsh); b->sh_ao0 = NEW(int, nsh); b->sh_nao = NEW(int, nsh);
  b->sh_np = NEW(int, nsh); b->sh_m = NEW(int, nsh);
  b->sh_alp = NEW(double, 6 * nsh); b->sh_cc = NEW(double, 6 * nsh);
  b->ao_at = NEW(int, nao); b->ao_sh = NEW(int, nao);
  int ish = 0, iao = 0;
  for (int i = 0; i < n; ++i) {
    const elem_t *e = &EL[at[i]];
    b->at_sh0[i] = ish; b->at_nsh[i] = e->nsh; b->at_ao0[i] = iao;
    for (int m = 0; m < e->nsh; ++m, ++ish) {
      int l = e->l[m];
      b->sh_at[ish] = i; b->sh_l[ish] = l; b->sh_ao0[ish] = iao; b->sh_nao[ish] = 2 * l + 1;
      b->sh_np[ish] = e->nprim[m]; b->sh_m[ish] = m;
      int info = slater_to_gauss(e->nprim[m], e->pqn[m], l, e->zeta[m], &b->sh_alp[6 * ish], &b->sh_cc[6 * ish]);
      if (info) return 2;
      for (int k = 0; k < 2 * l + 1; ++k, ++iao) { b->ao_at[iao] = i; b->ao_sh[iao] = ish; }
    }
    b->at_nao[i] = iao - b->at_ao0[i];
  }
  return 0;
}

/* ========================================================================
 * Gaussian multipole integrals over a shell pair
 * (get_multiints / multipole_3d / olapp: src/intgrad.f90:95-111,509-555,879-955;
 *  Cartesian->spherical: dtrf2 src/intgrad.f90:144-231)
 * ====================================================================== */
static const int CART[3][6][3] = {
  {{0, 0, 0}},
  {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
  {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}}};
static const int NCART[3] = {1, 3, 6};

/* cartesian d normalisation relative to xx: set_d_function trafo(5:10) */
static double dnorm(int l, int c) { return (l == 2 && c >= 3) ? sqrt(3.0) : 1.0; }

/* Gaussian moment integral E_m = int t^m exp(-g t^2) dt / sqrt(pi/g)  (olapp) */
static double gmoment(int m, double g) {
  static const double dftr[8] = {1., 1., 3., 15., 105., 945., 10395., 135135.};
  if (m & 1) return 0.0;
  int h = m / 2;
  return ipow(0.5 / g, h) * dftr[h];
}

/* 1D overlap-type integrals of (x-A)^a (x-B)^b  times 1, (x-C), (x-C)^2
 * (and one more angular momentum on A for derivatives) about product centre P. */
static void oned_moments(int la, int lb, double pa, double pb, double pc, double g,
                         double v[4][4][3]) {
  /* coefficients of (t+pa)^a and (t+pb)^b in t = x-P */
  double e[8];
  for (int m = 0; m < 8; ++m) e[m] = gmoment(m, g);
  for (int a = 0; a <= la; ++a)
    for (int bb = 0; bb <= lb; ++bb) {
      double ca[5] = {0}, cb[5] = {0}, cp[9] = {0};
      /* binomial expansion */
      for (int k = 0; k <= a; ++k) {
        double binom = 1.0; for (int q = 0; q < k; ++q) binom = binom * (a - q) / (q + 1);
        ca[k] = binom * ipow(pa, a - k);
      }
      for (int k = 0; k <= bb; ++k) {
        double binom = 1.0; for (int q = 0; q < k; ++q) binom = binom * (bb - q) / (q + 1);
        cb[k] = binom * ipow(pb, bb - k);
      }
      for (int i = 0; i <= a; ++i) for (int j = 0; j <= bb; ++j) cp[i + j] += ca[i] * cb[j];
      double s0 = 0, s1 = 0, s2 = 0;
      for (int t = 0; t <= a + bb; ++t) {
        s0 += cp[t] * e[t];
        s1 += cp[t] * (e[t + 1] + pc * e[t]);
        s2 += cp[t] * (e[t + 2] + 2.0 * pc * e[t + 1] + pc * pc * e[t]);
      }
      v[a][bb][0] = s0; v[a][bb][1] = s1; v[a][bb][2] = s2;
    }
}

/* spherical transform of the d side(s) of a cartesian block (dtrf2) */
static void sph_transform(int li, int lj, double blk[6][6]) {
  if (li < 2 && lj < 2) return;
  const double s3 = sqrt(3.0);
  /* columns m = 0..4 of the cartesian->spherical matrix T[c][m] */
  const double T[6][5] = {{0.5 * s3, 0.5, 0, 0, 0}, {-0.5 * s3, 0.5, 0, 0, 0}, {0, -1.0, 0, 0, 0},
                          {0, 0, 1, 0, 0}, {0, 0, 0, 1, 0}, {0, 0, 0, 0, 1}};
  double tmp[6][6];
  int ni = NCART[li], nj = NCART[lj];
  /* i side */
  if (li == 2) {
    for (int m = 0; m < 5; ++m) for (int j = 0; j < nj; ++j) {
      double s = 0; for (int c = 0; c < 6; ++c) s += T[c][m] * blk[c][j];
      tmp[m][j] = s;
    }
    ni = 5;
    for (int i = 0; i < ni; ++i) for (int j = 0; j < nj; ++j) blk[i][j] = tmp[i][j];
  }
  if (lj == 2) {
    for (int i = 0; i < ni; ++i) for (int m = 0; m < 5; ++m) {
      double s = 0; for (int c = 0; c < 6; ++c) s += blk[i][c] * T[c][m];
      tmp[i][m] = s;
    }
    for (int i = 0; i < ni; ++i) for (int m = 0; m < 5; ++m) blk[i][m] = tmp[i][m];
  }
}

/* S, D(3), Q(6: xx,yy,zz,xy,xz,yz) of shell pair (ish on ri, jsh on rj), moments
 * about rc, spherical AOs: out[i_ao][j_ao].  Also (grad != NULL) derivatives
 * w.r.t. the centre ri of S and of D/Q about rj (multipole_grad_3d). */
typedef struct { double s[6][6], d[3][6][6], q[6][6][6]; } sdq_blk;
typedef struct { double s[3][6][6], d[3][3][6][6], q[3][6][6][6]; } sdq_grad;

static void shell_pair_sdq(const basis_t *b, int ish, int jsh, const double ri[3], const double rj[3],
                           const double rc[3], double intcut, sdq_blk *out, sdq_grad *gout) {
  const int li = b->sh_l[ish], lj = b->sh_l[jsh];
  const int ni = NCART[li], nj = NCART[lj];
  memset(out, 0, sizeof(*out));
  if (gout) memset(gout, 0, sizeof(*gout));
  double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
  double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
  if (r2 > 2000.0) return; /* max_r2: src/intgrad.f90:904 */
  const double sqrtpi = sqrt(PI);
  for (int ip = 0; ip < b->sh_np[ish]; ++ip) {
    const double ai = b->sh_alp[6 * ish + ip];
    for (int jp = 0; jp < b->sh_np[jsh]; ++jp) {
      const double aj = b->sh_alp[6 * jsh + jp];
      double est = ai * aj * r2;
      if (est > intcut * (ai + aj)) continue;
      const double g = ai + aj, ab = 1.0 / g;
      est = est * ab;
      const double kab = exp(-est) * ipow(sqrtpi * sqrt(ab), 3);
      const double cc = kab * b->sh_cc[6 * ish + ip] * b->sh_cc[6 * jsh + jp];
      double v[3][4][4][3];
      double rp[3];
      for (int k = 0; k < 3; ++k) {
        rp[k] = (ai * ri[k] + aj * rj[k]) * ab;
        oned_moments(li + (gout ? 1 : 0), lj, rp[k] - ri[k], rp[k] - rj[k], rp[k] - rc[k], g, v[k]);
      }
      for (int ci = 0; ci < ni; ++ci)
        for (int cj = 0; cj < nj; ++cj) {
          const int *a = CART[li][ci], *bb = CART[lj][cj];
          const double f = cc * dnorm(li, ci) * dnorm(lj, cj);
          double val[3][3];
          for (int k = 0; k < 3; ++k)
            for (int m = 0; m < 3; ++m) val[k][m] = v[k][a[k]][bb[k]][m];
          out->s[ci][cj] += f * val[0][0] * val[1][0] * val[2][0];
          out->d[0][ci][cj] += f * val[0][1] * val[1][0] * val[2][0];
          out->d[1][ci][cj] += f * val[0][0] * val[1][1] * val[2][0];
          out->d[2][ci][cj] += f * val[0][0] * val[1][0] * val[2][1];
          out->q[0][ci][cj] += f * val[0][2] * val[1][0] * val[2][0];
          out->q[1][ci][cj] += f * val[0][0] * val[1][2] * val[2][0];
          out->q[2][ci][cj] += f * val[0][0] * val[1][0] * val[2][2];
          out->q[3][ci][cj] += f * val[0][1] * val[1][1] * val[2][0];
          out->q[4][ci][cj] += f * val[0][1] * val[1][0] * val[2][1];
          out->q[5][ci][cj] += f * val[0][0] * val[1][1] * val[2][1];
          if (gout) {
            /* d/dA of (x-A)^a exp(-ai (x-A)^2) = 2 ai (x-A)^(a+1) - a (x-A)^(a-1) */
            double gra[3][3];
            for (int k = 0; k < 3; ++k)
              for (int m = 0; m < 3; ++m) {
                double t = 2.0 * ai * v[k][a[k] + 1][bb[k]][m];
                if (a[k] > 0) t -= a[k] * v[k][a[k] - 1][bb[k]][m];
                gra[k][m] = t;
              }
            for (int k = 0; k < 3; ++k) {
              /* product of the three directions with direction k differentiated */
              double w[3][3];
              for (int q = 0; q < 3; ++q) for (int m = 0; m < 3; ++m) w[q][m] = (q == k) ? gra[q][m] : val[q][m];
              gout->s[k][ci][cj] += f * w[0][0] * w[1][0] * w[2][0];
              gout->d[k][0][ci][cj] += f * w[0][1] * w[1][0] * w[2][0];
              gout->d[k][1][ci][cj] += f * w[0][0] * w[1][1] * w[2][0];
              gout->d[k][2][ci][cj] += f * w[0][0] * w[1][0] * w[2][1];
              gout->q[k][0][ci][cj] += f * w[0][2] * w[1][0] * w[2][0];
              gout->q[k][1][ci][cj] += f * w[0][0] * w[1][2] * w[2][0];
              gout->q[k][2][ci][cj] += f * w[0][0] * w[1][0] * w[2][2];
              gout->q[k][3][ci][cj] += f * w[0][1] * w[1][1] * w[2][0];
              gout->q[k][4][ci][cj] += f * w[0][1] * w[1][0] * w[2][1];
              gout->q[k][5][ci][cj] += f * w[0][0] * w[1][1] * w[2][1];
            }
          }
        }
    }
  }
  sph_transform(li, lj, out->s);
  for (int k = 0; k < 3; ++k) sph_transform(li, lj, out->d[k]);
  for (int k = 0; k < 6; ++k) sph_transform(li, lj, out->q[k]);
  if (gout)
    for (int x = 0; x < 3; ++x) {
      sph_transform(li, lj, gout->s[x]);
      for (int k = 0; k < 3; ++k) sph_transform(li, lj, gout->d[x][k]);
      for (int k = 0; k < 6; ++k) sph_transform(li, lj, gout->q[x][k]);
    }
}

/* ========================================================================
 * Coordination numbers (src/disp/coordinationnumber.F90:304-599,
 * src/disp/ncoord.f90:469-617)
 * dcndr[(b*n + a)*3 + k] = dCN_b / dR_a,k  (reference dcndr(k, a, b))
 * ====================================================================== */
static double exp_count(double k, double r, double r0) { return 1.0 / (1.0 + exp(-k * (r0 / r - 1.0))); }
static double dexp_count(double k, double r, double r0) {
  double ex = exp(-k * (r0 / r - 1.0));
  return (-k * r0 * ex) / (r * r * ((ex + 1.0) * (ex + 1.0)));
}
static double erf_count(double k, double r, double r0) { return 0.5 * (1.0 + erf(-k * (r - r0) / r0)); }
static double derf_count(double k, double r, double r0) {
  return -k / sqrt(PI) / r0 * exp(-k * k * (r - r0) * (r - r0) / (r0 * r0));
}

/* cnType%gfn (kcn = 10) or cnType%cov (erf, kcn = 7.5, EN scaled); cutoff 40 bohr */
static void cn_lp(int n, const int *at, const double *xyz, int cov, double *cn, double *dcndr) {
  const double cutoff2 = 40.0 * 40.0;
  const double k4 = 4.10451, k5 = 19.08857, k6 = 2 * 11.28174 * 11.28174;
  memset(cn, 0, sizeof(double) * n);
  if (dcndr) memset(dcndr, 0, sizeof(double) * 3 * n * n);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j <= i; ++j) {
      double den = 1.0;
      if (cov) {
        double d = fabs(k_pauling_en[at[i] - 1] - k_pauling_en[at[j] - 1]) + k5;
        den = k4 * exp(-(d * d) / k6);
      }
      double rij[3] = {xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], xyz[3 * i + 2] - xyz[3 * j + 2]};
      double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
      if (r2 > cutoff2 || r2 < 1.0e-12) continue;
      double r1 = sqrt(r2);
      double rc = rcov_d3(at[i]) + rcov_d3(at[j]);
      double cf, df;
      if (cov) { cf = den * erf_count(7.5, r1, rc); df = den * derf_count(7.5, r1, rc); }
      else {
        cf = exp_count(10.0, r1, rc) * exp_count(20.0, r1, rc + 2);
        df = dexp_count(10.0, r1, rc) * exp_count(20.0, r1, rc + 2) + exp_count(10.0, r1, rc) * dexp_count(20.0, r1, rc + 2);
      }
      cn[i] += cf;
      if (i != j) cn[j] += cf;
      if (dcndr) {
        for (int k = 0; k < 3; ++k) {
          double cd = df * rij[k] / r1;
          dcndr[(i * n + i) * 3 + k] += cd;
          dcndr[(j * n + j) * 3 + k] -= cd;
          dcndr[(j * n + i) * 3 + k] += cd; /* dCN_j / dR_i */
          dcndr[(i * n + j) * 3 + k] -= cd; /* dCN_i / dR_j */
        }
      }
    }
  }
}

int om_cn_gfn(int n, const int *at, const double *xyz, double *cn, double *dcndr) {
  init_elements();
  cn_lp(n, at, xyz, 0, cn, dcndr);
  return 0;
}

/* ncoord_d4 (thr 1600 bohr^2) and ncoord_erf: src/disp/ncoord.f90:469-617 */
static void ncoord_erf_d4(int n, const int *at, const double *xyz, int enscale, double *cn) {
  const double k4 = 4.10451, k5 = 19.08857, k6 = 2 * 11.28174 * 11.28174, kn = 7.5;
  memset(cn, 0, sizeof(double) * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1], dz = xyz[3 * j + 2] - xyz[3 * i + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > 1600.0) continue;
      double r = sqrt(r2);
      double rco = rcov_nc(at[j]) + rcov_nc(at[i]);
      double tmp = erf_count(kn, r, rco);
      if (enscale) {
        double d = fabs(k_ncoord_en[at[i] - 1] - k_ncoord_en[at[j] - 1]) + k5;
        tmp = (k4 * exp(-(d * d) / k6)) * tmp;
      }
      cn[i] += tmp;
      cn[j] += tmp;
    }
}

/* ========================================================================
 * DFT-D4 model (src/disp/dftd4.F90:106-203,348-466,588-864,948-1147)
 * ====================================================================== */
static double zeta_fn(double a, double c, double qref, double qmod) {
  if (qmod < 0.0) return exp(a);
  return exp(a * (1.0 - exp(c * (1.0 - qref / qmod))));
}
static double dzeta_fn(double a, double c, double qref, double qmod) {
  if (qmod < 0.0) return 0.0;
  return -a * c * exp(c * (1.0 - qref / qmod)) * zeta_fn(a, c, qref, qmod) * qref / (qmod * qmod);
}
static double cngw(double wf, double cn, double cnref) {
  double val = -wf * (cn - cnref) * (cn - cnref);
  return (val < -200.0) ? 0.0 : exp(val);
}
static double trapzd(const double pol[23]) {
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

static const double D4_GA = 3.0, D4_GC = 2.0, D4_WF = 6.0; /* initDispersion gfn2.f90:762-774 */
static const double D4_S6 = 1.0, D4_S8 = 2.7, D4_S10 = 0.0, D4_A1 = 0.52, D4_A2 = 5.0, D4_S9 = 5.0;
static const int D4_ALP = 16;

typedef struct {
  int nref[D4_MAX_ELEM + 1];
  int ncount[D4_MAX_ELEM + 1][D4_MAX_REF];
  double cn[D4_MAX_ELEM + 1][D4_MAX_REF], q[D4_MAX_ELEM + 1][D4_MAX_REF];
  double alpha[D4_MAX_ELEM + 1][D4_MAX_REF][23];
  int have[D4_MAX_ELEM + 1];
} d4model_t;

/* newD4Model for the GFN2 charge reference (p_refq_gfn2xtb: secq = 0) */
static void d4_model_elem(d4model_t *m, int ia) {
  if (m->have[ia]) return;
  m->have[ia] = 1;
  int zi = ia - 1;
  int cncount[19];
  memset(cncount, 0, sizeof(cncount));
  cncount[0] = 1;
  int nref = k_d4_refn[zi];
  m->nref[ia] = nref;
  for (int j = 0; j < nref; ++j) {
    int is = k_d4_refsys[zi * D4_MAX_REF + j];
    double iz = k_d4_zeff[is - 1];
    double zfac = zeta_fn(D4_GA, k_chemical_hardness_d4[is - 1] * D4_GC, 0.0 + iz,
                          k_d4_refh[zi * D4_MAX_REF + j] + iz);
    m->cn[ia][j] = k_d4_refcovcn[zi * D4_MAX_REF + j];
    m->q[ia][j] = k_d4_refq[zi * D4_MAX_REF + j];
    int icn = (int)lround(k_d4_refcn[zi * D4_MAX_REF + j]);
    cncount[icn] += 1;
    for (int f = 0; f < 23; ++f) {
      double sec_al = k_d4_sscale[is - 1] * k_d4_secaiw[(is - 1) * 23 + f] * zfac;
      double v = k_d4_ascale[zi * D4_MAX_REF + j] *
                 (k_d4_alphaiw[(zi * D4_MAX_REF + j) * 23 + f] - k_d4_hcount[zi * D4_MAX_REF + j] * sec_al);
      m->alpha[ia][j][f] = v > 0.0 ? v : 0.0;
    }
  }
  for (int j = 0; j < nref; ++j) {
    int icn = cncount[(int)lround(k_d4_refcn[zi * D4_MAX_REF + j])];
    m->ncount[ia][j] = icn * (icn + 1) / 2;
  }
}

static double d4_refc6(const d4model_t *m, int ia, int ii, int ja, int jj) {
  double a[23];
  for (int f = 0; f < 23; ++f) a[f] = m->alpha[ia][ii][f] * m->alpha[ja][jj][f];
  return (3.0 / PI) * trapzd(a);
}

typedef struct {
  int ndim, *itbl;
  double *gw, *dispmat; /* [ndim], [ndim][ndim] */
  d4model_t *model;
} d4_t;

static void d4_free(d4_t *d) { free(d->itbl); free(d->gw); free(d->dispmat); free(d->model); memset(d, 0, sizeof(*d)); }

/* initDispersion (src/xtb/dispersion.f90:69-106): d4() + build_wdispmat() */
static void d4_setup(int n, const int *at, const double *xyz, d4_t *d) {
  d->model = NEW(d4model_t, 1);
  for (int i = 0; i < n; ++i) d4_model_elem(d->model, at[i]);
  d->itbl = NEW(int, n);
  int k = 0;
  for (int i = 0; i < n; ++i) { d->itbl[i] = k; k += d->model->nref[at[i]]; }
  d->ndim = k;
  d->gw = NEW(double, k);
  d->dispmat = NEW(double, (size_t)k * k);
  double *cn = NEW(double, n);
  ncoord_erf_d4(n, at, xyz, 1, cn);
  const d4model_t *m = d->model;
  for (int i = 0; i < n; ++i) {
    int ia = at[i];
    double norm = 0.0;
    for (int ii = 0; ii < m->nref[ia]; ++ii)
      for (int c = 1; c <= m->ncount[ia][ii]; ++c) norm += cngw(c * D4_WF, cn[i], m->cn[ia][ii]);
    norm = (norm > 1e-80) ? 1.0 / norm : 0.0;
    double maxcn = -1e300;
    for (int ii = 0; ii < m->nref[ia]; ++ii) if (m->cn[ia][ii] > maxcn) maxcn = m->cn[ia][ii];
    for (int ii = 0; ii < m->nref[ia]; ++ii) {
      double g = 0.0;
      for (int c = 1; c <= m->ncount[ia][ii]; ++c) g += cngw(c * D4_WF, cn[i], m->cn[ia][ii]) * norm;
      if (norm == 0.0) g = (fabs(maxcn - m->cn[ia][ii]) < 1e-12) ? 1.0 : 0.0;
      d->gw[d->itbl[i] + ii] = g;
    }
  }
  /* build_wdispmat: rthr = 72 bohr, gwcut = 1e-7 */
  const int nd = d->ndim;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      int ia = at[i], ja = at[j];
      int ki = d->itbl[i], kj = d->itbl[j];
      double r4r2ij = 3.0 * sqrt_zr4r2(ia) * sqrt_zr4r2(ja);
      double cut = D4_A1 * sqrt(r4r2ij) + D4_A2;
      double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1], dz = xyz[3 * j + 2] - xyz[3 * i + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > 72.0 * 72.0) continue;
      double oor6 = 1.0 / (r2 * r2 * r2 + ipow(cut, 6));
      double oor8 = 1.0 / (r2 * r2 * r2 * r2 + ipow(cut, 8));
      double oor10 = 1.0 / (ipow(r2, 5) + ipow(cut, 10));
      for (int ii = 0; ii < m->nref[ia]; ++ii)
        for (int jj = 0; jj < m->nref[ja]; ++jj) {
          double gwgw = d->gw[ki + ii] * d->gw[kj + jj];
          if (gwgw < 1.0e-7) continue;
          double c6 = d4_refc6(m, ia, ii, ja, jj);
          double c8 = r4r2ij * c6;
          double c10 = 49.0 / 40.0 * r4r2ij * r4r2ij * c6;
          double v = d->gw[ki + ii] * d->gw[kj + jj] * (-D4_S6 * (c6 * oor6) - D4_S8 * (c8 * oor8) - D4_S10 * (c10 * oor10));
          d->dispmat[(size_t)(ki + ii) * nd + kj + jj] = v;
          d->dispmat[(size_t)(kj + jj) * nd + ki + ii] = v;
        }
    }
  free(cn);
}

int om_d4_setup(int n, const int *at, const double *xyz, double *gw, double *dispmat) {
  init_elements();
  d4_t d;
  memset(&d, 0, sizeof(d));
  d4_setup(n, at, xyz, &d);
  if (gw) memcpy(gw, d.gw, sizeof(double) * d.ndim);
  if (dispmat) memcpy(dispmat, d.dispmat, sizeof(double) * d.ndim * d.ndim);
  d4_free(&d);
  return 0;
}

/* zeta vector (disppot / edisp_scc: src/disp/dftd4.F90:735-864) */
static void d4_zeta(const d4_t *d, int n, const int *at, const double *q, double *zv, double *dzv) {
  for (int i = 0; i < d->ndim; ++i) { zv[i] = 0.0; if (dzv) dzv[i] = 0.0; }
  for (int i = 0; i < n; ++i) {
    int ia = at[i];
    double iz = k_d4_zeff[ia - 1];
    double c = k_chemical_hardness_d4[ia - 1] * D4_GC;
    for (int ii = 0; ii < d->model->nref[ia]; ++ii) {
      int k = d->itbl[i] + ii;
      if (d->gw[k] < 1.0e-7) continue;
      zv[k] = zeta_fn(D4_GA, c, d->model->q[ia][ii] + iz, q[i] + iz);
      if (dzv) dzv[k] = dzeta_fn(D4_GA, c, d->model->q[ia][ii] + iz, q[i] + iz);
    }
  }
}

static void symv(int n, const double *a, const double *x, double *y, double alpha) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    const double *row = a + (size_t)i * n;
    for (int j = 0; j < n; ++j) s += row[j] * x[j];
    y[i] = alpha * s;
  }
}

static void d4_potential(const d4_t *d, int n, const int *at, const double *q, double *atom_shift) {
  double *zv = NEW(double, d->ndim), *dz = NEW(double, d->ndim), *dum = NEW(double, d->ndim);
  d4_zeta(d, n, at, q, zv, dz);
  symv(d->ndim, d->dispmat, zv, dum, 1.0);
  for (int i = 0; i < n; ++i)
    for (int ii = 0; ii < d->model->nref[at[i]]; ++ii) {
      int k = d->itbl[i] + ii;
      atom_shift[i] += dum[k] * dz[k];
    }
  free(zv); free(dz); free(dum);
}

static double d4_energy(const d4_t *d, int n, const int *at, const double *q) {
  double *zv = NEW(double, d->ndim), *dum = NEW(double, d->ndim);
  d4_zeta(d, n, at, q, zv, NULL);
  symv(d->ndim, d->dispmat, zv, dum, 0.5);
  double e = 0.0;
  for (int k = 0; k < d->ndim; ++k) e += dum[k] * zv[k];
  free(zv); free(dum);
  return e;
}

/* D4 gradient at fixed charges incl. ATM three-body energy
 * (d4_full_gradient_latp: src/disp/dftd4.F90:1624-1734, cutoff 60 / 40 bohr) */
static void deriv_atm_triple(double c6ij, double c6ik, double c6jk, double cij, double cjk, double cik,
                             double r2ij, double r2jk, double r2ik, double dc6ij, double dc6ji,
                             double dc6jk, double dc6kj, double dc6ik, double dc6ki,
                             const double rij[3], const double rjk[3], const double rik[3],
                             double *dE, double dG[3][3], double dCN[3]) {
  double c9 = -sqrt(c6ij * c6ik * c6jk);
  double ccc1 = cij * cjk * cik;
  double rrr2 = r2ij * r2jk * r2ik;
  double rrr1 = sqrt(rrr2);
  double rrr3 = rrr1 * rrr2;
  double ang_fact = 0.375 / (rrr2 * rrr3);
  double ang = ang_fact * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) * (-r2ij + r2jk + r2ik) + 1.0 / rrr3;
  double cralp = pow(ccc1 / rrr1, (double)D4_ALP / 3.0);
  double fdmp = 1.0 / (1.0 + 6.0 * cralp);
  double dfdmp = -(2.0 * D4_ALP * cralp) * fdmp * fdmp;
  *dE = -fdmp * ang * c9;
  double dang, dGr[3];
  dang = -ang_fact * (r2ij * r2ij * r2ij + r2ij * r2ij * (r2jk + r2ik) + r2ij * (3.0 * r2jk * r2jk + 2.0 * r2jk * r2ik + 3.0 * r2ik * r2ik) -
                      5.0 * (r2jk - r2ik) * (r2jk - r2ik) * (r2jk + r2ik));
  dGr[0] = (-dang * c9 * fdmp + dfdmp * c9 * ang) / r2ij;
  dang = -ang_fact * (r2ik * r2ik * r2ik + r2ik * r2ik * (r2jk + r2ij) + r2ik * (3.0 * r2jk * r2jk + 2.0 * r2jk * r2ij + 3.0 * r2ij * r2ij) -
                      5.0 * (r2jk - r2ij) * (r2jk - r2ij) * (r2jk + r2ij));
  dGr[1] = (-dang * c9 * fdmp + dfdmp * c9 * ang) / r2ik;
  dang = -ang_fact * (r2jk * r2jk * r2jk + r2jk * r2jk * (r2ik + r2ij) + r2jk * (3.0 * r2ik * r2ik + 2.0 * r2ik * r2ij + 3.0 * r2ij * r2ij) -
                      5.0 * (r2ik - r2ij) * (r2ik - r2ij) * (r2ik + r2ij));
  dGr[2] = (-dang * c9 * fdmp + dfdmp * c9 * ang) / r2jk;
  for (int k = 0; k < 3; ++k) {
    dG[0][k] = -dGr[0] * rij[k] - dGr[1] * rik[k];
    dG[1][k] = dGr[0] * rij[k] - dGr[2] * rjk[k];
    dG[2][k] = dGr[1] * rik[k] + dGr[2] * rjk[k];
  }
  double dc9;
  dc9 = 0.5 * c9 * (dc6ij / c6ij + dc6ik / c6ik);
  dCN[0] = -ang * fdmp * dc9;
  dc9 = 0.5 * c9 * (dc6ji / c6ij + dc6jk / c6jk);
  dCN[1] = -ang * fdmp * dc9;
  dc9 = 0.5 * c9 * (dc6ki / c6ik + dc6kj / c6jk);
  dCN[2] = -ang * fdmp * dc9;
}

static void d4_gradient(const d4_t *d, int n, const int *at, const double *xyz, const double *q,
                        double *grad, double *e3_out) {
  const d4model_t *m = d->model;
  double *cn = NEW(double, n), *dcndr = NEW(double, 3 * n * n);
  cn_lp(n, at, xyz, 1, cn, dcndr); /* cnType%cov */
  const int MR = D4_MAX_REF;
  double *zetavec = NEW(double, MR * n), *zerovec = NEW(double, MR * n), *zetadcn = NEW(double, MR * n);
  double *zerodcn = NEW(double, MR * n), *zetadq = NEW(double, MR * n);
  /* weight_references: src/disp/dftd4.F90:948-1065 */
  for (int i = 0; i < n; ++i) {
    int ia = at[i];
    double zi = k_d4_zeff[ia - 1], gi = D4_GC * k_chemical_hardness_d4[ia - 1];
    double norm = 0.0, dnorm = 0.0;
    for (int r = 0; r < m->nref[ia]; ++r)
      for (int c = 1; c <= m->ncount[ia][r]; ++c) {
        double twf = c * D4_WF, gw = cngw(twf, cn[i], m->cn[ia][r]);
        norm += gw;
        dnorm += 2 * twf * (m->cn[ia][r] - cn[i]) * gw;
      }
    norm = (norm > 1e-80) ? 1.0 / norm : 0.0;
    double maxcn = -1e300;
    for (int r = 0; r < m->nref[ia]; ++r) if (m->cn[ia][r] > maxcn) maxcn = m->cn[ia][r];
    for (int r = 0; r < m->nref[ia]; ++r) {
      double expw = 0.0, expd = 0.0;
      for (int c = 1; c <= m->ncount[ia][r]; ++c) {
        double twf = c * D4_WF, gw = cngw(twf, cn[i], m->cn[ia][r]);
        expw += gw;
        expd += 2 * twf * (m->cn[ia][r] - cn[i]) * gw;
      }
      double gwk = expw * norm;
      if (norm == 0.0) gwk = (fabs(maxcn - m->cn[ia][r]) < 1e-12) ? 1.0 : 0.0;
      double z1 = zeta_fn(D4_GA, gi, m->q[ia][r] + zi, q[i] + zi);
      double z0 = zeta_fn(D4_GA, gi, m->q[ia][r] + zi, zi);
      double dgwk = expd * norm - expw * dnorm * norm * norm;
      zetavec[i * MR + r] = z1 * gwk;
      zerovec[i * MR + r] = z0 * gwk;
      zetadcn[i * MR + r] = z1 * dgwk;
      zetadq[i * MR + r] = dzeta_fn(D4_GA, gi, m->q[ia][r] + zi, q[i] + zi) * gwk;
      zerodcn[i * MR + r] = z0 * dgwk;
    }
  }
  double *c6 = NEW(double, n * n), *dc6dcn = NEW(double, n * n);
  double *dEdcn = NEW(double, n);
  /* get_atomic_c6 (src/disp/dftd4.F90:1069-1147) with zetavec */
  for (int pass = 0; pass < 2; ++pass) {
    const double *zv = pass == 0 ? zetavec : zerovec;
    const double *zd = pass == 0 ? zetadcn : zerodcn;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        int ia = at[i], ja = at[j];
        double c = 0, dci = 0, dcj = 0;
        for (int r = 0; r < m->nref[ia]; ++r)
          for (int s = 0; s < m->nref[ja]; ++s) {
            double rc6 = d4_refc6(m, ia, r, ja, s);
            c += zv[i * MR + r] * zv[j * MR + s] * rc6;
            dci += zd[i * MR + r] * zv[j * MR + s] * rc6;
            dcj += zv[i * MR + r] * zd[j * MR + s] * rc6;
          }
        c6[i * n + j] = c6[j * n + i] = c;
        dc6dcn[i * n + j] = dci; /* dc6dcn(iat, jat) */
        dc6dcn[j * n + i] = dcj;
      }
    if (pass == 0) {
      /* disp_gradient_latp: src/disp/dftd4.F90:1832-1933, cutoff 60 */
      for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
          int ia = at[i], ja = at[j];
          double r4r2ij = 3 * sqrt_zr4r2(ia) * sqrt_zr4r2(ja);
          double r0 = D4_A1 * sqrt(r4r2ij) + D4_A2;
          double rij[3] = {xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], xyz[3 * i + 2] - xyz[3 * j + 2]};
          double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
          if (r2 > 3600.0 || r2 < 1.0e-10) continue;
          double t6 = 1. / (r2 * r2 * r2 + ipow(r0, 6)), t8 = 1. / (ipow(r2, 4) + ipow(r0, 8)), t10 = 1. / (ipow(r2, 5) + ipow(r0, 10));
          double d6 = -6 * r2 * r2 * t6 * t6, d8 = -8 * r2 * r2 * r2 * t8 * t8, d10 = -10 * ipow(r2, 4) * t10 * t10;
          double disp = D4_S6 * t6 + D4_S8 * r4r2ij * t8 + D4_S10 * 49.0 / 40.0 * r4r2ij * r4r2ij * t10;
          double ddisp = D4_S6 * d6 + D4_S8 * r4r2ij * d8 + D4_S10 * 49.0 / 40.0 * r4r2ij * r4r2ij * d10;
          dEdcn[i] -= dc6dcn[i * n + j] * disp;
          if (i != j) {
            dEdcn[j] -= dc6dcn[j * n + i] * disp;
            for (int k = 0; k < 3; ++k) {
              double dg = -c6[i * n + j] * ddisp * rij[k];
              grad[3 * i + k] += dg;
              grad[3 * j + k] -= dg;
            }
          }
        }
    }
  }
  /* atm_gradient_latp, molecular branch: src/disp/dftd4.F90:2121-2174, cutoff 40 */
  double e3 = 0.0;
  const double cut2 = 1600.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      double rij[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1], xyz[3 * j + 2] - xyz[3 * i + 2]};
      double r2ij = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
      if (r2ij > cut2) continue;
      for (int k = 0; k < j; ++k) {
        double rik[3] = {xyz[3 * k] - xyz[3 * i], xyz[3 * k + 1] - xyz[3 * i + 1], xyz[3 * k + 2] - xyz[3 * i + 2]};
        double r2ik = rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2];
        if (r2ik > cut2) continue;
        double rjk[3] = {xyz[3 * k] - xyz[3 * j], xyz[3 * k + 1] - xyz[3 * j + 1], xyz[3 * k + 2] - xyz[3 * j + 2]};
        double r2jk = rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2];
        if (r2jk > cut2) continue;
        int ia = at[i], ja = at[j], ka = at[k];
        double cij = D4_A1 * sqrt(3.0 * sqrt_zr4r2(ia) * sqrt_zr4r2(ja)) + D4_A2;
        double cik = D4_A1 * sqrt(3.0 * sqrt_zr4r2(ia) * sqrt_zr4r2(ka)) + D4_A2;
        double cjk = D4_A1 * sqrt(3.0 * sqrt_zr4r2(ja) * sqrt_zr4r2(ka)) + D4_A2;
        double dE, dG[3][3], dCN[3];
        deriv_atm_triple(c6[i * n + j], c6[i * n + k], c6[j * n + k], cij, cjk, cik, r2ij, r2jk, r2ik,
                         dc6dcn[i * n + j], dc6dcn[j * n + i], dc6dcn[j * n + k], dc6dcn[k * n + j],
                         dc6dcn[i * n + k], dc6dcn[k * n + i], rij, rjk, rik, &dE, dG, dCN);
        const double sc = D4_S9; /* triple_scale = 1 for distinct atoms */
        e3 += dE * sc;
        for (int x = 0; x < 3; ++x) {
          grad[3 * i + x] += dG[0][x] * sc;
          grad[3 * j + x] += dG[1][x] * sc;
          grad[3 * k + x] += dG[2][x] * sc;
        }
        dEdcn[i] += dCN[0] * sc;
        dEdcn[j] += dCN[1] * sc;
        dEdcn[k] += dCN[2] * sc;
      }
    }
  /* gradient += dcndr . dEdcn */
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int k = 0; k < 3; ++k) grad[3 * a + k] += dcndr[(b * n + a) * 3 + k] * dEdcn[b];
  *e3_out = e3;
  free(cn); free(dcndr); free(zetavec); free(zerovec); free(zetadcn); free(zerodcn); free(zetadq);
  free(c6); free(dc6dcn); free(dEdcn);
}

/* ========================================================================
 * Isotropic electrostatics (src/coulomb/klopmanohno.F90:257-312,471-573,
 * src/xtb/coulomb.f90:89-142, src/xtb/thirdorder.f90:100-164)
 * ====================================================================== */
static void coulomb_matrix(int n, const int *at, const double *xyz, const basis_t *b, double *jm) {
  const int ns = b->nsh;
  memset(jm, 0, sizeof(double) * ns * ns);
  const double gexp = 2.0;
  for (int i = 0; i < n; ++i) {
    int ii = b->at_sh0[i];
    const elem_t *ei = &EL[at[i]];
    for (int j = 0; j < i; ++j) {
      int jj = b->at_sh0[j];
      const elem_t *ej = &EL[at[j]];
      double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1], dz = xyz[3 * j + 2] - xyz[3 * i + 2];
      double r1 = sqrt(dx * dx + dy * dy + dz * dz);
      for (int is = 0; is < ei->nsh; ++is)
        for (int js = 0; js < ej->nsh; ++js) {
          double gij = 0.5 * (ei->hard[is] + ej->hard[js]);
          double v = 1.0 / pow(pow(r1, gexp) + pow(gij, -gexp), 1.0 / gexp);
          jm[(jj + js) * ns + ii + is] = v;
          jm[(ii + is) * ns + jj + js] = v;
        }
    }
    for (int is = 0; is < ei->nsh; ++is) {
      for (int js = 0; js < is; ++js) {
        double gij = 0.5 * (ei->hard[is] + ei->hard[js]);
        jm[(ii + js) * ns + ii + is] = gij;
        jm[(ii + is) * ns + ii + js] = gij;
      }
      jm[(ii + is) * ns + ii + is] = ei->hard[is];
    }
  }
}

int om_coulomb_matrix(int n, const int *at, const double *xyz, double *jmat) {
  basis_t b;
  if (basis_build(n, at, &b)) return 1;
  coulomb_matrix(n, at, xyz, &b, jmat);
  basis_free(&b);
  return 0;
}

/* ========================================================================
 * Repulsion (src/xtb/repulsion.F90:41-160)
 * ====================================================================== */
static double repulsion(int n, const int *at, const double *xyz, double *grad) {
  double e = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      const elem_t *ei = &EL[at[i]], *ej = &EL[at[j]];
      double alpha = sqrt(ei->rep_alpha * ej->rep_alpha);
      double zeff = ei->rep_zeff * ej->rep_zeff;
      double kexp = (at[i] > 2 || at[j] > 2) ? 1.5 : 1.0;
      double rij[3] = {xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], xyz[3 * i + 2] - xyz[3 * j + 2]};
      double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
      if (r2 > 1600.0 || r2 < 1.0e-8) continue;
      double r1 = sqrt(r2);
      double t16 = pow(r1, kexp), t26 = exp(-alpha * t16), t27 = r1; /* rExp = 1 */
      double dE = zeff * t26 / t27;
      e += dE;
      if (grad)
        for (int k = 0; k < 3; ++k) {
          double dg = -(alpha * t16 * kexp + 1.0) * dE * rij[k] / r2;
          grad[3 * i + k] += dg;
          grad[3 * j + k] -= dg;
        }
    }
  return e;
}

int om_repulsion(int n, const int *at, const double *xyz, double *energy, double *grad) {
  init_elements();
  if (grad) memset(grad, 0, sizeof(double) * 3 * n);
  *energy = repulsion(n, at, xyz, grad);
  return 0;
}

/* ========================================================================
 * Linear algebra helpers: GEMM, Cholesky, Householder/QL eigensolver,
 * LU solve.  Optionally LAPACK/BLAS via dlopen for the CPU baseline.
 * ====================================================================== */
typedef void (*dsygvd_t)(const int *, const char *, const char *, const int *, double *, const int *, double *,
                         const int *, double *, double *, const int *, int *, const int *, int *);
typedef void (*dgemm_t)(const char *, const char *, const int *, const int *, const int *, const double *,
                        const double *, const int *, const double *, const int *, const double *, double *,
                        const int *);
static dsygvd_t p_dsygvd = NULL;
static dgemm_t p_dgemm = NULL;
static int g_lapack_tried = 0;

static void lapack_load(void) {
  if (g_lapack_tried) return;
  g_lapack_tried = 1;
  const char *names[] = {getenv("XTB_ORACLE_LAPACK_LIB"), "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so",
                         "libmkl_rt.so.2", "libopenblas.so.0", "liblapack.so.3", NULL};
  for (int i = 0; i < 7; ++i) {
    if (!names[i]) continue;
    void *h = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL);
    if (!h) continue;
    p_dsygvd = (dsygvd_t)dlsym(h, "dsygvd_");
    p_dgemm = (dgemm_t)dlsym(h, "dgemm_");
    if (p_dsygvd && p_dgemm) return;
    p_dsygvd = NULL; p_dgemm = NULL;
  }
}

/* C[m][n] = alpha * op(A) op(B) + beta C, row-major, square n x n only */
static void gemm_nt(int n, const double *A, const double *B, double *C, int use_lib) {
  /* C = A * B^T */
  if (use_lib && p_dgemm) {
    /* row-major C = A B^T  <=>  col-major C^T = B A^T */
    const double one = 1.0, zero = 0.0;
    p_dgemm("T", "N", &n, &n, &n, &one, B, &n, A, &n, &zero, C, &n);
    return;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      const double *a = A + (size_t)i * n, *bb = B + (size_t)j * n;
      for (int k = 0; k < n; ++k) s += a[k] * bb[k];
      C[(size_t)i * n + j] = s;
    }
}

static void gemm_nn(int n, const double *A, const double *B, double *C, int use_lib) {
  if (use_lib && p_dgemm) {
    const double one = 1.0, zero = 0.0;
    p_dgemm("N", "N", &n, &n, &n, &one, B, &n, A, &n, &zero, C, &n);
    return;
  }
  memset(C, 0, sizeof(double) * n * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      double a = A[(size_t)i * n + k];
      const double *bb = B + (size_t)k * n;
      double *c = C + (size_t)i * n;
      for (int j = 0; j < n; ++j) c[j] += a * bb[j];
    }
}

/* upper Cholesky S = U^T U in place (U stored in upper triangle) */
static int cholesky_upper(int n, double *a) {
  for (int j = 0; j < n; ++j) {
    double s = a[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= a[(size_t)k * n + j] * a[(size_t)k * n + j];
    if (s <= 0.0) return 1;
    double ujj = sqrt(s);
    a[(size_t)j * n + j] = ujj;
    for (int i = j + 1; i < n; ++i) {
      double t = a[(size_t)j * n + i];
      for (int k = 0; k < j; ++k) t -= a[(size_t)k * n + j] * a[(size_t)k * n + i];
      a[(size_t)j * n + i] = t / ujj;
    }
  }
  for (int i = 0; i < n; ++i) for (int j = 0; j < i; ++j) a[(size_t)i * n + j] = 0.0;
  return 0;
}

static double hypot2(double a, double b) { return hypot(a, b); }

/* Householder tridiagonalisation with accumulated transform (z on exit holds
 * Q, column-wise), then implicit QL iterations; eigenvectors returned as rows
 * of vt (vt[k*n + i] = component i of vector k), eigenvalues ascending. */
static int sym_eigen(int n, const double *a_in, double *w, double *vt) {
  double *a = NEW(double, (size_t)n * n), *d = NEW(double, n), *e = NEW(double, n);
  memcpy(a, a_in, sizeof(double) * n * n);
#define A(i, j) a[(size_t)(i) * n + (j)]
  for (int i = n - 1; i > 0; --i) {
    int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += fabs(A(i, k));
      if (scale == 0.0) e[i] = A(i, l);
      else {
        for (int k = 0; k <= l; ++k) { A(i, k) /= scale; h += A(i, k) * A(i, k); }
        double f = A(i, l);
        double g = (f >= 0.0) ? -sqrt(h) : sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          A(j, i) = A(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
          e[j] = g / h;
          f += e[j] * A(i, j);
        }
        double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A(i, j);
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) A(j, k) -= (f * e[k] + g * A(i, k));
        }
      }
    } else e[i] = A(i, l);
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
        for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
      }
    }
    d[i] = A(i, i);
    A(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
  }
  /* transpose accumulated Q so that row k is the k-th basis vector */
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) vt[(size_t)j * n + i] = A(i, j);
#undef A
  /* implicit QL */
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  if (n > 0) e[n - 1] = 0.0;
  int fail = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) + dd == dd) break;
      }
      if (m != l) {
        if (iter++ == 200) { fail = 1; break; }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot2(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = hypot2(f, g));
          if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          double *zi = vt + (size_t)i * n, *zi1 = vt + (size_t)(i + 1) * n;
          for (int k = 0; k < n; ++k) {
            f = zi1[k];
            zi1[k] = s * zi[k] + c * f;
            zi[k] = c * zi[k] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
    if (fail) break;
  }
  /* sort ascending (selection sort, stable enough for distinct keys) */
  int *idx = NEW(int, n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  for (int i = 0; i < n; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j) if (d[idx[j]] < d[idx[k]]) k = j;
    int t = idx[i]; idx[i] = idx[k]; idx[k] = t;
  }
  double *tmp = NEW(double, (size_t)n * n);
  for (int i = 0; i < n; ++i) { w[i] = d[idx[i]]; memcpy(tmp + (size_t)i * n, vt + (size_t)idx[i] * n, sizeof(double) * n); }
  memcpy(vt, tmp, sizeof(double) * n * n);
  free(tmp); free(idx); free(a); free(d); free(e);
  return fail;
}

/* generalized HC = SCe: fact_solve (src/mctc/lapack/eigensolve.F90:218-246).
 * U: upper Cholesky factor of S (row-major, U[i][j] j>=i).  On exit C[i][k] =
 * AO i of MO k, e ascending. */
static int gen_eigen(int n, const double *H, const double *S, const double *U, double *C, double *e, int use_lib) {
  if (use_lib && p_dsygvd) {
    double *a = NEW(double, (size_t)n * n), *bm = NEW(double, (size_t)n * n);
    memcpy(a, H, sizeof(double) * n * n);
    memcpy(bm, S, sizeof(double) * n * n);
    int itype = 1, info = 0, lwork = -1, liwork = -1, iw;
    double wq;
    p_dsygvd(&itype, "V", "U", &n, a, &n, bm, &n, e, &wq, &lwork, &iw, &liwork, &info);
    lwork = (int)wq; liwork = iw;
    double *work = NEW(double, lwork);
    int *iwork = NEW(int, liwork);
    p_dsygvd(&itype, "V", "U", &n, a, &n, bm, &n, e, work, &lwork, iwork, &liwork, &info);
    /* column-major Z(i,k) = a[i + k*n] -> row-major C[i][k] */
    for (int i = 0; i < n; ++i) for (int k = 0; k < n; ++k) C[(size_t)i * n + k] = a[(size_t)k * n + i];
    free(work); free(iwork); free(a); free(bm);
    return info;
  }
  double *Y = NEW(double, (size_t)n * n), *A = NEW(double, (size_t)n * n), *vt = NEW(double, (size_t)n * n);
  /* Y = U^-T H : solve U^T Y = H column by column (forward substitution) */
  for (int c = 0; c < n; ++c)
    for (int i = 0; i < n; ++i) {
      double s = H[(size_t)i * n + c];
      for (int k = 0; k < i; ++k) s -= U[(size_t)k * n + i] * Y[(size_t)k * n + c];
      Y[(size_t)i * n + c] = s / U[(size_t)i * n + i];
    }
  /* A = Y U^-1 : solve A U = Y row by row */
  for (int r = 0; r < n; ++r)
    for (int j = 0; j < n; ++j) {
      double s = Y[(size_t)r * n + j];
      for (int k = 0; k < j; ++k) s -= A[(size_t)r * n + k] * U[(size_t)k * n + j];
      A[(size_t)r * n + j] = s / U[(size_t)j * n + j];
    }
  for (int i = 0; i < n; ++i) for (int j = 0; j < i; ++j) {
    double s = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
    A[(size_t)i * n + j] = A[(size_t)j * n + i] = s;
  }
  int fail = sym_eigen(n, A, e, vt);
  /* C = U^-1 V : back substitution, V[:,k] = vt row k */
  for (int k = 0; k < n; ++k)
    for (int i = n - 1; i >= 0; --i) {
      double s = vt[(size_t)k * n + i];
      for (int j = i + 1; j < n; ++j) s -= U[(size_t)i * n + j] * C[(size_t)j * n + k];
      C[(size_t)i * n + k] = s / U[(size_t)i * n + i];
    }
  free(Y); free(A); free(vt);
  return fail;
}

/* LU with partial pivoting: solve A x = b (A overwritten), returns 1 if singular */
static int lu_solve(int n, double *a, double *bvec) {
  int *piv = NEW(int, n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i) if (fabs(a[(size_t)i * n + k]) > fabs(a[(size_t)p * n + k])) p = i;
    if (a[(size_t)p * n + k] == 0.0) { free(piv); return 1; }
    piv[k] = p;
    if (p != k) {
      for (int j = 0; j < n; ++j) { double t = a[(size_t)k * n + j]; a[(size_t)k * n + j] = a[(size_t)p * n + j]; a[(size_t)p * n + j] = t; }
      double t = bvec[k]; bvec[k] = bvec[p]; bvec[p] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      double f = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      a[(size_t)i * n + k] = f;
      for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
      bvec[i] -= f * bvec[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = bvec[i];
    for (int j = i + 1; j < n; ++j) s -= a[(size_t)i * n + j] * bvec[j];
    bvec[i] = s / a[(size_t)i * n + i];
  }
  free(piv);
  return 0;
}

/* matrix inverse by Gauss-Jordan with partial pivoting (matinv: src/broyden.f90:158-184) */
static int mat_inverse(int n, double *a) {
  double *inv = NEW(double, (size_t)n * n);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i) if (fabs(a[(size_t)i * n + k]) > fabs(a[(size_t)p * n + k])) p = i;
    if (a[(size_t)p * n + k] == 0.0) { free(inv); return 1; }
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = a[(size_t)k * n + j]; a[(size_t)k * n + j] = a[(size_t)p * n + j]; a[(size_t)p * n + j] = t;
        t = inv[(size_t)k * n + j]; inv[(size_t)k * n + j] = inv[(size_t)p * n + j]; inv[(size_t)p * n + j] = t;
      }
    double piv = a[(size_t)k * n + k];
    for (int j = 0; j < n; ++j) { a[(size_t)k * n + j] /= piv; inv[(size_t)k * n + j] /= piv; }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      double f = a[(size_t)i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) { a[(size_t)i * n + j] -= f * a[(size_t)k * n + j]; inv[(size_t)i * n + j] -= f * inv[(size_t)k * n + j]; }
    }
  }
  memcpy(a, inv, sizeof(double) * n * n);
  free(inv);
  return 0;
}

/* ========================================================================
 * EEQ guess (src/eeq_model.f90:60-80,678-1015; src/charge_model.f90:147-245)
 * ====================================================================== */
int om_eeq_guess(int n, const int *at, const double *xyz, double charge, double *q) {
  init_elements();
  int m = n + 1;
  double *cn = NEW(double, n), *A = NEW(double, (size_t)m * m), *x = NEW(double, m);
  ncoord_erf_d4(n, at, xyz, 0, cn);
  const double sqrt2pi = sqrt(2.0 / PI);
  for (int i = 0; i < n; ++i) {
    int z = at[i] - 1;
    double tmp = k_eeq_kappa[z] / (sqrt(cn[i]) + 1e-14);
    x[i] = -k_eeq_en[z] + tmp * cn[i];
  }
  for (int i = 0; i < n; ++i) {
    int zi = at[i] - 1;
    for (int j = 0; j < i; ++j) {
      int zj = at[j] - 1;
      double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1], dz = xyz[3 * j + 2] - xyz[3 * i + 2];
      double r1 = sqrt(dx * dx + dy * dy + dz * dz);
      double gamij = 1.0 / sqrt(k_eeq_alpha[zi] * k_eeq_alpha[zi] + k_eeq_alpha[zj] * k_eeq_alpha[zj]);
      double v = erf(gamij * r1) / r1;
      A[(size_t)j * m + i] = A[(size_t)i * m + j] = v;
    }
    A[(size_t)i * m + i] = k_eeq_gam[zi] + sqrt2pi / k_eeq_alpha[zi];
  }
  for (int i = 0; i < m; ++i) { A[(size_t)n * m + i] = 1.0; A[(size_t)i * m + n] = 1.0; }
  A[(size_t)n * m + n] = 0.0;
  x[n] = charge;
  int bad = lu_solve(m, A, x);
  if (!bad) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) { q[i] = x[i]; s += x[i]; }
    if (fabs(s - charge) > 1.e-6) bad = 2;
  }
  free(cn); free(A); free(x);
  return bad;
}

/* ========================================================================
 * Anisotropic electrostatics (src/aespot.F90)
 * ====================================================================== */
static int lin3(int i, int j) { /* lin(l1,l2) for 1-based 1..3, returns 0-based packed index */
  int a = i > j ? i : j, b = i > j ? j : i;
  return b + a * (a - 1) / 2 - 1;
}

/* gab3/gab5 with CN-dependent radii: get_radcn + mmomgabzero (src/aespot.F90:1219-1281) */
static void aes_setup(int n, const int *at, const double *xyz, const double *cn, double *radcn, double *gab3, double *gab5) {
  const double shift = 1.2, expo = 4.0, rmax = 5.0;
  for (int i = 0; i < n; ++i) {
    const elem_t *e = &EL[at[i]];
    double t1 = cn[i] - e->valcn - shift;
    radcn[i] = e->mrad + (rmax - e->mrad) / (1.0 + exp(-expo * t1));
  }
  memset(gab3, 0, sizeof(double) * n * n);
  memset(gab5, 0, sizeof(double) * n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      double t2 = 0.0;
      for (int k = 0; k < 3; ++k) { double t = xyz[3 * j + k] - xyz[3 * i + k]; t2 += t * t; }
      double rinv = 1.0 / sqrt(t2);
      double rco = 0.5 * (radcn[i] + radcn[j]);
      double d3 = 1.0 / (1.0 + 6.0 * pow(rco * rinv, 3.0));
      double d5 = 1.0 / (1.0 + 6.0 * pow(rco * rinv, 4.0));
      gab3[j * n + i] = gab3[i * n + j] = d3 * pow(rinv, 3.0);
      gab5[j * n + i] = gab5[i * n + j] = d5 * pow(rinv, 5.0);
    }
}

/* setvsdq (src/aespot.F90:727-848) with shift terms, setdvsdq (:861-952) without */
static void aes_potential(int n, const int *at, const double *xyz, const double *q, const double *dipm, const double *qp,
                          const double *gab3, const double *gab5, int with_shift, double *vs, double *vd, double *vq) {
  for (int i = 0; i < n; ++i) {
    const double *ra = xyz + 3 * i;
    double stmp = 0.0, dtmp[3] = {0, 0, 0}, qtmp[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < n; ++j) {
      double g3 = gab3[j * n + i], g5 = gab5[j * n + i];
      const double *rb = xyz + 3 * j;
      double dra[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
      double dum3a, dum5a = 0.0, r2a = 0.0, r2ab = 0.0, t1a = 0.0, t2a = 0.0, t3a = 0.0;
      for (int l1 = 1; l1 <= 3; ++l1) {
        double a1 = ra[l1 - 1], d1 = dra[l1 - 1];
        r2a += a1 * a1;
        r2ab += d1 * d1;
        t1a += a1 * d1;
        t2a += dipm[3 * j + l1 - 1] * d1;
        t3a += a1 * dipm[3 * j + l1 - 1];
        for (int l2 = 1; l2 <= 3; ++l2) {
          int ll = lin3(l1, l2);
          double d2 = dra[l2 - 1], a2 = ra[l2 - 1];
          dum5a = dum5a - qp[6 * j + ll] * d1 * d2;
          if (with_shift) dum5a = dum5a - 1.50 * q[j] * d1 * d2 * a1 * a2;
          if (l2 >= l1) continue;
          int ki = l1 + l2 + 1 - 1;
          qtmp[ki] = qtmp[ki] - 3.0 * q[j] * g5 * d2 * d1;
        }
        qtmp[l1 - 1] = qtmp[l1 - 1] - 1.50 * q[j] * g5 * d1 * d1;
      }
      if (with_shift) {
        dum3a = -t1a * q[j] - t2a;
        dum5a = dum5a + t3a * r2ab - 3.0 * t1a * t2a + 0.50 * q[j] * r2a * r2ab;
      } else {
        dum3a = -t2a;
      }
      stmp = stmp + dum5a * g5 + dum3a * g3;
      for (int l1 = 0; l1 < 3; ++l1) {
        double d3a = dra[l1] * q[j];
        double d5a = 3.0 * dra[l1] * t2a - r2ab * dipm[3 * j + l1];
        if (with_shift) d5a = d5a - q[j] * r2ab * ra[l1] + 3.0 * q[j] * dra[l1] * t1a;
        dtmp[l1] = dtmp[l1] + d3a * g3 + d5a * g5;
        qtmp[l1] = qtmp[l1] + 0.50 * r2ab * q[j] * g5;
      }
    }
    vs[i] = stmp;
    for (int k = 0; k < 3; ++k) vd[3 * i + k] = dtmp[k];
    for (int k = 0; k < 6; ++k) vq[6 * i + k] = qtmp[k];
    /* CT correction terms */
    const elem_t *e = &EL[at[i]];
    double qs1 = e->dipk * 2.0, qs2 = e->quadk * 6.0;
    double t3a = 0.0, t2a = 0.0;
    for (int l1 = 1; l1 <= 3; ++l1) {
      if (with_shift) t3a = t3a + ra[l1 - 1] * dipm[3 * i + l1 - 1] * qs1;
      vd[3 * i + l1 - 1] = vd[3 * i + l1 - 1] - qs1 * dipm[3 * i + l1 - 1];
      for (int l2 = 1; l2 < l1; ++l2) {
        int ll = lin3(l1, l2), ki = l1 + l2 + 1 - 1;
        vq[6 * i + ki] = vq[6 * i + ki] - qp[6 * i + ll] * qs2;
        if (with_shift) {
          t3a = t3a - ra[l1 - 1] * ra[l2 - 1] * qp[6 * i + ll] * qs2;
          vd[3 * i + l1 - 1] = vd[3 * i + l1 - 1] + ra[l2 - 1] * qp[6 * i + ll] * qs2;
          vd[3 * i + l2 - 1] = vd[3 * i + l2 - 1] + ra[l1 - 1] * qp[6 * i + ll] * qs2;
        }
      }
      int ll = lin3(l1, l1);
      vq[6 * i + l1 - 1] = vq[6 * i + l1 - 1] - qp[6 * i + ll] * qs2 * 0.50;
      if (with_shift) {
        t3a = t3a - ra[l1 - 1] * ra[l1 - 1] * qp[6 * i + ll] * qs2 * 0.50;
        vd[3 * i + l1 - 1] = vd[3 * i + l1 - 1] + ra[l1 - 1] * qp[6 * i + ll] * qs2;
      }
      t2a = t2a + qp[6 * i + ll];
    }
    if (with_shift) vs[i] = vs[i] + t3a;
    t2a = t2a * e->quadk;
    for (int l1 = 0; l1 < 3; ++l1) {
      vq[6 * i + l1] = vq[6 * i + l1] + t2a;
      if (with_shift) {
        vd[3 * i + l1] = vd[3 * i + l1] - 2.0 * ra[l1] * t2a;
        vs[i] = vs[i] + t2a * ra[l1] * ra[l1];
      }
    }
  }
}

/* aniso_electro_cpu (src/aespot.F90:569-662) */
static void aes_energy(int n, const int *at, const double *xyz, const double *q, const double *dipm, const double *qp,
                       const double *gab3, const double *gab5, double *eaes, double *epol) {
  static const int idx[3][3] = {{0, 1, 3}, {1, 2, 4}, {3, 4, 5}};
  double ep = 0.0, e01 = 0.0, e02 = 0.0, e11 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double *dp1 = dipm + 3 * i, *qp1 = qp + 6 * i;
    double tt = dp1[0] * dp1[0] + dp1[1] * dp1[1] + dp1[2] * dp1[2];
    double tt3 = 0.0;
    for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) { double v = qp1[idx[l][k]]; tt3 += v * v; }
    ep += EL[at[i]].dipk * tt + tt3 * EL[at[i]].quadk;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      const double *dp1 = dipm + 3 * i, *qp1 = qp + 6 * i;
      double q1 = q[i];
      double rij[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1], xyz[3 * j + 2] - xyz[3 * i + 2]};
      double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
      double ed = 0, eq = 0, edd = 0;
      for (int k = 0; k < 3; ++k) {
        ed = ed + q[j] * dp1[k] * rij[k];
        ed = ed - dipm[3 * j + k] * q1 * rij[k];
        for (int l = 0; l < 3; ++l) {
          int kl = idx[l][k];
          double tt = rij[l] * rij[k], tt3 = 3.0 * tt;
          eq = eq + q[j] * qp1[kl] * tt;
          eq = eq + qp[6 * j + kl] * q1 * tt;
          edd = edd - dipm[3 * j + k] * dp1[l] * tt3;
        }
        edd = edd + dipm[3 * j + k] * dp1[k] * r2;
      }
      e01 += ed * gab3[j * n + i];
      e02 += eq * gab5[j * n + i];
      e11 += edd * gab5[j * n + i];
    }
  *eaes = e01 + e02 + e11;
  *epol = ep;
}

/* aniso_grad (src/aespot.F90:1070-1165); drad[(i*n + a)*3 + k] = dR_i/dX_a,k */
static void aes_gradient(int n, const double *xyz, const double *q, const double *dipm, const double *qp,
                         const double *radcn, const double *drad, const double *gab3, const double *gab5, double *g) {
  const double kdmp3 = 3.0, kdmp5 = 4.0;
  for (int i = 0; i < n; ++i) {
    double q1 = q[i];
    const double *rr = xyz + 3 * i, *dip = dipm + 3 * i, *qp1 = qp + 6 * i;
    double tmp2 = 0.0;
    for (int j = 0; j < n; ++j) {
      if (i == j) continue;
      double rij[3] = {xyz[3 * j] - rr[0], xyz[3 * j + 1] - rr[1], xyz[3 * j + 2] - rr[2]};
      double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
      double rabi = 1.0 / sqrt(r2);
      double rco = 0.5 * (radcn[i] + radcn[j]);
      double damp1 = 1.0 / (1.0 + 6.0 * pow(rco * rabi, kdmp3));
      double ddamp1 = -kdmp3 * rabi * (damp1 * damp1 - damp1);
      double dgab3 = (-3.0 * rabi * pow(rabi, 3.0)) * damp1 + ddamp1 * pow(rabi, 3.0);
      double damp2 = 1.0 / (1.0 + 6.0 * pow(rco * rabi, kdmp5));
      double ddamp2 = -kdmp5 * rabi * (damp2 * damp2 - damp2);
      double dgab5 = (-5.0 * rabi * pow(rabi, 5.0)) * damp2 + ddamp2 * pow(rabi, 5.0);
      double ed = 0, edd = 0, eq = 0;
      double g35 = gab5[j * n + i], g3 = gab3[j * n + i];
      for (int k = 1; k <= 3; ++k) {
        ed = ed + q[j] * dip[k - 1] * rij[k - 1];
        ed = ed - dipm[3 * j + k - 1] * q1 * rij[k - 1];
        double tt = q1 * dipm[3 * j + k - 1] - q[j] * dip[k - 1];
        g[3 * i + k - 1] += g3 * tt;
        double ddm2 = 0, ddm3a = 0, ddm3b = 0, qqa = 0, qqb = 0;
        for (int l = 1; l <= 3; ++l) {
          int kl = lin3(l, k);
          double t = rij[l - 1] * rij[k - 1], t3 = 3.0 * t;
          eq = eq + q[j] * qp1[kl] * t;
          eq = eq + qp[6 * j + kl] * q1 * t;
          edd = edd - dipm[3 * j + k - 1] * dip[l - 1] * t3;
          ddm2 = ddm2 + dipm[3 * j + l - 1] * dip[l - 1];
          ddm3a = ddm3a + dip[l - 1] * rij[l - 1];
          ddm3b = ddm3b + dipm[3 * j + l - 1] * rij[l - 1];
          qqa = qqa + rij[l - 1] * qp[6 * j + kl];
          qqb = qqb + rij[l - 1] * qp1[kl];
        }
        edd = edd + dipm[3 * j + k - 1] * dip[k - 1] * r2;
        g[3 * i + k - 1] += -2.0 * g35 * ddm2 * rij[k - 1];
        g[3 * i + k - 1] += 3.0 * g35 * ddm3a * dipm[3 * j + k - 1];
        g[3 * i + k - 1] += 3.0 * g35 * ddm3b * dip[k - 1];
        g[3 * i + k - 1] += -2.0 * g35 * qqa * q1;
        g[3 * i + k - 1] += -2.0 * g35 * qqb * q[j];
      }
      for (int k = 0; k < 3; ++k) {
        double dxi = rij[k] * rabi;
        g[3 * i + k] -= ed * dgab3 * dxi;
        g[3 * i + k] -= (eq + edd) * dgab5 * dxi;
      }
      double rab = 0.50 * (radcn[i] + radcn[j]);
      tmp2 += ed * kdmp3 * g3 * (damp1 / rab) * pow(rab * rabi, kdmp3);
      tmp2 += (eq + edd) * kdmp5 * g35 * (damp2 / rab) * pow(rab * rabi, kdmp5);
    }
    tmp2 = 3.0 * tmp2;
    for (int a = 0; a < n; ++a)
      for (int k = 0; k < 3; ++k) g[3 * a + k] -= tmp2 * drad[(i * n + a) * 3 + k];
  }
}

/* ========================================================================
 * Occupations and Fermi smearing (src/scc_core.f90:983-1064,1112-1170)
 * ====================================================================== */
static void occu(int ndim, int nel, int nopen, int *ihomoa, int *ihomob, double *focca, double *foccb) {
  int *focc = NEW(int, ndim);
  for (int i = 0; i < ndim; ++i) { focca[i] = foccb[i] = 0.0; }
  if (nel % 2 == 0) {
    int ihomo = nel / 2;
    for (int i = 0; i < ihomo && i < ndim; ++i) focc[i] = 2;
    if (nopen > 1)
      for (int i = 1; i <= nopen / 2; ++i) { focc[ihomo - i] -= 1; focc[ihomo + i - 1] += 1; }
  } else {
    int na = nel / 2 + (nopen - 1) / 2 + 1, nb = nel / 2 - (nopen - 1) / 2;
    for (int i = 0; i < na && i < ndim; ++i) focc[i] += 1;
    for (int i = 0; i < nb && i < ndim; ++i) focc[i] += 1;
  }
  for (int i = 0; i < ndim; ++i) {
    if (focc[i] == 2) { focca[i] = 1.0; foccb[i] = 1.0; }
    if (focc[i] == 1) focca[i] = 1.0;
  }
  *ihomoa = 0; *ihomob = 0;
  for (int i = 0; i < ndim; ++i) {
    if (focca[i] > 0.99) *ihomoa = i + 1;
    if (foccb[i] > 0.99) *ihomob = i + 1;
  }
  free(focc);
}

static void fermismear(int norbs, int nel, double t, const double *eig, double *occ, double *efermi, double *s_out) {
  const double boltz = KB * AUTOEV, thr = 1e-9;
  const double sqrttiny = sqrt(2.2250738585072014e-308);
  double bkt = boltz * t;
  double ef;
  if (nel + 1 > norbs) ef = eig[nel - 1];
  else if (nel == 0) ef = eig[0];
  else {
    ef = 0.5 * (eig[nel - 1] + eig[nel]);
    double occt = nel;
    for (int cyc = 0; cyc < 200; ++cyc) {
      double tot = 0.0, totd = 0.0;
      for (int i = 0; i < norbs; ++i) {
        double ff = 0.0, df;
        if ((eig[i] - ef) / bkt < 50) {
          double ex = exp((eig[i] - ef) / bkt);
          ff = 1.0 / (ex + 1.0);
          df = ex / (bkt * (ex + 1.0) * (ex + 1.0));
        } else df = 0.0;
        occ[i] = ff;
        tot += ff;
        totd += df;
      }
      (void)sqrttiny;
      double change = (occt - tot) / totd;
      ef += change;
      if (fabs(occt - tot) <= thr) break;
    }
  }
  double s = 0.0;
  for (int i = 0; i < norbs; ++i)
    if (occ[i] > thr && 1.0 - occ[i] > thr) s += occ[i] * log(occ[i]) + (1.0 - occ[i]) * log(1.0 - occ[i]);
  *efermi = ef;
  *s_out = s * KB * t;
}

/* ========================================================================
 * Broyden mixing (src/broyden.f90:37-153)
 * ====================================================================== */
typedef struct { int n, maxit; double *df, *u, *a, *omega, *qlast, *dqlast; } broyden_t;

static void broyden(broyden_t *B, double *q, const double *dq, int iter, double alpha) {
  const int n = B->n;
  const double omega0 = 0.01, minw = 1.0, maxw = 100000.0, wfac = 0.01;
  if (iter == 1) {
    memcpy(B->dqlast, dq, sizeof(double) * n);
    memcpy(B->qlast, q, sizeof(double) * n);
    for (int k = 0; k < n; ++k) q[k] = q[k] + alpha * dq[k];
    return;
  }
  int it1 = iter - 1, i1 = it1 - 1; /* 0-based row of the new history entry */
  double s = 0.0;
  for (int k = 0; k < n; ++k) s += dq[k] * dq[k];
  double w = sqrt(s);
  w = (w > wfac / maxw) ? wfac / w : maxw;
  if (w < minw) w = minw;
  B->omega[i1] = w;
  double *dfi = B->df + (size_t)i1 * n;
  double nn = 0.0;
  for (int k = 0; k < n; ++k) { dfi[k] = dq[k] - B->dqlast[k]; nn += dfi[k] * dfi[k]; }
  double inv = 1.0 / sqrt(nn);
  for (int k = 0; k < n; ++k) dfi[k] = inv * dfi[k];
  double *c = NEW(double, it1), *beta = NEW(double, (size_t)it1 * it1), *gam = NEW(double, it1);
  for (int i = 0; i < it1; ++i) {
    const double *dfk = B->df + (size_t)i * n;
    double a = 0.0, cc = 0.0;
    for (int k = 0; k < n; ++k) { a += dfk[k] * dfi[k]; cc += dfk[k] * dq[k]; }
    B->a[(size_t)i * B->maxit + i1] = a;
    B->a[(size_t)i1 * B->maxit + i] = a;
    c[i] = B->omega[i] * cc;
  }
  for (int i = 0; i < it1; ++i) {
    for (int j = 0; j < it1; ++j) beta[(size_t)j * it1 + i] = B->omega[j] * B->omega[i] * B->a[(size_t)j * B->maxit + i];
    beta[(size_t)i * it1 + i] += omega0 * omega0;
  }
  mat_inverse(it1, beta);
  for (int j = 0; j < it1; ++j) {
    double gsum = 0.0;
    for (int i = 0; i < it1; ++i) gsum += c[i] * beta[(size_t)i * it1 + j];
    gam[j] = gsum;
  }
  double *ui = B->u + (size_t)i1 * n;
  for (int k = 0; k < n; ++k) ui[k] = alpha * dfi[k] + inv * (q[k] - B->qlast[k]);
  memcpy(B->dqlast, dq, sizeof(double) * n);
  memcpy(B->qlast, q, sizeof(double) * n);
  for (int k = 0; k < n; ++k) q[k] = q[k] + alpha * dq[k];
  for (int i = 0; i < it1; ++i) {
    const double *uk = B->u + (size_t)i * n;
    double f = B->omega[i] * gam[i];
    for (int k = 0; k < n; ++k) q[k] = q[k] - f * uk[k];
  }
  free(c); free(beta); free(gam);
}

/* ========================================================================
 * Integrals + H0 (build_SDQH0: src/xtb/hamiltonian.F90:146-371) and the
 * self energies (getSelfEnergy2D: :98-141).  Dense full symmetric storage:
 * S[i][j], D[k][i][j], Q[k][i][j] (moments about the origin), H0 (eV).
 * ====================================================================== */
static double shell_poly(double ipoly, double jpoly, double irad, double jrad, const double *x1, const double *x2) {
  double dx = x1[0] - x2[0], dy = x1[1] - x2[1], dz = x1[2] - x2[2];
  double rab = sqrt(dx * dx + dy * dy + dz * dz);
  double rr = jrad + irad;
  double r = rab / rr;
  double rf1 = 1.0 + 0.01 * ipoly * sqrt(r);
  double rf2 = 1.0 + 0.01 * jpoly * sqrt(r);
  return rf1 * rf2;
}

static double h0_km(int izp, int jzp, int li, int lj) {
  double den = (EL[izp].en - EL[jzp].en) * (EL[izp].en - EL[jzp].en);
  double enpoly = 1.0 + 0.02 * den * (1.0 + 0.0 * den);
  return KSCALE[lj][li] * enpoly * 1.0;
}

static void build_sdqh0(int n, const int *at, const double *xyz, const basis_t *b, const double *selfe, double intcut,
                        double *S, double *D, double *Q, double *H0, double *H0n) {
  const int N = b->nao;
  const size_t NN = (size_t)N * N;
  memset(S, 0, sizeof(double) * NN);
  memset(D, 0, sizeof(double) * 3 * NN);
  memset(Q, 0, sizeof(double) * 6 * NN);
  memset(H0, 0, sizeof(double) * NN);
  memset(H0n, 0, sizeof(double) * NN);
  const double origin[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (int i = 0; i < n; ++i) {
    sdq_blk blk;
    for (int j = 0; j < i; ++j) {
      int izp = at[i], jzp = at[j];
      const double *ra = xyz + 3 * i, *rb = xyz + 3 * j;
      for (int is = 0; is < EL[izp].nsh; ++is) {
        int ish = b->at_sh0[i] + is, li = b->sh_l[ish];
        for (int js = 0; js < EL[jzp].nsh; ++js) {
          int jsh = b->at_sh0[j] + js, lj = b->sh_l[jsh];
          double hii = selfe[ish], hjj = selfe[jsh];
          double zi = EL[izp].zeta[is], zj = EL[jzp].zeta[js];
          double zetaij = pow(2 * sqrt(zi * zj) / (zi + zj), 0.5);
          double km = h0_km(izp, jzp, li, lj);
          double hav = 0.5 * km * (hii + hjj) * zetaij;
          double shpoly = shell_poly(EL[izp].shpoly_l[li], EL[jzp].shpoly_l[lj], EL[izp].rad, EL[jzp].rad, ra, rb);
          shell_pair_sdq(b, ish, jsh, ra, rb, origin, intcut, &blk, NULL);
          for (int ii = 0; ii < 2 * li + 1; ++ii)
            for (int jj = 0; jj < 2 * lj + 1; ++jj) {
              size_t ia = b->sh_ao0[ish] + ii, ja = b->sh_ao0[jsh] + jj;
              double s = blk.s[ii][jj];
              double h = hav * shpoly * s;
              H0[ia * N + ja] = H0[ja * N + ia] = h;
              H0n[ia * N + ja] = H0n[ja * N + ia] = hav * shpoly;
              S[ia * N + ja] = S[ja * N + ia] = s;
              for (int k = 0; k < 3; ++k) D[k * NN + ia * N + ja] = D[k * NN + ja * N + ia] = blk.d[k][ii][jj];
              for (int k = 0; k < 6; ++k) Q[k * NN + ia * N + ja] = Q[k * NN + ja * N + ia] = blk.q[k][ii][jj];
            }
        }
      }
    }
  }
  /* on-site blocks */
  for (int i = 0; i < n; ++i) {
    const double *ra = xyz + 3 * i;
    sdq_blk blk;
    for (int is = 0; is < EL[at[i]].nsh; ++is) {
      int ish = b->at_sh0[i] + is;
      for (int ii = 0; ii < b->sh_nao[ish]; ++ii) {
        size_t ia = b->sh_ao0[ish] + ii;
        S[ia * N + ia] = 1.0;
        H0[ia * N + ia] = selfe[ish];
        H0n[ia * N + ia] = selfe[ish];
      }
      for (int js = 0; js <= is; ++js) {
        int jsh = b->at_sh0[i] + js;
        shell_pair_sdq(b, ish, jsh, ra, ra, origin, intcut, &blk, NULL);
        for (int ii = 0; ii < b->sh_nao[ish]; ++ii)
          for (int jj = 0; jj < b->sh_nao[jsh]; ++jj) {
            size_t ia = b->sh_ao0[ish] + ii, ja = b->sh_ao0[jsh] + jj;
            for (int k = 0; k < 3; ++k) D[k * NN + ia * N + ja] = D[k * NN + ja * N + ia] = blk.d[k][ii][jj];
            for (int k = 0; k < 6; ++k) Q[k * NN + ia * N + ja] = Q[k * NN + ja * N + ia] = blk.q[k][ii][jj];
          }
      }
    }
  }
}

static void self_energies(int n, const int *at, const basis_t *b, const double *cn, double *selfe, double *dsedcn) {
  for (int i = 0; i < n; ++i) {
    const elem_t *e = &EL[at[i]];
    for (int m = 0; m < e->nsh; ++m) {
      int ish = b->at_sh0[i] + m;
      selfe[ish] = e->selfe[m] - e->kcn[m] * cn[i];
      dsedcn[ish] = -e->kcn[m];
    }
  }
}

int om_basis_maps(int n, const int *at, int *sh_at, int *sh_l, int *sh_ao0, int *ao_at, int *ao_sh, int *at_sh0, int *at_ao0) {
  basis_t b;
  if (basis_build(n, at, &b)) return 1;
  for (int s = 0; s < b.nsh; ++s) { sh_at[s] = b.sh_at[s]; sh_l[s] = b.sh_l[s]; sh_ao0[s] = b.sh_ao0[s]; }
  for (int a = 0; a < b.nao; ++a) { ao_at[a] = b.ao_at[a]; ao_sh[a] = b.ao_sh[a]; }
  for (int i = 0; i < n; ++i) { at_sh0[i] = b.at_sh0[i]; at_ao0[i] = b.at_ao0[i]; }
  basis_free(&b);
  return 0;
}

int om_dims(int n, const int *at, int *nsh, int *nao, int *ndim_d4) {
  init_elements();
  int s = 0, a = 0, d = 0;
  for (int i = 0; i < n; ++i) {
    if (at[i] < 1 || at[i] > GFN2_MAX_ELEM) return 1;
    s += EL[at[i]].nsh;
    for (int m = 0; m < EL[at[i]].nsh; ++m) a += 2 * EL[at[i]].l[m] + 1;
    d += k_d4_refn[at[i] - 1];
  }
  *nsh = s; *nao = a; *ndim_d4 = d;
  return 0;
}

int om_integrals(int n, const int *at, const double *xyz, double accuracy, double *S, double *D, double *Q,
                 double *H0, double *H0n, double *cn_out) {
  basis_t b;
  if (basis_build(n, at, &b)) return 1;
  double *cn = NEW(double, n), *selfe = NEW(double, b.nsh), *dse = NEW(double, b.nsh);
  cn_lp(n, at, xyz, 0, cn, NULL);
  self_energies(n, at, &b, cn, selfe, dse);
  double intcut = 25.0 - 10.0 * log10(accuracy);
  if (intcut < 20.0) intcut = 20.0;
  build_sdqh0(n, at, xyz, &b, selfe, intcut, S, D, Q, H0, H0n);
  if (cn_out) memcpy(cn_out, cn, sizeof(double) * n);
  free(cn); free(selfe); free(dse);
  basis_free(&b);
  return 0;
}

/* ========================================================================
 * The single point: scf (src/scf_module.F90:74-920) + scc (scc_core.f90:255-640)
 * ====================================================================== */
/* ---- implicit solvation: GBSA / ALPB -------------------------------------
 * Solvent parameters: src/solv/model.f90 (normalizeSolventName,
 * loadInternalParam, paramToModel); model set-up: src/solv/gbsa.f90:217-442;
 * Born radii: src/solv/born.f90:35-378; surface: src/solv/sasa.f90:42-205;
 * kernels: src/solv/kernel.f90; potential, energy and gradient:
 * src/solv/gbsa.f90:445-777, 991-1079.  GFN2 couples through the Mulliken
 * atomic charges (src/scf_module.F90:299-321). */
static const double SOLV_W = 0.3 * AATOAU;
#define SOLV_W3 (SOLV_W * (SOLV_W * SOLV_W))
#define SOLV_AH0 0.5
#define SOLV_AH1 (3.0 / (4.0 * SOLV_W))
#define SOLV_AH3 (-1.0 / (4.0 * SOLV_W3))
static const double KCALTOAU = 1.0 / 627.50947428;

typedef struct {
  int n, kernel, lhb;
  double keps, alpbet, gshift, gsasa, adet;
  double *vdwr, *rho, *svdw, *vdwsa, *wrp, *trj2, *gamsasa, *hbmag;
  double *brad, *brdr, *sasa, *dsdrt, *dsdr, *hbw, *dhbdw, *bmat;
} solv_t;

static void solv_free(solv_t *s) {
  free(s->vdwr); free(s->rho); free(s->svdw); free(s->vdwsa); free(s->wrp); free(s->trj2); free(s->gamsasa);
  free(s->hbmag); free(s->brad); free(s->brdr); free(s->sasa); free(s->dsdrt); free(s->dsdr); free(s->hbw);
  free(s->dhbdw); free(s->bmat);
  memset(s, 0, sizeof(*s));
}

/* normalizeSolventName (src/solv/model.f90) */
static void solv_normalize(const char *in, char *out, size_t cap) {
  size_t k = 0;
  for (; in[k] && k + 1 < cap; ++k) out[k] = (char)((in[k] >= 'A' && in[k] <= 'Z') ? in[k] - 'A' + 'a' : in[k]);
  out[k] = 0;
  static const char *alias[][2] = {
      {"methylcyanide", "acetonitrile"}, {"methylenechloride", "ch2cl2"}, {"dichloromethane", "ch2cl2"},
      {"dcm", "ch2cl2"}, {"chloroform", "chcl3"}, {"trichloromethane", "chcl3"}, {"tcm", "chcl3"},
      {"carbondisulfide", "cs2"}, {"diethylether", "ether"}, {"dimethylformamide", "dmf"},
      {"dimethylsulfoxide", "dmso"}, {"furan", "furane"}, {"nhexan", "hexane"}, {"n-hexan", "hexane"},
      {"nhexane", "hexane"}, {"n-hexane", "hexane"}, {"tetrahydrofuran", "thf"}, {"h2o", "water"}};
  for (size_t a = 0; a < sizeof(alias) / sizeof(alias[0]); ++a)
    if (!strcmp(out, alias[a][0])) { snprintf(out, cap, "%s", alias[a][1]); break; }
}

/* loadInternalParam: Still kernel -> GFN2/GBSA sets, P16 kernel -> GFN2/ALPB sets */
static const SolvParamSet *solv_lookup(const char *solvent, int kernel) {
  char nm[64];
  solv_normalize(solvent, nm, sizeof(nm));
  if (!strcmp(nm, "dichlormethane")) snprintf(nm, sizeof(nm), "ch2cl2");
  const SolvParamSet *tab = kernel == 1 ? kSolvGbsaGfn2 : kSolvAlpbGfn2;
  const int cnt = kernel == 1 ? kSolvGbsaGfn2Count : kSolvAlpbGfn2Count;
  if (kernel == 1) {
    if (!strcmp(nm, "water")) snprintf(nm, sizeof(nm), "h2o");
    if (!strcmp(nm, "hexane")) snprintf(nm, sizeof(nm), "nhexan");
  }
  for (int i = 0; i < cnt; ++i)
    if (!strcmp(tab[i].name, nm)) return &tab[i];
  return NULL;
}

static double det3(const double m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

/* shape descriptor: getADet / addADetDeriv (src/solv/gbsa.f90:661-777) */
static double solv_adet(int n, const double *xyz, const double *rad, double inertia[3][3], double center[3],
                        double *totrad3) {
  const double tof = 2.0 / 5.0;
  double tot = 0.0;
  center[0] = center[1] = center[2] = 0.0;
  for (int i = 0; i < n; ++i) {
    double r2 = rad[i] * rad[i], r3 = r2 * rad[i];
    tot += r3;
    for (int k = 0; k < 3; ++k) center[k] += xyz[3 * i + k] * r3;
  }
  for (int k = 0; k < 3; ++k) center[k] /= tot;
  memset(inertia, 0, sizeof(double) * 9);
  for (int i = 0; i < n; ++i) {
    double r2 = rad[i] * rad[i], r3 = r2 * rad[i], v[3], d2 = 0.0;
    for (int k = 0; k < 3; ++k) { v[k] = xyz[3 * i + k] - center[k]; d2 += v[k] * v[k]; }
    for (int k = 0; k < 3; ++k) inertia[k][k] += r3 * (d2 + tof * r2);
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) inertia[k][c] -= r3 * v[c] * v[k];
  }
  *totrad3 = tot;
  return sqrt(pow(det3((const double(*)[3])inertia), 1.0 / 3.0) / (tof * tot));
}

/* set-up and geometry update: paramToModel, initBorn, update (Born radii,
 * surface, H-bond term, shape descriptor, interaction matrix) */
static int solv_setup(solv_t *s, const char *solvent, int kernel, int alpb, int n, const int *at, const double *xyz) {
  const SolvParamSet *p = solv_lookup(solvent, kernel);
  if (!p) return 1;
  memset(s, 0, sizeof(*s));
  s->n = n;
  s->kernel = kernel;
  const double eps = p->epsv;
  s->gshift = p->gshift * KCALTOAU;   /* + state shift: 0 for the default gsolv state */
  const double c1 = p->c1, boff = p->soset * 0.1 * AATOAU, probe = p->rprobe * AATOAU;
  int anyhb = 0;
  for (int k = 0; k < 94; ++k) if (fabs(p->hbond[k]) > 1.e-3) anyhb = 1;
  s->alpbet = alpb ? 0.571412 / eps : 0.0;
  s->keps = (1.0 / eps - 1.0) / (1.0 + s->alpbet);
  s->vdwr = NEW(double, n); s->rho = NEW(double, n); s->svdw = NEW(double, n); s->vdwsa = NEW(double, n);
  s->wrp = NEW(double, n); s->trj2 = NEW(double, 2 * n); s->gamsasa = NEW(double, n); s->hbmag = NEW(double, n);
  double maxsa = 0.0;
  for (int i = 0; i < n; ++i) {
    const int z = at[i];
    if (z < 1 || z > 94) return 2;
    s->vdwr[i] = AATOAU * kVdwRadD3Aa[z - 1];
    s->rho[i] = s->vdwr[i] * p->sx[z - 1];
    s->svdw[i] = s->vdwr[i] - boff;
    const double sa = s->vdwr[i] + probe;
    s->vdwsa[i] = sa;
    s->trj2[2 * i] = (sa - SOLV_W) * (sa - SOLV_W);
    s->trj2[2 * i + 1] = (sa + SOLV_W) * (sa + SOLV_W);
    double r = sa + SOLV_W;
    double w = (0.25 / SOLV_W + 3.0 * SOLV_AH3 * (0.2 * r * r - 0.5 * r * sa + sa * sa / 3.0)) * r * r * r;
    r = sa - SOLV_W;
    w = w - (0.25 / SOLV_W + 3.0 * SOLV_AH3 * (0.2 * r * r - 0.5 * r * sa + sa * sa / 3.0)) * r * r * r;
    s->wrp[i] = w;
    s->gamsasa[i] = p->gamscale[z - 1] * (4.0 * PI) * 1.0e-5;
    s->hbmag[i] = anyhb ? -(p->hbond[z - 1] * p->hbond[z - 1]) * KCALTOAU : 0.0;
    if (s->hbmag[i] < 0.0) s->lhb = 1;
    if (sa > maxsa) maxsa = sa;
  }
  const double lrcut = 35.0 * AATOAU, srcut = 2.0 * (SOLV_W + maxsa) + 2.0 * AATOAU;
  const double lr2 = lrcut * lrcut, sr2 = srcut * srcut;
  const size_t nn3 = (size_t)3 * n * n;

  /* Born radii: compute_psi + compute_bornr (OBC alpha 1, beta 0.8, gamma 4.85);
   * psi derivatives dpsi[(i*n + j)*3] = d psi_i / d R_j */
  double *psi = NEW(double, n), *dpsi = NEW(double, nn3), *dtr = NEW(double, 3 * n);
  for (int ii = 0; ii < n; ++ii)
    for (int jj = 0; jj < ii; ++jj) {
      double dr[3], d2 = 0.0;
      for (int k = 0; k < 3; ++k) { dr[k] = xyz[3 * ii + k] - xyz[3 * jj + k]; d2 += dr[k] * dr[k]; }
      if (!(d2 < lr2)) continue;
      const double r = sqrt(d2), r1 = 1.0 / r;
      const double rhoi = s->rho[ii], rhoj = s->rho[jj], rvdwi = s->vdwr[ii], rvdwj = s->vdwr[jj];
      const int ovij = r >= (rvdwi + rhoj) ? 0 : 1, ovji = r >= (rhoi + rvdwj) ? 0 : 1;
      /* contribution of sphere (radius rho_o) around the other atom to the atom with radius rvdw */
      double gi = 0.0, dgi = 0.0, gj = 0.0, dgj = 0.0;
      int has_i = 0, has_j = 0;
      /* non-overlapping form */
#define PSI_FAR(rh, g, dg)                                   \
      do {                                                   \
        const double ap = r + (rh), am = r - (rh), ab = ap * am; \
        const double rhab = (rh) / ab, lnab = 0.5 * log(am / ap) * r1; \
        g = rhab + lnab;                                     \
        dg = -2.0 * rhab / ab + (rhab - lnab) * r1 * r1;     \
      } while (0)
      /* overlapping form */
#define PSI_NEAR(rh, rv, g, dg)                                                                      \
      do {                                                                                           \
        const double r12 = 0.5 * r1, r24 = r12 * r12, ap = r + (rh), am = r - (rh);                  \
        const double rh1 = 1.0 / (rv), rhr1 = 1.0 / ap, aprh1 = ap * rh1, lnab = log(aprh1);         \
        g = rh1 - rhr1 + r12 * (0.5 * am * (rhr1 - rh1 * aprh1) - lnab);                             \
        dg = rhr1 * rhr1 * (1.0 - 0.25 * am * r1 * (1.0 + aprh1 * aprh1)) + (rh) * r24 * (rhr1 - rh1 * aprh1) + \
             r12 * (r1 * lnab - rhr1);                                                               \
        dg = dg * r1;                                                                                \
      } while (0)
      if (!ovij) { PSI_FAR(rhoj, gi, dgi); has_i = 1; }
      else if ((r + rhoj) > rvdwi) { PSI_NEAR(rhoj, rvdwi, gi, dgi); has_i = 1; }
      if (!ovij && !ovji && fabs(rhoi - rhoj) < 1.e-8) { gj = gi; dgj = dgi; has_j = 1; }
      else if (!ovji) { PSI_FAR(rhoi, gj, dgj); has_j = 1; }
      else if ((r + rhoi) > rvdwj) { PSI_NEAR(rhoi, rvdwj, gj, dgj); has_j = 1; }
#undef PSI_FAR
#undef PSI_NEAR
      if (has_i) {
        psi[ii] += gi;
        for (int k = 0; k < 3; ++k) {
          const double d = dgi * dr[k];
          dtr[3 * ii + k] += d;
          dpsi[((size_t)ii * n + jj) * 3 + k] -= d;
        }
      }
      if (has_j) {
        psi[jj] += gj;
        for (int k = 0; k < 3; ++k) {
          const double d = dgj * dr[k];
          dtr[3 * jj + k] -= d;
          dpsi[((size_t)jj * n + ii) * 3 + k] += d;
        }
      }
    }
  for (int i = 0; i < n; ++i) for (int k = 0; k < 3; ++k) dpsi[((size_t)i * n + i) * 3 + k] = dtr[3 * i + k];
  s->brad = NEW(double, n);
  s->brdr = dpsi;
  for (int i = 0; i < n; ++i) {
    const double alp = 1.0, bet = 0.8, gam = 4.85;
    double br = psi[i];
    const double svdwi = s->svdw[i], vdwri = s->vdwr[i], s1 = 1.0 / svdwi, v1 = 1.0 / vdwri, s2 = 0.5 * svdwi;
    br = br * s2;
    double arg2 = br * (gam * br - bet);
    const double arg = br * (alp + arg2);
    arg2 = 2.0 * arg2 + alp + gam * br * br;
    const double th = tanh(arg), ch = cosh(arg);
    br = 1.0 / (s1 - v1 * th);
    br = c1 * br;
    double dp = ch * (s1 - v1 * th);
    dp = s2 * v1 * arg2 / (dp * dp);
    dp = c1 * dp;
    s->brad[i] = br;
    for (int j = 0; j < n; ++j) for (int k = 0; k < 3; ++k) s->brdr[((size_t)i * n + j) * 3 + k] *= dp;
  }
  free(psi); free(dtr);

  /* solvent accessible surface: compute_numsa / compute_w_sp; neighbours in
   * ascending index order within srcut; dsdrt[(i*n + j)*3] = d sasa_i / d R_j */
  s->sasa = NEW(double, n);
  s->dsdrt = NEW(double, nn3);
  {
    int *nb = NEW(int, n), *gi = NEW(int, n);
    double *gd = NEW(double, 3 * n);
    for (int i = 0; i < n; ++i) {
      int nno = 0;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) { double d = xyz[3 * i + k] - xyz[3 * j + k]; d2 += d * d; }
        if (d2 < lr2 && d2 < sr2) nb[nno++] = j;
      }
      const double rsas = s->vdwsa[i], wr = s->wrp[i];
      double sasai = 0.0;
      double *grads = s->dsdrt + (size_t)i * n * 3;
      for (int ip = 0; ip < 230; ++ip) {
        double xp[3];
        for (int k = 0; k < 3; ++k) xp[k] = xyz[3 * i + k] + rsas * kLebedev230[ip][k];
        double sp = 1.0;
        int nni = 0;
        for (int m = 0; m < nno; ++m) {
          const int a = nb[m];
          double tj[3], tj2;
          for (int k = 0; k < 3; ++k) tj[k] = xp[k] - xyz[3 * a + k];
          tj2 = tj[0] * tj[0] + tj[1] * tj[1] + tj[2] * tj[2];
          if (tj2 < s->trj2[2 * a + 1]) {
            if (tj2 <= s->trj2[2 * a]) { sp = 0.0; break; }
            const double sqtj = sqrt(tj2), uj = sqtj - s->vdwsa[a], ah3uj2 = SOLV_AH3 * uj * uj;
            double ds = SOLV_AH1 + 3.0 * ah3uj2;
            double sij = SOLV_AH0 + (SOLV_AH1 + ah3uj2) * uj;
            sij = fmin(sij, 1.0);
            sij = fmax(0.0, sij);
            if (sij < 2.2250738585072014e-308) { sp = 0.0; break; }
            sp = sp * sij;
            ds = ds / (sij * sqtj);
            gi[nni] = a;
            for (int k = 0; k < 3; ++k) gd[3 * nni + k] = ds * tj[k];
            ++nni;
          }
        }
        if (sp > 1.e-6) {
          const double wsa = kLebedev230[ip][3] * wr * sp;
          sasai += wsa;
          for (int m = 0; m < nni; ++m)
            for (int k = 0; k < 3; ++k) {
              const double d = wsa * gd[3 * m + k];
              grads[3 * i + k] += d;
              grads[3 * gi[m] + k] -= d;
            }
        }
      }
      s->sasa[i] = sasai;
    }
    free(nb); free(gi); free(gd);
  }
  s->dsdr = NEW(double, 3 * n);
  s->gsasa = 0.0;
  for (int i = 0; i < n; ++i) {
    for (size_t x = 0; x < (size_t)3 * n; ++x) s->dsdr[x] += s->dsdrt[(size_t)i * n * 3 + x] * s->gamsasa[i];
    s->gsasa += s->sasa[i] * s->gamsasa[i];
  }
  /* hydrogen bond term: compute_fhb */
  s->hbw = NEW(double, n);
  s->dhbdw = NEW(double, n);
  if (s->lhb)
    for (int i = 0; i < n; ++i) {
      const double smaxd = 1.0 / (s->vdwsa[i] * s->vdwsa[i]);
      s->hbw[i] = s->hbmag[i] * (s->sasa[i] * smaxd);
      s->dhbdw[i] = s->hbmag[i] * smaxd;
    }
  if (s->alpbet > 0.0) {
    double inertia[3][3], center[3], tot;
    s->adet = solv_adet(n, xyz, s->vdwr, inertia, center, &tot);
  }
  /* interaction matrix: addBornMatStill / addBornMatP16 + H-bond + ALPB */
  s->bmat = NEW(double, (size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) { double d = xyz[3 * i + k] - xyz[3 * j + k]; d2 += d * d; }
      const double r1 = sqrt(d2);
      double f;
      if (kernel == 1) {
        const double r2 = r1 * r1, aa = s->brad[i] * s->brad[j], dd = 0.25 * r2 / aa, expd = exp(-dd);
        f = 1.0 / sqrt(r2 + aa * expd);
      } else {
        const double ab = sqrt(s->brad[i] * s->brad[j]);
        double arg = ab / (ab + (1.028 / 16.0) * r1);
        arg = arg * arg; arg = arg * arg; arg = arg * arg; arg = arg * arg;
        f = 1.0 / (r1 + ab * arg);
      }
      s->bmat[(size_t)i * n + j] = s->keps * f + s->bmat[(size_t)i * n + j];
      s->bmat[(size_t)j * n + i] = s->keps * f + s->bmat[(size_t)j * n + i];
    }
  for (int i = 0; i < n; ++i) s->bmat[(size_t)i * n + i] += s->keps * (1.0 / s->brad[i]);
  if (s->lhb) for (int i = 0; i < n; ++i) s->bmat[(size_t)i * n + i] += 2 * s->hbw[i];
  if (s->alpbet > 0.0) {
    const double c = s->keps * s->alpbet / s->adet;
    for (size_t x = 0; x < (size_t)n * n; ++x) s->bmat[x] += c;
  }
  return 0;
}

/* solvation energy at charges q: getEnergy */
static double solv_energy(const solv_t *s, const double *q, double *work) {
  symv(s->n, s->bmat, q, work, 0.5);
  double e = 0.0;
  for (int i = 0; i < s->n; ++i) e += q[i] * work[i];
  return e + s->gsasa + s->gshift;
}

/* gradient: addGradientStill / addGradientP16, surface, H-bond, ALPB shape */
static void solv_gradient(const solv_t *s, const double *xyz, const double *q, double *grad) {
  const int n = s->n;
  const double keps = s->keps;
  double *grddb = NEW(double, n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      double v[3], d2 = 0.0;
      for (int k = 0; k < 3; ++k) { v[k] = xyz[3 * i + k] - xyz[3 * j + k]; d2 += v[k] * v[k]; }
      const double r = sqrt(d2), qq = q[i] * q[j];
      if (s->kernel == 1) {
        const double r2 = r * r, aa = s->brad[i] * s->brad[j], dd = 0.25 * r2 / aa, expd = exp(-dd);
        const double fgb2 = r2 + aa * expd, dfgb2 = 1.0 / fgb2, dfgb = sqrt(dfgb2), dfgb3 = dfgb2 * dfgb * keps;
        const double ap = (1.0 - 0.25 * expd) * dfgb3;
        for (int k = 0; k < 3; ++k) {
          const double dr = ap * v[k];
          grad[3 * i + k] = grad[3 * i + k] - dr * qq;
          grad[3 * j + k] = grad[3 * j + k] + dr * qq;
        }
        const double bp = -0.5 * expd * (1.0 + dd) * dfgb3;
        grddb[i] = grddb[i] + (s->brad[j] * bp) * qq;
        grddb[j] = grddb[j] + (s->brad[i] * bp) * qq;
      } else {
        const double z = 1.028, ab = sqrt(s->brad[i] * s->brad[j]);
        const double arg1 = ab / (ab + (z / 16.0) * r);
        double a16 = arg1 * arg1;
        a16 = a16 * a16; a16 = a16 * a16; a16 = a16 * a16;
        const double fgb = r + ab * a16, dfgb = 1.0 / fgb, dfgb2 = dfgb * dfgb;
        const double ap = (1.0 - z * arg1 * a16) * dfgb2;
        for (int k = 0; k < 3; ++k) {
          const double dG = ap * v[k] * keps / r * qq;
          grad[3 * i + k] = grad[3 * i + k] - dG;
          grad[3 * j + k] = grad[3 * j + k] + dG;
        }
        const double bp = -0.5 * (r * z / ab * arg1 + 1.0) / ab * a16 * dfgb2;
        grddb[i] = grddb[i] + s->brad[j] * bp * keps * qq;
        grddb[j] = grddb[j] + s->brad[i] * bp * keps * qq;
      }
    }
  for (int i = 0; i < n; ++i) {
    const double bp = 1.0 / s->brad[i], qq = q[i] * bp;
    grddb[i] = grddb[i] + (-0.5 * keps * qq * bp) * q[i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < 3; ++k) grad[3 * j + k] += s->brdr[((size_t)i * n + j) * 3 + k] * grddb[i];
  free(grddb);
  for (int x = 0; x < 3 * n; ++x) grad[x] += s->dsdr[x];
  if (s->lhb)
    for (int i = 0; i < n; ++i) {
      double dh = s->dhbdw[i];
      if (fabs(dh) <= 0.0) continue;
      dh = dh * q[i] * q[i];
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < 3; ++k) grad[3 * j + k] += s->dsdrt[((size_t)i * n + j) * 3 + k] * dh;
    }
  if (s->alpbet > 0.0) {
    double I[3][3], center[3], tot, qt = 0.0;
    for (int i = 0; i < n; ++i) qt += q[i];
    const double adet = solv_adet(n, xyz, s->vdwr, I, center, &tot);
    const double ke = s->keps * s->alpbet;
    const double f = (250.0 / (48.0 * tot * tot * tot * pow(adet, 5))) * (-0.5 * ke * qt * qt / (adet * adet));
    double A[3][3];
    A[0][0] = I[0][0] * (I[1][1] + I[2][2]) - I[0][1] * I[0][1] - I[0][2] * I[0][2];
    A[1][0] = A[0][1] = I[0][1] * I[2][2] - I[0][2] * I[1][2];
    A[2][0] = A[0][2] = I[0][2] * I[1][1] - I[0][1] * I[2][1];
    A[1][1] = I[1][1] * (I[0][0] + I[2][2]) - I[0][1] * I[0][1] - I[1][2] * I[1][2];
    A[2][1] = A[1][2] = I[0][0] * I[1][2] - I[0][1] * I[0][2];
    A[2][2] = I[2][2] * (I[0][0] + I[1][1]) - I[0][2] * I[0][2] - I[1][2] * I[1][2];
    for (int i = 0; i < n; ++i) {
      const double r3 = s->vdwr[i] * s->vdwr[i] * s->vdwr[i];
      double v[3];
      for (int k = 0; k < 3; ++k) v[k] = xyz[3 * i + k] - center[k];
      for (int k = 0; k < 3; ++k) grad[3 * i + k] += r3 * ((A[k][0] * f) * v[0] + (A[k][1] * f) * v[1] + (A[k][2] * f) * v[2]);
    }
  }
}

int om_singlepoint(const om_options *opt_in, int n, const int *at, const double *xyz, double charge, int uhf, om_result *res) {
  init_elements();
  om_options opt = {1.0, 300.0, 250, 0, 0, NULL, 1, 0, NULL, NULL, NULL};
  if (opt_in) opt = *opt_in;
  if (opt.lapack) lapack_load();
  const int use_lib = opt.lapack && p_dsygvd && p_dgemm;
  basis_t b;
  if (n < 1) return 1;
  if (basis_build(n, at, &b)) return 1;
  const int N = b.nao, NS = b.nsh;
  const size_t NN = (size_t)N * N;
  int ret = 0;

  /* thresholds: src/scf_module.F90:264-277 */
  const double acc = opt.accuracy;
  double intcut = 25.0 - 10.0 * log10(acc);
  if (intcut < 20.0) intcut = 20.0;
  const double neglect = 10.0e-9 * acc;
  const double scfconv = 1.e-6 * acc;
  const double qconv = 1.e-4 * acc;
  const double broydamp = 0.40;
  const int broy = n > 1;

  /* electrons and occupations: src/scf_module.F90:332-354 */
  double zsum = 0.0;
  for (int i = 0; i < n; ++i) zsum += EL[at[i]].zval;
  int nel = (int)lround(zsum - charge);
  int nopen = uhf;
  if (nopen == 0 && nel % 2 != 0) nopen = 1;
  if (nel > 2 * N) { basis_free(&b); return 3; }
  double *zsh = NEW(double, NS);
  double eatoms = 0.0;
  for (int i = 0; i < n; ++i) { /* setzshell: src/scc_core.f90:1509-1542 */
    const elem_t *e = &EL[at[i]];
    double ntot = -1.0e-6;
    for (int m = 0; m < e->nsh; ++m) {
      int s = b.at_sh0[i] + m;
      zsh[s] = e->refocc[m];
      ntot += zsh[s];
      if (ntot > e->zval) zsh[s] = 0.0;
      eatoms += e->selfe[m] * zsh[s];
    }
  }
  double *focca = NEW(double, N), *foccb = NEW(double, N), *focc = NEW(double, N);
  int ihomoa = 0, ihomob = 0;
  occu(N, nel, nopen, &ihomoa, &ihomob, focca, foccb);
  const int ihomo = ihomoa;

  /* initial charges: EEQ (newWavefunction src/xtb/calculator.f90:432-496) or SAD */
  double *q = NEW(double, n), *qsh = NEW(double, NS);
  if (opt.guess == 0) {
    if (om_eeq_guess(n, at, xyz, charge, q)) for (int i = 0; i < n; ++i) q[i] = charge / n;
  } else for (int i = 0; i < n; ++i) q[i] = charge / n;
  for (int i = 0; i < n; ++i) { /* iniqshell: src/scc_core.f90:1475-1506 */
    const elem_t *e = &EL[at[i]];
    double ntot = -1.0e-6;
    for (int m = 0; m < e->nsh; ++m) {
      double zs = e->refocc[m];
      ntot += zs;
      if (ntot > e->zval) zs = 0.0;
      qsh[b.at_sh0[i] + m] = (zs / e->zval) * q[i];
    }
  }
  double *dipm = NEW(double, 3 * n), *qp = NEW(double, 6 * n);
  if (opt.init_qsh) { /* restart from a previous wavefunction (src/api/interface.f90:83-101) */
    memcpy(qsh, opt.init_qsh, sizeof(double) * NS);
    for (int i = 0; i < n; ++i) q[i] = 0.0;
    for (int s = 0; s < NS; ++s) q[b.sh_at[s]] += qsh[s];
    if (opt.init_dipm) memcpy(dipm, opt.init_dipm, sizeof(double) * 3 * n);
    if (opt.init_qp) memcpy(qp, opt.init_qp, sizeof(double) * 6 * n);
  }

  /* classical terms */
  double *grad = NEW(double, 3 * n);
  double ep = repulsion(n, at, xyz, grad);
  double *cn = NEW(double, n), *dcndr = NEW(double, 3 * n * n);
  cn_lp(n, at, xyz, 0, cn, dcndr);
  d4_t d4;
  memset(&d4, 0, sizeof(d4));
  d4_setup(n, at, xyz, &d4);
  double *jmat = NEW(double, (size_t)NS * NS);
  coulomb_matrix(n, at, xyz, &b, jmat);
  double *shgam = NEW(double, NS);
  for (int i = 0; i < n; ++i) for (int m = 0; m < EL[at[i]].nsh; ++m) shgam[b.at_sh0[i] + m] = EL[at[i]].third[m];
  /* implicit solvation (src/scf_module.F90:299-321) */
  solv_t sv;
  memset(&sv, 0, sizeof(sv));
  double *svwork = NEW(double, n);
  if (opt.solvent && opt.solvent[0]) {
    if (solv_setup(&sv, opt.solvent, opt.solv_kernel == 2 ? 2 : 1, opt.solv_alpb, n, at, xyz)) {
      solv_free(&sv);
      free(svwork);
      free(shgam); free(jmat); d4_free(&d4); free(cn); free(dcndr); free(grad);
      free(dipm); free(qp); free(q); free(qsh); free(focca); free(foccb); free(focc); free(zsh);
      basis_free(&b);
      return 6;
    }
  }

  /* integrals */
  double *selfe = NEW(double, NS), *dsedcn = NEW(double, NS);
  self_energies(n, at, &b, cn, selfe, dsedcn);
  double *S = NEW(double, NN), *D = NEW(double, 3 * NN), *Q = NEW(double, 6 * NN);
  double *H0 = NEW(double, NN), *H0n = NEW(double, NN);
  build_sdqh0(n, at, xyz, &b, selfe, intcut, S, D, Q, H0, H0n);
  /* dipole/quadrupole lists: setdqlist (src/aespot.F90:34-85), thr2 = neglect*1e-2 */
  unsigned char *dmask = NEW(unsigned char, NN), *qmask = NEW(unsigned char, NN);
  {
    const double thr2 = neglect * 1.0e-2;
    for (size_t i = 0; i < (size_t)N; ++i)
      for (size_t j = 0; j <= i; ++j) {
        double t1 = 0.0, t2 = 0.0;
        for (int k = 0; k < 3; ++k) { double d = D[k * NN + i * N + j]; t1 += d * d; }
        for (int k = 0; k < 3; ++k) { double v = Q[k * NN + i * N + j]; t2 -= v * v; }
        for (int k = 0; k < 6; ++k) { double v = Q[k * NN + i * N + j]; t2 += 2.0 * v * v; }
        dmask[i * N + j] = dmask[j * N + i] = t1 > thr2;
        qmask[i * N + j] = qmask[j * N + i] = t2 > thr2;
      }
  }
  double *radcn = NEW(double, n), *gab3 = NEW(double, n * n), *gab5 = NEW(double, n * n);
  aes_setup(n, at, xyz, cn, radcn, gab3, gab5);

  /* Cholesky of S (src/scc_core.f90:411-413) */
  double *U = NEW(double, NN);
  memcpy(U, S, sizeof(double) * NN);
  if (!use_lib && cholesky_upper(N, U)) { ret = 4; goto cleanup_early; }

  /* ---- SCC iterations ------------------------------------------------- */
  {
    const int thisiter = opt.maxiter;
    const int nbr = NS + 9 * n;
    broyden_t B;
    B.n = nbr; B.maxit = thisiter;
    B.df = NEW(double, (size_t)thisiter * nbr); B.u = NEW(double, (size_t)thisiter * nbr);
    B.a = NEW(double, (size_t)thisiter * thisiter); B.omega = NEW(double, thisiter);
    B.qlast = NEW(double, nbr); B.dqlast = NEW(double, nbr);
    double *q_in = NEW(double, nbr), *dq = NEW(double, nbr);
    double *H = NEW(double, NN), *C = NEW(double, NN), *P = NEW(double, NN), *emo = NEW(double, N);
    double *shift = NEW(double, NS), *ashift = NEW(double, n);
    double *vs = NEW(double, n), *vd = NEW(double, 3 * n), *vq = NEW(double, 6 * n);
    double *tmp = NEW(double, NN), *jq = NEW(double, NS);
    double eel = 0.0, eold, ees = 0.0, eaes = 0.0, epol = 0.0, ed = 0.0, egap = 0.0, ga = 0.0, gb = 0.0;
    int converged = 0, lastdiag = 0, iter;
    double damp = 0.20; /* damp0: src/scf_module.F90:274 */
    double efa = 0, efb = 0;
    for (iter = 1; iter <= thisiter; ++iter) {
      /* potentials: ies%addShift, setvsdq, scD4%addShift, addToShellShift */
      for (int i = 0; i < n; ++i) ashift[i] = 0.0;
      for (int s = 0; s < NS; ++s) shift[s] = qsh[s] * qsh[s] * shgam[s];
      symv(NS, jmat, qsh, jq, 1.0);
      for (int s = 0; s < NS; ++s) shift[s] += jq[s];
      aes_potential(n, at, xyz, q, dipm, qp, gab3, gab5, 1, vs, vd, vq);
      if (sv.n) { /* solvation%addShift (src/solv/gbsa.f90:495-521) */
        symv(n, sv.bmat, q, svwork, 1.0);
        for (int i = 0; i < n; ++i) ashift[i] = ashift[i] + svwork[i];
      }
      d4_potential(&d4, n, at, q, ashift);
      for (int s = 0; s < NS; ++s) shift[s] += ashift[b.sh_at[s]];
      /* Fock matrix: buildIsoAnisotropicH1 (src/scc_core.f90:154-251), eV */
      for (size_t i = 0; i < (size_t)N; ++i)
        for (size_t j = 0; j <= i; ++j) {
          double dum = S[i * N + j] * AUTOEV * 0.5;
          int ii = b.ao_sh[i], jj = b.ao_sh[j];
          double h = H0[i * N + j] - dum * (shift[ii] + shift[jj]);
          int ia = b.ao_at[i], ja = b.ao_at[j];
          h += dum * (vs[ia] + vs[ja]);
          if (dmask[i * N + j]) {
            double e1 = 0.0;
            for (int l = 0; l < 3; ++l) e1 += D[l * NN + i * N + j] * (vd[3 * ia + l] + vd[3 * ja + l]);
            h += 0.50 * e1 * AUTOEV;
          }
          if (qmask[i * N + j]) {
            double e1 = 0.0;
            for (int l = 0; l < 6; ++l) e1 += Q[l * NN + i * N + j] * (vq[6 * ia + l] + vq[6 * ja + l]);
            h += 0.50 * e1 * AUTOEV;
          }
          H[i * N + j] = H[j * N + i] = h;
        }
      if (gen_eigen(N, H, S, U, C, emo, use_lib)) { ret = 5; break; }
      if (ihomo + 1 <= N && ihomo >= 1) egap = emo[ihomo] - emo[ihomo - 1];
      /* Fermi smearing */
      if (opt.etemp > 0.1) {
        int ha, hb;
        if (nel > 0) occu(N, nel, nopen, &ha, &hb, focca, foccb);
        else { for (int i = 0; i < N; ++i) focca[i] = foccb[i] = 0.0; ha = hb = 0; }
        fermismear(N, ha, opt.etemp, emo, focca, &efa, &ga);
        fermismear(N, hb, opt.etemp, emo, foccb, &efb, &gb);
      } else { ga = gb = 0.0; }
      for (int i = 0; i < N; ++i) focc[i] = focca[i] + foccb[i];
      /* save input charges */
      for (int s = 0; s < NS; ++s) q_in[s] = qsh[s];
      for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) q_in[NS + 9 * i + k] = dipm[3 * i + k];
        for (int k = 0; k < 6; ++k) q_in[NS + 9 * i + 3 + k] = qp[6 * i + k];
      }
      /* density: dmat (src/scc_core.f90:1177-1205) */
      for (size_t i = 0; i < (size_t)N; ++i) for (int k = 0; k < N; ++k) tmp[i * N + k] = C[i * N + k] * focc[k];
      gemm_nt(N, C, tmp, P, use_lib);
      /* Mulliken shell populations: mpopsh (src/scc_core.f90:1408-1429) */
      for (int s = 0; s < NS; ++s) qsh[s] = 0.0;
      for (size_t i = 0; i < (size_t)N; ++i) {
        int ii = b.ao_sh[i];
        for (size_t j = 0; j < i; ++j) {
          double ps = P[i * N + j] * S[i * N + j];
          qsh[ii] += ps;
          qsh[b.ao_sh[j]] += ps;
        }
        qsh[ii] += P[i * N + i] * S[i * N + i];
      }
      for (int s = 0; s < NS; ++s) qsh[s] = zsh[s] - qsh[s];
      for (int i = 0; i < n; ++i) q[i] = 0.0;
      for (int s = 0; s < NS; ++s) q[b.sh_at[s]] += qsh[s];
      /* energy: electro (src/scc_core.f90:686-722) */
      eold = eel;
      {
        double e3 = 0.0;
        for (int s = 0; s < NS; ++s) e3 += qsh[s] * qsh[s] * qsh[s] * shgam[s] / 3.0;
        symv(NS, jmat, qsh, jq, 1.0);
        double ej = 0.0;
        for (int s = 0; s < NS; ++s) ej += jq[s] * qsh[s];
        ees = 0.5 * ej + e3;
        double h = 0.0;
        for (size_t i = 0; i < (size_t)N; ++i) {
          for (size_t j = 0; j < i; ++j) h += P[i * N + j] * H0[i * N + j];
          h += P[i * N + i] * H0[i * N + i] * 0.5;
        }
        eel = ees + 2.0 * h * EVTOAU;
      }
      /* CAMM: mmompop (src/aespot.F90:289-427) */
      memset(dipm, 0, sizeof(double) * 3 * n);
      memset(qp, 0, sizeof(double) * 6 * n);
      for (size_t i = 0; i < (size_t)N; ++i)
        for (size_t j = 0; j <= i; ++j) {
          int ii = b.ao_at[i], jj = b.ao_at[j];
          const double *ra = xyz + 3 * ii, *rb = xyz + 3 * jj;
          double pij = P[i * N + j], ps = pij * S[i * N + j];
          int diag = (i == j);
          for (int k = 1; k <= 3; ++k) {
            double xk1 = ra[k - 1], xk2 = rb[k - 1];
            double pdmk = pij * D[(k - 1) * NN + i * N + j];
            dipm[3 * ii + k - 1] += xk1 * ps - pdmk;
            if (!diag) dipm[3 * jj + k - 1] += xk2 * ps - pdmk;
            for (int l = 1; l < k; ++l) {
              int kl = k * (k - 1) / 2 + l - 1, kj = k + l + 1 - 1;
              double xl1 = ra[l - 1], xl2 = rb[l - 1];
              double pdml = pij * D[(l - 1) * NN + i * N + j];
              double pqm = pij * Q[kj * NN + i * N + j];
              qp[6 * ii + kl] += pdmk * xl1 + pdml * xk1 - xl1 * xk1 * ps - pqm;
              if (!diag) qp[6 * jj + kl] += pdmk * xl2 + pdml * xk2 - xl2 * xk2 * ps - pqm;
            }
            int kl = k * (k + 1) / 2 - 1;
            double pqm = pij * Q[(k - 1) * NN + i * N + j];
            qp[6 * ii + kl] += 2.0 * pdmk * xk1 - xk1 * xk1 * ps - pqm;
            if (!diag) qp[6 * jj + kl] += 2.0 * pdmk * xk2 - xk2 * xk2 * ps - pqm;
          }
        }
      for (int i = 0; i < n; ++i) {
        double tii = 0.50 * (qp[6 * i] + qp[6 * i + 2] + qp[6 * i + 5]);
        for (int k = 0; k < 6; ++k) qp[6 * i + k] *= 1.50;
        qp[6 * i] -= tii; qp[6 * i + 2] -= tii; qp[6 * i + 5] -= tii;
      }
      aes_energy(n, at, xyz, q, dipm, qp, gab3, gab5, &eaes, &epol);
      eel += eaes + epol;
      ed = d4_energy(&d4, n, at, q);
      eel += ed;
      if (sv.n) eel = eel + solv_energy(&sv, q, svwork);   /* src/scc_core.f90:544-548 */
      eel += ga + gb;
      int econv = fabs(eel - eold) < scfconv;
      for (int s = 0; s < NS; ++s) dq[s] = qsh[s] - q_in[s];
      for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) dq[NS + 9 * i + k] = dipm[3 * i + k] - q_in[NS + 9 * i + k];
        for (int k = 0; k < 6; ++k) dq[NS + 9 * i + 3 + k] = qp[6 * i + k] - q_in[NS + 9 * i + 3 + k];
      }
      double rmsq = 0.0;
      for (int k = 0; k < nbr; ++k) rmsq += dq[k] * dq[k];
      rmsq = sqrt(rmsq / (double)n);
      int qconv_ok = rmsq < qconv;
      if (!broy) {
        /* simple damping for a single atom: src/scc_core.f90:574-602 */
        for (int s = 0; s < NS; ++s) qsh[s] = damp * qsh[s] + (1.0 - damp) * q_in[s];
        for (int i = 0; i < n; ++i) {
          for (int k = 0; k < 3; ++k) dipm[3 * i + k] = damp * dipm[3 * i + k] + (1.0 - damp) * q_in[NS + 9 * i + k];
          for (int k = 0; k < 6; ++k) qp[6 * i + k] = damp * qp[6 * i + k] + (1.0 - damp) * q_in[NS + 9 * i + 3 + k];
        }
        damp = (eel - eold < 0.0) ? damp * 1.15 : 0.20;
        if (damp > 1.0) damp = 1.0;
        if (egap < 1.0 && damp > 0.5) damp = 0.5;
      } else {
        broyden(&B, q_in, dq, iter, broydamp);
        for (int s = 0; s < NS; ++s) qsh[s] = q_in[s];
        for (int i = 0; i < n; ++i) {
          for (int k = 0; k < 3; ++k) dipm[3 * i + k] = q_in[NS + 9 * i + k];
          for (int k = 0; k < 6; ++k) qp[6 * i + k] = q_in[NS + 9 * i + 3 + k];
        }
      }
      for (int i = 0; i < n; ++i) q[i] = 0.0;
      for (int s = 0; s < NS; ++s) q[b.sh_at[s]] += qsh[s];
      if (econv && qconv_ok) {
        converged = 1;
        if (lastdiag) break;
        lastdiag = 1;
      }
    }
    int jter = iter <= thisiter ? iter : thisiter;

    /* ---- gradient (src/scf_module.F90:665-800) ------------------------ */
    double *Pew = NEW(double, NN);
    for (size_t i = 0; i < (size_t)N; ++i) for (int k = 0; k < N; ++k) tmp[i * N + k] = C[i * N + k] * (focc[k] * emo[k] * EVTOAU);
    gemm_nt(N, C, tmp, Pew, use_lib);
    double *ves = shift; /* last SCC shell shift */
    aes_potential(n, at, xyz, q, dipm, qp, gab3, gab5, 0, vs, vd, vq);
    double *dhdcn = NEW(double, n);
    /* build_dSDQH0_noreset (src/xtb/hamiltonian.F90:575-792) */
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) {
        int izp = at[i], jzp = at[j];
        const double *ri = xyz + 3 * i, *rj = xyz + 3 * j;
        double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
        double rij2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
        if (rij2 > 1600.0) continue;
        for (int is = 0; is < EL[izp].nsh; ++is) {
          int ish = b.at_sh0[i] + is, li = b.sh_l[ish];
          for (int js = 0; js < EL[jzp].nsh; ++js) {
            int jsh = b.at_sh0[j] + js, lj = b.sh_l[jsh];
            double hii = selfe[ish], hjj = selfe[jsh];
            /* dshellPoly (src/grad_core.f90:40-72) */
            double rab = sqrt(rij2), rr = rab / (EL[izp].rad + EL[jzp].rad);
            double k1 = EL[izp].shpoly_l[li] * 0.01, k2 = EL[jzp].shpoly_l[lj] * 0.01;
            double t14 = sqrt(rr), t15 = k1 * t14, t17 = 1.0 / rij2, t23 = k2 * t14;
            double shpoly = (1.0 + t15) * (1.0 + k2 * t14);
            double dshp[3];
            for (int k = 0; k < 3; ++k) dshp[k] = (t15 * (1.0 + t23) + (1.0 + t15) * k2 * t14) * 0.5 * t17 * rij[k];
            double hav = 0.5 * (hii + hjj);
            sdq_blk blk;
            sdq_grad gb2;
            shell_pair_sdq(&b, ish, jsh, ri, rj, rj, intcut, &blk, &gb2);
            double gxyz[3] = {0, 0, 0}, dcn = 0.0;
            for (int ii = 0; ii < 2 * li + 1; ++ii)
              for (int jj = 0; jj < 2 * lj + 1; ++jj) {
                size_t ia = b.sh_ao0[ish] + ii, ja = b.sh_ao0[jsh] + jj;
                double pij = P[ja * N + ia];
                double hij = H0n[ja * N + ia] * EVTOAU;
                double hpij = hij * pij;
                double sij = S[ja * N + ia];
                double fac = 2 * hpij - 2 * Pew[ja * N + ia] - pij * (ves[ish] + ves[jsh]) + pij * (vs[i] + vs[j]);
                for (int x = 0; x < 3; ++x) {
                  /* moments about R_i from those about R_j (shiftintg) */
                  double s0 = blk.s[ii][jj];
                  double dj[3], qj[6];
                  for (int k = 0; k < 3; ++k) dj[k] = blk.d[k][ii][jj];
                  for (int k = 0; k < 6; ++k) qj[k] = blk.q[k][ii][jj];
                  double gs = gb2.s[x][ii][jj];
                  double gdj[3], gqj[6], gdi[3], gqi[6];
                  for (int k = 0; k < 3; ++k) gdj[k] = gb2.d[x][k][ii][jj];
                  for (int k = 0; k < 6; ++k) gqj[k] = gb2.q[x][k][ii][jj];
                  const double *r = rij;
                  for (int k = 0; k < 3; ++k) gdi[k] = gdj[k] - r[k] * gs;
                  gqi[0] = gqj[0] - 2 * r[0] * gdj[0] + r[0] * r[0] * gs;
                  gqi[1] = gqj[1] - 2 * r[1] * gdj[1] + r[1] * r[1] * gs;
                  gqi[2] = gqj[2] - 2 * r[2] * gdj[2] + r[2] * r[2] * gs;
                  gqi[3] = gqj[3] - r[0] * gdj[1] - r[1] * gdj[0] + r[0] * r[1] * gs;
                  gqi[4] = gqj[4] - r[0] * gdj[2] - r[2] * gdj[0] + r[0] * r[2] * gs;
                  gqi[5] = gqj[5] - r[1] * gdj[2] - r[2] * gdj[1] + r[1] * r[2] * gs;
                  /* explicit dependence of the operator centre R_i */
                  if (x == 0) { gdi[0] -= s0; gqi[0] += -2 * dj[0] + 2 * r[0] * s0; gqi[3] += -dj[1] + r[1] * s0; gqi[4] += -dj[2] + r[2] * s0; }
                  if (x == 1) { gdi[1] -= s0; gqi[1] += -2 * dj[1] + 2 * r[1] * s0; gqi[3] += -dj[0] + r[0] * s0; gqi[5] += -dj[2] + r[2] * s0; }
                  if (x == 2) { gdi[2] -= s0; gqi[2] += -2 * dj[2] + 2 * r[2] * s0; gqi[4] += -dj[0] + r[0] * s0; gqi[5] += -dj[1] + r[1] * s0; }
                  (void)qj;
                  double gv = 2 * hpij * sij * dshp[x] / shpoly + gs * fac;
                  double dtmp = 0.0, qtmp = 0.0;
                  for (int k = 0; k < 3; ++k) dtmp += gdi[k] * vd[3 * i + k] + gdj[k] * vd[3 * j + k];
                  for (int k = 0; k < 6; ++k) qtmp += gqi[k] * vq[6 * i + k] + gqj[k] * vq[6 * j + k];
                  gxyz[x] += gv + pij * dtmp + pij * qtmp;
                }
                dcn += hpij / hav * sij;
              }
            dhdcn[i] += dcn * dsedcn[ish];
            dhdcn[j] += dcn * dsedcn[jsh];
            for (int x = 0; x < 3; ++x) { grad[3 * i + x] += gxyz[x]; grad[3 * j + x] -= gxyz[x]; }
          }
        }
      }
    for (int i = 0; i < n; ++i)
      for (int m = 0; m < EL[at[i]].nsh; ++m) {
        int ish = b.at_sh0[i] + m;
        for (int k = 0; k < b.sh_nao[ish]; ++k) {
          size_t a = b.sh_ao0[ish] + k;
          dhdcn[i] += P[a * N + a] * dsedcn[ish] * EVTOAU;
        }
      }
    for (int a = 0; a < n; ++a)
      for (int bb = 0; bb < n; ++bb)
        for (int k = 0; k < 3; ++k) grad[3 * a + k] += dcndr[(bb * n + a) * 3 + k] * dhdcn[bb];
    /* AES gradient: setdvsdq, dradcn, aniso_grad */
    {
      const double shift_ = 1.2, expo = 4.0, rmax = 5.0;
      double *drad = NEW(double, 3 * n * n);
      for (int i = 0; i < n; ++i) {
        const elem_t *e = &EL[at[i]];
        double t1 = exp(-expo * (cn[i] - e->valcn - shift_));
        double t2 = (rmax - e->mrad) / (1.0 + 2.0 * t1 + t1 * t1);
        t2 = t2 * expo * t1;
        for (int a = 0; a < n; ++a) for (int k = 0; k < 3; ++k) drad[(i * n + a) * 3 + k] = dcndr[(i * n + a) * 3 + k] * t2;
      }
      aes_gradient(n, xyz, q, dipm, qp, radcn, drad, gab3, gab5, grad);
      free(drad);
    }
    /* D4 gradient + ATM energy */
    double embd = 0.0;
    d4_gradient(&d4, n, at, xyz, q, grad, &embd);
    /* solvation gradient (src/scf_module.F90:761-777) */
    if (sv.n) solv_gradient(&sv, xyz, q, grad);
    /* Coulomb derivatives: getCoulombDerivsCluster + gemv(djdr, qsh) */
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) {
        const elem_t *ei = &EL[at[i]], *ej = &EL[at[j]];
        double vec[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1], xyz[3 * j + 2] - xyz[3 * i + 2]};
        double r1 = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
        for (int is = 0; is < ei->nsh; ++is)
          for (int js = 0; js < ej->nsh; ++js) {
            double gij = 0.5 * (ei->hard[is] + ej->hard[js]);
            double g1 = 1.0 / (pow(r1, 2.0) + pow(gij, -2.0));
            double qi = qsh[b.at_sh0[i] + is], qj = qsh[b.at_sh0[j] + js];
            for (int k = 0; k < 3; ++k) {
              double dG = -vec[k] * pow(r1, 0.0) * g1 * pow(g1, 0.5);
              grad[3 * i + k] -= dG * qi * qj;
              grad[3 * j + k] += dG * qj * qi;
            }
          }
      }
    /* energies: src/scf_module.F90:890-913 */
    double energy = eel + ep + embd;
    /* Wiberg bond orders (get_wiberg src/scc_core.f90:1208-1245) and dipole */
    if (res) {
      res->n = n; res->nsh = NS; res->nao = N; res->ndim_d4 = d4.ndim;
      res->iterations = jter; res->converged = converged; res->ihomo = ihomo;
      res->e_total = energy; res->e_elec = eel + embd; res->e_es = ees; res->e_aes = eaes; res->e_axc = epol;
      res->e_disp = ed + embd; res->e_rep = ep; res->e_atm = embd; res->hl_gap_ev = egap;
      res->e_solv = sv.n ? solv_energy(&sv, q, svwork) : 0.0;
      double gn = 0.0;
      for (int k = 0; k < 3 * n; ++k) gn += grad[k] * grad[k];
      res->gnorm = sqrt(gn);
      double dip[3] = {0, 0, 0};
      for (int i = 0; i < n; ++i) for (int k = 0; k < 3; ++k) dip[k] += xyz[3 * i + k] * EL[at[i]].zval;
      for (size_t i = 0; i < (size_t)N; ++i) {
        for (size_t j = 0; j < i; ++j) for (int k = 0; k < 3; ++k) dip[k] -= 2.0 * P[i * N + j] * D[k * NN + i * N + j];
        for (int k = 0; k < 3; ++k) dip[k] -= P[i * N + i] * D[k * NN + i * N + i];
      }
      for (int k = 0; k < 3; ++k) res->dipole[k] = dip[k];
      if (res->gradient) memcpy(res->gradient, grad, sizeof(double) * 3 * n);
      if (res->q) memcpy(res->q, q, sizeof(double) * n);
      if (res->qsh) memcpy(res->qsh, qsh, sizeof(double) * NS);
      if (res->dipm) memcpy(res->dipm, dipm, sizeof(double) * 3 * n);
      if (res->qp) memcpy(res->qp, qp, sizeof(double) * 6 * n);
      if (res->emo) memcpy(res->emo, emo, sizeof(double) * N);
      if (res->focc) memcpy(res->focc, focc, sizeof(double) * N);
      if (res->density) memcpy(res->density, P, sizeof(double) * NN);
      if (res->coeff) memcpy(res->coeff, C, sizeof(double) * NN);
      if (res->wbo) {
        /* closed shell: get_wiberg, sum over the AO pairs of two atoms of
         * (PS)_km (PS)_mk; open shell: get_unrestricted_wiberg
         * (src/scc_core.f90:1248-1295), 2 [(Pa S)_km (Pa S)_mk + (Pb S)_km (Pb S)_mk]
         * with the spin densities Pa = C fa C^T, Pb = C fb C^T (src/scf_module.F90:856-873) */
        const int nspin = nopen > 0 ? 2 : 1;
        double *PS = NEW(double, NN * nspin);
        if (nopen > 0) {
          const double *fs[2] = {focca, foccb};
          double *Ps = NEW(double, NN);
          for (int sp = 0; sp < 2; ++sp) {
            for (size_t i = 0; i < (size_t)N; ++i) for (int k = 0; k < N; ++k) tmp[i * N + k] = C[i * N + k] * fs[sp][k];
            gemm_nt(N, C, tmp, Ps, use_lib);
            gemm_nn(N, Ps, S, PS + (size_t)sp * NN, use_lib);
          }
          free(Ps);
        } else {
          gemm_nn(N, P, S, PS, use_lib);
        }
        memset(res->wbo, 0, sizeof(double) * n * n);
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < i; ++j) {
            double r2 = 0.0;
            for (int k = 0; k < 3; ++k) { double t = xyz[3 * i + k] - xyz[3 * j + k]; r2 += t * t; }
            double xs = 0.0;
            if (r2 < 100.0)
              for (int k = b.at_ao0[i]; k < b.at_ao0[i] + b.at_nao[i]; ++k)
                for (int m = b.at_ao0[j]; m < b.at_ao0[j] + b.at_nao[j]; ++m)
                  for (int sp = 0; sp < nspin; ++sp)
                    xs += PS[sp * NN + (size_t)k * N + m] * PS[sp * NN + (size_t)m * N + k];
            if (nopen > 0) xs *= 2.0;
            res->wbo[i * n + j] = res->wbo[j * n + i] = xs;
          }
        free(PS);
      }
    }
    (void)eatoms;
    free(Pew); free(dhdcn);
    free(B.df); free(B.u); free(B.a); free(B.omega); free(B.qlast); free(B.dqlast);
    free(q_in); free(dq); free(H); free(C); free(P); free(emo); free(shift); free(ashift);
    free(vs); free(vd); free(vq); free(tmp); free(jq);
  }
cleanup_early:
  free(U); free(radcn); free(gab3); free(gab5); free(dmask); free(qmask);
  free(S); free(D); free(Q); free(H0); free(H0n); free(selfe); free(dsedcn);
  free(jmat); free(shgam); d4_free(&d4); free(cn); free(dcndr); free(grad);
  solv_free(&sv); free(svwork);
  free(dipm); free(qp); free(q); free(qsh); free(focca); free(foccb); free(focc); free(zsh);
  basis_free(&b);
  return ret;
}
