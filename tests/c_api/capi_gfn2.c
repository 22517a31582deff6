/* Conformance test of the xtb.h C API on the GFN2-xTB path.
 *
 * Expected values are the reference's own (test/api/c_api_example.c:105-112
 * propyne; test/unit/test_gfn2.f90:140-164 water, with the EEQ guess of the
 * C API).  Also checks the error convention, warm restarts from a reused
 * results object, getters before a single point and unsupported calls.
 * Exit status 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xtb.h"

static int failures = 0;
#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      fprintf(stderr, "FAIL: " __VA_ARGS__); \
      fprintf(stderr, "\n");              \
      failures++;                         \
    }                                     \
  } while (0)

int main(void) {
  const int natoms = 7;
  const int attyp[7] = {6, 6, 6, 1, 1, 1, 1};
  const double coord[3 * 7] = {
      +0.00000000000000, +0.00000000000000, -1.79755622305860, +0.00000000000000, +0.00000000000000,
      +0.95338756106749, +0.00000000000000, +0.00000000000000, +3.22281255790261, -0.96412815539807,
      -1.66991895015711, -2.53624948351102, -0.96412815539807, +1.66991895015711, -2.53624948351102,
      +1.92825631079613, +0.00000000000000, -2.53624948351102, +0.00000000000000, +0.00000000000000,
      +5.23010455462158};
  double energy = 0.0, dipole[3], q[7], wbo[49], grad[21], virial[9];
  char buffer[512];
  int buffersize = 512, nao = -1;

  CHECK(xtb_getAPIVersion() == XTB_API_VERSION, "API version");
  xtb_TEnvironment env = xtb_newEnvironment();
  xtb_TCalculator calc = xtb_newCalculator();
  xtb_TResults res = xtb_newResults();
  xtb_TMolecule mol = xtb_newMolecule(env, &natoms, attyp, coord, NULL, NULL, NULL, NULL);
  CHECK(mol != NULL && !xtb_checkEnvironment(env), "newMolecule");

  /* getter before any single point: error in the log, cleared by getError */
  xtb_getEnergy(env, res, &energy);
  CHECK(xtb_checkEnvironment(env) == 1, "getEnergy before singlepoint must fail");
  xtb_getError(env, buffer, &buffersize);
  CHECK(strstr(buffer, "Energy is not available") != NULL, "error text: %s", buffer);
  CHECK(xtb_checkEnvironment(env) == 0, "getError empties the log");
  xtb_getNao(env, res, &nao);
  CHECK(nao == 0, "nao before singlepoint");

  /* single point without a loaded calculator */
  xtb_singlepoint(env, mol, calc, res);
  CHECK(xtb_checkEnvironment(env) == 1, "singlepoint without calculator must fail");
  xtb_getError(env, buffer, &buffersize);

  xtb_setVerbosity(env, XTB_VERBOSITY_MUTED);
  xtb_loadGFN2xTB(env, mol, calc, NULL);
  xtb_setAccuracy(env, calc, 1.0);
  xtb_setElectronicTemp(env, calc, 300.0);
  xtb_setMaxIter(env, calc, 30);
  CHECK(!xtb_checkEnvironment(env), "load/setters");

  xtb_singlepoint(env, mol, calc, res);
  CHECK(!xtb_checkEnvironment(env), "singlepoint");
  xtb_getEnergy(env, res, &energy);
  xtb_getCharges(env, res, q);
  xtb_getDipole(env, res, dipole);
  xtb_getBondOrders(env, res, wbo);
  xtb_getGradient(env, res, grad);
  xtb_getVirial(env, res, virial);
  xtb_getNao(env, res, &nao);
  CHECK(!xtb_checkEnvironment(env), "getters");
  CHECK(fabs(energy - -8.3824793849585) < 1.0e-9, "energy %.12f", energy);
  CHECK(fabs(q[5] - 0.05184019996829) < 1.0e-8, "charge %.12f", q[5]);
  CHECK(fabs(dipole[2] - -0.298279305689518) < 1.0e-6, "dipole %.12f", dipole[2]);
  CHECK(fabs(wbo[9] - 2.89823984265213) < 1.0e-8, "wbo %.12f", wbo[9]);
  CHECK(nao == 16, "nao %d", nao);
  for (int i = 0; i < 9; ++i) CHECK(virial[i] == 0.0, "virial");

  /* orbital data: eigenvalues ascending (Hartree), occupations sum to the electron count */
  double emo[16], focc[16], cmo[256];
  xtb_getOrbitalEigenvalues(env, res, emo);
  xtb_getOrbitalOccupations(env, res, focc);
  xtb_getOrbitalCoefficients(env, res, cmo);
  CHECK(!xtb_checkEnvironment(env), "orbital getters");
  double nel = 0.0;
  for (int i = 0; i < 16; ++i) nel += focc[i];
  CHECK(fabs(nel - 16.0) < 1e-10, "electron count %.12f", nel);
  for (int i = 1; i < 16; ++i) CHECK(emo[i] >= emo[i - 1], "emo ascending");

  /* warm restart: the same results object reproduces the converged energy */
  double e2 = 0.0;
  xtb_singlepoint(env, mol, calc, res);
  xtb_getEnergy(env, res, &e2);
  CHECK(!xtb_checkEnvironment(env), "restart singlepoint");
  CHECK(fabs(e2 - energy) < 1.0e-8, "restart energy %.12f vs %.12f", e2, energy);

  /* copy of the results object */
  xtb_TResults res2 = xtb_copyResults(res);
  double e3 = 0.0;
  xtb_getEnergy(env, res2, &e3);
  CHECK(e3 == e2, "copyResults");
  xtb_delete(res2);
  CHECK(res2 == NULL, "xtb_delete sets NULL");

  /* numerical Hessian from a fresh results object (reference c_api_example.c:242-261) */
  xtb_delete(res);
  res = xtb_newResults();
  double* hess = (double*)calloc(9 * 49, sizeof(double));
  xtb_hessian(env, mol, calc, res, hess, NULL, NULL, NULL, NULL);
  CHECK(!xtb_checkEnvironment(env), "hessian");
  CHECK(fabs(hess[0] - 0.4790088649) < 1.0e-9, "Hessian[0,0] %.12f", hess[0]);
  CHECK(fabs(hess[3] - -0.0528761190) < 1.0e-9, "Hessian[0,3] %.12f", hess[3]);
  CHECK(fabs(hess[3] - hess[63]) < 1.0e-9, "Hessian symmetry");
  CHECK(fabs(hess[9 * 49 - 1] - 0.3636571159) < 1.0e-9, "Hessian[21,21] %.12f", hess[9 * 49 - 1]);
  free(hess);

  /* implicit solvent, the reference example's sequence on one results object:
   * gas phase, then GBSA, then ALPB, each restarted from the previous
   * wavefunction (reference test/api/c_api_example.c:114-158) */
  char solvent[] = "h2o", gbsa[] = "gbsa", alpb[] = "alpb", cosmo[] = "cosmo", nosuch[] = "nosuchsolvent";
  xtb_TResults rs = xtb_newResults();
  xtb_singlepoint(env, mol, calc, rs);
  xtb_setSolvent(env, calc, solvent, NULL, NULL, NULL, gbsa);
  CHECK(!xtb_checkEnvironment(env), "setSolvent gbsa");
  xtb_singlepoint(env, mol, calc, rs);
  xtb_getEnergy(env, rs, &energy);
  xtb_getCharges(env, rs, q);
  xtb_getDipole(env, rs, dipole);
  xtb_getBondOrders(env, rs, wbo);
  CHECK(!xtb_checkEnvironment(env), "GBSA singlepoint");
  CHECK(fabs(energy - -8.38393864716134) < 1.0e-9, "GBSA energy %.12f", energy);
  CHECK(fabs(q[5] - 0.06090868805034) < 1.0e-8, "GBSA charge %.12f", q[5]);
  CHECK(fabs(dipole[2] - -0.35455233974705) < 1.0e-6, "GBSA dipole %.12f", dipole[2]);
  CHECK(fabs(wbo[9] - 2.89453979224265) < 1.0e-8, "GBSA bond order %.12f", wbo[9]);
  xtb_releaseSolvent(env, calc);
  xtb_setSolvent(env, calc, solvent, NULL, NULL, NULL, alpb);
  CHECK(!xtb_checkEnvironment(env), "setSolvent alpb");
  xtb_singlepoint(env, mol, calc, rs);
  xtb_getEnergy(env, rs, &energy);
  xtb_getCharges(env, rs, q);
  xtb_getDipole(env, rs, dipole);
  xtb_getBondOrders(env, rs, wbo);
  CHECK(!xtb_checkEnvironment(env), "ALPB singlepoint");
  CHECK(fabs(energy - -8.384076843892) < 1.0e-9, "ALPB energy %.12f", energy);
  CHECK(fabs(q[5] - 0.0644849340) < 1.0e-8, "ALPB charge %.12f", q[5]);
  CHECK(fabs(dipole[2] - -0.3641866008) < 1.0e-6, "ALPB dipole %.12f", dipole[2]);
  CHECK(fabs(wbo[9] - 2.8932146955) < 1.0e-8, "ALPB bond order %.12f", wbo[9]);
  /* as in the reference, only CPCM-X results carry a separate solvation energy */
  xtb_getSolvationEnergy(env, rs, &e3);
  CHECK(xtb_checkEnvironment(env) == 1, "getSolvationEnergy outside CPCM-X must fail");
  xtb_getError(env, buffer, &buffersize);
  xtb_releaseSolvent(env, calc);
  xtb_delete(rs);

  /* unknown solvents and unsupported models leave an error */
  xtb_setSolvent(env, calc, nosuch, NULL, NULL, NULL, gbsa);
  CHECK(xtb_checkEnvironment(env) == 1, "unknown solvent must fail");
  xtb_getError(env, buffer, &buffersize);
  CHECK(strstr(buffer, "Could not add solvation model") != NULL, "unknown solvent text: %s", buffer);
  xtb_setSolvent(env, calc, solvent, NULL, NULL, NULL, cosmo);
  CHECK(xtb_checkEnvironment(env) == 1, "setSolvent cosmo must fail");
  xtb_getError(env, buffer, &buffersize);
  CHECK(strstr(buffer, "not supported") != NULL, "unsupported text: %s", buffer);

  /* overlapping atoms are rejected */
  double bad[6] = {0, 0, 0, 0, 0, 0};
  const int two = 2, hh[2] = {1, 1};
  xtb_TMolecule badmol = xtb_newMolecule(env, &two, hh, bad, NULL, NULL, NULL, NULL);
  CHECK(badmol == NULL && xtb_checkEnvironment(env) == 1, "overlapping atoms");
  xtb_getError(env, buffer, &buffersize);

  /* NULL environment: silent no-op */
  xtb_getEnergy(NULL, res, &energy);

  xtb_delete(res);
  xtb_delete(mol);
  xtb_delete(calc);
  xtb_delete(env);
  if (failures) {
    fprintf(stderr, "%d failure(s)\n", failures);
    return 1;
  }
  printf("capi_gfn2: all checks passed (E = %.12f)\n", energy);
  return 0;
}
