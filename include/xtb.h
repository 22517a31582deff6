/*
 * xtb C API (version 2.0.0) as exported by libxtb_amd.so.
 *
 * Drop-in for the reference's include/xtb.h: same names, signatures, handle
 * semantics and error convention (errors are logged in the environment and
 * polled with xtb_checkEnvironment / xtb_getError; NULL environments make
 * every call a no-op).  The implementation is the Fortran layer in
 * xtb_amd/fortran/xtb_capi.f90 on top of the MI355X engine.
 *
 * Supported: GFN2-xTB single points (energy + gradient) of molecules, in the
 * gas phase or in implicit solvent (xtb_setSolvent models "gbsa" and "alpb");
 * xtb_hessian runs its 6n displaced single points as one GPU batch.
 * GFN0/GFN1/GFN-FF, COSMO/CPCM-X and external charges set an environment
 * error "... is not supported by this GFN2-xTB implementation".
 * Units: bohr, Hartree, Hartree/bohr, e; orbital energies in Hartree.
 */
#pragma once

/* API version 10000*major + 100*minor + patch (reference include/xtb.h:25) */
#define XTB_API_VERSION 20000

/* print levels (reference include/xtb.h:30-33) */
#define XTB_VERBOSITY_FULL 2
#define XTB_VERBOSITY_MINIMAL 1
#define XTB_VERBOSITY_MUTED 0

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h>
#endif

/* opaque handles (reference include/xtb.h:55-64) */
typedef struct _xtb_TEnvironment* xtb_TEnvironment;
typedef struct _xtb_TMolecule* xtb_TMolecule;
typedef struct _xtb_TCalculator* xtb_TCalculator;
typedef struct _xtb_TResults* xtb_TResults;

/* reference include/xtb.h:72 */
int xtb_getAPIVersion(void);

/* environment: reference include/xtb.h:79-114 */
xtb_TEnvironment xtb_newEnvironment(void);
void xtb_delEnvironment(xtb_TEnvironment* env);
int xtb_checkEnvironment(xtb_TEnvironment env);
void xtb_showEnvironment(xtb_TEnvironment env, const char* message);
void xtb_getError(xtb_TEnvironment env, char* buffer, const int* buffersize);
void xtb_setOutput(xtb_TEnvironment env, const char* filename);
void xtb_releaseOutput(xtb_TEnvironment env);
void xtb_setVerbosity(xtb_TEnvironment env, int verbosity);

/* molecule (bohr): reference include/xtb.h:120-140 */
xtb_TMolecule xtb_newMolecule(xtb_TEnvironment env, const int* natoms, const int* numbers,
                              const double* positions, const double* charge, const int* uhf,
                              const double* lattice, const bool* periodic);
void xtb_delMolecule(xtb_TMolecule* mol);
void xtb_updateMolecule(xtb_TEnvironment env, xtb_TMolecule mol, const double* positions,
                        const double* lattice);

/* calculator: reference include/xtb.h:146-228 */
xtb_TCalculator xtb_newCalculator(void);
void xtb_delCalculator(xtb_TCalculator* calc);
void xtb_loadGFN0xTB(xtb_TEnvironment env, xtb_TMolecule mol, xtb_TCalculator calc, char* filename);
void xtb_loadGFN1xTB(xtb_TEnvironment env, xtb_TMolecule mol, xtb_TCalculator calc, char* filename);
void xtb_loadGFN2xTB(xtb_TEnvironment env, xtb_TMolecule mol, xtb_TCalculator calc, char* filename);
void xtb_loadGFNFF(xtb_TEnvironment env, xtb_TMolecule mol, xtb_TCalculator calc, char* filename);
void xtb_setSolvent(xtb_TEnvironment env, xtb_TCalculator calc, char* solvent, int* state, double* temp,
                    int* grid, char* model);
void xtb_releaseSolvent(xtb_TEnvironment env, xtb_TCalculator calc);
void xtb_setExternalCharges(xtb_TEnvironment env, xtb_TCalculator calc, int* n, int* numbers, double* charges,
                            double* positions);
void xtb_releaseExternalCharges(xtb_TEnvironment env, xtb_TCalculator calc);
void xtb_setAccuracy(xtb_TEnvironment env, xtb_TCalculator calc, double accuracy);
void xtb_setMaxIter(xtb_TEnvironment env, xtb_TCalculator calc, int iterations);
void xtb_setElectronicTemp(xtb_TEnvironment env, xtb_TCalculator calc, double temperature);

/* calculations: reference include/xtb.h:230-253 */
void xtb_cpcmx_calc(xtb_TEnvironment env, xtb_TMolecule mol, xtb_TCalculator calc, xtb_TResults res);
void xtb_singlepoint(xtb_TEnvironment env, xtb_TMolecule mol, xtb_TCalculator calc, xtb_TResults res);
/* hessian [3n][3n] (symmetric), step_size in bohr (NULL: 0.005), atom_index_list
 * 1-based (NULL: all atoms; else up to natoms entries, ended by an entry <= 0),
 * dipole_gradient [3n][3], polarizability_gradient [3n][6] (zero for GFN2-xTB).
 * The argument order after `hessian` is the reference implementation's
 * (src/api/interface.f90:310-312: step before the list); its header lists them
 * the other way round as int pointers (include/xtb.h:244-252). */
void xtb_hessian(xtb_TEnvironment env, xtb_TMolecule mol, xtb_TCalculator calc, xtb_TResults res, double* hessian,
                 const double* step_size, const int* atom_index_list, double* dipole_gradient,
                 double* polarizability_gradient);

/* results: reference include/xtb.h:259-342 */
xtb_TResults xtb_newResults(void);
void xtb_delResults(xtb_TResults* res);
xtb_TResults xtb_copyResults(xtb_TResults res);
void xtb_getEnergy(xtb_TEnvironment env, xtb_TResults res, double* energy);
void xtb_getSolvationEnergy(xtb_TEnvironment env, xtb_TResults res, double* energy);
void xtb_getGradient(xtb_TEnvironment env, xtb_TResults res, double* gradient);
void xtb_getPCGradient(xtb_TEnvironment env, xtb_TResults res, double* gradient);
void xtb_getVirial(xtb_TEnvironment env, xtb_TResults res, double* virial);
void xtb_getDipole(xtb_TEnvironment env, xtb_TResults res, double* dipole);
void xtb_getCharges(xtb_TEnvironment env, xtb_TResults res, double* charges);
void xtb_getBondOrders(xtb_TEnvironment env, xtb_TResults res, double* wbo);
void xtb_getNao(xtb_TEnvironment env, xtb_TResults res, int* nao);
void xtb_getOrbitalEigenvalues(xtb_TEnvironment env, xtb_TResults res, double* emo);
void xtb_getOrbitalOccupations(xtb_TEnvironment env, xtb_TResults res, double* focc);
void xtb_getOrbitalCoefficients(xtb_TEnvironment env, xtb_TResults res, double* c);

#ifdef __cplusplus
}
#endif

/* delete a handle of any type and set it to NULL (reference include/xtb.h:36-41) */
#ifndef __cplusplus
#define xtb_delete(ptr)                                                                                     \
  _Generic((ptr), xtb_TEnvironment: xtb_delEnvironment, xtb_TMolecule: xtb_delMolecule,                     \
           xtb_TCalculator: xtb_delCalculator, xtb_TResults: xtb_delResults)(&ptr)
#endif
