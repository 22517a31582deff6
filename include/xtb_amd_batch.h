/*
 * Batched / served GFN2-xTB single points on one MI355X (gfx950).
 *
 * This interface has no counterpart in xtb's C API (include/xtb.h): it is the
 * extension through which many single points run as one batch.  Every molecule
 * of a batch is computed exactly as xtb_singlepoint would compute it
 * (src/api/interface.f90:40-156 -> scf, src/scf_module.F90:74-920) with a
 * fresh results object, i.e. from the EEQ guess.
 *
 * Units: bohr, Hartree, Hartree/bohr; orbital energies in eV.
 * Errors: functions return 0 (or a non-negative count) on success and a
 * negative value on failure; the message is available from
 * xtbamd_batch_error().
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

typedef struct xtbamd_batch_s* xtbamd_batch;

typedef struct {
  int natoms, nshells, nao, iterations, converged, failed, ihomo;
  double energy, gnorm, hl_gap_ev;
  double e_elec, e_es, e_aes, e_axc, e_disp, e_rep, e_atm;
  double e_solv;      /* solvation free energy (0 in the gas phase) */
  double dipole[3];
  /* caller-owned output arrays, may be NULL */
  double* gradient;   /* [natoms][3] */
  double* charges;    /* [natoms] */
  double* qsh;        /* [nshells] */
  double* dipm;       /* [natoms][3] */
  double* qp;         /* [natoms][6] xx,xy,yy,xz,yz,zz */
  double* wbo;        /* [natoms][natoms] */
  double* emo;        /* [nao] eV */
  double* focc;       /* [nao] */
  double* coeff;      /* [nao][nao] row = AO, column = MO */
  double* density;    /* [nao][nao] */
} xtbamd_result;

/* numbers/positions/charges/uhf are concatenated over the nmol molecules;
 * natoms[m] gives the size of molecule m.  guess: 0 = EEQ, 1 = SAD. */
xtbamd_batch xtbamd_batch_new(int nmol, const int* natoms, const int* numbers, const double* positions,
                              const double* charges, const int* uhf, double accuracy, double etemp,
                              int maxiter, int guess);
/* last error message of the calling thread (batch creation) or of the batch */
const char* xtbamd_last_error(void);
/* restart guess of molecule m (call before setup/run): shell charges [nshells],
 * atomic dipoles [natoms][3], quadrupoles [natoms][6] of a previous result;
 * mirrors the wavefunction reuse of xtb_singlepoint (src/api/interface.f90:83-101) */
int xtbamd_batch_set_guess(xtbamd_batch b, int m, const double* qsh, const double* dipm, const double* qp);
/* new positions [natoms][3] (bohr) of molecule m, same atoms: the device
 * copy of xtb_updateMolecule (src/api/molecule.f90:114-159) for a batch that
 * is kept across single points; the next setup/run recomputes everything
 * that depends on them */
int xtbamd_batch_set_positions(xtbamd_batch b, int m, const double* positions);
/* on != 0: every run after the first starts each molecule from the shell
 * charges, atomic dipoles and quadrupoles the previous run left on the device
 * (no EEQ guess), the resident form of the wavefunction reuse of a results
 * object (src/api/interface.f90:83-101) */
int xtbamd_batch_keep_wavefunction(xtbamd_batch b, int on);
/* implicit solvent for every molecule of the batch (call before setup/run;
 * NULL or "" returns to the gas phase).  kernel 1: Still kernel with the
 * GFN2/GBSA parameter sets; 2: P16 kernel with the GFN2/ALPB sets; alpb adds
 * the ALPB shape correction; state 1 gsolv, 2 reference, 3 1 bar / 1 M;
 * temperature in K.  xtb_setSolvent's "gbsa" is (1, 0), "alpb" is (2, 1)
 * (src/api/calculator.f90:341-440, src/solv/model.f90).  Returns -1 for a
 * solvent without parameters. */
int xtbamd_batch_set_solvent(xtbamd_batch b, const char* solvent, int kernel, int alpb, int state,
                             double temperature);
/* 1 when the solvent has a parameter set for the kernel, else 0 */
int xtbamd_solvent_known(const char* solvent, int kernel);
int xtbamd_batch_setup(xtbamd_batch b);
/* one SCC iteration of every active molecule; returns the number still active */
int xtbamd_batch_iterate(xtbamd_batch b);
int xtbamd_batch_gradient(xtbamd_batch b);
/* setup + SCC to convergence + gradient */
int xtbamd_batch_run(xtbamd_batch b);
int xtbamd_batch_size(xtbamd_batch b);
/* work counters of the last run: SCC steps launched, block-Jacobi rounds */
int xtbamd_batch_counters(xtbamd_batch b, int* scc_steps, long long* jacobi_rounds);
int xtbamd_batch_get(xtbamd_batch b, int m, xtbamd_result* out);
/* current atomic charges [natoms] of molecule m: after xtbamd_batch_setup the
 * initial guess (EEQ, src/eeq_model.f90:678-1015 via newWavefunction,
 * src/xtb/calculator.f90:462-480; SAD; or the restart guess), after an
 * iteration the Mulliken charges of that iteration */
int xtbamd_batch_charges(xtbamd_batch b, int m, double* q);
void xtbamd_batch_free(xtbamd_batch b);
/* HIP-event timing of kernel families on the streams they run on, from this
 * call on.  kernels: bit 0 = the batched FP64 GEMM (k_gemm), bit 1 = the
 * block-Jacobi pivot kernel (k_jac_pairsR), bit 2 = the block-Jacobi tile
 * updates (k_jac_apply_R); 0 = off.  profile_result(kernel = 1 GEMM, 2
 * pivots, 3 tile updates) gives the number of launches, their summed
 * duration and their summed algorithmic work (GEMM: 2 N^3 flops per molecule
 * processed; pivots: 9216 flops per inner round of a 32x32 pivot problem;
 * tile updates: compulsory HBM bytes, each touched element of A and W read
 * and written once).  The Jacobi families are timed on a sample of the
 * rounds and scaled to all of them. */
int xtbamd_batch_profile(xtbamd_batch b, int kernels);
int xtbamd_batch_profile_result(xtbamd_batch b, int kernel, int* launches, double* total_ms, double* work);
/* kernels 2 (pivots) and 3 (tile updates): the summed execution spans of the sampled
 * launches (device wall clock from the start of the launch's first wave to the
 * end of its last, what a kernel trace records as its duration), scaled to all
 * launches; the HIP-event times of profile_result also hold the dispatch gaps
 * in the stream */
int xtbamd_batch_profile_span(xtbamd_batch b, int kernel, double* span_ms);
/* ---- continuous batching (serving) ----
 * Requests are queued by xtbamd_server_submit and admitted into the running
 * batch whenever its packed device arrays (sized from hbm_gb for molecules of
 * up to max_atoms atoms, in up to max_slots slots) have room; each
 * xtbamd_server_step admits, runs one SCC iteration of every resident
 * molecule and finishes (gradient, results to the host) those that stopped
 * iterating.  A request's result is bitwise the same alone and in any mix. */
typedef struct xtbamd_server_s* xtbamd_server;
xtbamd_server xtbamd_server_new(double hbm_gb, int max_atoms, int max_slots, double accuracy, double etemp,
                                int maxiter);
/* returns the request id (>= 0), -1 on invalid input */
long long xtbamd_server_submit(xtbamd_server s, int natoms, const int* numbers, const double* positions,
                               double charge, int uhf);
/* returns the number of requests not yet finished (queued + running) */
int xtbamd_server_step(xtbamd_server s);
/* ids of requests finished since the last call (at most max_ids; the rest on the next call) */
int xtbamd_server_finished(xtbamd_server s, long long* ids, int max_ids);
/* copy a finished result; with all array pointers NULL only sizes and scalars
 * are filled and the result is kept, otherwise it is released */
int xtbamd_server_result(xtbamd_server s, long long id, xtbamd_result* out);
int xtbamd_server_pending(xtbamd_server s, int* queued, int* running);
/* implicit solvent of every request admitted from now on (arguments as
 * xtbamd_batch_set_solvent); only while no request is running */
int xtbamd_server_set_solvent(xtbamd_server s, const char* solvent, int kernel, int alpb, int state,
                              double temperature);
void xtbamd_server_free(xtbamd_server s);

/* ---- single point that allocates nothing ----
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Computes what xtbamd_batch_new (one molecule, EEQ
 * guess) + xtbamd_batch_run + xtbamd_batch_get compute, with every device
 * array carved from a 64 MiB scratch area that the library holds on each GPU
 * it runs on until the process ends (molecules of up to about 450 AOs; a
 * larger one returns -1 with a message naming the limit).  No hipMalloc,
 * hipFree or allocating library call is made; the first call on a device
 * sets up the scratch address and four events; the work runs on the null
 * stream.  Calls on one device share its scratch area: calls from threads of
 * one process are serialised, and the area keeps the last call's contents.  Inputs and
 * outputs are host arrays as for the batch interface (out's array pointers
 * may be NULL). */
int xtbamd_singlepoint_noalloc(int natoms, const int* numbers, const double* positions, double charge, int uhf,
                               double accuracy, double etemp, int maxiter, xtbamd_result* out);

/* Symmetric eigen-decomposition of nprob matrices by the blocked FP64 solver
 * that the engine uses for its largest eigenproblems (Householder
 * tridiagonalisation, divide and conquer, back-transformation: the steps of
 * dsyevd, called by the reference at src/mctc/lapack/eigensolve.F90:218-246
 * through src/scc_core.f90:475).  a: the matrices, row-major, dimension n[k]
 * (>= 2), concatenated; evals: eigenvalues ascending, concatenated; evecs:
 * eigenvectors as rows (row j of problem k = eigenvector j), concatenated;
 * tri (may be NULL): per problem the diagonal (n[k]) and the off-diagonal
 * (n[k], last entry 0) of the tridiagonal form.  Allocates and frees its own
 * device memory. */
int xtbamd_eigh(int nprob, const int* n, const double* a, double* evals, double* evecs, double* tri);

/* copies `bytes` bytes device to device with one plain copy kernel (k_copy):
 * the reference point for hardware byte counters */
int xtbamd_copy_probe(long long bytes);
/* device / runtime information: number of visible GPUs */
int xtbamd_device_count(void);

#ifdef __cplusplus
}
#endif
