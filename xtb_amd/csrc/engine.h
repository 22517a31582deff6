// Batched GFN2-xTB SCC engine for MI355X: device data layout and launchers.
//
// A Batch holds M independent molecules.  Every per-molecule quantity is
// stored packed: atom-, shell-, AO- and D4-reference-indexed vectors at the
// offsets of MolDesc, dense matrices at per-molecule offsets.  All kernels of
// one SCC step take the whole batch; a molecule's arithmetic never depends on
// the other molecules of the batch (fixed per-molecule reduction trees, no
// floating-point atomics), so a result is bitwise the same alone and in a mix.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "gfn2_common.h"

namespace xtbamd {

struct MolDesc {
  int n, nsh, nao, nd4, ne;     // atoms, shells, AOs, D4 references, EEQ dim (n+1)
  int a0, s0, o0, d0;           // offsets of atom/shell/AO/D4 vectors
  long long m0;                 // offset of the nao x nao matrices
  long long j0;                 // nsh x nsh
  long long x0;                 // nd4 x nd4
  long long p0;                 // n x n
  long long e0;                 // ne x ne
  long long v0;                 // Broyden vectors (nbr)
  long long h0;                 // Broyden history (maxiter x nbr)
  long long g0;                 // Broyden small matrices (maxiter x maxiter)
  int nbr, nel, nopen, maxiter;
  double charge;
  int mol;                      // index of the molecule (slot) in the batch
  int npsp;                     // leading pairs of the pair list with s and p shells only
  long long y0;                 // offset of the EEQ vectors (n+1)
};

// Per-molecule SCC state machine (src/scc_core.f90:434-640).
struct MolState {
  int iter;        // iterations done
  int converged;   // criteria met at least once
  int lastdiag;
  int active;      // still iterating
  int failed;      // error code (0 = fine)
  int ihomoa, ihomob, ihomo;
  double eel, eold, egap, ees, eaes, epol, ed, ga, gb, rmsq, efa, efb;
  double eh0;      // Tr(P H0) in eV
  double damp;     // simple mixing factor (single atoms)
  double ep, embd, etotal, gnorm;
  double dipole[3];
};

struct Options {
  double accuracy = 1.0;
  double etemp = 300.0;
  int maxiter = 250;
  int guess = 0;   // 0 EEQ, 1 SAD
};

struct MoleculeInput {
  std::vector<int> numbers;
  std::vector<double> xyz;   // bohr, [n][3]
  double charge = 0.0;
  int uhf = 0;
};

// Host-side index maps of one molecule (local indices).
struct HostMol {
  std::vector<int> at, sh_at, sh_l, sh_ao0, sh_m, at_sh0, at_ao0, ao_at, ao_sh, d4_at, d4_ref;
  std::vector<double> xyz;
  // atom pairs (i, j), i > j, grouped by element pair (so that a wave of the
  // pair kernels runs one code path), n (n - 1) / 2 entries
  std::vector<int2> pairs;
};

// First-fit allocator of ranges in [0, cap) (one per packed device array).
struct Arena {
  long long cap = 0;
  std::vector<std::pair<long long, long long>> free_;   // sorted [begin, end)
  void init(long long c);
  long long take(long long n);                          // -1 when no range fits
  void give(long long off, long long n);
};

// Capacities of the packed arrays (elements), and the number of slots.
struct Capacity {
  int slots = 0;
  long long n = 0, nsh = 0, nao = 0, nd4 = 0, mm = 0, jj = 0, xx = 0, pp = 0, ee = 0, ne = 0, v = 0, h = 0, g = 0;
};

// Per-molecule results copied back to the host.
struct MolResult {
  int n = 0, nsh = 0, nao = 0, iterations = 0, converged = 0, failed = 0, ihomo = 0;
  double energy = 0, gnorm = 0, hl_gap_ev = 0, e_elec = 0, e_es = 0, e_aes = 0, e_axc = 0;
  double e_disp = 0, e_rep = 0, e_atm = 0, e_solv = 0;
  double dipole[3] = {0, 0, 0};
  std::vector<double> gradient, q, qsh, dipm, qp, wbo, emo, focc, coeff, density;
};

// Jacobi eigenproblem set: each problem is a dim x dim symmetric matrix at
// a_off in A, eigenvectors accumulate at w_off in W (column k = vector k).
struct JacobiProb {
  int dim, nb, nbp, rounds;   // 32-blocks, padded even count, rounds per sweep
  long long a_off, w_off, q_off;
  int ev_off;                 // eigenvalue offset
  int group;                  // lane (stream) of the two-lane round schedule
  int ld;                     // row stride of the problem's A and W (dim, or dim rounded up to 16)
  long long src_off;          // the matrix in the caller's packed layout (row stride dim)
  int maxsw;                  // sweeps before the problem is given up (failed)
  int pad_;
};

// stream lanes of the block-Jacobi rounds (at most)
constexpr int kMaxLanes = 3;
// Counters and clocks of sampled launches are spread over kProfLines 128-byte
// lines, workgroup b using line b mod kProfLines: same-address device atomics
// from thousands of workgroups serialise (about 20 ns each), which stretched a
// sampled launch to ten times its length.  The host reduces over the lines.
constexpr int kProfLines = 64, kProfLineU64 = 16;

// A set of Jacobi problems (one per molecule) with its device work lists.
struct JacSet {
  std::vector<JacobiProb> probs;
  JacobiProb* d_probs = nullptr;
  int nsmall = 0, npairs = 0, nta = 0, ntw = 0, max_rounds = 0;
  int max_rounds_lanes = 0;   // the same without the blocked eigensolver's problems
  int* d_small = nullptr;
  int2* d_pairs = nullptr;
  int4 *d_ta = nullptr, *d_tw = nullptr, *d_to = nullptr;
  int nto = 0;
  // the work lists hold lane 0's items first: lane g owns pairs[gpairs[g],
  // gpairs[g+1]) and likewise for the tile lists
  int gpairs[kMaxLanes + 1] = {}, gta[kMaxLanes + 1] = {}, gtw[kMaxLanes + 1] = {}, gto[kMaxLanes + 1] = {};
  // XCD-ordered lists of the lanes: pivots (m, k) and tiles (m, y, z, kind:
  // 0 A tile (k, l), 1 W tile (rb, l), 2 skip); lane g owns [gxp[g], gxp[g+1])
  int2* d_xp = nullptr;
  int4* d_xt = nullptr;
  int gxp[kMaxLanes + 1] = {}, gxt[kMaxLanes + 1] = {};
  int* d_big = nullptr;   // problems above 64 (tridiagonal QL path)
  // padded working copies of A and W (row stride ld, every row 128-byte aligned)
  double *ja = nullptr, *jw = nullptr;
  long long jcap = 0;
  int nbig = 0, big_max_dim = 0, big_max_ld = 0;   // big_max_ld: largest ld of any problem
  unsigned long long* offmax = nullptr;
  double* qbuf = nullptr;
  int* rot = nullptr;
  int* state = nullptr;
  double* scale = nullptr;
  int* perm = nullptr;
  long long qsize = 0;
};

struct SolvDev;   // implicit-solvent model on the device (k_solv.hip)
struct EigProb;   // blocked eigensolver (k_eig.hip)
struct DcNode;
struct GemmJob;

// The problems of a batch solved by the blocked eigensolver (dimension at
// least Batch::eig_blk_min): work arrays, the divide and conquer trees and
// the product job lists, built by Batch::rebuild().
struct EigSet {
  int np = 0, maxn = 0, npanel = 0, maxdepth = 0;
  std::vector<int> mols;                 // batch slot of each problem
  std::vector<int> zfin;                 // buffer (0/1) of each problem's final eigenvectors
  std::vector<long long> v_off;          // each problem's vector block in vec
  std::vector<int> ld;                   // each problem's row stride
  EigProb* d_probs = nullptr;
  int* act = nullptr;                    // per problem: solved in this call
  int* zf = nullptr;                     // zfin on the device
  double *Aw = nullptr, *Vh = nullptr, *Z0 = nullptr, *Z1 = nullptr, *U = nullptr;
  double *X = nullptr, *P = nullptr, *P2 = nullptr, *Tf = nullptr, *vec = nullptr;
  int* ivec = nullptr;
  double* part2 = nullptr;    // partial sums of the tridiagonalisation's row kernel
  double* G = nullptr;        // Gram matrices of the back-transformation's reflector blocks
  int nblk = 0;               // back-transformation blocks of the largest problem
  long long cap[20] = {};     // capacities (elements) of the arrays above and below
  long long nmat = 0, nx = 0;   // elements in use of the n x ld arrays and of X
  // product launches: jobs [job0, job0 + njob), tiles [tile0, tile0 + ntile)
  struct Launch {
    int job0 = 0, njob = 0;
    long long tile0 = 0, ntile = 0;
  };
  std::vector<Launch> upd, bt1, bt2, bt3, dcg;   // per panel (upd), per block (bt*), per tree level (dcg)
  Launch gram;                                   // the Gram matrices of every block
  GemmJob* d_jobs = nullptr;
  int2* d_tiles = nullptr;
  // divide and conquer: leaves (prob, start, size, -), all merges (for the
  // splitting), merges by level (height 1..maxdepth)
  int4* d_leaves = nullptr;
  int nleaves = 0;
  DcNode* d_nodes = nullptr;
  int nnodes = 0;
  std::vector<int> lvl0, lvln, lvlmax;   // per level: first node, node count, largest node
  // the whole solve captured once as a HIP graph and replayed (use_graph):
  // its launches and their arguments are fixed from the build on, so a
  // replay is the eager sequence without the host's per-launch work; the
  // graph is re-captured when the arguments of a call change
  bool use_graph = false;
  hipGraphExec_t graph = nullptr;
  const void* graph_key[5] = {};
};

// Device memory, streams, events and a host word buffer lent to a Batch: a
// batch built on a loan calls no allocator (its device arrays are carved from
// [base, base + bytes) in 256-byte aligned pieces) and releases nothing.
struct Loan {
  char* base = nullptr;
  size_t bytes = 0;
  hipStream_t streams[4] = {};   // main stream, then the second, small-problem and third-lane streams
  hipEvent_t events[4] = {};     // fork, join, join3, join4 (timing disabled)
  int* host_words = nullptr;     // 64 ints of host memory for the active-count reads
};

// device memory for an EigSet's arrays
struct EigMem {
  virtual void* get(size_t bytes) = 0;
  virtual void put(void* p) = 0;
  virtual ~EigMem() = default;
};
// work arrays, trees and product job lists of es's problems (es.mols set,
// dims their dimensions)
void eig_set_build(EigSet& es, const std::vector<int>& dims, EigMem& mem);
// the blocked eigensolver on es's problems: problem k's matrix is read from
// Am at jprobs[es.mols[k]] (a_off, ld), its eigenvalues are written to Am's
// diagonal and its eigenvectors as rows of Wm (w_off, ld); jstate gets the
// done / failed flags (kJacState ints per problem); mstate (optional) skips
// molecules that no longer iterate
void eig_set_run(EigSet& es, const JacobiProb* jprobs, int* jstate, const MolState* mstate, double* Am, double* Wm,
                 hipStream_t s);

class Batch {
 public:
  // a batch of fixed molecules (packed layout)
  Batch(const std::vector<MoleculeInput>& mols, const Options& opt, hipStream_t stream = nullptr);
  // the same on lent memory and streams (allocates nothing; throws when the
  // molecules do not fit loan.bytes)
  Batch(const std::vector<MoleculeInput>& mols, const Options& opt, const Loan& loan);
  // an empty batch of cap.slots slots for continuous batching (admit/release)
  Batch(const Capacity& cap, const Options& opt, hipStream_t stream = nullptr);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // full single point of every molecule (setup, SCC loop, gradient)
  void run();
  // molecules the next setup()/gradient() work on (default: all)
  void select(const std::vector<int>& mols);
  void select_all();
  void setup();
  // returns true while any molecule is active; queue_next: queue the next
  // iteration's first part before waiting for the host's count
  bool scc_iteration(bool queue_next = false);
  void scc_prologue();
  bool prologue_queued_ = false;
  void gradient();
  // restart guess for molecule m: shell charges, atomic dipoles [n][3] and
  // quadrupoles [n][6] of a previous wavefunction (used by the next setup)
  void set_guess(int m, const double* qsh_in, const double* dipm_in, const double* qp_in);
  // new positions [n][3] (bohr) of molecule m, same atoms; the next run
  // recomputes everything that depends on them
  void set_positions(int m, const double* xyz);
  // on: every run after the first starts each molecule from the wavefunction
  // (shell charges, atomic dipoles and quadrupoles) the previous run left on
  // the device instead of an initial guess
  void keep_wavefunction(bool on) { keep_wfn_ = on; }
  // implicit solvent for every molecule of the batch (empty name: gas phase);
  // kernel 1 Still (GBSA parameter sets) or 2 P16 (ALPB parameter sets), alpb
  // adds the ALPB shape term, state 1 gsolv / 2 reference / 3 1 bar.  Throws
  // for a solvent without parameters.  Takes effect at the next setup().
  void set_solvent(const std::string& name, int kernel, bool alpb, int state = 1, double temperature = 298.15);
  bool solvated() const { return solv_on_; }
  MolResult result(int m, bool matrices = false) const;
  // current atomic charges of molecule m (after setup(): the initial guess)
  void charges(int m, double* out) const;
  int size() const { return nmol; }
  // sizes and index maps of one molecule; throws for invalid input
  static void describe(const MoleculeInput& mi, int maxiter, MolDesc& d, HostMol& hm, MolState& st);
  // footprint of one molecule in each packed array
  static Capacity footprint(const MolDesc& d);
  // place a molecule into a free slot (device arrays, index maps); -1 when
  // a packed array has no room.  rebuild() must follow before the next step.
  int admit(const MoleculeInput& mi);
  void release(int m);
  void rebuild();
  bool occupied(int m) const { return occupied_[m] != 0; }
  const std::string& error() const { return error_; }

  // --- layout (public for the kernel launchers) ---
  Options opt;
  hipStream_t stream;
  // lanes of the block-Jacobi rounds: the problems are split between
  // jac_lanes streams (stream, stream2, stream4) that run their rounds side by
  // side; stream2 also carries other work beside stream outside the rounds
  hipStream_t stream2 = nullptr;
  hipStream_t stream3 = nullptr;   // the small (dim <= 64) eigenproblems beside the lanes
  hipStream_t stream4 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join3 = nullptr, ev_join4 = nullptr;
  hipEvent_t ev_atm = nullptr;   // the ATM dispersion terms of the set-up (stream3) are done
  int jac_lanes = 3;
  // the block-Jacobi problems work on padded copies (rows 128-byte aligned);
  // the register path only (no refinement, banded rounds, LDS kernels or QL)
  bool jac_padded = true;
  int nmol = 0, max_n = 0, max_nsh = 0, max_nao = 0, max_nd4 = 0, max_nbr = 0;
  double intcut = 25.0, neglect = 1e-8, scfconv = 1e-6, qconv = 1e-4;
  std::vector<MolDesc> desc;
  MolDesc* d_desc = nullptr;
  MolDesc* d_sub = nullptr;     // descriptors of the selected molecules
  int nsub = 0;
  int sub_max_n = 0, sub_max_nsh = 0, sub_max_nao = 0, sub_max_nd4 = 0;
  long long sub_max_psp = 0, sub_max_pd = 0;   // largest s/p-only and d-shell parts of the pair lists
  int* phase_act = nullptr;     // 1 for selected molecules (per slot)
  int* open_act = nullptr;      // 1 for selected open-shell molecules (per slot)
  int nsub_open = 0;
  std::vector<int> sel_;
  MolState* d_state = nullptr;
  MolState* d_state_stage = nullptr;   // staging for setup()'s one upload of initial states
  Capacity cap_;
  Arena ar_[13];
  std::vector<char> occupied_;
  std::vector<HostMol> hmol_;
  std::vector<char> elem_present_;
  long long tiles_cap_ = 0, small_cap_ = 0, pairs_cap_ = 0, ta_cap_ = 0, tw_cap_ = 0, to_cap_ = 0, big_cap_ = 0;
  long long xp_cap_ = 0, xt_cap_ = 0;
  void allocate(const Capacity& c);
  void init_common();
  void place(int m, const MolDesc& d, const HostMol& hm, const MolState& st, bool upload);
  void update_c6ref();
  std::string error_;

  // atom vectors
  int* at = nullptr;
  double *xyz = nullptr, *cn = nullptr, *cn_d4 = nullptr, *cn_eeq = nullptr, *cn_cov = nullptr;
  double *q = nullptr, *dipm = nullptr, *qp = nullptr, *radcn = nullptr;
  double *vs = nullptr, *vd = nullptr, *vq = nullptr, *ashift = nullptr;
  double *grad = nullptr, *dhdcn = nullptr, *dedcn = nullptr, *tmp2 = nullptr, *eat = nullptr;
  double *grad_atm = nullptr, *dedcn_atm = nullptr;   // ATM gradient and dE/dCN terms of the set-up
  double *ae01 = nullptr;   // per-atom AES energy partials (3 per atom)
  // shell vectors
  int *sh_at = nullptr, *sh_l = nullptr, *sh_ao0 = nullptr, *sh_m = nullptr;
  int *at_sh0 = nullptr, *at_ao0 = nullptr;
  double *qsh = nullptr, *zsh = nullptr, *shift = nullptr, *shgam = nullptr, *selfe = nullptr;
  double *dsedcn = nullptr, *jq = nullptr;
  // AO vectors
  int *ao_at = nullptr, *ao_sh = nullptr;
  double *emo = nullptr, *focc = nullptr, *focca = nullptr, *foccb = nullptr, *ev = nullptr;
  double *row_q = nullptr, *row_e = nullptr, *row_d = nullptr, *row_qp = nullptr, *row_pd = nullptr;
  // D4
  int *d4_at = nullptr, *d4_ref = nullptr;
  double *gw = nullptr, *zv = nullptr, *dzv = nullptr, *d4dum = nullptr, *c6ref = nullptr;
  double *zetavec = nullptr, *zerovec = nullptr, *zetadcn = nullptr, *zerodcn = nullptr;
  // matrices (nao^2)
  double *S = nullptr, *H0 = nullptr, *H0n = nullptr, *D = nullptr, *Q = nullptr;
  double *H = nullptr, *C = nullptr, *T = nullptr, *A = nullptr, *W = nullptr, *P = nullptr;
  // other matrices
  double *J = nullptr, *dispmat = nullptr, *gab3 = nullptr, *gab5 = nullptr;
  double *c6 = nullptr, *dc6 = nullptr, *c60 = nullptr, *dc60 = nullptr;
  double *E = nullptr, *xE = nullptr;
  // Broyden
  double *q_in = nullptr, *dq = nullptr, *qlast = nullptr, *dqlast = nullptr;
  double *bdf = nullptr, *bu = nullptr, *ba = nullptr, *bomega = nullptr;
  double *yE = nullptr, *erep = nullptr, *wbo = nullptr;
  double* gpair = nullptr;
  int2* pair_list = nullptr;   // the molecules' HostMol::pairs at p0   // per atom pair H0 gradient terms: 5 doubles at 5 * (p0 + i n + j)
  // Jacobi bookkeeping
  JacSet jac_n;
  int* jac_nactive = nullptr;
  int* perm = nullptr;
  int* eeq_failed = nullptr;
  int* eeq_big = nullptr;        // sub-list indices of EEQ systems factorised column by column
  // implicit solvation: device model, Born radii and chain factors, surface
  // (per atom and per grid point), H-bond weights, potential, dE/dbrad, the
  // interaction matrix (n x n at p0) and per-molecule surface energy, shape
  // descriptor and q^T B q / 2
  bool solv_on_ = false;
  SolvDev* d_solv = nullptr;
  double *sv_brad = nullptr, *sv_dpsi = nullptr, *sv_sasa = nullptr, *sv_hbw = nullptr, *sv_pot = nullptr;
  double *sv_grddb = nullptr, *sv_w = nullptr, *sv_B = nullptr, *sv_gsasa = nullptr, *sv_adet = nullptr;
  double *sv_qbq = nullptr, sv_gshift = 0.0;
  int* sv_failed = nullptr;
  double* ns_c = nullptr;
  unsigned long long* ns_res = nullptr;
  int *ns_stage = nullptr, *ns_act = nullptr, *ns_failed = nullptr, *ns_fin = nullptr;
  int* h_pinned = nullptr;
  long long jacobi_rounds_ = 0;
  long long oa_steps_ = 0;
  const Element* elements_dev = nullptr;
  int ntiles = 0, ntiles_sym = 0;   // d_tiles: all tiles, then the b >= a tiles of symmetric products
  int4* d_tiles = nullptr;
  // problems of at least this dimension go to the blocked eigensolver
  // (tridiagonalisation, divide and conquer, back-transformation) instead of
  // the block-Jacobi rounds; register path only.  Off unless the tuning file
  // or XTBAMD_EIG_BLK_MIN sets it.
  int eig_blk_min = 1 << 30;
  // the blocked solve as a captured HIP graph (XTBAMD_GRAPH=1); off by
  // default: replaying it measured no faster than the eager launches (one
  // 4200-AO single point 1.523 against 1.514 s, its launches are bound by
  // their own latency, not by the host), and the profiler's kernel trace of
  // a replayed graph of this size ended in a host fault
  bool eig_graph = false;
  // SCC iterations (from the first) whose eigenproblems above eig_blk_min go
  // to the blocked eigensolver; later ones, with the previous orbitals as the
  // start, to the block-Jacobi lanes
  int eig_blk_iters = 1 << 30;
  int jac_period_ = 1;   // rounds between the lanes' progress tests
  EigSet eig_;
  void eig_build();
  // the blocked eigensolver on eig_'s problems: A from Am (Jacobi problem
  // layout), eigenvalues onto Am's diagonal, eigenvector rows into Wm
  void eig_blocked(const JacSet& js, double* Am, double* Wm, bool check_active, hipStream_t s);
  // returns true when W comes out as rows (row j = eigenvector j)
  void jacobi_lanes(JacSet& js, double* Am, double* Wm);
  bool jacobi(JacSet& js, double* Am, double* Wm, double* evout, bool check_active, bool keep_w = false,
              int banded = 0);
  int banded_passes = 2;   // banded rounds pairs after a warm-start refinement
  // eigenvectors as rows of W: LDS Jacobi up to 64, tridiagonal QL above
  void eigh_rows(JacSet& js, double* Am, double* Wm, double* evout);
  bool use_ql = false;     // tridiagonal QL (rows of W) instead of block Jacobi (columns of W)
  int* eig_failed = nullptr;
  // refine W (init I) towards the eigenvectors of A; leaves A <- W^T A W
  void refine(double* Am, double* Wm);
  int oa_max_steps = 0;     // refinement steps before the Jacobi sweeps (0 = off)
  int inner_sweeps = 1;    // sweeps of each pivot problem per block-Jacobi round
  double jac_tol = 1e-12;   // Jacobi convergence: max |A_ij| <= jac_tol * max(|A_ii|, 1)
  // before a molecule's final diagonalization the tolerance may follow its
  // SCC progress: jac_tol_q x the last RMS charge change, between jac_tol and
  // jac_tol_mid; the first iteration (orbitals from the initial guess) takes
  // jac_tol_mid.  Default: only the first iteration loose (1e-8; jac_tol_q
  // 0): on the 512-molecule step 1.3 ms less, energies within 3e-12 Eh, no
  // iteration count changed, every golden value of the reference's tests
  // kept.  Looser later iterations (jac_tol_q 1e-6: 7% faster) move the
  // trajectory enough to shift HOMO-LUMO gaps by up to 3e-7 eV and
  // finite-difference Hessians beyond the reference's test tolerances.
  double jac_tol_mid = 1e-8;
  double jac_tol_q = 0.0;
  int jac_max_sweeps = 40;   // sweeps before a Jacobi problem is given up (the molecule fails)
  int jac_cross = 1;   // most outer rounds' pivots rotate only the pairs across their two blocks
  bool trace_rounds = false;   // print the Jacobi rounds of each SCC iteration to stderr
  bool pivot_lds = false;  // 32x32 pivots: LDS kernel instead of the register-resident one
  bool apply_lds = false;  // 32x32 pivots: LDS tile updates with W as columns instead of register ones with W as rows
  int jac_p = 32;          // block-Jacobi pivot size (64: 32-wide blocks, 32: 16-wide blocks)
  bool use_mfma = true;   // FP64 matrix-core GEMM (else FP64 VALU FMA)
  // mixed precision: the Fock transforms (H C, C^T H C) and the density
  // products (C f C^T, P S) on the FP32 matrix cores with an FP64 residual
  // correction (k_gemm_mix); the eigenvector update C W stays FP64
  bool mixed = false;
  bool mixed_here_ = false;   // set around the products of an SCC iteration that may run mixed
  double* Sm = nullptr;
  double* G = nullptr;
  int* oa_act = nullptr;
  int* oa_ok = nullptr;
  unsigned long long* oa_stat = nullptr;
  // sym: the product is symmetric (computed on the b >= a tiles, mirrored)
  void gemm(int kind, const double* a, const double* b, double* c, const double* scale, const int* perm,
            bool check_active, double alpha = 1.0, const int* mask = nullptr, bool sym = false);
  void orthogonalizer();   // C = S^(-1/2) by Newton-Schulz iterations
  void eeq_solve();        // EEQ charges by Gaussian elimination

  // HIP-event timing of kernel families over a run (roofline measurement):
  // bit 1 = k_gemm, bit 2 = the block-Jacobi pivots, bit 3 = the block-Jacobi
  // tile updates (k_jac_apply_R).  Algorithmic work is counted on the device:
  // 2 N^3 flops per molecule a GEMM launch processes; 9216 flops per inner
  // round of a 32x32 pivot problem (2048 elements rotated on both sides of A
  // and 1024 in Q, 3 flops each); the tile updates' compulsory HBM bytes
  // (each touched element of A and W read once and written once, A through
  // one of its mirrored tiles).  The Jacobi families are timed on a sample of
  // the rounds.
  static constexpr int kProfKinds = 4;   // 1 GEMM, 2 Jacobi pivots, 3 Jacobi tile updates
  // sampled launches per family (pivots, tile updates) with device wall-clock
  // spans; each holds a clock pair in each of its kProfLines lines
  static constexpr int kProfTsSlots = 1024, kProfSlotU64 = kProfLines * kProfLineU64;
  unsigned long long* prof_ts_ = nullptr;
  int prof_ts_used_[2] = {0, 0};   // pivots, tile updates
  double prof_span(int kind);
  int prof_mask = 0;
  std::vector<hipEvent_t> prof_ev_[kProfKinds];
  int prof_used_[kProfKinds] = {};
  int prof_all_[kProfKinds] = {};   // every launch of a family timed on a sample
  long long prof_round_ = 0;       // block-Jacobi rounds since prof_enable (the sample clock)
  unsigned long long* prof_work_ = nullptr;   // [kProfKinds]
  void prof_begin(int kind, hipStream_t s = nullptr);
  void prof_end(int kind, const int* mask, int mstride, hipStream_t s = nullptr);
  int prof_result(int kind, int* launches, double* total_ms, double* work);
  void prof_enable(int mask);
  std::vector<void*> allocs_;
  Loan loan_;                  // lent memory and streams (loan_.base != nullptr)
  size_t loan_used_ = 0;       // bytes of the loan carved so far
  bool on_loan() const { return loan_.base != nullptr; }
  void build_fixed(const std::vector<MoleculeInput>& mols);
  std::vector<MolState> h_state0_;   // initial SCC state, restored by every setup()
  std::vector<MolState> state_stage_;
  std::vector<std::vector<double>> guess_qsh_, guess_dipm_, guess_qp_;
  bool setup_done_ = false;
  bool keep_wfn_ = false;    // start from the resident wavefunction when there is one
  bool wfn_valid_ = false;   // a completed run left its wavefunction on the device
  int iterations_ = 0;
  template <class T_>
  T_* alloc(size_t n);
};

// Continuous batching: requests are admitted into batch slots as packed-array
// room (the HBM budget) frees up; each step runs one SCC iteration of every
// resident molecule, whatever its iteration; finished molecules get their
// gradient, leave their results on the host and free their slots.
class Server {
 public:
  Server(const Capacity& cap, const Options& opt);
  long long submit(const MoleculeInput& mi);
  // admission + one SCC iteration + completion; returns requests not finished
  int step();
  // ids finished since the last call
  std::vector<long long> take_finished();
  bool result(long long id, MolResult& out, bool erase = true);
  int running() const;
  int queued() const { return (int)queue_.size(); }
  Batch batch;

 private:
  std::deque<std::pair<long long, MoleculeInput>> queue_;
  std::vector<long long> slot_req_;
  std::map<long long, MolResult> done_;
  std::vector<long long> finished_;
  std::vector<int> active_host_;
  int* d_active_ = nullptr;
  long long next_id_ = 0;
};

// capacity for continuous batching from an HBM budget: the packed arrays hold
// hbm_bytes worth of molecules of up to max_atoms atoms (at most 4 AOs,
// 2 shells and 7 D4 references per atom), in up to max_slots slots
Capacity server_capacity(double hbm_bytes, int max_atoms, int max_slots, int maxiter);

}  // namespace xtbamd
