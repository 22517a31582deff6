// Host side of the batched GFN2-xTB engine: layout, work lists, the SCC
// driver and the C ABI of the batch interface (include/xtb_amd_batch.h).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cctype>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "engine.h"
#include "gfn2_common.h"
// kernels (single translation unit)
#include "k_aes.hip"
#include "k_d4.hip"
#include "k_grad.hip"
#include "k_linalg.hip"
#include "k_eig.hip"
#include "k_scc.hip"
#include "k_setup.hip"
#include "k_solv.hip"

#define HIPCHECK(x)                                                                              \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess)                                                                        \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace xtbamd {

namespace {

__global__ void k_fill(double* p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_int(int* p, long long n, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// n (start, end) clock pairs, one per 128-byte line
__global__ void k_ts_init(unsigned long long* ts, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    ts[(long long)kProfLineU64 * i] = ~0ull;
    ts[(long long)kProfLineU64 * i + 1] = 0ull;
  }
}
__global__ void k_copy(const double* a, double* b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// EEQ systems with at least this many unknowns are factorised column by column
// across the whole device instead of inside one workgroup
constexpr int kEeqBigNe = 256;

// EEQ charges: Gaussian elimination with partial pivoting of the constrained
// (n+1) x (n+1) system E y = x in place (the reference solves it with dsysv,
// src/eeq_model.f90:678-1015), then the shell charges (iniqshell:
// src/scc_core.f90:1475-1506); uniform charges if the solve fails
// (src/xtb/calculator.f90:479-484).  One workgroup per molecule.
// Systems of at least big_ne unknowns arrive factorised by k_eeq_lu_pivot /
// k_eeq_lu_update (the same elimination spread over the whole device); this
// kernel then only substitutes.
__global__ void __launch_bounds__(256) k_eeq_gesv(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                  const int* __restrict__ at, const int* __restrict__ at_sh0, double* E,
                                                  double* xE, double* yE, double* q, double* qsh,
                                                  const double* __restrict__ zsh, int* failed, int big_ne) {
  __shared__ double sh[4];
  __shared__ double shv[4];
  __shared__ int shi[4];
  __shared__ int piv;
  const MolDesc md = desc[blockIdx.x];
  const int ne = md.ne, n = md.n;
  double* M = E + md.e0;
  double* b = xE + md.y0;
  double* y = yE + md.y0;
  const int tid = threadIdx.x;
  const bool factored = ne >= big_ne;
  bool ok = !(factored && failed[md.mol]);
  for (int k = 0; k < (factored ? 0 : ne); ++k) {
    // pivot: largest |M[i][k]|, i >= k, smallest index on ties
    double best = -1.0;
    int bi = k;
    for (int i = k + tid; i < ne; i += blockDim.x) {
      const double v = fabs(M[(long long)i * ne + k]);
      if (v > best) { best = v; bi = i; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { shv[tid >> 6] = best; shi[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      double bv = shv[0];
      int bx = shi[0];
      for (int w = 1; w < 4; ++w)
        if (shv[w] > bv || (shv[w] == bv && shi[w] < bx)) { bv = shv[w]; bx = shi[w]; }
      piv = (bv > 0.0) ? bx : -1;
    }
    __syncthreads();
    const int p = piv;
    if (p < 0) { ok = false; break; }
    if (p != k) {
      for (int j = tid; j < ne; j += blockDim.x) {
        const double t = M[(long long)k * ne + j];
        M[(long long)k * ne + j] = M[(long long)p * ne + j];
        M[(long long)p * ne + j] = t;
      }
      if (tid == 0) { const double t = b[k]; b[k] = b[p]; b[p] = t; }
      __syncthreads();
    }
    // multipliers l_ik = M_ik / M_kk stored below the diagonal, then the
    // rank-1 update of the trailing block and of the right-hand side
    const double dkk = M[(long long)k * ne + k];
    for (int i = k + 1 + tid; i < ne; i += blockDim.x) M[(long long)i * ne + k] /= dkk;
    __syncthreads();
    const int m = ne - k - 1;
    for (long long e = tid; e < (long long)m * m; e += blockDim.x) {
      const int i = k + 1 + (int)(e / m), j = k + 1 + (int)(e % m);
      M[(long long)i * ne + j] -= M[(long long)i * ne + k] * M[(long long)k * ne + j];
    }
    for (int i = k + 1 + tid; i < ne; i += blockDim.x) b[i] -= M[(long long)i * ne + k] * b[k];
    __syncthreads();
  }
  if (ok) {
    for (int k = ne - 1; k >= 0; --k) {
      double s = 0.0;
      for (int j = k + 1 + tid; j < ne; j += blockDim.x) s += M[(long long)k * ne + j] * y[j];
      s = block_sum<256>(s, sh);
      if (tid == 0) y[k] = (b[k] - s) / M[(long long)k * ne + k];
      __syncthreads();
    }
    double tot = 0.0;
    for (int i = tid; i < n; i += blockDim.x) {
      q[md.a0 + i] = y[i];
      tot += y[i];
    }
    tot = block_sum<256>(tot, sh);
    if (!(fabs(tot - md.charge) <= 1.e-6)) ok = false;
  }
  if (!ok) {
    for (int i = tid; i < n; i += blockDim.x) q[md.a0 + i] = md.charge / n;
    if (tid == 0) failed[md.mol] = 1;
  }
  __syncthreads();
  for (int i = tid; i < n; i += blockDim.x) {
    const Element& e = el[at[md.a0 + i]];
    for (int s = 0; s < e.nsh; ++s) {
      const int k = md.s0 + at_sh0[md.a0 + i] + s;
      qsh[k] = (zsh[k] / e.zval) * q[md.a0 + i];
    }
  }
}

__global__ void k_eeq_reset(const MolDesc* __restrict__ desc, int ns, int* failed) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ns) failed[desc[k].mol] = 0;
}

// Failure codes of MolState::failed (0 = fine).
constexpr int kFailOrtho = 2;   // S^(-1/2) iterations did not converge
constexpr int kFailDiag = 3;    // eigensolver gave up ("Diagonalization of Hamiltonian failed")
constexpr int kFailEeq = 4;     // EEQ system singular
constexpr int kFailSolv = 7;    // solvent surface

// A molecule whose solver flagged a failure (flags[m * stride] != 0, per
// slot) gets the code and stops iterating, like the early return of the
// reference's SCC after a failed diagonalization (src/scc_core.f90:475-480);
// its results are still filled.  The first failure's code is kept.
__global__ void k_fail_fold(MolState* st, int nmol, const int* __restrict__ flags, int stride, int code,
                            const int* __restrict__ sel) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol || (sel && !sel[m]) || !flags[(long long)m * stride]) return;
  if (!st[m].failed) st[m].failed = code;
  st[m].active = 0;
}

// Column k of the LU factorisation with partial pivoting of large EEQ systems
// (one workgroup per listed molecule): pivot search with the same tie rule as
// k_eeq_gesv, row swap, multipliers and the right-hand-side update.
__global__ void __launch_bounds__(256) k_eeq_lu_pivot(const MolDesc* __restrict__ desc, const int* __restrict__ list,
                                                      int k, double* E, double* xE, int* failed) {
  __shared__ double shv[4];
  __shared__ int shi[4];
  __shared__ int piv;
  const MolDesc md = desc[list[blockIdx.x]];
  const int ne = md.ne;
  if (k >= ne || failed[md.mol]) return;
  double* M = E + md.e0;
  double* b = xE + md.y0;
  const int tid = threadIdx.x;
  double best = -1.0;
  int bi = k;
  for (int i = k + tid; i < ne; i += blockDim.x) {
    const double v = fabs(M[(long long)i * ne + k]);
    if (v > best) { best = v; bi = i; }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(best, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
  }
  if ((tid & 63) == 0) { shv[tid >> 6] = best; shi[tid >> 6] = bi; }
  __syncthreads();
  if (tid == 0) {
    double bv = shv[0];
    int bx = shi[0];
    for (int w = 1; w < 4; ++w)
      if (shv[w] > bv || (shv[w] == bv && shi[w] < bx)) { bv = shv[w]; bx = shi[w]; }
    piv = (bv > 0.0) ? bx : -1;
    if (piv < 0) failed[md.mol] = 1;
  }
  __syncthreads();
  const int p = piv;
  if (p < 0) return;
  if (p != k) {
    for (int j = tid; j < ne; j += blockDim.x) {
      const double t = M[(long long)k * ne + j];
      M[(long long)k * ne + j] = M[(long long)p * ne + j];
      M[(long long)p * ne + j] = t;
    }
    if (tid == 0) { const double t = b[k]; b[k] = b[p]; b[p] = t; }
    __syncthreads();
  }
  const double dkk = M[(long long)k * ne + k], bk = b[k];
  for (int i = k + 1 + tid; i < ne; i += blockDim.x) {
    const double l = M[(long long)i * ne + k] / dkk;
    M[(long long)i * ne + k] = l;
    b[i] -= l * bk;
  }
}

// rank-1 update of the trailing block after column k: one workgroup per row
__global__ void __launch_bounds__(256) k_eeq_lu_update(const MolDesc* __restrict__ desc, const int* __restrict__ list,
                                                       int k, double* E, const int* __restrict__ failed) {
  const MolDesc md = desc[list[blockIdx.y]];
  const int ne = md.ne;
  const int i = k + 1 + blockIdx.x;
  if (i >= ne || failed[md.mol]) return;
  double* M = E + md.e0;
  const double l = M[(long long)i * ne + k];
  double* Mi = M + (long long)i * ne;
  const double* Mk = M + (long long)k * ne;
  for (int j = k + 1 + threadIdx.x; j < ne; j += blockDim.x) Mi[j] -= l * Mk[j];
}

// zero the atomic dipoles and quadrupoles (or the gradient) of the selected molecules
// the ATM gradient and CN derivative terms computed beside the SCC, added
// where the gradient takes them
__global__ void k_add_atm(const MolDesc* __restrict__ desc, const double* __restrict__ gatm,
                          const double* __restrict__ datm, double* grad, double* dedcn) {
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.n) return;
  for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + i) + k] += gatm[3 * (md.a0 + i) + k];
  dedcn[md.a0 + i] += datm[md.a0 + i];
}

__global__ void k_zero_atoms(const MolDesc* __restrict__ desc, double* v3, double* v6) {
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.n) return;
  for (int k = 0; k < 3; ++k) v3[3 * (md.a0 + i) + k] = 0.0;
  if (v6)
    for (int k = 0; k < 6; ++k) v6[6 * (md.a0 + i) + k] = 0.0;
}

// atomic charges from shell charges, each atom's shells summed in ascending
// order from zero (the order of the host-side restart guess)
__global__ void k_q_from_qsh(const MolDesc* __restrict__ desc, const int* __restrict__ sh_at,
                             const double* __restrict__ qsh, double* q) {
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.n) return;
  double acc = 0.0;
  for (int s = 0; s < md.nsh; ++s)
    if (sh_at[md.s0 + s] == i) acc += qsh[md.s0 + s];
  q[md.a0 + i] = acc;
}

// Newton-Schulz bookkeeping for the selected molecules: stage 0 (running),
// the others stay done (stage 2)
__global__ void k_ns_begin(int nmol, const int* __restrict__ sel, int* stage, int* act, int* failed,
                           unsigned long long* res, double* cbound) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  stage[m] = sel[m] ? 0 : 2;
  act[m] = sel[m] ? 1 : 0;
  if (sel[m]) {
    failed[m] = 0;
    cbound[m] = 0.0;
  }
  res[m] = 0ull;
}

// SAD guess: zero atomic charges except the total charge spread evenly
__global__ void k_sad_guess(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                            const int* __restrict__ at, const int* __restrict__ at_sh0, double* q, double* qsh,
                            const double* __restrict__ zsh) {
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.n) return;
  const double qi = md.charge / md.n;
  q[md.a0 + i] = qi;
  const Element& e = el[at[md.a0 + i]];
  for (int s = 0; s < e.nsh; ++s) {
    const int k = md.s0 + at_sh0[md.a0 + i] + s;
    qsh[k] = (zsh[k] / e.zval) * qi;
  }
}

// --- Newton-Schulz iteration for S^(-1/2) (coupled form, Higham, "Functions
// of Matrices", 2008, eq. 6.35): Y0 = S/c, Z0 = I,
// T = (3I - Z Y)/2, Y <- Y T, Z <- T Z; Z -> (S/c)^(-1/2).
// c = max row sum |S_ij| bounds the spectrum of S from above.
// grid (row chunks of 64, molecules); half the maximum over
// the molecule's rows by an atomic max on the bit patterns of cbound (cleared
// by k_ns_begin; non-negative doubles order like their bit patterns).  The
// maximum absolute row sum bounds the largest eigenvalue of S (Gershgorin),
// so the eigenvalues mu of Y0 = S / cbound lie in (0, 2]; the iteration maps
// each mu to mu (3 - mu)^2 / 4 and converges for mu in (0, 3), and the small
// ones, which set the number of iterations (each multiplies them by about
// 9/4), start twice as large as with the full bound.
__global__ void __launch_bounds__(256) k_ns_bound(const MolDesc* __restrict__ desc, const double* __restrict__ S,
                                                  double* cbound) {
  // 64 rows per workgroup, one wave per row at a time with its lanes along
  // the row (coalesced reads); one atomic per workgroup (same-address
  // atomics from every row had serialised)
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  const int N = md.nao;
  const int i0 = blockIdx.x * 64;
  if (i0 >= N) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double m = 0.0;
  for (int i = i0 + w; i < min(N, i0 + 64); i += 4) {
    const double* row = S + md.m0 + (long long)i * N;
    double r = 0.0;
    for (int j = lane; j < N; j += 64) r += fabs(row[j]);
    m = fmax(m, wave_sum(r));
  }
  m = block_max<256>(m, sh);
  if (threadIdx.x == 0 && m > 0.0)
    atomicMax((unsigned long long*)&cbound[md.mol], (unsigned long long)__double_as_longlong(0.5 * m));
}

__global__ void k_ns_init(const MolDesc* __restrict__ desc, const double* __restrict__ S,
                          const double* __restrict__ cbound, double* Y, double* Z) {
  const MolDesc md = desc[blockIdx.y];
  const long long N = md.nao;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N * N; t += (long long)gridDim.x * blockDim.x) {
    Y[md.m0 + t] = S[md.m0 + t] / cbound[md.mol];
    Z[md.m0 + t] = (t / N == t % N) ? 1.0 : 0.0;
  }
}

// T = (3I - P)/2 with P = Z Y, and max |I - P| per molecule
__global__ void __launch_bounds__(256) k_ns_form(const MolDesc* __restrict__ desc, const int* __restrict__ act,
                                                 double* P, unsigned long long* res) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  if (!act[md.mol]) return;
  const long long N = md.nao;
  if ((long long)blockIdx.x * blockDim.x >= N * N) return;
  double r = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N * N; t += (long long)gridDim.x * blockDim.x) {
    const double d = (t / N == t % N) ? 1.0 : 0.0;
    const double p = P[md.m0 + t];
    r = fmax(r, fabs(d - p));
    P[md.m0 + t] = 0.5 * (3.0 * d - p);
  }
  r = block_max<256>(r, sh);
  if (threadIdx.x == 0 && r > 0.0) atomicMax(&res[md.mol], (unsigned long long)__double_as_longlong(r));
}

// per-molecule convergence: residual < 1e-10, then one more update (quadratic
// convergence takes it to rounding level); act = 0 once done, and fin = the
// parity of the iteration whose buffers hold the molecule's final Y, Z
__global__ void __launch_bounds__(256) k_ns_mark(int nmol, int* stage, int* act, unsigned long long* res, int iter,
                                                 int maxiter, int* failed, int* nactive, int* fin) {
  __shared__ double sh[4];
  double c = 0.0;
  for (int m = threadIdx.x; m < nmol; m += blockDim.x) {
    const double r = __longlong_as_double((long long)res[m]);
    res[m] = 0ull;
    const int before = stage[m];
    if (stage[m] == 1) stage[m] = 2;
    else if (stage[m] == 0 && r < 1e-10) stage[m] = 1;
    if (stage[m] == 0 && iter >= maxiter) { stage[m] = 2; failed[m] = 1; }
    if (before < 2 && stage[m] == 2) fin[m] = iter & 1;
    act[m] = stage[m] < 2 ? 1 : 0;
    c += act[m];
  }
  c = block_sum<256>(c, sh);
  if (threadIdx.x == 0) nactive[0] = (int)c;
}

__global__ void k_ident_active(const MolDesc* __restrict__ desc, const MolState* __restrict__ st, double* W) {
  const MolDesc md = desc[blockIdx.y];
  if (!st[blockIdx.y].active) return;
  const long long N = md.nao;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N * N) W[md.m0 + t] = (t / N == t % N) ? 1.0 : 0.0;
}

// algorithmic flops of one batched GEMM launch: 2 N^3 per molecule processed
__global__ void k_count_gemm(const MolDesc* __restrict__ desc, int nmol, const int* __restrict__ mask, int mstride,
                             unsigned long long* work) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  if (mask && !mask[(long long)m * mstride]) return;
  const unsigned long long N = (unsigned long long)desc[m].nao;
  prof_add(work, 2ull * N * N * N);
}

__global__ void k_copy_state_active(const MolState* __restrict__ st, int nmol, int* act) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nmol) act[m] = st[m].active;
}

__global__ void k_copy_mask(const MolDesc* __restrict__ desc, const int* __restrict__ act, const double* __restrict__ a,
                            double* b) {
  const MolDesc md = desc[blockIdx.y];
  if (!act[md.mol]) return;
  const long long n2 = (long long)md.nao * md.nao;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n2) b[md.m0 + t] = a[md.m0 + t];
}

__global__ void k_ns_final(const MolDesc* __restrict__ desc, const double* __restrict__ Z0,
                           const double* __restrict__ Z1, const int* __restrict__ fin,
                           const double* __restrict__ cbound, double* C) {
  const MolDesc md = desc[blockIdx.y];
  const long long N = md.nao;
  const double* Z = fin[md.mol] ? Z1 : Z0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < N * N; t += (long long)gridDim.x * blockDim.x) C[md.m0 + t] = Z[md.m0 + t] / sqrt(cbound[md.mol]);
}

// total energy, gradient norm and dipole (src/scf_module.F90:882-913,
// calc_dipole src/dipole.f90:94-128)
__global__ void __launch_bounds__(256) k_final(const MolDesc* __restrict__ desc, MolState* state,
                                               const Element* __restrict__ el, const int* __restrict__ at,
                                               const double* __restrict__ xyz, const double* __restrict__ grad,
                                               const double* __restrict__ erep, const double* __restrict__ eatm,
                                               const double* __restrict__ row_d) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.x];
  MolState& st = state[md.mol];
  double ep = 0, eb = 0, g2 = 0, dx = 0, dy = 0, dz = 0;
  for (int i = threadIdx.x; i < md.n; i += blockDim.x) {
    ep += erep[md.a0 + i];
    eb += eatm[md.a0 + i];
    for (int k = 0; k < 3; ++k) g2 += grad[3 * (md.a0 + i) + k] * grad[3 * (md.a0 + i) + k];
    const double z = el[at[md.a0 + i]].zval;
    dx += xyz[3 * (md.a0 + i)] * z;
    dy += xyz[3 * (md.a0 + i) + 1] * z;
    dz += xyz[3 * (md.a0 + i) + 2] * z;
  }
  for (int mu = threadIdx.x; mu < md.nao; mu += blockDim.x) {
    dx -= row_d[3 * (md.o0 + mu)];
    dy -= row_d[3 * (md.o0 + mu) + 1];
    dz -= row_d[3 * (md.o0 + mu) + 2];
  }
  ep = block_sum<256>(ep, sh);
  eb = block_sum<256>(eb, sh);
  g2 = block_sum<256>(g2, sh);
  dx = block_sum<256>(dx, sh);
  dy = block_sum<256>(dy, sh);
  dz = block_sum<256>(dz, sh);
  if (threadIdx.x == 0) {
    st.ep = ep;
    st.embd = eb;
    st.etotal = st.eel + ep + eb;
    st.gnorm = sqrt(g2);
    st.dipole[0] = dx;
    st.dipole[1] = dy;
    st.dipole[2] = dz;
  }
}

inline unsigned grid1(long long n, int b) { return (unsigned)((n + b - 1) / b); }
// blocks per molecule of the grid-stride kernels over n x n matrices (256
// threads each): a launch needs no more blocks than this however large the
// largest molecule, so the small molecules' empty blocks cost nothing
inline unsigned gstride(long long n2) { return (unsigned)std::min<long long>(grid1(n2, 256), 64); }

// Aufbau occupation counts (occu: src/scc_core.f90:1112-1170)
void host_occu(int ndim, int nel, int nopen, int& ihomoa, int& ihomob) {
  std::vector<int> f(ndim, 0);
  if (nel % 2 == 0) {
    const int ih = nel / 2;
    for (int i = 0; i < ih && i < ndim; ++i) f[i] = 2;
    if (nopen > 1)
      for (int i = 1; i <= nopen / 2; ++i) {
        if (ih - i >= 0 && ih - i < ndim) f[ih - i] -= 1;
        if (ih + i - 1 < ndim) f[ih + i - 1] += 1;
      }
  } else {
    const int na = nel / 2 + (nopen - 1) / 2 + 1, nb = nel / 2 - (nopen - 1) / 2;
    for (int i = 0; i < na && i < ndim; ++i) f[i] += 1;
    for (int i = 0; i < nb && i < ndim; ++i) f[i] += 1;
  }
  ihomoa = ihomob = 0;
  for (int i = 0; i < ndim; ++i) {
    if (f[i] >= 1) ihomoa = i + 1;
    if (f[i] == 2) ihomob = i + 1;
  }
}

}  // namespace

template <class T_>
T_* Batch::alloc(size_t n) {
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T_);
  if (on_loan()) {
    // carve from the lent area, every piece 256-byte aligned
    const size_t off = (loan_used_ + 255) / 256 * 256;
    if (off + bytes > loan_.bytes)
      throw std::runtime_error("molecule needs more device memory than the " +
                               std::to_string(loan_.bytes >> 20) + " MiB scratch area holds");
    loan_used_ = off + bytes;
    return (T_*)(loan_.base + off);
  }
  void* p = nullptr;
  HIPCHECK(hipMalloc(&p, bytes));
  allocs_.push_back(p);
  return (T_*)p;
}

// --------------------------------------------------------------------------
// packed-array allocator
void Arena::init(long long c) {
  cap = c;
  free_.clear();
  if (c > 0) free_.push_back({0, c});
}

long long Arena::take(long long n) {
  if (n <= 0) return 0;
  for (size_t i = 0; i < free_.size(); ++i) {
    const long long b = free_[i].first, e = free_[i].second;
    if (e - b >= n) {
      if (e - b == n) free_.erase(free_.begin() + i);
      else free_[i].first = b + n;
      return b;
    }
  }
  return -1;
}

void Arena::give(long long off, long long n) {
  if (n <= 0) return;
  auto it = free_.begin();
  while (it != free_.end() && it->first < off) ++it;
  it = free_.insert(it, {off, off + n});
  // merge with the next and the previous range
  if (it + 1 != free_.end() && it->second == (it + 1)->first) {
    it->second = (it + 1)->second;
    free_.erase(it + 1);
  }
  if (it != free_.begin() && (it - 1)->second == it->first) {
    (it - 1)->second = it->second;
    free_.erase(it);
  }
}

// --------------------------------------------------------------------------
void Batch::describe(const MoleculeInput& mi, int maxiter, MolDesc& d, HostMol& hm, MolState& st) {
  const Element* el = host_elements();
  std::memset(&d, 0, sizeof(d));
  hm = HostMol();
  d.n = (int)mi.numbers.size();
  if (d.n < 1) throw std::runtime_error("molecule without atoms");
  if ((int)mi.xyz.size() != 3 * d.n) throw std::runtime_error("coordinate array has the wrong size");
  double zsum = 0.0;
  for (int i = 0; i < d.n; ++i) {
    const int z = mi.numbers[i];
    if (z < 1 || z > XTB_MAX_Z) throw std::runtime_error("element not parametrised in GFN2-xTB");
    const Element& e = el[z];
    hm.at.push_back(z);
    hm.at_sh0.push_back(d.nsh);
    hm.at_ao0.push_back(d.nao);
    for (int s = 0; s < e.nsh; ++s) {
      hm.sh_at.push_back(i);
      hm.sh_l.push_back(e.l[s]);
      hm.sh_ao0.push_back(d.nao);
      hm.sh_m.push_back(s);
      for (int k = 0; k < 2 * e.l[s] + 1; ++k) {
        hm.ao_at.push_back(i);
        hm.ao_sh.push_back(d.nsh);
      }
      d.nsh += 1;
      d.nao += 2 * e.l[s] + 1;
    }
    for (int r = 0; r < e.nref; ++r) {
      hm.d4_at.push_back(i);
      hm.d4_ref.push_back(r);
    }
    d.nd4 += e.nref;
    zsum += e.zval;
    for (int k = 0; k < 3; ++k) hm.xyz.push_back(mi.xyz[3 * i + k]);
  }
  // atom pairs i > j grouped by element pair: first the groups of two
  // elements with s and p shells only, then those with a d shell (each part
  // in the order of the groups' first pairs in lower-triangle order,
  // lower-triangle order inside a group; a counting sort, linear in the
  // number of pairs).  The pair kernels take the first npsp pairs with code
  // for s and p shells only (far fewer registers) and the rest apart.
  {
    const Element* elh = host_elements();
    auto has_d = [&](int z) {
      for (int s = 0; s < elh[z].nsh; ++s)
        if (elh[z].l[s] == 2) return true;
      return false;
    };
    std::vector<int> group(((size_t)XTB_MAX_Z + 1) * (XTB_MAX_Z + 1), -1);
    std::vector<long long> count;
    std::vector<char> gd;
    for (int i = 1; i < d.n; ++i)
      for (int j = 0; j < i; ++j) {
        int& g = group[(size_t)hm.at[i] * (XTB_MAX_Z + 1) + hm.at[j]];
        if (g < 0) {
          g = (int)count.size();
          count.push_back(0);
          gd.push_back(has_d(hm.at[i]) || has_d(hm.at[j]));
        }
        count[g] += 1;
      }
    std::vector<long long> start(count.size(), 0);
    long long off = 0;
    for (size_t g = 0; g < count.size(); ++g)
      if (!gd[g]) {
        start[g] = off;
        off += count[g];
      }
    d.npsp = (int)off;
    for (size_t g = 0; g < count.size(); ++g)
      if (gd[g]) {
        start[g] = off;
        off += count[g];
      }
    hm.pairs.resize((size_t)d.n * (d.n - 1) / 2);
    for (int i = 1; i < d.n; ++i)
      for (int j = 0; j < i; ++j) {
        const int g = group[(size_t)hm.at[i] * (XTB_MAX_Z + 1) + hm.at[j]];
        hm.pairs[start[g]++] = make_int2(i, j);
      }
  }
  d.ne = d.n + 1;
  d.charge = mi.charge;
  d.nel = (int)std::lround(zsum - mi.charge);
  d.nopen = mi.uhf;
  if (d.nopen == 0 && d.nel % 2 != 0) d.nopen = 1;
  if (d.nel > 2 * d.nao) throw std::runtime_error("Not enough basis functions for filling orbitals");
  d.nbr = d.nsh + 9 * d.n;
  d.maxiter = maxiter;
  std::memset(&st, 0, sizeof(st));
  st.active = 1;
  st.damp = 0.20;
  host_occu(d.nao, d.nel, d.nopen, st.ihomoa, st.ihomob);
  st.ihomo = st.ihomoa;
}

Capacity Batch::footprint(const MolDesc& d) {
  Capacity c;
  const long long mx = d.maxiter;
  c.slots = 1;
  c.n = d.n;
  c.nsh = d.nsh;
  c.nao = d.nao;
  c.nd4 = d.nd4;
  c.mm = (long long)d.nao * d.nao;
  c.jj = (long long)d.nsh * d.nsh;
  c.xx = (long long)d.nd4 * d.nd4;
  c.pp = (long long)d.n * d.n;
  c.ee = (long long)d.ne * d.ne;
  c.ne = d.ne;
  c.v = d.nbr;
  c.h = mx * d.nbr;
  c.g = 3 * mx * mx + 4 * mx;
  return c;
}

static long long* cap_field(Capacity& c, int k) {
  long long* f[13] = {&c.n, &c.nsh, &c.nao, &c.nd4, &c.mm, &c.jj, &c.xx, &c.pp, &c.ee, &c.ne, &c.v, &c.h, &c.g};
  return f[k];
}

// The project autotuner's choices (python -m xtb_amd.tune writes them to
// tuning_gfx950.txt next to this library; XTBAMD_TUNING names another file,
// "off" none): "key value" lines, '#' comments.  Read once per process.
namespace {
struct Tuning {
  int eig_blk_min = -1, eig_blk_iters = -1, jac_lanes = -1, jac_cross = -1, gemm_valu = -1;
};
const Tuning& tuning() {
  static const Tuning t = [] {
    Tuning r;
    std::string path;
    if (const char* e = std::getenv("XTBAMD_TUNING")) {
      if (std::string(e) == "off") return r;
      path = e;
    } else {
      Dl_info info;
      if (dladdr((void*)&tuning, &info) && info.dli_fname) {
        std::string f(info.dli_fname);
        const size_t k = f.rfind('/');
        path = (k == std::string::npos ? std::string(".") : f.substr(0, k)) + "/tuning_gfx950.txt";
      }
    }
    FILE* fp = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!fp) return r;
    char key[64];
    char val[64];
    char line[256];
    while (std::fgets(line, sizeof(line), fp)) {
      if (line[0] == '#' || std::sscanf(line, "%63s %63s", key, val) != 2) continue;
      const std::string k(key), v(val);
      if (k == "eig_blk_min") r.eig_blk_min = std::atoi(val);
      else if (k == "eig_blk_iters") r.eig_blk_iters = std::atoi(val);
      else if (k == "jac_lanes") r.jac_lanes = std::atoi(val);
      else if (k == "jac_cross") r.jac_cross = std::atoi(val);
      else if (k == "gemm") r.gemm_valu = v == "valu" ? 1 : 0;
    }
    std::fclose(fp);
    return r;
  }();
  return t;
}
}  // namespace

void Batch::init_common() {
  // the autotuner's choices first; environment variables override them
  const Tuning& tu = tuning();
  if (tu.eig_blk_min > 0) eig_blk_min = std::max(65, tu.eig_blk_min);
  if (tu.eig_blk_iters >= 0) eig_blk_iters = tu.eig_blk_iters;
  if (tu.jac_lanes > 0) jac_lanes = std::min(kMaxLanes, tu.jac_lanes);
  if (tu.jac_cross >= 0) jac_cross = tu.jac_cross != 0;
  if (tu.gemm_valu >= 0) use_mfma = tu.gemm_valu == 0;
  if (const char* g = std::getenv("XTBAMD_GEMM")) use_mfma = std::string(g) != "valu";
  if (const char* g = std::getenv("XTBAMD_MIXED")) mixed = std::atoi(g) != 0;
  if (const char* g = std::getenv("XTBAMD_GRAPH")) eig_graph = std::atoi(g) != 0;
  if (const char* r = std::getenv("XTBAMD_REFINE_STEPS")) oa_max_steps = std::atoi(r);
  if (const char* r = std::getenv("XTBAMD_INNER_SWEEPS")) inner_sweeps = std::max(1, std::atoi(r));
  if (const char* r = std::getenv("XTBAMD_PIVOT")) jac_p = std::atoi(r) == 64 ? 64 : 32;
  if (const char* r = std::getenv("XTBAMD_PIVOT_KERNEL")) pivot_lds = std::strcmp(r, "lds") == 0;
  if (const char* r = std::getenv("XTBAMD_APPLY_KERNEL")) apply_lds = std::strcmp(r, "lds") == 0;
  if (const char* r = std::getenv("XTBAMD_BANDED")) banded_passes = std::max(0, std::atoi(r));
  if (const char* r = std::getenv("XTBAMD_JAC_TOL")) jac_tol = std::atof(r);
  jac_tol_mid = std::max(jac_tol, jac_tol_mid);
  if (const char* r = std::getenv("XTBAMD_JAC_TOL_MID")) jac_tol_mid = std::max(jac_tol, std::atof(r));
  if (const char* r = std::getenv("XTBAMD_JAC_TOL_Q")) jac_tol_q = std::max(0.0, std::atof(r));
  if (const char* r = std::getenv("XTBAMD_JAC_CROSS")) jac_cross = std::atoi(r) != 0;
  // bound on the sweeps of one diagonalization; a problem still above the
  // tolerance after it fails its molecule
  if (const char* r = std::getenv("XTBAMD_JAC_MAX_SWEEPS")) jac_max_sweeps = std::min(40, std::max(1, std::atoi(r)));
  trace_rounds = std::getenv("XTBAMD_TRACE_ROUNDS") != nullptr;
  if (const char* r = std::getenv("XTBAMD_JAC_LANES")) jac_lanes = std::min(kMaxLanes, std::max(1, std::atoi(r)));
  if (const char* r = std::getenv("XTBAMD_EIG_BLK_MIN")) eig_blk_min = std::max(65, std::atoi(r));
  if (const char* r = std::getenv("XTBAMD_EIG_BLK_ITERS")) eig_blk_iters = std::max(0, std::atoi(r));
  if (on_loan()) {
    // the lender's streams and events (created and kept by the lender)
    stream = loan_.streams[0];
    stream2 = loan_.streams[1];
    stream3 = loan_.streams[2];
    stream4 = loan_.streams[3];
    ev_fork = loan_.events[0];
    ev_join = loan_.events[1];
    ev_join3 = loan_.events[2];
    ev_join4 = loan_.events[3];
  } else {
    // default priority (as every stream here), no implicit ordering with the null stream
    HIPCHECK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&stream3, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&stream4, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_join3, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_atm, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_join4, hipEventDisableTiming));
  }
  // XTBAMD_EIG=ql selects the tridiagonal QL eigensolver instead of block Jacobi
  if (const char* r = std::getenv("XTBAMD_EIG")) use_ql = std::string(r) == "ql";
  // padded working copies on the register path; banded rounds only follow a
  // warm-start refinement (oa_max_steps > 0), which keeps the packed layout
  jac_padded = jac_p == 32 && !apply_lds && !pivot_lds && oa_max_steps == 0 && !use_ql;
  if (const char* r = std::getenv("XTBAMD_JAC_PAD")) jac_padded = jac_padded && std::atoi(r) != 0;
  intcut = std::max(20.0, 25.0 - 10.0 * std::log10(opt.accuracy));
  neglect = 10.0e-9 * opt.accuracy;
  scfconv = 1.e-6 * opt.accuracy;
  qconv = 1.e-4 * opt.accuracy;
}

void Batch::allocate(const Capacity& c) {
  cap_ = c;
  nmol = c.slots;
  for (int k = 0; k < 13; ++k) ar_[k].init(*cap_field(cap_, k));
  const long long tn = std::max(1LL, c.n), tsh = std::max(1LL, c.nsh), tao = std::max(1LL, c.nao);
  const long long tmm = std::max(1LL, c.mm);
  const int mx = opt.maxiter;
  desc.assign(nmol, MolDesc());
  for (int m = 0; m < nmol; ++m) {
    std::memset(&desc[m], 0, sizeof(MolDesc));
    desc[m].mol = m;
  }
  occupied_.assign(nmol, 0);
  hmol_.assign(nmol, HostMol());
  elem_present_.assign(XTB_MAX_Z + 1, 0);
  h_state0_.assign(nmol, MolState());
  for (auto& st : h_state0_) std::memset(&st, 0, sizeof(st));
  d_desc = alloc<MolDesc>(nmol);
  d_state = alloc<MolState>(nmol);
  at = alloc<int>(tn);
  xyz = alloc<double>(3 * tn);
  cn = alloc<double>(tn); cn_d4 = alloc<double>(tn); cn_eeq = alloc<double>(tn); cn_cov = alloc<double>(tn);
  q = alloc<double>(tn); dipm = alloc<double>(3 * tn); qp = alloc<double>(6 * tn); radcn = alloc<double>(tn);
  vs = alloc<double>(tn); vd = alloc<double>(3 * tn); vq = alloc<double>(6 * tn); ashift = alloc<double>(tn);
  grad = alloc<double>(3 * tn); dhdcn = alloc<double>(tn); dedcn = alloc<double>(tn);
  grad_atm = alloc<double>(3 * tn); dedcn_atm = alloc<double>(tn);
  tmp2 = alloc<double>(tn); eat = alloc<double>(tn); ae01 = alloc<double>(2 * tn);
  sh_at = alloc<int>(tsh); sh_l = alloc<int>(tsh); sh_ao0 = alloc<int>(tsh); sh_m = alloc<int>(tsh);
  at_sh0 = alloc<int>(tn); at_ao0 = alloc<int>(tn);
  qsh = alloc<double>(tsh); zsh = alloc<double>(tsh); shift = alloc<double>(tsh);
  shgam = alloc<double>(tsh); selfe = alloc<double>(tsh); dsedcn = alloc<double>(tsh);
  jq = alloc<double>(tsh);
  ao_at = alloc<int>(tao); ao_sh = alloc<int>(tao);
  emo = alloc<double>(tao); focc = alloc<double>(tao); focca = alloc<double>(tao);
  foccb = alloc<double>(tao); ev = alloc<double>(tao);
  row_q = alloc<double>(tao); row_e = alloc<double>(tao); row_d = alloc<double>(3 * tao);
  row_qp = alloc<double>(6 * tao); row_pd = alloc<double>(tao);
  d4_at = alloc<int>(std::max(1LL, c.nd4)); d4_ref = alloc<int>(std::max(1LL, c.nd4));
  gw = alloc<double>(std::max(1LL, c.nd4)); zv = alloc<double>(std::max(1LL, c.nd4));
  dzv = alloc<double>(std::max(1LL, c.nd4)); d4dum = alloc<double>(std::max(1LL, c.nd4));
  c6ref = alloc<double>((size_t)(XTB_MAX_Z + 1) * (XTB_MAX_Z + 1) * 49);
  HIPCHECK(hipMemset(c6ref, 0, (size_t)(XTB_MAX_Z + 1) * (XTB_MAX_Z + 1) * 49 * sizeof(double)));
  zetavec = alloc<double>(XTB_MAX_REF * tn); zerovec = alloc<double>(XTB_MAX_REF * tn);
  zetadcn = alloc<double>(XTB_MAX_REF * tn); zerodcn = alloc<double>(XTB_MAX_REF * tn);
  S = alloc<double>(tmm); H0 = alloc<double>(tmm); H0n = alloc<double>(tmm);
  D = alloc<double>(3 * tmm); Q = alloc<double>(6 * tmm);
  H = alloc<double>(tmm); C = alloc<double>(tmm); T = alloc<double>(tmm);
  A = alloc<double>(tmm); W = alloc<double>(tmm); P = alloc<double>(tmm);
  J = alloc<double>(std::max(1LL, c.jj)); dispmat = alloc<double>(std::max(1LL, c.xx));
  const long long tpp = std::max(1LL, c.pp);
  gab3 = alloc<double>(tpp); gab5 = alloc<double>(tpp);
  c6 = alloc<double>(tpp); dc6 = alloc<double>(tpp); c60 = alloc<double>(tpp); dc60 = alloc<double>(tpp);
  wbo = alloc<double>(tpp);
  gpair = alloc<double>(5 * tpp);
  pair_list = alloc<int2>(tpp);
  E = alloc<double>(std::max(1LL, c.ee));
  xE = alloc<double>(std::max(1LL, c.ne)); yE = alloc<double>(std::max(1LL, c.ne));
  const long long tv = std::max(1LL, c.v);
  q_in = alloc<double>(tv); dq = alloc<double>(tv); qlast = alloc<double>(tv); dqlast = alloc<double>(tv);
  bdf = alloc<double>(std::max(1LL, c.h)); bu = alloc<double>(std::max(1LL, c.h)); ba = alloc<double>(std::max(1LL, c.g));
  bomega = alloc<double>((long long)nmol * mx);
  erep = alloc<double>(tn);
  perm = alloc<int>(tao);
  Sm = alloc<double>(tmm);
  G = alloc<double>(tmm);
  jac_nactive = alloc<int>(kMaxLanes);
  d_state_stage = alloc<MolState>(nmol);
  eeq_failed = alloc<int>(nmol);
  eeq_big = alloc<int>(nmol);
  eig_failed = alloc<int>(nmol);
  HIPCHECK(hipMemset(eig_failed, 0, nmol * sizeof(int)));
  d_sub = alloc<MolDesc>(nmol);
  phase_act = alloc<int>(nmol);
  open_act = alloc<int>(nmol);
  ns_c = alloc<double>(nmol);
  oa_act = alloc<int>(nmol);
  oa_ok = alloc<int>(nmol);
  oa_stat = alloc<unsigned long long>(4LL * nmol);
  ns_res = alloc<unsigned long long>(nmol);
  ns_stage = alloc<int>(nmol);
  ns_fin = alloc<int>(nmol);
  ns_act = alloc<int>(nmol);
  ns_failed = alloc<int>(nmol);
  // Jacobi set: per-slot problem table and state; work lists grow on demand
  JacSet& js = jac_n;
  js.probs.assign(nmol, JacobiProb());
  js.d_probs = alloc<JacobiProb>(nmol);
  js.offmax = alloc<unsigned long long>(2LL * nmol);   // two sweep-parity buffers
  js.state = alloc<int>((long long)nmol * kJacState);
  js.scale = alloc<double>(nmol);
  // pivot-factor buffer: at most (nao/64 + 1) pairs per molecule
  js.qsize = std::max(4096LL, (c.nao / 64 + 2LL * nmol) * 4096LL);
  js.qbuf = alloc<double>(js.qsize);
  js.rot = alloc<int>(js.qsize / 1024 + 1);
  js.perm = nullptr;
  Element* d_el = alloc<Element>(XTB_MAX_Z + 1);
  HIPCHECK(hipMemcpy(d_el, host_elements(), (XTB_MAX_Z + 1) * sizeof(Element), hipMemcpyHostToDevice));
  elements_dev = d_el;
  // every slot starts empty and inactive
  HIPCHECK(hipMemset(d_state, 0, nmol * sizeof(MolState)));
  if (on_loan()) h_pinned = loan_.host_words;
  else HIPCHECK(hipHostMalloc((void**)&h_pinned, 64 * sizeof(int), hipHostMallocDefault));
}

// reference C6 of all element pairs present (D4 reference tables)
void Batch::update_c6ref() {
  const Element* el = host_elements();
  std::vector<double> tab((size_t)(XTB_MAX_Z + 1) * (XTB_MAX_Z + 1) * 49, 0.0);
  for (int za = 1; za <= XTB_MAX_Z; ++za) {
    if (!elem_present_[za]) continue;
    for (int zb = 1; zb <= XTB_MAX_Z; ++zb) {
      if (!elem_present_[zb]) continue;
      for (int r = 0; r < el[za].nref; ++r)
        for (int s2 = 0; s2 < el[zb].nref; ++s2)
          tab[((size_t)za * (XTB_MAX_Z + 1) + zb) * 49 + r * 7 + s2] = host_ref_c6(za, r, zb, s2);
    }
  }
  HIPCHECK(hipMemcpy(c6ref, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
}

// slot m takes the molecule: offsets from the packed-array allocators
void Batch::place(int m, const MolDesc& d0, const HostMol& hm, const MolState& st, bool upload) {
  MolDesc d = d0;
  Capacity f = footprint(d);
  long long off[13];
  for (int k = 0; k < 13; ++k) {
    off[k] = ar_[k].take(*cap_field(f, k));
    if (off[k] < 0) {
      for (int j = 0; j < k; ++j) ar_[j].give(off[j], *cap_field(f, j));
      throw std::runtime_error("no room");
    }
  }
  d.a0 = (int)off[0];
  d.s0 = (int)off[1];
  d.o0 = (int)off[2];
  d.d0 = (int)off[3];
  d.m0 = off[4];
  d.j0 = off[5];
  d.x0 = off[6];
  d.p0 = off[7];
  d.e0 = off[8];
  d.y0 = off[9];
  d.v0 = off[10];
  d.h0 = off[11];
  d.g0 = off[12];
  d.mol = m;
  desc[m] = d;
  hmol_[m] = hm;
  h_state0_[m] = st;
  occupied_[m] = 1;
  if (upload) {
    auto up_i = [&](int* base, long long o, const std::vector<int>& v) {
      if (!v.empty())
        HIPCHECK(hipMemcpyAsync(base + o, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    };
    up_i(at, d.a0, hm.at); up_i(at_sh0, d.a0, hm.at_sh0); up_i(at_ao0, d.a0, hm.at_ao0);
    up_i(sh_at, d.s0, hm.sh_at); up_i(sh_l, d.s0, hm.sh_l); up_i(sh_ao0, d.s0, hm.sh_ao0); up_i(sh_m, d.s0, hm.sh_m);
    up_i(ao_at, d.o0, hm.ao_at); up_i(ao_sh, d.o0, hm.ao_sh);
    up_i(d4_at, d.d0, hm.d4_at); up_i(d4_ref, d.d0, hm.d4_ref);
    HIPCHECK(hipMemcpyAsync(xyz + 3LL * d.a0, hm.xyz.data(), hm.xyz.size() * sizeof(double), hipMemcpyHostToDevice,
                            stream));
    if (!hm.pairs.empty())
      HIPCHECK(hipMemcpyAsync(pair_list + d.p0, hm.pairs.data(), hm.pairs.size() * sizeof(int2), hipMemcpyHostToDevice,
                              stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  for (int z : hm.at) elem_present_[z] = 1;
}

int Batch::admit(const MoleculeInput& mi) {
  MolDesc d;
  HostMol hm;
  MolState st;
  describe(mi, opt.maxiter, d, hm, st);
  int slot = -1;
  for (int m = 0; m < nmol; ++m)
    if (!occupied_[m]) { slot = m; break; }
  if (slot < 0) return -1;
  // all packed arrays must have room
  Capacity f = footprint(d);
  for (int k = 0; k < 13; ++k) {
    const long long need = *cap_field(f, k);
    bool fits = need <= 0;
    for (auto& r : ar_[k].free_)
      if (r.second - r.first >= need) { fits = true; break; }
    if (!fits) return -1;
  }
  bool newel = false;
  for (int z : hm.at) newel = newel || !elem_present_[z];
  place(slot, d, hm, st, true);
  if (newel) update_c6ref();
  if (!guess_qsh_.empty()) {
    guess_qsh_[slot].clear();
    guess_dipm_[slot].clear();
    guess_qp_[slot].clear();
  }
  return slot;
}

void Batch::release(int m) {
  if (m < 0 || m >= nmol || !occupied_[m]) return;
  Capacity f = footprint(desc[m]);
  const MolDesc& d = desc[m];
  const long long off[13] = {d.a0, d.s0, d.o0, d.d0, d.m0, d.j0, d.x0, d.p0, d.e0, d.y0, d.v0, d.h0, d.g0};
  for (int k = 0; k < 13; ++k) ar_[k].give(off[k], *cap_field(f, k));
  std::memset(&desc[m], 0, sizeof(MolDesc));
  desc[m].mol = m;
  occupied_[m] = 0;
  hmol_[m] = HostMol();
  MolState z;
  std::memset(&z, 0, sizeof(z));
  h_state0_[m] = z;
  HIPCHECK(hipMemcpyAsync(d_state + m, &h_state0_[m], sizeof(MolState), hipMemcpyHostToDevice, stream));
  HIPCHECK(hipStreamSynchronize(stream));
}

// work lists, grid bounds and descriptor table of the occupied slots
void Batch::rebuild() {
  max_n = max_nsh = max_nao = max_nd4 = max_nbr = 0;
  for (int m = 0; m < nmol; ++m) {
    if (!occupied_[m]) continue;
    const MolDesc& d = desc[m];
    max_n = std::max(max_n, d.n);
    max_nsh = std::max(max_nsh, d.nsh);
    max_nao = std::max(max_nao, d.nao);
    max_nd4 = std::max(max_nd4, d.nd4);
    max_nbr = std::max(max_nbr, d.nbr);
  }
  HIPCHECK(hipMemcpyAsync(d_desc, desc.data(), nmol * sizeof(MolDesc), hipMemcpyHostToDevice, stream));
  auto grow = [&](auto*& ptr, long long& capv, long long need) {
    using T_ = typename std::remove_reference<decltype(*ptr)>::type;
    if (need <= capv && ptr) return;
    if (ptr) {
      HIPCHECK(hipStreamSynchronize(stream));
      auto it = std::find(allocs_.begin(), allocs_.end(), (void*)ptr);
      if (it != allocs_.end()) allocs_.erase(it);
      if (!on_loan()) HIPCHECK(hipFree(ptr));   // a loan's pieces are never released
    }
    capv = std::max<long long>(need, 2 * capv);
    ptr = alloc<T_>(capv);
  };
  // GEMM tiles over (molecule, 64x64 tile)
  std::vector<int4> tl, ts;
  for (int m = 0; m < nmol; ++m) {
    if (!occupied_[m]) continue;
    const int nt = (desc[m].nao + 63) / 64;
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        tl.push_back(make_int4(m, a, b, 0));
        if (b >= a) ts.push_back(make_int4(m, a, b, 1));
      }
  }
  ntiles = (int)tl.size();
  ntiles_sym = (int)ts.size();
  tl.insert(tl.end(), ts.begin(), ts.end());
  grow(d_tiles, tiles_cap_, std::max<long long>(1, ntiles + ntiles_sym));
  if (ntiles)
    HIPCHECK(hipMemcpyAsync(d_tiles, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice, stream));
  // Jacobi problems: dim nao, matrices at m0, eigenvalues at o0
  JacSet& js = jac_n;
  long long qoff = 0;
  std::vector<int> small, big, blk;
  std::vector<int2> pairs;
  std::vector<int4> ta, tw, to;
  // the blocked eigensolver takes the largest problems on the register path
  const bool blk_path = jac_padded && jac_p == 32 && !use_ql;
  js.max_rounds = 0;
  js.max_rounds_lanes = 0;
  js.big_max_dim = 0;
  js.big_max_ld = 0;
  long long poff = 0;
  for (int m = 0; m < nmol; ++m) {
    JacobiProb& p = js.probs[m];
    std::memset(&p, 0, sizeof(p));
    if (!occupied_[m]) continue;
    p.dim = desc[m].nao;
    p.nb = (p.dim + jac_p / 2 - 1) / (jac_p / 2);
    p.nbp = p.nb + (p.nb & 1);
    p.rounds = std::max(1, p.nbp - 1);
    p.src_off = desc[m].m0;
    p.ld = jac_padded ? (p.dim + 15) / 16 * 16 : p.dim;
    p.a_off = jac_padded ? poff : desc[m].m0;
    p.w_off = p.a_off;
    if (jac_padded) poff += (long long)p.ld * p.ld;
    p.ev_off = desc[m].o0;
    p.q_off = qoff;
    p.maxsw = jac_max_sweeps;
    js.big_max_ld = std::max(js.big_max_ld, p.ld);
    if (p.dim <= 64) {
      small.push_back(m);
      continue;
    }
    // the largest problems also go to the blocked eigensolver, which takes
    // them for the first eig_blk_iters SCC iterations (the lanes after)
    const bool to_blk = blk_path && p.dim >= eig_blk_min;
    if (to_blk) blk.push_back(m);
    big.push_back(m);
    js.big_max_dim = std::max(js.big_max_dim, p.dim);
    qoff += (long long)(p.nbp / 2) * jac_p * jac_p;
    js.max_rounds = std::max(js.max_rounds, p.rounds);
    if (!to_blk) js.max_rounds_lanes = std::max(js.max_rounds_lanes, p.rounds);
  }
  // lanes of equal tile-update bytes per round (dim^2): largest problems
  // first, each to the lane with least so far
  {
    std::vector<int> order(big);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return js.probs[x].dim > js.probs[y].dim; });
    double load[kMaxLanes] = {};
    for (int m : order) {
      int g = 0;
      for (int h = 1; h < jac_lanes; ++h)
        if (load[h] < load[g]) g = h;
      js.probs[m].group = g;
      load[g] += (double)js.probs[m].dim * js.probs[m].dim;
    }
  }
  for (int g = 0; g < kMaxLanes; ++g) {
    js.gpairs[g] = (int)pairs.size();
    js.gta[g] = (int)ta.size();
    js.gtw[g] = (int)tw.size();
    js.gto[g] = (int)to.size();
    for (int m : big) {
      const JacobiProb& p = js.probs[m];
      if (p.group != g) continue;
      const int np = p.nbp / 2;
      for (int k = 0; k < np; ++k) pairs.push_back(make_int2(m, k));
      for (int k = 0; k < np; ++k)
        for (int l = k; l < np; ++l) ta.push_back(make_int4(m, k, l, 0));
      const int nwb = (p.dim + jac_p - 1) / jac_p;
      for (int rb = 0; rb < nwb; ++rb)
        for (int l = 0; l < np; ++l) tw.push_back(make_int4(m, rb, l, 0));
      const int nrb = (p.dim + 63) / 64;
      for (int a = 0; a < nrb; ++a)
        for (int b = a; b < nrb; ++b) to.push_back(make_int4(m, a, b, 0));
    }
  }
  js.gpairs[kMaxLanes] = (int)pairs.size();
  js.gta[kMaxLanes] = (int)ta.size();
  js.gtw[kMaxLanes] = (int)tw.size();
  js.gto[kMaxLanes] = (int)to.size();
  // XCD-aware round lists of the lanes: workgroup b of a launch runs on XCD
  // b mod kXcds, so a problem placed on XCD x has its pivots and its A and W
  // tiles at positions = x (mod kXcds) and its pivot factors Q are fetched by
  // one XCD's L2 only.  Problems go to the least loaded XCD (tile count),
  // largest first; short rows are padded with skip entries (m = -1).
  std::vector<int2> xp;
  std::vector<int4> xt;
  for (int g = 0; g < kMaxLanes; ++g) {
    js.gxp[g] = (int)xp.size();
    js.gxt[g] = (int)xt.size();
    std::vector<int> pl;
    for (int m : big)
      if (js.probs[m].group == g) pl.push_back(m);
    auto ntile = [&](int m) {
      const JacobiProb& p = js.probs[m];
      const long long np = p.nbp / 2, nwb = (p.dim + jac_p - 1) / jac_p;
      return np * (np + 1) / 2 + nwb * np;
    };
    std::stable_sort(pl.begin(), pl.end(), [&](int x, int y) { return ntile(x) > ntile(y); });
    std::vector<std::vector<int2>> qp(kXcds);
    std::vector<std::vector<int4>> qt(kXcds);
    std::vector<long long> xl(kXcds, 0);
    long long lane_tiles = 0;
    for (int m : pl) lane_tiles += ntile(m);
    for (int m : pl) {
      const JacobiProb& p = js.probs[m];
      const int np = p.nbp / 2, nwb = (p.dim + jac_p - 1) / jac_p;
      // a problem with more than half an XCD's share of the lane's tiles is
      // spread over all XCDs (pivot k and the tiles of pair l on XCD
      // k, l mod kXcds), the others stay on one XCD each
      const bool spread = 2 * kXcds * ntile(m) > lane_tiles;
      const int x = (int)(std::min_element(xl.begin(), xl.end()) - xl.begin());
      auto home = [&](int l) { return spread ? l % kXcds : x; };
      for (int k = 0; k < np; ++k) qp[home(k)].push_back(make_int2(m, k));
      for (int k = 0; k < np; ++k)
        for (int l = k; l < np; ++l) {
          qt[home(l)].push_back(make_int4(m, k, l, 0));
          xl[home(l)] += 1;
        }
      for (int rb = 0; rb < nwb; ++rb)
        for (int l = 0; l < np; ++l) {
          qt[home(l)].push_back(make_int4(m, rb, l, 1));
          xl[home(l)] += 1;
        }
    }
    size_t mp = 0, mt = 0;
    for (int x = 0; x < kXcds; ++x) {
      mp = std::max(mp, qp[x].size());
      mt = std::max(mt, qt[x].size());
    }
    for (size_t j = 0; j < mp; ++j)
      for (int x = 0; x < kXcds; ++x) xp.push_back(j < qp[x].size() ? qp[x][j] : make_int2(-1, 0));
    for (size_t j = 0; j < mt; ++j)
      for (int x = 0; x < kXcds; ++x) xt.push_back(j < qt[x].size() ? qt[x][j] : make_int4(-1, 0, 0, 2));
  }
  js.gxp[kMaxLanes] = (int)xp.size();
  js.gxt[kMaxLanes] = (int)xt.size();
  grow(js.d_xp, xp_cap_, std::max<long long>(1, (long long)xp.size()));
  grow(js.d_xt, xt_cap_, std::max<long long>(1, (long long)xt.size()));
  if (!xp.empty()) HIPCHECK(hipMemcpyAsync(js.d_xp, xp.data(), xp.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
  if (!xt.empty()) HIPCHECK(hipMemcpyAsync(js.d_xt, xt.data(), xt.size() * sizeof(int4), hipMemcpyHostToDevice, stream));
  if (qoff > js.qsize) throw std::runtime_error("Jacobi pivot buffer too small");
  if (jac_padded) {
    long long c0 = js.jcap;
    grow(js.ja, c0, std::max<long long>(1, poff));
    grow(js.jw, js.jcap, std::max<long long>(1, poff));
  }
  HIPCHECK(hipMemcpyAsync(js.d_probs, js.probs.data(), nmol * sizeof(JacobiProb), hipMemcpyHostToDevice, stream));
  js.nsmall = (int)small.size();
  js.npairs = (int)pairs.size();
  js.nta = (int)ta.size();
  js.ntw = (int)tw.size();
  js.nto = (int)to.size();
  grow(js.d_small, small_cap_, std::max<long long>(1, js.nsmall));
  grow(js.d_pairs, pairs_cap_, std::max<long long>(1, js.npairs));
  grow(js.d_ta, ta_cap_, std::max<long long>(1, js.nta));
  grow(js.d_tw, tw_cap_, std::max<long long>(1, js.ntw));
  grow(js.d_to, to_cap_, std::max<long long>(1, js.nto));
  eig_.mols = blk;
  eig_build();
  js.nbig = (int)big.size();
  grow(js.d_big, big_cap_, std::max<long long>(1, js.nbig));
  if (!big.empty()) HIPCHECK(hipMemcpyAsync(js.d_big, big.data(), big.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  if (!small.empty())
    HIPCHECK(hipMemcpyAsync(js.d_small, small.data(), small.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  if (!pairs.empty())
    HIPCHECK(hipMemcpyAsync(js.d_pairs, pairs.data(), pairs.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
  if (!ta.empty()) HIPCHECK(hipMemcpyAsync(js.d_ta, ta.data(), ta.size() * sizeof(int4), hipMemcpyHostToDevice, stream));
  if (!tw.empty()) HIPCHECK(hipMemcpyAsync(js.d_tw, tw.data(), tw.size() * sizeof(int4), hipMemcpyHostToDevice, stream));
  if (!to.empty()) HIPCHECK(hipMemcpyAsync(js.d_to, to.data(), to.size() * sizeof(int4), hipMemcpyHostToDevice, stream));
  HIPCHECK(hipStreamSynchronize(stream));
}

Batch::Batch(const std::vector<MoleculeInput>& mols, const Options& o, hipStream_t s) : opt(o), stream(s) {
  init_common();
  build_fixed(mols);
}

Batch::Batch(const std::vector<MoleculeInput>& mols, const Options& o, const Loan& loan)
    : opt(o), stream(loan.streams[0]), loan_(loan) {
  if (!loan.base || !loan.host_words) throw std::runtime_error("invalid loan");
  init_common();
  build_fixed(mols);
}

void Batch::build_fixed(const std::vector<MoleculeInput>& mols) {
  if (mols.empty()) throw std::runtime_error("empty batch");
  const int nm = (int)mols.size();
  std::vector<MolDesc> ds(nm);
  std::vector<HostMol> hms(nm);
  std::vector<MolState> sts(nm);
  Capacity tot;
  tot.slots = nm;
  for (int m = 0; m < nm; ++m) {
    describe(mols[m], opt.maxiter, ds[m], hms[m], sts[m]);
    Capacity f = footprint(ds[m]);
    for (int k = 0; k < 13; ++k) *cap_field(tot, k) += *cap_field(f, k);
  }
  allocate(tot);
  for (int m = 0; m < nm; ++m) place(m, ds[m], hms[m], sts[m], false);
  // bulk upload: the packed layout is the concatenation of the molecules
  auto cat_i = [&](std::vector<int> HostMol::*f) {
    std::vector<int> v;
    for (const HostMol& h : hms) v.insert(v.end(), (h.*f).begin(), (h.*f).end());
    return v;
  };
  auto up_i = [&](int* d_, const std::vector<int>& h) {
    if (!h.empty()) HIPCHECK(hipMemcpy(d_, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
  };
  up_i(at, cat_i(&HostMol::at)); up_i(at_sh0, cat_i(&HostMol::at_sh0)); up_i(at_ao0, cat_i(&HostMol::at_ao0));
  up_i(sh_at, cat_i(&HostMol::sh_at)); up_i(sh_l, cat_i(&HostMol::sh_l)); up_i(sh_ao0, cat_i(&HostMol::sh_ao0));
  up_i(sh_m, cat_i(&HostMol::sh_m)); up_i(ao_at, cat_i(&HostMol::ao_at)); up_i(ao_sh, cat_i(&HostMol::ao_sh));
  up_i(d4_at, cat_i(&HostMol::d4_at)); up_i(d4_ref, cat_i(&HostMol::d4_ref));
  std::vector<double> xall;
  for (const HostMol& h : hms) xall.insert(xall.end(), h.xyz.begin(), h.xyz.end());
  HIPCHECK(hipMemcpy(xyz, xall.data(), xall.size() * sizeof(double), hipMemcpyHostToDevice));
  {
    std::vector<int2> pall((size_t)std::max(1LL, cap_.pp), make_int2(0, 0));
    for (int m = 0; m < nm; ++m) std::copy(hms[m].pairs.begin(), hms[m].pairs.end(), pall.begin() + desc[m].p0);
    HIPCHECK(hipMemcpy(pair_list, pall.data(), pall.size() * sizeof(int2), hipMemcpyHostToDevice));
  }
  update_c6ref();
  rebuild();
}

Batch::Batch(const Capacity& cap, const Options& o, hipStream_t s) : opt(o), stream(s) {
  init_common();
  if (cap.slots <= 0) throw std::runtime_error("empty batch");
  allocate(cap);
  rebuild();
}

void Batch::set_solvent(const std::string& name, int kernel, bool alpb, int state, double temperature) {
  if (name.empty()) {
    solv_on_ = false;
    setup_done_ = false;
    return;
  }
  if (kernel != 1 && kernel != 2) throw std::runtime_error("unknown Born kernel");
  const SolvParamSet* ps = solv_param_set(name, kernel);
  if (!ps) throw std::runtime_error("solvent: '" + name + "' is not parametrized");
  for (int z = 1; z <= XTB_MAX_Z; ++z)
    if (elem_present_[z] && kVdwRadD3Aa[z - 1] <= 0.0) throw std::runtime_error("element without solvation radius");
  const SolvDev model = solv_model(*ps, kernel, alpb, state, temperature);
  if (!d_solv) {
    const long long tn = std::max(1LL, cap_.n), tpp = std::max(1LL, cap_.pp);
    d_solv = alloc<SolvDev>(1);
    sv_brad = alloc<double>(tn); sv_dpsi = alloc<double>(tn); sv_sasa = alloc<double>(tn);
    sv_hbw = alloc<double>(tn); sv_pot = alloc<double>(tn); sv_grddb = alloc<double>(tn);
    sv_w = alloc<double>(tn * kSolvGrid);
    sv_B = alloc<double>(tpp);
    sv_gsasa = alloc<double>(nmol); sv_adet = alloc<double>(nmol); sv_qbq = alloc<double>(nmol);
    sv_failed = alloc<int>(nmol);
    HIPCHECK(hipMemset(sv_failed, 0, nmol * sizeof(int)));
    HIPCHECK(hipMemset(sv_qbq, 0, nmol * sizeof(double)));
    HIPCHECK(hipMemset(sv_gsasa, 0, nmol * sizeof(double)));
  }
  HIPCHECK(hipMemcpy(d_solv, &model, sizeof(SolvDev), hipMemcpyHostToDevice));
  sv_gshift = model.gshift;
  solv_on_ = true;
  setup_done_ = false;
}

Batch::~Batch() {
  for (auto& v : prof_ev_)
    for (hipEvent_t e : v) (void)hipEventDestroy(e);
  if (eig_.graph) (void)hipGraphExecDestroy(eig_.graph);
  if (on_loan()) {
    // the lender owns the memory, streams and events; only wait for the work
    (void)hipStreamSynchronize(stream);
    (void)hipStreamSynchronize(stream2);
    (void)hipStreamSynchronize(stream3);
    (void)hipStreamSynchronize(stream4);
    return;
  }
  for (void* p : allocs_) (void)hipFree(p);
  if (h_pinned) (void)hipHostFree(h_pinned);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (ev_join3) (void)hipEventDestroy(ev_join3);
  if (ev_join4) (void)hipEventDestroy(ev_join4);
  if (ev_atm) (void)hipEventDestroy(ev_atm);
  if (stream2) (void)hipStreamDestroy(stream2);
  if (stream3) (void)hipStreamDestroy(stream3);
  if (stream4) (void)hipStreamDestroy(stream4);
}

// ---------------------------------------------------------------------------
bool Batch::jacobi(JacSet& js, double* Am, double* Wm, double* evout, bool check_active, bool keep_w, int banded) {
  const int nb = nmol;
  const bool rows = jac_p == 32 && !apply_lds && !keep_w && banded == 0;
  const MolState* act = check_active ? (const MolState*)d_state : (const MolState*)nullptr;
  double* Wout = Wm;
  if (jac_padded) {
    if (!rows || pivot_lds) throw std::runtime_error("padded Jacobi problems need the register path");
    // A into the padded working copy; the eigenvectors come back to Wm at the end
    hipLaunchKernelGGL(k_jac_pad, dim3(gstride((long long)js.big_max_ld * js.big_max_ld), nb), dim3(256), 0, stream,
                       js.d_probs, (const double*)Am, js.ja, act, 1);
    Am = js.ja;
    Wm = js.jw;
  }
  // W = I over the whole device (k_jac_init only sets the scales and states)
  hipLaunchKernelGGL(k_jac_init, dim3(nb), dim3(256), 0, stream, js.d_probs, (const double*)Am, Wm, js.scale, js.state,
                     check_active ? (const MolState*)d_state : (const MolState*)nullptr, js.offmax, 1,
                     jac_tol_mid / jac_tol, jac_tol_q, jac_tol);
  if (!keep_w)
    hipLaunchKernelGGL(k_jac_ident, dim3(gstride((long long)js.big_max_ld * js.big_max_ld), nb), dim3(256), 0, stream,
                       js.d_probs, Wm, check_active ? (const MolState*)d_state : (const MolState*)nullptr);
  // lanes: the problems of lanes 1 and 2 run their rounds on stream2 and
  // stream4 beside those of lane 0; the problems of a lane never touch another
  // lane's data, so the results do not depend on the lanes
  const bool lanes = rows && !pivot_lds && js.npairs > 0;
  // the largest problems go to the blocked eigensolver on stream3 (after the
  // small ones), beside the lanes
  const bool blk = rows && !pivot_lds && eig_.np > 0 && iterations_ < eig_blk_iters;
  // the blocked path's problems are marked done for the lanes
  if (blk)
    hipLaunchKernelGGL(k_eig_claim, dim3(1), dim3(256), 0, stream, (const EigProb*)eig_.d_probs, eig_.np, js.state);
  // the lanes test their progress once per period of their largest problem
  jac_period_ = blk ? std::max(1, js.max_rounds_lanes) : js.max_rounds;
  const bool forked = lanes || blk;
  // the problems of dimension <= 64 (one workgroup each, k_jac_small) run
  // on stream3 beside the lanes
  if (forked) {
    HIPCHECK(hipEventRecord(ev_fork, stream));
    HIPCHECK(hipStreamWaitEvent(stream2, ev_fork, 0));
    HIPCHECK(hipStreamWaitEvent(stream3, ev_fork, 0));
    HIPCHECK(hipStreamWaitEvent(stream4, ev_fork, 0));
  }
  if (js.nsmall)
    hipLaunchKernelGGL(k_jac_small, dim3(js.nsmall), dim3(256), 0, forked ? stream3 : stream, js.d_probs, js.d_small,
                       Am, Wm, (const double*)js.scale, js.state, keep_w ? 1 : 0, rows ? 1 : 0, jac_tol);
  if (blk) eig_blocked(js, Am, Wm, check_active, stream3);
  if (forked) {
    if (lanes) jacobi_lanes(js, Am, Wm);
    HIPCHECK(hipEventRecord(ev_join, stream2));
    HIPCHECK(hipStreamWaitEvent(stream, ev_join, 0));
    HIPCHECK(hipEventRecord(ev_join3, stream3));
    HIPCHECK(hipStreamWaitEvent(stream, ev_join3, 0));
    HIPCHECK(hipEventRecord(ev_join4, stream4));
    HIPCHECK(hipStreamWaitEvent(stream, ev_join4, 0));
  } else if (js.npairs) {
    // banded rounds first (neighbouring blocks only), then cyclic sweeps;
    // every round starts with the convergence test of the problems at a
    // sweep boundary; the host reads the number still running once per
    // max_rounds rounds
    std::vector<int> sched;
    for (int b = 0; b < 2 * banded; ++b) sched.push_back(kBanded + b);
    const int rmax = (jac_max_sweeps + 1) * js.max_rounds;
    for (int r = 0; r <= rmax; ++r) sched.push_back(r);
    for (size_t si = 0; si < sched.size(); ++si) {
      const int r = sched[si];
      hipLaunchKernelGGL(k_jac_offmax, dim3(js.nto), dim3(256), 0, stream, js.d_probs, js.d_to, (const double*)Am,
                         (const int*)js.state, js.offmax, r);
      hipLaunchKernelGGL(k_jac_mark, dim3(1), dim3(256), 0, stream, js.d_probs, nb, (const double*)js.scale,
                         js.state, js.offmax, r, jac_nactive, jac_tol, -1);
      if (r < kBanded && r % js.max_rounds == 0) {
        HIPCHECK(hipMemcpyAsync(h_pinned, jac_nactive, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        if (h_pinned[0] <= 0) break;
      }
      unsigned long long* jw = (prof_mask & 4) ? prof_work_ + 2 : (unsigned long long*)nullptr;
      if (jac_p == 32) {
        prof_begin(2);
        if (pivot_lds)
          hipLaunchKernelGGL((k_jac_pairsP<32, 64>), dim3(js.npairs), dim3(64), 0, stream, js.d_probs, js.d_pairs, r,
                             (const double*)Am, js.qbuf, js.rot, (const double*)js.scale, (const int*)js.state,
                             inner_sweeps, jw);
        else
          hipLaunchKernelGGL(k_jac_pairsR<false>, dim3(js.npairs), dim3(64), 0, stream, js.d_probs, js.d_pairs, r,
                             (const double*)Am, js.qbuf, js.rot, (const double*)js.scale, js.state, inner_sweeps, jw,
                             jac_tol, (unsigned long long*)nullptr, 0, (unsigned long long*)nullptr);
        prof_end(2, nullptr, 0);
        if (rows) {
          hipLaunchKernelGGL(k_jac_apply_aR, dim3(js.nta), dim3(64), 0, stream, js.d_probs, js.d_ta, r, Am,
                             (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
          hipLaunchKernelGGL(k_jac_apply_wR, dim3(js.ntw), dim3(64), 0, stream, js.d_probs, js.d_tw, r, Wm,
                             (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
        } else {
          hipLaunchKernelGGL((k_jac_apply_aP<32>), dim3(js.nta), dim3(64), 0, stream, js.d_probs, js.d_ta, r, Am,
                             (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
          hipLaunchKernelGGL((k_jac_apply_wP<32>), dim3(js.ntw), dim3(64), 0, stream, js.d_probs, js.d_tw, r, Wm,
                             (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
        }
        jacobi_rounds_ += 1;
        continue;
      }
      prof_begin(2);
      hipLaunchKernelGGL(k_jac_pairs, dim3(js.npairs), dim3(256), 0, stream, js.d_probs, js.d_pairs, r,
                         (const double*)Am, js.qbuf, js.rot, (const double*)js.scale, (const int*)js.state,
                         inner_sweeps, jw);
      prof_end(2, nullptr, 0);
      if (use_mfma) {
        hipLaunchKernelGGL(k_jac_apply_a_mfma, dim3(js.nta), dim3(256), 0, stream, js.d_probs, js.d_ta, r, Am,
                           (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
        hipLaunchKernelGGL(k_jac_apply_w_mfma, dim3(js.ntw), dim3(256), 0, stream, js.d_probs, js.d_tw, r, Wm,
                           (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
      } else {
        hipLaunchKernelGGL(k_jac_apply_a, dim3(js.nta), dim3(256), 0, stream, js.d_probs, js.d_ta, r, Am,
                           (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
        hipLaunchKernelGGL(k_jac_apply_w, dim3(js.ntw), dim3(256), 0, stream, js.d_probs, js.d_tw, r, Wm,
                           (const double*)js.qbuf, (const int*)js.rot, (const int*)js.state);
      }
      jacobi_rounds_ += 1;
    }
  }
  if (jac_padded)
    hipLaunchKernelGGL(k_jac_pad, dim3(gstride((long long)js.big_max_ld * js.big_max_ld), nb), dim3(256), 0, stream,
                       js.d_probs, (const double*)Wm, Wout, act, 0);
  if (evout)
    hipLaunchKernelGGL(k_jac_sort, dim3(grid1(js.big_max_ld, 256), nb), dim3(256), 0, stream, js.d_probs,
                       (const double*)Am, evout, perm,
                       check_active ? (const MolState*)d_state : (const MolState*)nullptr, 1 << 30);
  return rows;
}

// ---------------------------------------------------------------------------
// blocked eigensolver: work arrays, trees and product job lists of eig_'s
// problems (built once per rebuild; the device addresses are fixed)
void Batch::eig_build() {
  // device memory of the batch (a loan's pieces are never released)
  struct BatchMem : EigMem {
    Batch* b;
    explicit BatchMem(Batch* bb) : b(bb) {}
    void* get(size_t bytes) override { return b->alloc<char>(bytes); }
    void put(void* p) override {
      auto it = std::find(b->allocs_.begin(), b->allocs_.end(), p);
      if (it != b->allocs_.end()) b->allocs_.erase(it);
      if (!b->on_loan()) HIPCHECK(hipFree(p));
    }
  } mem(this);
  std::vector<int> dims;
  for (int m : eig_.mols) dims.push_back(desc[m].nao);
  eig_set_build(eig_, dims, mem);
  // captured on stream3 (the null stream of a loan cannot be captured)
  eig_.use_graph = eig_graph && !on_loan();
}

void eig_set_build(EigSet& es, const std::vector<int>& dims, EigMem& mem) {
  if (es.graph) {
    // the captured launches name the old work arrays
    HIPCHECK(hipGraphExecDestroy(es.graph));
    es.graph = nullptr;
  }
  es.np = (int)es.mols.size();
  if (es.np == 0) return;
  auto grow = [&](auto*& ptr, long long& capv, long long need) {
    using T_ = typename std::remove_reference<decltype(*ptr)>::type;
    if (need <= capv && ptr) return;
    if (ptr) {
      HIPCHECK(hipDeviceSynchronize());
      mem.put((void*)ptr);
    }
    capv = std::max<long long>(need, capv);
    ptr = (T_*)mem.get((size_t)std::max(1LL, capv) * sizeof(T_));
  };
  std::vector<EigProb> probs(es.np);
  long long mat = 0, vv = 0, iv = 0, xx = 0, pp = 0, pp2 = 0, tt = 0, gg = 0;
  es.v_off.clear();
  es.ld.clear();
  es.maxn = 0;
  es.maxdepth = 0;
  for (int k = 0; k < es.np; ++k) {
    EigProb& e = probs[k];
    std::memset(&e, 0, sizeof(e));
    e.n = dims[k];
    e.ld = (e.n + 15) / 16 * 16;
    e.mol = es.mols[k];
    e.m_off = mat;
    e.v_off = vv;
    es.v_off.push_back(vv);
    es.ld.push_back(e.ld);
    e.i_off = iv;
    e.x_off = xx;
    e.nchunk = std::max(1, (e.n - 1 + kBtChunk - 1) / kBtChunk);
    e.p_off = pp;
    e.p_off2 = pp2;
    e.t_off = tt;
    e.g_off = gg;
    pp2 += (long long)((e.ld + 255) / 256) * kTrdPart;
    const int nblk_e = std::max(1, (e.n - 1 + kBtNb - 1) / kBtNb);
    tt += (long long)nblk_e * kBtNb * e.nchunk * kBtNb;
    gg += (long long)nblk_e * kBtNb * kBtNb;
    mat += (long long)e.n * e.ld;
    vv += (long long)kEigVecs * e.ld;
    iv += (long long)kEigInts * e.ld;
    xx += 4LL * kEigNb * e.ld;
    pp += (long long)e.n * kBtNb * e.nchunk;
    int depth = 0;
    while (((e.n + (1 << depth) - 1) >> depth) > kDcLeaf) ++depth;
    e.depth = depth;
    es.maxn = std::max(es.maxn, e.n);
    es.maxdepth = std::max(es.maxdepth, depth);
  }
  es.npanel = (es.maxn - 1 + kEigNb - 1) / kEigNb;
  es.nblk = (es.maxn - 1 + kBtNb - 1) / kBtNb;
  grow(es.Aw, es.cap[0], mat);
  grow(es.Vh, es.cap[1], mat);
  grow(es.Z0, es.cap[2], mat);
  grow(es.Z1, es.cap[3], mat);
  grow(es.U, es.cap[4], mat);
  grow(es.vec, es.cap[5], vv);
  grow(es.ivec, es.cap[6], iv);
  grow(es.X, es.cap[7], xx);
  grow(es.P, es.cap[8], pp);
  grow(es.P2, es.cap[9], pp);
  grow(es.Tf, es.cap[10], tt);
  grow(es.d_probs, es.cap[11], es.np);
  grow(es.act, es.cap[12], es.np);
  grow(es.zf, es.cap[13], es.np);
  grow(es.part2, es.cap[18], pp2);
  grow(es.G, es.cap[19], gg);
  es.nmat = mat;
  es.nx = xx;
  // divide and conquer trees: every leaf at the problem's depth
  std::vector<int4> leaves;
  std::vector<std::vector<DcNode>> lv(es.maxdepth + 1);
  es.zfin.assign(es.np, 0);
  for (int k = 0; k < es.np; ++k) {
    const int D = probs[k].depth;
    es.zfin[k] = D & 1;
    struct Item { int start, size, depth; };
    std::vector<Item> st{{0, probs[k].n, 0}};
    while (!st.empty()) {
      const Item it = st.back();
      st.pop_back();
      if (it.depth == D) {
        leaves.push_back(make_int4(k, it.start, it.size, 0));
        continue;
      }
      const int n1 = it.size / 2;
      DcNode nd;
      nd.prob = k;
      nd.start = it.start;
      nd.n1 = n1;
      nd.n2 = it.size - n1;
      lv[D - it.depth].push_back(nd);
      st.push_back({it.start, n1, it.depth + 1});
      st.push_back({it.start + n1, it.size - n1, it.depth + 1});
    }
  }
  std::vector<DcNode> nodes;
  es.lvl0.assign(es.maxdepth + 1, 0);
  es.lvln.assign(es.maxdepth + 1, 0);
  es.lvlmax.assign(es.maxdepth + 1, 0);
  for (int h = 1; h <= es.maxdepth; ++h) {
    es.lvl0[h] = (int)nodes.size();
    es.lvln[h] = (int)lv[h].size();
    for (const DcNode& nd : lv[h]) {
      nodes.push_back(nd);
      es.lvlmax[h] = std::max(es.lvlmax[h], nd.n1 + nd.n2);
    }
  }
  es.nleaves = (int)leaves.size();
  es.nnodes = (int)nodes.size();
  grow(es.d_leaves, es.cap[14], std::max(1, es.nleaves));
  grow(es.d_nodes, es.cap[15], std::max(1, es.nnodes));
  // product jobs
  std::vector<GemmJob> jobs;
  std::vector<int2> tiles;
  auto add = [&](std::vector<EigSet::Launch>& lst, const std::vector<GemmJob>& js_) {
    EigSet::Launch L;
    L.job0 = (int)jobs.size();
    L.tile0 = (long long)tiles.size();
    for (GemmJob g : js_) {
      g.tm = (g.M + 63) / 64;
      g.tn = (g.N + 63) / 64;
      g.tile0 = (long long)tiles.size() - L.tile0;
      const int jid = (int)jobs.size() - L.job0;
      for (int t = 0; t < g.tm * g.tn; ++t) tiles.push_back(make_int2(jid, t));
      jobs.push_back(g);
    }
    L.njob = (int)jobs.size() - L.job0;
    L.ntile = (long long)tiles.size() - L.tile0;
    lst.push_back(L);
  };
  auto job = [](int M, int N, int K, const double* A, long long lda, const double* B, long long ldb, double* C,
                long long ldc, double alpha, double beta, const int* act) {
    GemmJob g;
    std::memset(&g, 0, sizeof(g));
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.B = B; g.C = C;
    g.lda = lda; g.ldb = ldb; g.ldc = ldc;
    g.alpha = alpha; g.beta = beta;
    g.act = act;
    return g;
  };
  es.upd.clear(); es.bt1.clear(); es.bt2.clear(); es.bt3.clear(); es.dcg.clear();
  for (int pi = 0; pi < es.npanel; ++pi) {
    const int j0 = pi * kEigNb;
    std::vector<GemmJob> u;
    for (int k = 0; k < es.np; ++k) {
      const EigProb& e = probs[k];
      const int n = e.n, ld = e.ld;
      if (j0 >= n - 1) continue;
      double* Xp = es.X + e.x_off;
      double* Yp = Xp + 2LL * kEigNb * ld;
      const int s0 = j0 + kEigNb;
      // trailing update A22 -= X^T Y on [s0, n) x [s0, n)
      if (s0 <= n - 1)
        u.push_back(job(n - s0, n - s0, 2 * kEigNb, Xp + s0, ld, Yp + s0, ld, es.Aw + e.m_off + (long long)s0 * ld + s0,
                        ld, -1.0, 1.0, es.act + k));
    }
    add(es.upd, u);
  }
  // back-transformation of the eigenvector rows Z, blocks of kBtNb reflectors
  // (rows j0.. of Vh, each zero before its own start and 1 there): the Gram
  // matrices G = V V^T of all blocks in one launch, then per block
  // P = Z V^T in K chunks (column block c of P), P2 = [P_1 .. P_C] [T .. T]^T,
  // Z -= P2 V
  {
    std::vector<GemmJob> gr;
    for (int bi = 0; bi < es.nblk; ++bi) {
      const int j0 = bi * kBtNb;
      std::vector<GemmJob> b1, b2, b3;
      for (int k = 0; k < es.np; ++k) {
        const EigProb& e = probs[k];
        const int n = e.n, ld = e.ld;
        if (j0 >= n - 1) continue;
        const int jb = std::min(kBtNb, n - 1 - j0);
        const int* act = es.act + k;
        double* Zf = (es.zfin[k] ? es.Z1 : es.Z0) + e.m_off;
        const int k0 = j0 + 1;
        const double* Vp = es.Vh + e.m_off + (long long)j0 * ld + k0;
        double* P = es.P + e.p_off;
        double* P2 = es.P2 + e.p_off;
        const int tw = e.nchunk * kBtNb;
        const double* T = es.Tf + e.t_off + (long long)bi * kBtNb * tw;
        gr.push_back(job(jb, jb, n - k0, Vp, ld, Vp, ld, es.G + e.g_off + (long long)bi * kBtNb * kBtNb, kBtNb, 1.0,
                         0.0, act));
        for (int c = 0; c < e.nchunk; ++c) {
          const int ks = k0 + c * kBtChunk, kl = std::max(0, std::min(kBtChunk, n - ks));
          b1.push_back(job(n, jb, kl, Zf + std::min(ks, n - 1), ld, Vp + std::min(ks - k0, n - 1 - k0), ld,
                           P + c * kBtNb, tw, 1.0, 0.0, act));
        }
        b2.push_back(job(n, jb, tw, P, tw, T, tw, P2, kBtNb, 1.0, 0.0, act));
        b3.push_back(job(n, n - k0, jb, P2, kBtNb, Vp, ld, Zf + k0, ld, -1.0, 1.0, act));
      }
      add(es.bt1, b1);
      add(es.bt2, b2);
      add(es.bt3, b3);
    }
    std::vector<EigSet::Launch> g1;
    add(g1, gr);
    es.gram = g1[0];
  }
  for (int h = 1; h <= es.maxdepth; ++h) {
    std::vector<GemmJob> g;
    for (int x = 0; x < es.lvln[h]; ++x) {
      const DcNode& nd = nodes[es.lvl0[h] + x];
      const EigProb& e = probs[nd.prob];
      const int m = nd.n1 + nd.n2, ld = e.ld;
      const double* Zin = ((h - 1) & 1 ? es.Z1 : es.Z0) + e.m_off;
      double* Zout = (h & 1 ? es.Z1 : es.Z0) + e.m_off;
      int* I = es.ivec + e.i_off;
      GemmJob j = job(m, m, m, es.U + e.m_off + (long long)nd.start * ld, ld, Zin + nd.start, ld, Zout + nd.start, ld,
                      1.0, 0.0, es.act + nd.prob);
      j.bk = I + (long long)kIrow * ld + nd.start;
      j.crow = I + (long long)kIrot * ld + nd.start;
      j.kdev = I + (long long)kIcnt * ld + nd.start;
      g.push_back(j);
    }
    add(es.dcg, g);
  }
  grow(es.d_jobs, es.cap[16], std::max<long long>(1, (long long)jobs.size()));
  grow(es.d_tiles, es.cap[17], std::max<long long>(1, (long long)tiles.size()));
  HIPCHECK(hipMemcpy(es.d_probs, probs.data(), es.np * sizeof(EigProb), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(es.zf, es.zfin.data(), es.np * sizeof(int), hipMemcpyHostToDevice));
  if (!leaves.empty())
    HIPCHECK(hipMemcpy(es.d_leaves, leaves.data(), leaves.size() * sizeof(int4), hipMemcpyHostToDevice));
  if (!nodes.empty())
    HIPCHECK(hipMemcpy(es.d_nodes, nodes.data(), nodes.size() * sizeof(DcNode), hipMemcpyHostToDevice));
  if (!jobs.empty()) HIPCHECK(hipMemcpy(es.d_jobs, jobs.data(), jobs.size() * sizeof(GemmJob), hipMemcpyHostToDevice));
  if (!tiles.empty()) HIPCHECK(hipMemcpy(es.d_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice));
  // P's columns beyond a partial panel's width meet zero columns of T: keep
  // them finite
  HIPCHECK(hipMemset(es.P, 0, (size_t)std::max(1LL, pp) * sizeof(double)));
}

void Batch::eig_blocked(const JacSet& js, double* Am, double* Wm, bool check_active, hipStream_t s) {
  eig_set_run(eig_, js.d_probs, js.state, check_active ? (const MolState*)d_state : (const MolState*)nullptr, Am, Wm,
              s);
}

static void eig_set_launch(EigSet& es, const JacobiProb* jprobs, int* jstate, const MolState* mstate, double* Am,
                           double* Wm, hipStream_t s);

void eig_set_run(EigSet& es, const JacobiProb* jprobs, int* jstate, const MolState* mstate, double* Am, double* Wm,
                 hipStream_t s) {
  if (!es.use_graph || es.np == 0) {
    eig_set_launch(es, jprobs, jstate, mstate, Am, Wm, s);
    return;
  }
  const void* key[5] = {jprobs, jstate, mstate, Am, Wm};
  if (!es.graph || !std::equal(key, key + 5, es.graph_key)) {
    if (es.graph) HIPCHECK(hipGraphExecDestroy(es.graph));
    es.graph = nullptr;
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    eig_set_launch(es, jprobs, jstate, mstate, Am, Wm, s);
    HIPCHECK(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&es.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHECK(e);
    std::copy(key, key + 5, es.graph_key);
  }
  HIPCHECK(hipGraphLaunch(es.graph, s));
}

static void eig_set_launch(EigSet& es, const JacobiProb* jprobs, int* jstate, const MolState* mstate, double* Am,
                           double* Wm, hipStream_t s) {
  const int np = es.np;
  auto gemm_launch = [&](const EigSet::Launch& L, bool ta, bool tb) {
    if (L.ntile == 0) return;
    const dim3 g((unsigned)L.ntile), b(256);
    const GemmJob* jb = es.d_jobs + L.job0;
    const int2* tl = es.d_tiles + L.tile0;
    if (ta) hipLaunchKernelGGL((k_bgemm<true, false>), g, b, 0, s, jb, L.njob, tl);
    else if (tb) hipLaunchKernelGGL((k_bgemm<false, true>), g, b, 0, s, jb, L.njob, tl);
    else hipLaunchKernelGGL((k_bgemm<false, false>), g, b, 0, s, jb, L.njob, tl);
  };
  const long long nn = (long long)es.maxn * es.maxn;
  hipLaunchKernelGGL(k_eig_in, dim3(gstride(nn), np), dim3(256), 0, s, (const EigProb*)es.d_probs, jprobs, mstate,
                     es.act, (const double*)Am, es.Aw);
  // 1. tridiagonalisation
  const int nmax = es.maxn;
  for (int pi = 0; pi < es.npanel; ++pi) {
    const int j0 = pi * kEigNb;
    const int j1 = std::min(j0 + kEigNb, nmax - 1);   // this panel's columns [j0, j1)
    hipLaunchKernelGGL(k_fill, dim3(grid1(es.nx, 256)), dim3(256), 0, s, es.X, es.nx, 0.0);
    hipLaunchKernelGGL(k_trd_row, dim3(grid1(nmax - j0, 256), np), dim3(256), 0, s, (const EigProb*)es.d_probs,
                       (const int*)es.act, (const double*)es.Aw, es.X, es.vec, es.part2, -1, j0, j0);
    for (int j = j0; j < j1; ++j) {
      hipLaunchKernelGGL(k_trd_col, dim3(grid1(nmax - 1 - j, 4), np), dim3(256), 0, s, (const EigProb*)es.d_probs,
                         (const int*)es.act, (const double*)es.Aw, es.Vh, es.X, es.vec, (const double*)es.part2, j,
                         j0);
      const int next = j + 1 < j0 + kEigNb ? j + 1 : -1;
      hipLaunchKernelGGL(k_trd_row, dim3(grid1(nmax - 1 - j, 256), np), dim3(256), 0, s, (const EigProb*)es.d_probs,
                         (const int*)es.act, (const double*)es.Aw, es.X, es.vec, es.part2, j, next, j0);
    }
    gemm_launch(es.upd[pi], true, false);
  }
  hipLaunchKernelGGL(k_trd_save, dim3(grid1(nmax, 256), np), dim3(256), 0, s, (const EigProb*)es.d_probs,
                     (const int*)es.act, es.vec);
  // 2. divide and conquer
  hipLaunchKernelGGL(k_fill, dim3(grid1(es.nmat, 256)), dim3(256), 0, s, es.Z0, es.nmat, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(grid1(es.nmat, 256)), dim3(256), 0, s, es.Z1, es.nmat, 0.0);
  if (es.nnodes)
    hipLaunchKernelGGL(k_dc_split, dim3(grid1(es.nnodes, 64)), dim3(64), 0, s, (const EigProb*)es.d_probs,
                       (const int*)es.act, (const DcNode*)es.d_nodes, es.nnodes, es.vec);
  hipLaunchKernelGGL(k_dc_leaf, dim3(es.nleaves), dim3(256), 0, s, (const EigProb*)es.d_probs, (const int*)es.act,
                     (const int4*)es.d_leaves, es.vec, es.Z0);
  for (int h = 1; h <= es.maxdepth; ++h) {
    const int nl = es.lvln[h], mm = es.lvlmax[h];
    if (nl == 0) continue;
    const DcNode* nd = es.d_nodes + es.lvl0[h];
    double* Zin = (h - 1) & 1 ? es.Z1 : es.Z0;
    double* Zout = h & 1 ? es.Z1 : es.Z0;
    hipLaunchKernelGGL(k_dc_deflate, dim3(nl), dim3(256), 0, s, (const EigProb*)es.d_probs, (const int*)es.act, nd,
                       es.vec, es.ivec, Zin);
    hipLaunchKernelGGL(k_dc_secular, dim3(grid1(mm, 4), nl), dim3(256), 0, s, (const EigProb*)es.d_probs,
                       (const int*)es.act, nd, es.vec, es.ivec);
    hipLaunchKernelGGL(k_dc_zhat, dim3(grid1(mm, 64), nl), dim3(64), 0, s, (const EigProb*)es.d_probs,
                       (const int*)es.act, nd, es.vec, (const int*)es.ivec);
    hipLaunchKernelGGL(k_dc_place, dim3(grid1(mm, 256), nl), dim3(256), 0, s, (const EigProb*)es.d_probs,
                       (const int*)es.act, nd, es.vec, es.ivec);
    hipLaunchKernelGGL(k_dc_rows, dim3(mm, nl), dim3(256), 0, s, (const EigProb*)es.d_probs, (const int*)es.act, nd,
                       (const double*)es.vec, es.ivec, (const double*)Zin, Zout);
    hipLaunchKernelGGL(k_dc_vectors, dim3(grid1(mm, 4), nl), dim3(256), 0, s, (const EigProb*)es.d_probs,
                       (const int*)es.act, nd, es.vec, (const int*)es.ivec, es.U);
    gemm_launch(es.dcg[h - 1], false, false);
    hipLaunchKernelGGL(k_dc_lam, dim3(grid1(mm, 256), nl), dim3(256), 0, s, (const EigProb*)es.d_probs,
                       (const int*)es.act, nd, es.vec);
  }
  // 3. back-transformation, last block first
  gemm_launch(es.gram, false, true);
  hipLaunchKernelGGL(k_trd_tblock, dim3(std::max(1, es.nblk), np), dim3(256), 0, s, (const EigProb*)es.d_probs,
                     (const int*)es.act, (const double*)es.G, (const double*)es.vec, es.Tf);
  for (int bi = es.nblk - 1; bi >= 0; --bi) {
    gemm_launch(es.bt1[bi], false, true);
    gemm_launch(es.bt2[bi], false, true);
    gemm_launch(es.bt3[bi], false, false);
  }
  hipLaunchKernelGGL(k_eig_out, dim3(gstride(nn), np), dim3(256), 0, s, (const EigProb*)es.d_probs, jprobs,
                     (const int*)es.act, (const double*)es.vec, (const double*)es.Z0, (const double*)es.Z1,
                     (const int*)es.zf, Am, Wm, jstate);
}

// block-Jacobi rounds of the two lanes (cyclic sweeps, register pivots and
// tile updates).  Lane g's kernels go to its stream with its part of the work
// lists, so the launch gaps of one lane overlap the other lane's work.  A
// round is two launches per lane: the pivots (which first test the problems
// at a sweep boundary against the off-diagonal maximum the tile updates
// measured at the end of the sweep before) and the A and W tile updates.
// Once per period (the rounds of a sweep of the lanes' largest problem),
// after the pivots, the lanes count their problems still running; the host
// reads the counts once the round's tile updates and the next round's pivots
// are queued, with the GPU still busy, and stops a lane that has none left.
constexpr int kProfEvery = 51;

void Batch::jacobi_lanes(JacSet& js, double* Am, double* Wm) {
  const hipStream_t ls[kMaxLanes] = {stream, stream2, stream4};
  bool live[kMaxLanes];
  for (int g = 0; g < kMaxLanes; ++g) live[g] = js.gpairs[g] < js.gpairs[g + 1];
  unsigned long long* jw = (prof_mask & 4) ? prof_work_ + 2 : (unsigned long long*)nullptr;
  // the initial off-diagonal maxima go to the buffer the round-0 test reads
  for (int g = 0; g < kMaxLanes; ++g) {
    const int nto = js.gto[g + 1] - js.gto[g];
    if (live[g] && nto > 0)
      hipLaunchKernelGGL(k_jac_offmax, dim3(nto), dim3(256), 0, ls[g], js.d_probs, js.d_to + js.gto[g],
                         (const double*)Am, (const int*)js.state, js.offmax + nmol, 0);
  }
  const int period = std::max(1, jac_period_);
  const int rmax = (jac_max_sweeps + 1) * js.max_rounds;
  // the pivots of round rr on every live lane; timing events, device spans
  // and work counts on every kProfEvery-th round only, counted over all
  // eigensolves (51 is prime to the cross-round period 8, so the sample meets
  // every kind of round alike): an event pair per launch would cost more host
  // time than the launch
  auto launch_pivots = [&](int rr) {
    const bool sample = (prof_round_ % kProfEvery) == 0;
    for (int g = 0; g < kMaxLanes; ++g) {
      if (!live[g]) continue;
      if (prof_mask & 4) prof_all_[2] += 1;
      if (sample) prof_begin(2, ls[g]);
      unsigned long long* pts = nullptr;
      if (sample && (prof_mask & 4) && prof_ts_used_[0] < kProfTsSlots)
        pts = prof_ts_ + (size_t)kProfSlotU64 * (prof_ts_used_[0]++);
      if (jac_cross)
        hipLaunchKernelGGL(k_jac_pairsR<true>, dim3(js.gxp[g + 1] - js.gxp[g]), dim3(64), 0, ls[g], js.d_probs,
                           js.d_xp + js.gxp[g], rr, (const double*)Am, js.qbuf, js.rot, (const double*)js.scale,
                           js.state, inner_sweeps, sample ? jw : nullptr, jac_tol, js.offmax, nmol, pts);
      else
        hipLaunchKernelGGL(k_jac_pairsR<false>, dim3(js.gxp[g + 1] - js.gxp[g]), dim3(64), 0, ls[g], js.d_probs,
                           js.d_xp + js.gxp[g], rr, (const double*)Am, js.qbuf, js.rot, (const double*)js.scale,
                           js.state, inner_sweeps, sample ? jw : nullptr, jac_tol, js.offmax, nmol, pts);
      if (sample) prof_end(2, nullptr, 0, ls[g]);
    }
  };
  // Once per period the lanes count their problems still running and copy the
  // counts to the host on their own streams (no join between the lanes).  The
  // host reads them kCheckLag rounds later, after queueing that round's tile
  // updates and the next round's pivots, so that the GPU works while the host
  // waits; a lane found empty has only finished problems, which the launches
  // queued meanwhile skip (the pivots and tile updates test the problem state
  // first).
  constexpr int kCheckLag = 0;   // 0-2 measured alike on the 512-molecule batch
  const hipEvent_t lane_ev[kMaxLanes] = {ev_join3, ev_join, ev_join4};
  bool pivots_queued = false;   // round r's pivots went out before the host read the counts
  int pending = -1;             // round of the counts on their way to the host
  bool waiting[kMaxLanes] = {};   // lanes whose counts are on their way
  for (int r = 0; r <= rmax; ++r) {
    const bool sample = (prof_round_ % kProfEvery) == 0;
    if (!pivots_queued) launch_pivots(r);
    pivots_queued = false;
    if (r % period == 0 && pending < 0) {
      for (int g = 0; g < kMaxLanes; ++g) {
        if (!live[g]) continue;
        hipLaunchKernelGGL(k_jac_count, dim3(1), dim3(256), 0, ls[g], js.d_probs, nmol, (const int*)js.state, g,
                           jac_nactive + g);
        HIPCHECK(hipMemcpyAsync(h_pinned + g, jac_nactive + g, sizeof(int), hipMemcpyDeviceToHost, ls[g]));
        HIPCHECK(hipEventRecord(lane_ev[g], ls[g]));
        waiting[g] = true;
      }
      pending = r;
    }
    unsigned long long* tw = (sample && (prof_mask & 8)) ? prof_work_ + 3 : (unsigned long long*)nullptr;
    for (int g = 0; g < kMaxLanes; ++g) {
      if (!live[g]) continue;
      if (prof_mask & 8) prof_all_[3] += 1;
      if (sample) prof_begin(3, ls[g]);
      unsigned long long* ts = nullptr;
      if (sample && (prof_mask & 8) && prof_ts_used_[1] < kProfTsSlots)
        ts = prof_ts_ + (size_t)kProfSlotU64 * (kProfTsSlots + prof_ts_used_[1]++);
      hipLaunchKernelGGL(k_jac_apply_R, dim3(js.gxt[g + 1] - js.gxt[g]), dim3(64), 0, ls[g], js.d_probs,
                         js.d_xt + js.gxt[g], r, Am, Wm, (const double*)js.qbuf, (const int*)js.rot,
                         (const int*)js.state, js.offmax, nmol, tw, ts);
      if (sample) prof_end(3, nullptr, 0, ls[g]);
    }
    jacobi_rounds_ += 1;
    prof_round_ += 1;
    if (pending >= 0 && (r == pending + kCheckLag || r == rmax)) {
      if (r < rmax) {
        launch_pivots(r + 1);
        pivots_queued = true;
      }
      bool any = false;
      for (int g = 0; g < kMaxLanes; ++g) {
        if (waiting[g]) {
          HIPCHECK(hipEventSynchronize(lane_ev[g]));
          live[g] = live[g] && h_pinned[g] > 0;
          waiting[g] = false;
        }
        any = any || live[g];
      }
      pending = -1;
      if (!any) break;
    }
  }
}

void Batch::eigh_rows(JacSet& js, double* Am, double* Wm, double* evout) {
  const MolState* st = d_state;
  // state and scale of the small problems (W is written whole below)
  hipLaunchKernelGGL(k_jac_init, dim3(nmol), dim3(256), 0, stream, js.d_probs, (const double*)Am, Wm, js.scale,
                     js.state, st, js.offmax, 1, 1.0, 0.0, jac_tol);
  if (js.nsmall) {
    hipLaunchKernelGGL(k_jac_small, dim3(js.nsmall), dim3(256), 0, stream, js.d_probs, js.d_small, Am, Wm,
                       (const double*)js.scale, js.state, 0, 1, jac_tol);
    hipLaunchKernelGGL(k_jac_sort, dim3(1, nmol), dim3(256), 0, stream, js.d_probs, (const double*)Am, evout, perm, st,
                       64);
  }
  if (js.nbig) {
    const int cap = std::min(js.big_max_dim, 1600);
    prof_begin(2);
    hipLaunchKernelGGL(k_tred_ql, dim3(js.nbig), dim3(256), (size_t)5 * cap * sizeof(double), stream, js.d_probs,
                       (const int*)js.d_big, Am, Wm, G, evout, perm, st, cap, eig_failed,
                       (prof_mask & 4) ? prof_work_ + 2 : (unsigned long long*)nullptr);
    prof_end(2, nullptr, 0);
  }
}

void Batch::refine(double* Am, double* Wm) {
  const dim3 gmm(grid1((long long)max_nao * max_nao, 256), nmol);
  // refining molecules = SCC-active ones
  hipLaunchKernelGGL(k_copy_state_active, dim3(grid1(nmol, 64)), dim3(64), 0, stream, (const MolState*)d_state, nmol,
                     oa_act);
  hipLaunchKernelGGL(k_fill_int, dim3(grid1(nmol, 64)), dim3(64), 0, stream, oa_ok, (long long)nmol, 0);
  hipLaunchKernelGGL(k_fill, dim3(grid1(4LL * nmol, 256)), dim3(256), 0, stream, (double*)oa_stat, 4LL * nmol, 0.0);
  hipLaunchKernelGGL(k_copy_mask, gmm, dim3(256), 0, stream, d_desc, (const int*)oa_act, (const double*)Am, Sm);
  for (int step = 0; step <= oa_max_steps; ++step) {
    if (step > 0) {
      gemm(1, Wm, Wm, G, nullptr, nullptr, false, 1.0, oa_act);   // X^T X
      gemm(0, Am, Wm, T, nullptr, nullptr, false, 1.0, oa_act);   // A X
      gemm(1, Wm, T, Sm, nullptr, nullptr, false, 1.0, oa_act);   // X^T A X
    }
    hipLaunchKernelGGL(k_oa_e, gmm, dim3(256), 0, stream, d_desc, (const int*)oa_act, (const double*)Sm,
                       step > 0 ? (const double*)G : (const double*)nullptr, H, oa_stat);
    hipLaunchKernelGGL(k_oa_mark, dim3(1), dim3(256), 0, stream, nmol, (const MolDesc*)d_desc, oa_act, oa_ok, oa_stat,
                       step, oa_max_steps, jac_nactive);
    HIPCHECK(hipMemcpyAsync(h_pinned, jac_nactive, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (h_pinned[0] <= 0) break;
    if (step == 0) {
      hipLaunchKernelGGL(k_oa_update, gmm, dim3(256), 0, stream, d_desc, (const int*)oa_act, (const double*)H, Wm, 1);
    } else {
      gemm(0, Wm, H, P, nullptr, nullptr, false, 1.0, oa_act);    // X E
      hipLaunchKernelGGL(k_oa_update, gmm, dim3(256), 0, stream, d_desc, (const int*)oa_act, (const double*)P, Wm, 0);
    }
    oa_steps_ += 1;
  }
  hipLaunchKernelGGL(k_oa_finish, gmm, dim3(256), 0, stream, d_desc, (const MolState*)d_state, (const int*)oa_ok,
                     (const double*)Sm, Am, Wm);
}

void Batch::set_guess(int m, const double* qsh_in, const double* dipm_in, const double* qp_in) {
  if (m < 0 || m >= nmol) throw std::runtime_error("invalid molecule index");
  if (guess_qsh_.empty()) {
    guess_qsh_.resize(nmol);
    guess_dipm_.resize(nmol);
    guess_qp_.resize(nmol);
  }
  const MolDesc& d = desc[m];
  guess_qsh_[m].assign(qsh_in, qsh_in + d.nsh);
  guess_dipm_[m].assign(dipm_in, dipm_in + 3 * d.n);
  guess_qp_[m].assign(qp_in, qp_in + 6 * d.n);
}

void Batch::gemm(int kind, const double* a, const double* b, double* c, const double* scale_, const int* perm_,
                 bool check_active, double alpha, const int* mask, bool sym) {
  const int4* tiles = sym ? d_tiles + ntiles : d_tiles;
  const int nt = sym ? ntiles_sym : ntiles;
  if (nt <= 0) return;
  // SCC-active molecules: the 'active' field of MolState, read with its stride
  const int* m = mask;
  int ms = 1;
  if (check_active) {
    m = &d_state[0].active;
    ms = (int)(sizeof(MolState) / sizeof(int));
  }
  prof_begin(1);
  // the mixed-precision option: the Fock transforms and the density product
  // (kinds 0-2) on the FP32 matrix cores with the FP64 residual correction
  if (mixed && mixed_here_ && kind <= 2) {
    if (kind == 0)
      hipLaunchKernelGGL((k_gemm_mix<false, false, false>), dim3(nt), dim3(256), 0, stream, tiles, d_desc, a, b, c,
                         scale_, alpha, m, ms);
    else if (kind == 1)
      hipLaunchKernelGGL((k_gemm_mix<true, false, false>), dim3(nt), dim3(256), 0, stream, tiles, d_desc, a, b, c,
                         scale_, alpha, m, ms);
    else
      hipLaunchKernelGGL((k_gemm_mix<false, true, true>), dim3(nt), dim3(256), 0, stream, tiles, d_desc, a, b, c,
                         scale_, alpha, m, ms);
    prof_end(1, m, ms);
    return;
  }
#define XTBAMD_GEMM(K, TA, TB, SC, PM)                                                                              \
  case K:                                                                                                          \
    if (use_mfma)                                                                                                  \
      hipLaunchKernelGGL((k_gemm_mfma<TA, TB, SC, PM>), dim3(nt), dim3(256), 0, stream, tiles, d_desc, a, b, c,  \
                         scale_, perm_, alpha, m, ms);                                                              \
    else                                                                                                           \
      hipLaunchKernelGGL((k_gemm<TA, TB, SC, PM>), dim3(nt), dim3(256), 0, stream, tiles, d_desc, a, b, c, scale_, \
                         perm_, alpha, m, ms);                                                                      \
    break;
  switch (kind) {
    XTBAMD_GEMM(0, false, false, false, false)
    XTBAMD_GEMM(1, true, false, false, false)
    XTBAMD_GEMM(2, false, true, true, false)
    XTBAMD_GEMM(4, false, true, false, true)
    default:
    XTBAMD_GEMM(3, false, false, false, true)
  }
#undef XTBAMD_GEMM
  prof_end(1, m, ms);
}

void Batch::prof_enable(int mask) {
  prof_mask = mask;
  for (int k = 0; k < kProfKinds; ++k) {
    prof_used_[k] = prof_all_[k] = 0;
    // events created up front: creating them inside a timed loop costs host time
    if ((mask & (1 << k)) && prof_ev_[k].size() < 4096) {
      const size_t old = prof_ev_[k].size();
      prof_ev_[k].resize(4096);
      for (size_t i = old; i < prof_ev_[k].size(); ++i) HIPCHECK(hipEventCreate(&prof_ev_[k][i]));
    }
  }
  prof_round_ = 0;
  // two families of kProfTsSlots sampled launches, kProfLines clock pairs each
  if (!prof_ts_) prof_ts_ = alloc<unsigned long long>((size_t)2 * kProfTsSlots * kProfSlotU64);
  prof_ts_used_[0] = prof_ts_used_[1] = 0;
  hipLaunchKernelGGL(k_ts_init, dim3(grid1(2 * kProfTsSlots * kProfLines, 256)), dim3(256), 0, stream, prof_ts_,
                     2 * kProfTsSlots * kProfLines);
  // the work counters: kind k at offset k of each of the kProfLines lines
  if (!prof_work_) prof_work_ = alloc<unsigned long long>((size_t)kProfLines * kProfLineU64);
  hipLaunchKernelGGL(k_fill, dim3(grid1(kProfLines * kProfLineU64, 256)), dim3(256), 0, stream, (double*)prof_work_,
                     (long long)kProfLines * kProfLineU64, 0.0);
  HIPCHECK(hipStreamSynchronize(stream));
}

void Batch::prof_begin(int kind, hipStream_t s) {
  if (!(prof_mask & (1 << kind))) return;
  std::vector<hipEvent_t>& ev = prof_ev_[kind];
  if (prof_used_[kind] + 2 > (int)ev.size()) {
    const size_t old = ev.size();
    ev.resize(old + 1024);
    for (size_t i = old; i < ev.size(); ++i) HIPCHECK(hipEventCreate(&ev[i]));
  }
  HIPCHECK(hipEventRecord(ev[prof_used_[kind]], s ? s : stream));
}

void Batch::prof_end(int kind, const int* mask, int mstride, hipStream_t s) {
  if (!(prof_mask & (1 << kind))) return;
  HIPCHECK(hipEventRecord(prof_ev_[kind][prof_used_[kind] + 1], s ? s : stream));
  prof_used_[kind] += 2;
  if (kind == 1)
    hipLaunchKernelGGL(k_count_gemm, dim3(grid1(nmol, 256)), dim3(256), 0, stream, d_desc, nmol, mask, mstride,
                       prof_work_ + 1);
}

double Batch::prof_span(int kind) {
  // the tile updates' sampled execution spans (device wall clock), scaled
  // like the event times to all launches of the family
  if ((kind != 2 && kind != 3) || !prof_ts_) return 0.0;
  const int fam = kind - 2, nused = prof_ts_used_[fam];
  if (nused == 0) return 0.0;
  HIPCHECK(hipDeviceSynchronize());
  std::vector<unsigned long long> h((size_t)kProfSlotU64 * nused);
  HIPCHECK(hipMemcpy(h.data(), prof_ts_ + (size_t)kProfSlotU64 * fam * kProfTsSlots,
                     h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  int khz = 0;
  HIPCHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
  double ticks = 0.0;
  int used = 0;
  for (int i = 0; i < nused; ++i) {
    // earliest start and latest end over the slot's lines
    unsigned long long t0 = ~0ull, t1 = 0ull;
    for (int l = 0; l < kProfLines; ++l) {
      const unsigned long long* c = h.data() + (size_t)kProfSlotU64 * i + (size_t)kProfLineU64 * l;
      t0 = std::min(t0, c[0]);
      t1 = std::max(t1, c[1]);
    }
    if (t0 != ~0ull && t1 >= t0) {
      ticks += (double)(t1 - t0);
      ++used;
    }
  }
  if (used == 0 || khz <= 0) return 0.0;
  const double ms = ticks / (double)khz;   // ticks / (kHz) = ms
  return ms * (double)prof_all_[kind] / used;
}

int Batch::prof_result(int kind, int* launches, double* total_ms, double* work) {
  if (kind < 1 || kind >= kProfKinds) throw std::runtime_error("unknown kernel family");
  HIPCHECK(hipStreamSynchronize(stream));
  double t = 0.0;
  for (int i = 0; i + 1 < prof_used_[kind]; i += 2) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, prof_ev_[kind][i], prof_ev_[kind][i + 1]));
    t += ms;
  }
  unsigned long long w = 0;
  if (prof_work_) {
    std::vector<unsigned long long> h((size_t)kProfLines * kProfLineU64);
    HIPCHECK(hipMemcpy(h.data(), prof_work_, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int l = 0; l < kProfLines; ++l) w += h[(size_t)kProfLineU64 * l + kind];
  }
  // a sampled family (timed on a subset of its launches) is reported for all
  // of its launches: the sample's time and work scaled by all / sampled
  const int timed = prof_used_[kind] / 2;
  double f = 1.0;
  if (prof_all_[kind] > timed && timed > 0) f = (double)prof_all_[kind] / timed;
  if (launches) *launches = f > 1.0 ? prof_all_[kind] : timed;
  if (total_ms) *total_ms = t * f;
  if (work) *work = (double)w * f;
  return 0;
}

void Batch::select(const std::vector<int>& mols) {
  sel_ = mols;
  nsub = (int)mols.size();
  std::vector<MolDesc> sub(nsub);
  std::vector<int> flag(2 * nmol, 0);
  nsub_open = 0;
  for (int k = 0; k < nsub; ++k) {
    if (mols[k] < 0 || mols[k] >= nmol) throw std::runtime_error("invalid molecule index");
    sub[k] = desc[mols[k]];
    flag[mols[k]] = 1;
    if (sub[k].nopen > 0) {
      flag[nmol + mols[k]] = 1;
      nsub_open += 1;
    }
  }
  if (nsub) HIPCHECK(hipMemcpyAsync(d_sub, sub.data(), nsub * sizeof(MolDesc), hipMemcpyHostToDevice, stream));
  HIPCHECK(hipMemcpyAsync(phase_act, flag.data(), nmol * sizeof(int), hipMemcpyHostToDevice, stream));
  HIPCHECK(hipMemcpyAsync(open_act, flag.data() + nmol, nmol * sizeof(int), hipMemcpyHostToDevice, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  sub_max_n = sub_max_nsh = sub_max_nao = sub_max_nd4 = 0;
  sub_max_psp = sub_max_pd = 0;
  for (const MolDesc& d : sub) {
    sub_max_n = std::max(sub_max_n, d.n);
    sub_max_nsh = std::max(sub_max_nsh, d.nsh);
    sub_max_nao = std::max(sub_max_nao, d.nao);
    sub_max_nd4 = std::max(sub_max_nd4, d.nd4);
    sub_max_psp = std::max<long long>(sub_max_psp, d.npsp);
    sub_max_pd = std::max<long long>(sub_max_pd, (long long)d.n * (d.n - 1) / 2 - d.npsp);
  }
}

void Batch::select_all() {
  std::vector<int> all(nmol);
  for (int m = 0; m < nmol; ++m) all[m] = m;
  select(all);
}

void Batch::orthogonalizer() {
  // C = S^(-1/2) for the selected molecules
  const dim3 gmm(gstride((long long)sub_max_nao * sub_max_nao), nsub);
  hipLaunchKernelGGL(k_ns_begin, dim3(grid1(nmol, 64)), dim3(64), 0, stream, nmol, (const int*)phase_act, ns_stage,
                     ns_act, ns_failed, ns_res, ns_c);
  hipLaunchKernelGGL(k_ns_bound, dim3(grid1(sub_max_nao, 64), nsub), dim3(256), 0, stream, d_sub, (const double*)S,
                     ns_c);
  hipLaunchKernelGGL(k_ns_init, gmm, dim3(256), 0, stream, d_sub, (const double*)S, (const double*)ns_c, A, W);
  const int maxit = 60;
  // Y and Z alternate between two buffers each (A/H and W/T) by iteration
  // parity; a molecule's final pair is at the parity recorded when it stops
  double* Yb[2] = {A, H};
  double* Zb[2] = {W, T};
  // The per-molecule state machine (k_ns_mark) and the activity mask of the
  // products live on the device, so an iteration queued after every molecule
  // has stopped changes nothing.  The host therefore reads iteration it's
  // count of running molecules only after iteration it + 1 is queued: the GPU
  // never waits for the host's read.
  for (int it = 0; it <= maxit; ++it) {
    double *Y = Yb[it & 1], *Z = Zb[it & 1];
    // Y and Z are polynomials in S: they commute, and Z Y, Y T and T Z are symmetric
    gemm(0, Z, Y, P, nullptr, nullptr, false, 1.0, ns_act, true);     // P = Z Y
    hipLaunchKernelGGL(k_ns_form, gmm, dim3(256), 0, stream, d_sub, (const int*)ns_act, P, ns_res);
    hipLaunchKernelGGL(k_ns_mark, dim3(1), dim3(256), 0, stream, nmol, ns_stage, ns_act, ns_res, it, maxit, ns_failed,
                       jac_nactive, ns_fin);
    HIPCHECK(hipMemcpyAsync(h_pinned + (it & 1), jac_nactive, sizeof(int), hipMemcpyDeviceToHost, stream));
    gemm(0, Y, P, Yb[(it + 1) & 1], nullptr, nullptr, false, 1.0, ns_act, true);   // Y T
    gemm(0, P, Z, Zb[(it + 1) & 1], nullptr, nullptr, false, 1.0, ns_act, true);   // T Z
    if (it > 0) {
      HIPCHECK(hipEventSynchronize(ev_join3));   // iteration it - 1's count is on the host
      if (h_pinned[(it - 1) & 1] <= 0) break;
    }
    HIPCHECK(hipEventRecord(ev_join3, stream));
  }
  hipLaunchKernelGGL(k_ns_final, gmm, dim3(256), 0, stream, d_sub, (const double*)Zb[0], (const double*)Zb[1],
                     (const int*)ns_fin, (const double*)ns_c, C);
}

void Batch::setup() {
  if (nsub == 0) return;
  const Element* el = elements_dev;
  const int ns = nsub;
  // every single point starts from the initial SCC state (iteration 0,
  // active): one upload of the selected molecules' states, scattered to
  // their slots on the device
  state_stage_.resize(ns);
  for (int k = 0; k < ns; ++k) state_stage_[k] = h_state0_[sel_[k]];
  HIPCHECK(hipMemcpyAsync(d_state_stage, state_stage_.data(), ns * sizeof(MolState), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_state_scatter, dim3(grid1(ns, 64)), dim3(64), 0, stream, (const MolDesc*)d_sub,
                     (const MolState*)d_state_stage, d_state, ns);
  HIPCHECK(hipStreamSynchronize(stream));
  iterations_ = 0;
  jacobi_rounds_ = 0;
  oa_steps_ = 0;
  const dim3 gat(sub_max_n, ns);
  hipLaunchKernelGGL(k_cn, gat, dim3(64), 0, stream, d_sub, el, at, xyz, cn, cn_cov, cn_d4, cn_eeq);
  hipLaunchKernelGGL(k_shell_setup, dim3(grid1(sub_max_nsh, 64), ns), dim3(64), 0, stream, d_sub, el, at, sh_at,
                     sh_m, cn, selfe, dsedcn, shgam, zsh);
  // the three-body (ATM) dispersion takes the zero-charge C6 only, which the
  // geometry and the CN fix: its energy and gradient run on stream3 beside
  // the SCC (into their own buffers, added where the gradient takes them)
  {
    const bool fork = stream3 != stream;
    if (fork) {
      HIPCHECK(hipEventRecord(ev_atm, stream));
      HIPCHECK(hipStreamWaitEvent(stream3, ev_atm, 0));
      if (stream4 != stream) HIPCHECK(hipStreamWaitEvent(stream4, ev_atm, 0));
    }
    hipLaunchKernelGGL(k_d4_refweights, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream3, d_sub, el, at, cn_cov,
                       (const double*)nullptr, zetavec, zerovec, zetadcn, zerodcn);
    hipLaunchKernelGGL(k_d4_c6<1>, dim3(gstride((long long)sub_max_n * (sub_max_n + 1) / 2), ns), dim3(256), 0,
                       stream3, d_sub, el, at, c6ref, zetavec, zerovec, zetadcn, zerodcn, c6, dc6, c60, dc60);
    hipLaunchKernelGGL(k_d4_atm, dim3(sub_max_n, ns), dim3(256), 0, stream3, d_sub, el, at, xyz, (const double*)c60,
                       (const double*)dc60, grad_atm, dedcn_atm, eat);
    if (fork) HIPCHECK(hipEventRecord(ev_atm, stream3));
  }
  // the matrices and factors that depend on the geometry and CN only
  // (Coulomb, D4 dispersion matrix, AES damping, implicit solvent) on
  // stream4 beside the integrals and S^(-1/2); joined before the SCC
  const hipStream_t s4 = stream4 != stream ? stream4 : stream;
  hipLaunchKernelGGL(k_coulomb, dim3(grid1((long long)sub_max_nsh * sub_max_nsh, 256), ns), dim3(256), 0, s4,
                     d_sub, el, at, xyz, sh_at, sh_m, J);
  hipLaunchKernelGGL(k_d4_weights, dim3(grid1(sub_max_nd4, 64), ns), dim3(64), 0, s4, d_sub, el, at, d4_at, d4_ref,
                     cn_d4, gw);
  hipLaunchKernelGGL(k_dispmat, dim3(grid1((long long)sub_max_nd4 * sub_max_nd4, 256), ns), dim3(256), 0, s4,
                     d_sub, el, at, xyz, d4_at, d4_ref, gw, c6ref, dispmat);
  hipLaunchKernelGGL(k_aes_setup, dim3(grid1((long long)sub_max_n * sub_max_n, 256), ns), dim3(256), 0, s4, d_sub,
                     el, at, xyz, cn, radcn, gab3, gab5);
  // implicit solvent: Born radii, surface, H-bond weights, shape, matrix B
  if (solv_on_ && sub_max_n > 0) {
    hipLaunchKernelGGL(k_eeq_reset, dim3(grid1(ns, 64)), dim3(64), 0, s4, d_sub, ns, sv_failed);
    hipLaunchKernelGGL(k_solv_born, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, s4, d_sub,
                       (const SolvDev*)d_solv, at, xyz, sv_brad, sv_dpsi);
    hipLaunchKernelGGL(k_solv_sasa, dim3(sub_max_n, ns), dim3(64), 0, s4, d_sub, (const SolvDev*)d_solv, at, xyz,
                       sv_w, sv_sasa, sv_failed);
    hipLaunchKernelGGL(k_solv_mol, dim3(ns), dim3(256), 0, s4, d_sub, (const SolvDev*)d_solv, at, xyz,
                       (const double*)sv_sasa, sv_hbw, sv_gsasa, sv_adet);
    hipLaunchKernelGGL(k_solv_bmat, dim3(grid1((long long)sub_max_n * sub_max_n, 256), ns), dim3(256), 0, s4,
                       d_sub, (const SolvDev*)d_solv, xyz, (const double*)sv_brad, (const double*)sv_hbw,
                       (const double*)sv_adet, sv_B);
  }
  // a resident wavefunction: every selected molecule starts from the shell
  // charges and multipole moments the previous run left on the device, as
  // the reference starts from the wavefunction of a reused results object
  // (src/api/interface.f90:83-101), so no initial guess is computed
  const bool resident = keep_wfn_ && wfn_valid_;
  const bool eeq = opt.guess == 0 && !resident;
  wfn_valid_ = false;
  // stream2 (the EEQ solve below) may start once CN and zsh exist; its
  // launches are enqueued after the integrals so those start first.
  if (eeq) {
    HIPCHECK(hipEventRecord(ev_fork, stream));
    HIPCHECK(hipStreamWaitEvent(stream2, ev_fork, 0));
  }
  // the s/p-only pairs, then the pairs with a d shell
  if (sub_max_psp > 0)
    hipLaunchKernelGGL(k_sdqh0<1>, dim3(grid1(sub_max_psp, 64), ns), dim3(64), 0, stream, d_sub, el, at, xyz, at_sh0,
                       at_ao0, selfe, intcut, (const int2*)pair_list, S, H0, H0n, D, Q);
  if (sub_max_pd > 0)
    hipLaunchKernelGGL(k_sdqh0<2>, dim3(grid1(sub_max_pd, 64), ns), dim3(64), 0, stream, d_sub, el, at, xyz, at_sh0,
                       at_ao0, selfe, intcut, (const int2*)pair_list, S, H0, H0n, D, Q);
  hipLaunchKernelGGL(k_sdq_onsite, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub, el, at, xyz, at_sh0,
                     at_ao0, selfe, intcut, S, H0, H0n, D, Q);
  {
    // the integral kernels wrote the lower triangles: mirror them, S first
    // (S^(-1/2) needs it next), the other 11 on stream4 after the
    // geometry-only matrices (nothing reads them before the SCC)
    const int nt = (sub_max_nao + 31) / 32;
    const int ntp = std::max(1, nt * (nt + 1) / 2);
    if (s4 != stream) {
      HIPCHECK(hipEventRecord(ev_join3, stream));
      HIPCHECK(hipStreamWaitEvent(s4, ev_join3, 0));
    }
    hipLaunchKernelGGL(k_sym_lower, dim3(ntp, ns, 1), dim3(256), 0, stream, d_sub, S, H0, H0n, D, Q, 0);
    hipLaunchKernelGGL(k_sym_lower, dim3(ntp, ns, 11), dim3(256), 0, s4, d_sub, S, H0, H0n, D, Q, 1);
  }
  if (s4 != stream) HIPCHECK(hipEventRecord(ev_join4, s4));
  // initial charges: EEQ (default) or SAD.  The EEQ solve runs on stream2
  // beside the integrals and S^(-1/2) (it reads only CN, zsh and geometry,
  // writes only its own system and the charges); joined below.
  std::vector<int> big;
  if (resident) {
    hipLaunchKernelGGL(k_eeq_reset, dim3(grid1(ns, 64)), dim3(64), 0, stream, d_sub, ns, eeq_failed);
    hipLaunchKernelGGL(k_q_from_qsh, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub, sh_at,
                       (const double*)qsh, q);
  } else if (eeq) {
    const long long ne = sub_max_n + 1;
    hipLaunchKernelGGL(k_eeq_matrix, dim3(grid1(ne * ne, 256), ns), dim3(256), 0, stream2, d_sub, el, at, xyz,
                       cn_eeq, E, xE);
    hipLaunchKernelGGL(k_eeq_reset, dim3(grid1(ns, 64)), dim3(64), 0, stream2, d_sub, ns, eeq_failed);
    // large systems: one launch pair per column over the whole device
    int big_max = 0;
    for (int k = 0; k < ns; ++k)
      if (desc[sel_[k]].ne >= kEeqBigNe) {
        big.push_back(k);
        big_max = std::max(big_max, desc[sel_[k]].ne);
      }
    if (!big.empty()) {
      HIPCHECK(hipMemcpyAsync(eeq_big, big.data(), big.size() * sizeof(int), hipMemcpyHostToDevice, stream2));
      const int nb = (int)big.size();
      for (int k = 0; k < big_max; ++k) {
        hipLaunchKernelGGL(k_eeq_lu_pivot, dim3(nb), dim3(256), 0, stream2, d_sub, (const int*)eeq_big, k, E, xE,
                           eeq_failed);
        if (big_max - k - 1 > 0)
          hipLaunchKernelGGL(k_eeq_lu_update, dim3(big_max - k - 1, nb), dim3(256), 0, stream2, d_sub,
                             (const int*)eeq_big, k, E, (const int*)eeq_failed);
      }
    }
    hipLaunchKernelGGL(k_eeq_gesv, dim3(ns), dim3(256), 0, stream2, d_sub, el, at, at_sh0, E, xE, yE, q, qsh,
                       (const double*)zsh, eeq_failed, kEeqBigNe);
  } else {
    hipLaunchKernelGGL(k_sad_guess, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub, el, at, at_sh0, q,
                       qsh, (const double*)zsh);
  }
  if (!resident)
    hipLaunchKernelGGL(k_zero_atoms, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub, dipm, qp);
  // S-orthonormal initial orbital basis C = S^(-1/2)
  orthogonalizer();
  if (eeq) {
    HIPCHECK(hipEventRecord(ev_join, stream2));
    HIPCHECK(hipStreamWaitEvent(stream, ev_join, 0));
  }
  if (stream4 != stream) HIPCHECK(hipStreamWaitEvent(stream, ev_join4, 0));   // the geometry-only set-up
  // molecules whose S^(-1/2) or EEQ guess failed do not iterate
  hipLaunchKernelGGL(k_fail_fold, dim3(grid1(nmol, 64)), dim3(64), 0, stream, d_state, nmol, (const int*)ns_failed, 1,
                     kFailOrtho, (const int*)phase_act);
  if (eeq)
    hipLaunchKernelGGL(k_fail_fold, dim3(grid1(nmol, 64)), dim3(64), 0, stream, d_state, nmol, (const int*)eeq_failed,
                       1, kFailEeq, (const int*)phase_act);
  if (use_ql) HIPCHECK(hipMemsetAsync(eig_failed, 0, nmol * sizeof(int), stream));
  // restart guesses replace the initial charges (src/api/interface.f90:83-101, src/scf_module.F90:292)
  for (int m : sel_) {
    if (guess_qsh_.empty() || guess_qsh_[m].empty()) continue;
    const MolDesc& d = desc[m];
    std::vector<double> qa(d.n, 0.0);
    for (int s = 0; s < d.nsh; ++s) qa[hmol_[m].sh_at[s]] += guess_qsh_[m][s];
    HIPCHECK(hipMemcpyAsync(qsh + d.s0, guess_qsh_[m].data(), d.nsh * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(q + d.a0, qa.data(), d.n * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dipm + 3 * d.a0, guess_dipm_[m].data(), 3 * d.n * sizeof(double), hipMemcpyHostToDevice,
                            stream));
    HIPCHECK(hipMemcpyAsync(qp + 6 * d.a0, guess_qp_[m].data(), 6 * d.n * sizeof(double), hipMemcpyHostToDevice,
                            stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  HIPCHECK(hipGetLastError());
  setup_done_ = true;
}

// The first part of an SCC iteration: potentials, the Fock matrix and its
// transform to the current orbital basis.  Every kernel and product in it
// skips molecules that are no longer active (MolState on the device), so it
// may be queued before the host knows whether any molecule still is.
void Batch::scc_prologue() {
  const Element* el = elements_dev;
  const double thr2 = neglect * 1.0e-2;
  // the AES potentials (read only charges and moments, write only vs, vd,
  // vq) on stream2 beside the shell and D4 potentials; joined before the Fock build
  HIPCHECK(hipEventRecord(ev_fork, stream));
  HIPCHECK(hipStreamWaitEvent(stream2, ev_fork, 0));
  hipLaunchKernelGGL(k_aes_pot, dim3(std::max(1, max_n), nmol), dim3(64), 0, stream2, d_desc, el, at, xyz, q, dipm, qp,
                     gab3, gab5, 1, vs, vd, vq, (const MolState*)d_state);
  hipLaunchKernelGGL(k_shell_pot, dim3(grid1(max_nsh, 4), nmol), dim3(256), 0, stream, d_desc, (const double*)J,
                     (const double*)qsh, (const double*)shgam, shift, 1, (const MolState*)d_state);
  hipLaunchKernelGGL(k_d4_zeta, dim3(grid1(max_nd4, 64), nmol), dim3(64), 0, stream, d_desc, el, at, d4_at, d4_ref,
                     gw, q, zv, dzv, (const MolState*)d_state);
  hipLaunchKernelGGL(k_d4_symv, dim3(grid1(max_nd4, 4), nmol), dim3(256), 0, stream, d_desc, (const double*)dispmat,
                     (const double*)zv, d4dum, (const MolState*)d_state);
  if (solv_on_)
    hipLaunchKernelGGL(k_solv_pot, dim3(max_n, nmol), dim3(64), 0, stream, d_desc, (const double*)sv_B,
                       (const double*)q, sv_pot, (const MolState*)d_state);
  hipLaunchKernelGGL(k_shift_final, dim3(grid1(max_nsh, 64), nmol), dim3(64), 0, stream, d_desc, el, at, sh_at,
                     (const double*)d4dum, (const double*)dzv, solv_on_ ? (const double*)sv_pot : nullptr, shift,
                     (const MolState*)d_state);
  HIPCHECK(hipEventRecord(ev_join, stream2));
  HIPCHECK(hipStreamWaitEvent(stream, ev_join, 0));
  hipLaunchKernelGGL(k_fock, dim3(std::max(1, ((max_nao + 31) / 32) * ((max_nao + 31) / 32 + 1) / 2), nmol), dim3(256),
                     0, stream, d_desc, S, H0,
                     D, Q, shift, vs, vd, vq, ao_sh, ao_at, thr2, H, (const MolState*)d_state);
  // generalized eigenproblem in the current orbital basis C (S-orthonormal)
  mixed_here_ = true;   // the Fock transforms and the density product may run mixed
  gemm(0, H, C, T, nullptr, nullptr, true);   // T = H C
  gemm(1, C, T, A, nullptr, nullptr, true, 1.0, nullptr, true);   // A = C^T H C (symmetric)
  mixed_here_ = false;
}

bool Batch::scc_iteration(bool queue_next) {
  const long long rounds_before = jacobi_rounds_;
  const Element* el = elements_dev;
  if (prologue_queued_) prologue_queued_ = false;
  else scc_prologue();
  // C holds the previous eigenvectors from the second iteration on: A is
  // nearly diagonal, refine W = I towards its eigenvectors with GEMMs first
  const bool warm = iterations_ > 0 && oa_max_steps > 0;
  if (warm) {
    hipLaunchKernelGGL(k_ident_active, dim3(grid1((long long)max_nao * max_nao, 256), nmol), dim3(256), 0, stream,
                       d_desc, (const MolState*)d_state, W);
    refine(A, W);
  }
  // a molecule whose eigensolve gave up fails and stops iterating here
  auto fold_diag = [&]() {
    hipLaunchKernelGGL(k_fail_fold, dim3(grid1(nmol, 64)), dim3(64), 0, stream, d_state, nmol,
                       (const int*)jac_n.state + 1, kJacState, kFailDiag, (const int*)nullptr);
  };
  if (use_ql && !warm) {
    eigh_rows(jac_n, A, W, emo);
    fold_diag();
    hipLaunchKernelGGL(k_fail_fold, dim3(grid1(nmol, 64)), dim3(64), 0, stream, d_state, nmol,
                       (const int*)eig_failed, 1, kFailDiag, (const int*)nullptr);
    gemm(4, C, W, T, nullptr, perm, true);    // T = C W(perm, :)^T
  } else {
    const bool rows = jacobi(jac_n, A, W, emo, true, warm, warm ? banded_passes : 0);
    fold_diag();
    if (rows)
      gemm(4, C, W, T, nullptr, perm, true);  // T = C W(perm, :)^T
    else
      gemm(3, C, W, T, nullptr, perm, true);  // T = C W(:, perm)
  }
  hipLaunchKernelGGL(k_copy_active, dim3(gstride((long long)max_nao * max_nao), nmol), dim3(256), 0, stream,
                     d_desc, (const MolState*)d_state, (const double*)T, C);
  hipLaunchKernelGGL(k_occ_fermi, dim3(nmol), dim3(256), 0, stream, d_desc, d_state, (const double*)emo, focca, foccb,
                     focc, opt.etemp, (const double*)qsh, (const double*)dipm, (const double*)qp, q_in);
  mixed_here_ = true;
  gemm(2, C, C, P, focc, nullptr, true, 1.0, nullptr, true);      // P = C diag(f) C^T (symmetric)
  mixed_here_ = false;
  hipLaunchKernelGGL(k_population, dim3(grid1(max_nao, 4), nmol), dim3(256), 0, stream, d_desc, (const double*)P,
                     (const double*)S, (const double*)H0, (const double*)D, (const double*)Q, row_q, row_e, row_d,
                     row_qp, (const MolState*)d_state, 0);
  hipLaunchKernelGGL(k_pop_reduce, dim3(grid1(max_n, 64), nmol), dim3(64), 0, stream, d_desc, el, at, xyz, at_sh0,
                     at_ao0, (const double*)zsh, (const double*)row_q, (const double*)row_d, (const double*)row_qp,
                     qsh, q, dipm, qp, (const MolState*)d_state, 0);
  // the AES energy terms on stream2 beside the shell and D4 energy terms
  HIPCHECK(hipEventRecord(ev_fork, stream));
  HIPCHECK(hipStreamWaitEvent(stream2, ev_fork, 0));
  hipLaunchKernelGGL(k_aes_energy, dim3(std::max(1, max_n), nmol), dim3(64), 0, stream2, d_desc, el, at, xyz, q, dipm,
                     qp, gab3, gab5, ae01, (const MolState*)d_state);
  hipLaunchKernelGGL(k_shell_pot, dim3(grid1(max_nsh, 4), nmol), dim3(256), 0, stream, d_desc, (const double*)J,
                     (const double*)qsh, (const double*)shgam, jq, 0, (const MolState*)d_state);
  hipLaunchKernelGGL(k_d4_zeta, dim3(grid1(max_nd4, 64), nmol), dim3(64), 0, stream, d_desc, el, at, d4_at, d4_ref,
                     gw, q, zv, (double*)nullptr, (const MolState*)d_state);
  hipLaunchKernelGGL(k_d4_symv, dim3(grid1(max_nd4, 4), nmol), dim3(256), 0, stream, d_desc, (const double*)dispmat,
                     (const double*)zv, d4dum, (const MolState*)d_state);
  if (solv_on_)
    hipLaunchKernelGGL(k_solv_energy, dim3(nmol), dim3(256), 0, stream, d_desc, (const double*)sv_B, (const double*)q,
                       sv_qbq, (const MolState*)d_state);
  HIPCHECK(hipEventRecord(ev_join, stream2));
  HIPCHECK(hipStreamWaitEvent(stream, ev_join, 0));
  hipLaunchKernelGGL(k_scc_control, dim3(nmol), dim3(256), 0, stream, d_desc, d_state, (const double*)jq,
                     (const double*)shgam, (const double*)row_e, (const double*)ae01, (const double*)zv,
                     (const double*)d4dum, solv_on_ ? (const double*)sv_qbq : nullptr, (const double*)sv_gsasa,
                     sv_gshift, sh_at, qsh, q, dipm, qp, q_in, dq, qlast, dqlast, bdf, bu, ba, bomega, scfconv, qconv);
  HIPCHECK(hipGetLastError());
  iterations_ += 1;
  if (trace_rounds)
    std::fprintf(stderr, "scc iteration %d: %lld Jacobi rounds\n", iterations_, jacobi_rounds_ - rounds_before);
  // any molecule still active?
  hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, stream, (const MolState*)d_state, nmol, jac_nactive);
  HIPCHECK(hipMemcpyAsync(h_pinned, jac_nactive, sizeof(int), hipMemcpyDeviceToHost, stream));
  if (queue_next) {
    // the next iteration's first part keeps the GPU busy while the host
    // waits for the count (it changes nothing if no molecule is active)
    HIPCHECK(hipEventRecord(ev_join3, stream));
    scc_prologue();
    prologue_queued_ = true;
    HIPCHECK(hipEventSynchronize(ev_join3));
  } else {
    HIPCHECK(hipStreamSynchronize(stream));
  }
  return h_pinned[0] > 0;
}

void Batch::gradient() {
  if (nsub == 0) return;
  const Element* el = elements_dev;
  const int ns = nsub;
  // energy-weighted density into T
  hipLaunchKernelGGL(k_pew_scale, dim3(grid1(sub_max_nao, 64), ns), dim3(64), 0, stream, d_sub, (const double*)focc,
                     (const double*)emo, row_pd);
  gemm(2, C, C, T, row_pd, nullptr, false, 1.0, phase_act, true);   // symmetric
  hipLaunchKernelGGL(k_aes_pot, dim3(std::max(1, sub_max_n), ns), dim3(64), 0, stream, d_sub, el, at, xyz, q, dipm, qp,
                     gab3, gab5, 0, vs, vd, vq, (const MolState*)nullptr);
  hipLaunchKernelGGL(k_zero_atoms, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub, grad,
                     (double*)nullptr);
  const dim3 gat(sub_max_n, ns);
  // the H0 pair pass (writes only the pair buffer) on stream2 beside the AES
  // and D4 gradients; its per-atom sums follow them
  HIPCHECK(hipEventRecord(ev_fork, stream));
  HIPCHECK(hipStreamWaitEvent(stream2, ev_fork, 0));
  if (sub_max_psp > 0)
    hipLaunchKernelGGL(k_grad_h0_pairs<1>, dim3(grid1(sub_max_psp, 64), ns), dim3(64), 0, stream2, d_sub, el, at, xyz,
                       at_sh0, at_ao0, (const double*)selfe, (const double*)dsedcn, (const double*)P,
                       (const double*)T, (const double*)H0n, (const double*)S, (const double*)shift, (const double*)vs,
                       (const double*)vd, (const double*)vq, intcut, (const int2*)pair_list, gpair);
  if (sub_max_pd > 0)
    hipLaunchKernelGGL(k_grad_h0_pairs<2>, dim3(grid1(sub_max_pd, 64), ns), dim3(64), 0, stream2, d_sub, el, at, xyz,
                       at_sh0, at_ao0, (const double*)selfe, (const double*)dsedcn, (const double*)P,
                       (const double*)T, (const double*)H0n, (const double*)S, (const double*)shift, (const double*)vs,
                       (const double*)vd, (const double*)vq, intcut, (const int2*)pair_list, gpair);
  hipLaunchKernelGGL(k_aes_grad, gat, dim3(64), 0, stream, d_sub, xyz, q, dipm, qp, radcn, grad, tmp2);
  // the ATM terms from the set-up are done (and the weights below rewrite
  // their zero-charge inputs with the same values)
  if (stream3 != stream) HIPCHECK(hipStreamWaitEvent(stream, ev_atm, 0));
  hipLaunchKernelGGL(k_d4_refweights, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub, el, at, cn_cov, q,
                     zetavec, zerovec, zetadcn, zerodcn);
  hipLaunchKernelGGL(k_d4_c6<2>, dim3(gstride((long long)sub_max_n * (sub_max_n + 1) / 2), ns), dim3(256), 0, stream,
                     d_sub, el, at, c6ref, zetavec, zerovec, zetadcn, zerodcn, c6, dc6, c60, dc60);
  hipLaunchKernelGGL(k_d4_grad2, gat, dim3(64), 0, stream, d_sub, el, at, xyz, (const double*)c6, (const double*)dc6,
                     grad, dedcn);
  hipLaunchKernelGGL(k_add_atm, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub,
                     (const double*)grad_atm, (const double*)dedcn_atm, grad, dedcn);
  HIPCHECK(hipEventRecord(ev_join, stream2));
  HIPCHECK(hipStreamWaitEvent(stream, ev_join, 0));
  hipLaunchKernelGGL(k_grad_h0, gat, dim3(64), 0, stream, d_sub, el, at, xyz, at_sh0, at_ao0, (const double*)dsedcn,
                     (const double*)P, (const double*)gpair, grad, dhdcn);
  hipLaunchKernelGGL(k_grad_pairs, gat, dim3(64), 0, stream, d_sub, el, at, xyz, at_sh0, (const double*)qsh, grad,
                     erep);
  hipLaunchKernelGGL(k_cn_chain, gat, dim3(64), 0, stream, d_sub, el, at, xyz, (const double*)cn,
                     (const double*)dhdcn, (const double*)tmp2, (const double*)dedcn, grad);
  // implicit solvent: kernel, Born radius, surface, H-bond and shape terms, and
  // the solvation free energy at the final charges
  if (solv_on_) {
    hipLaunchKernelGGL(k_solv_grddb, dim3(grid1(sub_max_n, 64), ns), dim3(64), 0, stream, d_sub,
                       (const SolvDev*)d_solv, xyz, (const double*)q, (const double*)sv_brad, sv_grddb);
    hipLaunchKernelGGL(k_solv_grad, gat, dim3(64), 0, stream, d_sub, (const SolvDev*)d_solv, at, xyz, (const double*)q,
                       (const double*)sv_brad, (const double*)sv_dpsi, (const double*)sv_grddb, (const double*)sv_w,
                       grad, (const int*)nullptr);
    hipLaunchKernelGGL(k_solv_alpb_grad, dim3(ns), dim3(64), 0, stream, d_sub, (const SolvDev*)d_solv, at, xyz,
                       (const double*)q, grad, (const int*)nullptr);
    hipLaunchKernelGGL(k_solv_energy, dim3(ns), dim3(256), 0, stream, d_sub, (const double*)sv_B, (const double*)q,
                       sv_qbq, (const MolState*)nullptr);
  }
  // bond orders from P S; open-shell molecules from the spin densities
  // Pa = C fa C^T and Pb = C fb C^T (src/scf_module.F90:856-873)
  gemm(0, P, S, A, nullptr, nullptr, false, 1.0, phase_act);
  hipLaunchKernelGGL(k_wbo, dim3(grid1((long long)sub_max_n * sub_max_n, 256), ns), dim3(256), 0, stream, d_sub, xyz,
                     at_ao0, at, el, (const double*)A, wbo, 0, (const int*)nullptr);
  if (nsub_open > 0)
    for (int spin = 0; spin < 2; ++spin) {
      gemm(2, C, C, T, spin ? foccb : focca, nullptr, false, 1.0, open_act, true);   // spin density
      gemm(0, T, S, A, nullptr, nullptr, false, 1.0, open_act);
      hipLaunchKernelGGL(k_wbo, dim3(grid1((long long)sub_max_n * sub_max_n, 256), ns), dim3(256), 0, stream, d_sub,
                         xyz, at_ao0, at, el, (const double*)A, wbo, 1 + spin, (const int*)open_act);
    }
  if (stream3 != stream) HIPCHECK(hipStreamWaitEvent(stream, ev_atm, 0));   // the ATM energies
  hipLaunchKernelGGL(k_final, dim3(ns), dim3(256), 0, stream, d_sub, d_state, el, at, xyz, (const double*)grad,
                     (const double*)erep, (const double*)eat, (const double*)row_d);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream));
}

void Batch::run() {
  select_all();
  setup();
  int guard = 0;
  prologue_queued_ = false;
  // (no iteration is queued ahead of the guard's stop)
  while (scc_iteration(guard + 1 <= opt.maxiter + 2)) {
    if (++guard > opt.maxiter + 2) break;
  }
  prologue_queued_ = false;
  gradient();
  // the final shell charges and moments of every molecule are on the device
  wfn_valid_ = true;
}

void Batch::set_positions(int m, const double* xyz_in) {
  if (m < 0 || m >= nmol || !occupied_[m]) throw std::runtime_error("invalid molecule index");
  const MolDesc& d = desc[m];
  HostMol& hm = hmol_[m];
  hm.xyz.assign(xyz_in, xyz_in + 3 * (size_t)d.n);
  HIPCHECK(hipStreamSynchronize(stream));
  HIPCHECK(hipMemcpy(xyz + 3 * (size_t)d.a0, hm.xyz.data(), 3 * (size_t)d.n * sizeof(double),
                     hipMemcpyHostToDevice));
  setup_done_ = false;
}

void Batch::charges(int m, double* out) const {
  HIPCHECK(hipStreamSynchronize(stream));
  HIPCHECK(hipMemcpy(out, q + desc[m].a0, desc[m].n * sizeof(double), hipMemcpyDeviceToHost));
}

MolResult Batch::result(int m, bool matrices) const {
  MolResult r;
  const MolDesc& d = desc[m];
  MolState st;
  HIPCHECK(hipMemcpy(&st, d_state + m, sizeof(MolState), hipMemcpyDeviceToHost));
  r.n = d.n;
  r.nsh = d.nsh;
  r.nao = d.nao;
  r.iterations = st.iter;
  r.converged = st.converged;
  r.failed = st.failed;
  r.ihomo = st.ihomo;
  r.energy = st.etotal;
  r.gnorm = st.gnorm;
  r.hl_gap_ev = st.egap;
  r.e_elec = st.eel + st.embd;
  r.e_es = st.ees;
  r.e_aes = st.eaes;
  r.e_axc = st.epol;
  r.e_disp = st.ed + st.embd;
  r.e_rep = st.ep;
  r.e_atm = st.embd;
  if (solv_on_) {
    double qbq = 0.0, gs = 0.0;
    int bad = 0;
    HIPCHECK(hipMemcpy(&qbq, sv_qbq + m, sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(&gs, sv_gsasa + m, sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(&bad, sv_failed + m, sizeof(int), hipMemcpyDeviceToHost));
    r.e_solv = (qbq + gs) + sv_gshift;
    if (bad && !r.failed) r.failed = kFailSolv;
  }
  for (int k = 0; k < 3; ++k) r.dipole[k] = st.dipole[k];
  auto get = [&](std::vector<double>& v, const double* src, long long off, long long cnt) {
    v.resize(cnt);
    HIPCHECK(hipMemcpy(v.data(), src + off, cnt * sizeof(double), hipMemcpyDeviceToHost));
  };
  get(r.gradient, grad, 3LL * d.a0, 3LL * d.n);
  get(r.q, q, d.a0, d.n);
  get(r.qsh, qsh, d.s0, d.nsh);
  get(r.dipm, dipm, 3LL * d.a0, 3LL * d.n);
  get(r.qp, qp, 6LL * d.a0, 6LL * d.n);
  get(r.wbo, wbo, d.p0, (long long)d.n * d.n);
  get(r.emo, emo, d.o0, d.nao);
  get(r.focc, focc, d.o0, d.nao);
  if (matrices) {
    get(r.coeff, C, d.m0, (long long)d.nao * d.nao);
    get(r.density, P, d.m0, (long long)d.nao * d.nao);
  }
  return r;
}

// ---------------------------------------------------------------------------
// continuous batching
__global__ void k_active_flags(const MolState* __restrict__ st, int nmol, int* out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nmol) out[m] = st[m].active;
}

Capacity server_capacity(double hbm_bytes, int max_atoms, int max_slots, int maxiter) {
  MolDesc d;
  std::memset(&d, 0, sizeof(d));
  d.n = max_atoms;
  d.nsh = 2 * max_atoms;
  d.nao = 4 * max_atoms;
  d.nd4 = 7 * max_atoms;
  d.ne = d.n + 1;
  d.nbr = d.nsh + 9 * d.n;
  d.maxiter = maxiter;
  Capacity f = Batch::footprint(d);
  // bytes per element of each packed array (the buffers allocate() places there)
  const double per[13] = {8.0 * (3 + 4 + 1 + 3 + 6 + 1 + 1 + 3 + 6 + 1 + 3 + 3 + 2 + 4 + 4 * XTB_MAX_REF + 1) + 4.0 * 3,
                          8.0 * 7 + 4.0 * 4, 8.0 * 16 + 4.0 * 3, 8.0 * 4 + 4.0 * 2, 8.0 * 22, 8.0, 8.0,
                          8.0 * 13, 8.0, 16.0, 32.0, 16.0, 8.0};
  double bytes = 0.0;
  for (int k = 0; k < 13; ++k) bytes += per[k] * (double)*cap_field(f, k);
  const long long copies = std::max(1LL, (long long)(hbm_bytes / bytes));
  Capacity c;
  for (int k = 0; k < 13; ++k) *cap_field(c, k) = copies * *cap_field(f, k);
  c.slots = max_slots;
  return c;
}

Server::Server(const Capacity& cap, const Options& opt) : batch(cap, opt) {
  slot_req_.assign(batch.nmol, -1);
  active_host_.assign(batch.nmol, 0);
  batch.oa_max_steps = 0;   // per-batch warm-start refinement is not used here
  HIPCHECK(hipMalloc((void**)&d_active_, batch.nmol * sizeof(int)));
  batch.allocs_.push_back(d_active_);
}

long long Server::submit(const MoleculeInput& mi) {
  // validate now so that a bad request fails its caller, not a later step
  MolDesc d;
  HostMol hm;
  MolState st;
  Batch::describe(mi, batch.opt.maxiter, d, hm, st);
  // a request that can never fit the packed arrays is refused here instead of
  // waiting in the queue for room that will not come
  const Capacity f = Batch::footprint(d);
  const long long need[13] = {f.n, f.nsh, f.nao, f.nd4, f.mm, f.jj, f.xx, f.pp, f.ee, f.ne, f.v, f.h, f.g};
  const long long have[13] = {batch.cap_.n, batch.cap_.nsh, batch.cap_.nao, batch.cap_.nd4, batch.cap_.mm,
                              batch.cap_.jj, batch.cap_.xx, batch.cap_.pp, batch.cap_.ee, batch.cap_.ne,
                              batch.cap_.v, batch.cap_.h, batch.cap_.g};
  for (int k = 0; k < 13; ++k)
    if (need[k] > have[k]) throw std::runtime_error("request larger than the HBM budget of the server");
  const long long id = next_id_++;
  queue_.emplace_back(id, mi);
  return id;
}

int Server::running() const {
  int c = 0;
  for (long long r : slot_req_) c += r >= 0 ? 1 : 0;
  return c;
}

int Server::step() {
  // 1. admission: first fit over the queue (a large request does not block
  //    smaller ones behind it; the queue order is kept otherwise)
  std::vector<int> fresh;
  for (auto it = queue_.begin(); it != queue_.end();) {
    const int slot = batch.admit(it->second);
    if (slot >= 0) {
      slot_req_[slot] = it->first;
      fresh.push_back(slot);
      it = queue_.erase(it);
    } else {
      ++it;
    }
  }
  if (fresh.empty() && running() == 0 && !queue_.empty())
    throw std::runtime_error("request larger than the HBM budget of the server");
  if (!fresh.empty()) {
    batch.rebuild();
    batch.select(fresh);
    batch.setup();
  }
  if (running() == 0) return queued();
  // 2. one SCC iteration of every resident molecule
  batch.scc_iteration();
  // 3. completion: gradient and results of the molecules that stopped iterating
  hipLaunchKernelGGL(k_active_flags, dim3((batch.nmol + 255) / 256), dim3(256), 0, batch.stream,
                     (const MolState*)batch.d_state, batch.nmol, d_active_);
  HIPCHECK(hipMemcpyAsync(active_host_.data(), d_active_, batch.nmol * sizeof(int), hipMemcpyDeviceToHost,
                          batch.stream));
  HIPCHECK(hipStreamSynchronize(batch.stream));
  std::vector<int> fin;
  for (int m = 0; m < batch.nmol; ++m)
    if (slot_req_[m] >= 0 && !active_host_[m]) fin.push_back(m);
  if (!fin.empty()) {
    batch.select(fin);
    batch.gradient();
    for (int m : fin) {
      done_[slot_req_[m]] = batch.result(m, false);
      finished_.push_back(slot_req_[m]);
      slot_req_[m] = -1;
      batch.release(m);
    }
    batch.rebuild();
  }
  return running() + queued();
}

std::vector<long long> Server::take_finished() {
  std::vector<long long> f;
  f.swap(finished_);
  return f;
}

bool Server::result(long long id, MolResult& out, bool erase) {
  auto it = done_.find(id);
  if (it == done_.end()) return false;
  out = it->second;
  if (erase) done_.erase(it);
  return true;
}

}  // namespace xtbamd

// ===========================================================================
// C ABI of the batch interface (include/xtb_amd_batch.h)
#include "../../include/xtb_amd_batch.h"

struct xtbamd_batch_s {
  xtbamd::Batch* batch = nullptr;
};

namespace {
thread_local std::string g_last_error;
int fail(const std::exception& e) {
  g_last_error = e.what();
  return -1;
}

// Scratch area of xtbamd_singlepoint_noalloc: every device the library runs
// on holds its own copy, for the life of the process.  The name is only a
// host-side handle; its device address comes from hipGetSymbolAddress.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kNoAllocDevices = 64;
struct NoAllocDevice {
  std::mutex mu;          // calls on one device run one at a time
  bool ready = false;
  xtbamd::Loan loan;
  int words[64] = {};     // host words of the active-count reads
};
NoAllocDevice g_noalloc[kNoAllocDevices];

void copy_result(const xtbamd::MolResult& r, xtbamd_result* out) {
  out->natoms = r.n;
  out->nshells = r.nsh;
  out->nao = r.nao;
  out->iterations = r.iterations;
  out->converged = r.converged;
  out->failed = r.failed;
  out->ihomo = r.ihomo;
  out->energy = r.energy;
  out->gnorm = r.gnorm;
  out->hl_gap_ev = r.hl_gap_ev;
  out->e_elec = r.e_elec;
  out->e_es = r.e_es;
  out->e_aes = r.e_aes;
  out->e_axc = r.e_axc;
  out->e_disp = r.e_disp;
  out->e_rep = r.e_rep;
  out->e_atm = r.e_atm;
  out->e_solv = r.e_solv;
  for (int k = 0; k < 3; ++k) out->dipole[k] = r.dipole[k];
  auto put = [](double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  put(out->gradient, r.gradient);
  put(out->charges, r.q);
  put(out->qsh, r.qsh);
  put(out->dipm, r.dipm);
  put(out->qp, r.qp);
  put(out->wbo, r.wbo);
  put(out->emo, r.emo);
  put(out->focc, r.focc);
  put(out->coeff, r.coeff);
  put(out->density, r.density);
}
}  // namespace

extern "C" {

const char* xtbamd_last_error(void) { return g_last_error.c_str(); }

int xtbamd_copy_probe(long long bytes) {
  try {
    const long long n = std::max(1LL, bytes / (long long)sizeof(double));
    double *a = nullptr, *b = nullptr;
    HIPCHECK(hipMalloc(&a, n * sizeof(double)));
    HIPCHECK(hipMalloc(&b, n * sizeof(double)));
    hipLaunchKernelGGL(xtbamd::k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, n, 1.0);
    hipLaunchKernelGGL(xtbamd::k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (const double*)a, b, n);
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipFree(a));
    HIPCHECK(hipFree(b));
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

xtbamd_batch xtbamd_batch_new(int nmol, const int* natoms, const int* numbers, const double* positions,
                              const double* charges, const int* uhf, double accuracy, double etemp, int maxiter,
                              int guess) {
  try {
    if (nmol <= 0 || !natoms || !numbers || !positions) throw std::runtime_error("invalid batch arguments");
    std::vector<xtbamd::MoleculeInput> mols(nmol);
    long long off = 0;
    for (int m = 0; m < nmol; ++m) {
      const int n = natoms[m];
      if (n <= 0) throw std::runtime_error("molecule without atoms");
      mols[m].numbers.assign(numbers + off, numbers + off + n);
      mols[m].xyz.assign(positions + 3 * off, positions + 3 * (off + n));
      mols[m].charge = charges ? charges[m] : 0.0;
      mols[m].uhf = uhf ? uhf[m] : 0;
      off += n;
    }
    xtbamd::Options o;
    o.accuracy = accuracy;
    o.etemp = etemp;
    o.maxiter = maxiter;
    o.guess = guess;
    auto* h = new xtbamd_batch_s;
    h->batch = new xtbamd::Batch(mols, o);
    return h;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return nullptr;
  }
}

int xtbamd_batch_set_guess(xtbamd_batch b, int m, const double* qsh, const double* dipm, const double* qp) {
  try {
    if (!b || !qsh || !dipm || !qp) throw std::runtime_error("invalid guess arguments");
    b->batch->set_guess(m, qsh, dipm, qp);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_set_solvent(xtbamd_batch b, const char* solvent, int kernel, int alpb, int state,
                             double temperature) {
  try {
    if (!b) throw std::runtime_error("invalid batch");
    if (state < 1 || state > 3) throw std::runtime_error("unknown solution state");
    b->batch->set_solvent(solvent ? std::string(solvent) : std::string(), kernel, alpb != 0, state, temperature);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_solvent_known(const char* solvent, int kernel) {
  if (!solvent || (kernel != 1 && kernel != 2)) return 0;
  return xtbamd::solv_param_set(solvent, kernel) ? 1 : 0;
}

int xtbamd_batch_set_positions(xtbamd_batch b, int m, const double* positions) {
  try {
    if (!b || !positions) throw std::runtime_error("invalid position arguments");
    b->batch->set_positions(m, positions);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_keep_wavefunction(xtbamd_batch b, int on) {
  if (!b) return -1;
  b->batch->keep_wavefunction(on != 0);
  return 0;
}

int xtbamd_batch_setup(xtbamd_batch b) {
  try {
    b->batch->select_all();
    b->batch->setup();
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_iterate(xtbamd_batch b) {
  try {
    const bool any = b->batch->scc_iteration();
    return any ? 1 : 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_gradient(xtbamd_batch b) {
  try {
    b->batch->select_all();
    b->batch->gradient();
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_run(xtbamd_batch b) {
  try { b->batch->run(); return 0; } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_counters(xtbamd_batch b, int* scc_steps, long long* jacobi_rounds) {
  if (!b) return -1;
  if (scc_steps) *scc_steps = b->batch->iterations_;
  if (jacobi_rounds) *jacobi_rounds = b->batch->jacobi_rounds_;
  return 0;
}

int xtbamd_batch_profile(xtbamd_batch b, int kernels) {
  try {
    if (!b || kernels < 0 || kernels > 7) throw std::runtime_error("unknown kernel family");
    b->batch->prof_enable(kernels << 1);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_profile_span(xtbamd_batch b, int kernel, double* span_ms) {
  try {
    if (!b || !span_ms) throw std::runtime_error("invalid batch");
    *span_ms = b->batch->prof_span(kernel);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_profile_result(xtbamd_batch b, int kernel, int* launches, double* total_ms, double* work) {
  try {
    if (!b) throw std::runtime_error("invalid batch");
    return b->batch->prof_result(kernel, launches, total_ms, work);
  } catch (const std::exception& e) { return fail(e); }
}

struct xtbamd_server_s {
  xtbamd::Server* server = nullptr;
  std::vector<long long> finished;   // ids finished but not yet returned by xtbamd_server_finished
};

xtbamd_server xtbamd_server_new(double hbm_gb, int max_atoms, int max_slots, double accuracy, double etemp,
                                int maxiter) {
  try {
    if (hbm_gb <= 0 || max_atoms <= 0 || max_slots <= 0) throw std::runtime_error("invalid server arguments");
    xtbamd::Options o;
    o.accuracy = accuracy;
    o.etemp = etemp;
    o.maxiter = maxiter;
    xtbamd::Capacity c = xtbamd::server_capacity(hbm_gb * 1073741824.0, max_atoms, max_slots, maxiter);
    auto* h = new xtbamd_server_s;
    h->server = new xtbamd::Server(c, o);
    return h;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return nullptr;
  }
}

long long xtbamd_server_submit(xtbamd_server s, int natoms, const int* numbers, const double* positions,
                               double charge, int uhf) {
  try {
    if (!s || natoms <= 0 || !numbers || !positions) throw std::runtime_error("invalid request");
    xtbamd::MoleculeInput mi;
    mi.numbers.assign(numbers, numbers + natoms);
    mi.xyz.assign(positions, positions + 3 * natoms);
    mi.charge = charge;
    mi.uhf = uhf;
    return s->server->submit(mi);
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return -1;
  }
}

int xtbamd_server_step(xtbamd_server s) {
  try {
    if (!s) throw std::runtime_error("invalid server");
    return s->server->step();
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_server_finished(xtbamd_server s, long long* ids, int max_ids) {
  try {
    if (!s) throw std::runtime_error("invalid server");
    std::vector<long long>& pending = s->finished;
    std::vector<long long> f = s->server->take_finished();
    pending.insert(pending.end(), f.begin(), f.end());
    const int n = std::min<int>(max_ids, (int)pending.size());
    for (int i = 0; i < n; ++i) ids[i] = pending[i];
    pending.erase(pending.begin(), pending.begin() + n);
    return n;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_server_result(xtbamd_server s, long long id, xtbamd_result* out) {
  try {
    if (!s || !out) throw std::runtime_error("invalid result request");
    xtbamd::MolResult r;
    const bool sizes_only = !out->gradient && !out->charges && !out->qsh && !out->dipm && !out->qp && !out->wbo &&
                            !out->emo && !out->focc;
    if (!s->server->result(id, r, !sizes_only)) throw std::runtime_error("no finished result with this id");
    copy_result(r, out);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_server_pending(xtbamd_server s, int* queued, int* running) {
  if (!s) return -1;
  if (queued) *queued = s->server->queued();
  if (running) *running = s->server->running();
  return 0;
}

int xtbamd_server_set_solvent(xtbamd_server s, const char* solvent, int kernel, int alpb, int state,
                              double temperature) {
  try {
    if (!s) throw std::runtime_error("invalid server");
    if (s->server->running() > 0) throw std::runtime_error("the solvent can only change while no request is running");
    if (state < 1 || state > 3) throw std::runtime_error("unknown solution state");
    s->server->batch.set_solvent(solvent ? std::string(solvent) : std::string(), kernel, alpb != 0, state,
                                 temperature);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

void xtbamd_server_free(xtbamd_server s) {
  if (!s) return;
  delete s->server;
  delete s;
}

int xtbamd_singlepoint_noalloc(int natoms, const int* numbers, const double* positions, double charge, int uhf,
                               double accuracy, double etemp, int maxiter, xtbamd_result* out) {
  try {
    if (natoms <= 0 || !numbers || !positions || !out || maxiter <= 0)
      throw std::runtime_error("invalid single point arguments");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) throw std::runtime_error("device ordinal out of range");
    NoAllocDevice& nd = g_noalloc[dev];
    std::lock_guard<std::mutex> lock(nd.mu);
    if (!nd.ready) {
      // set-up, once per device with the device current: the device address
      // of this device's copy of the scratch array and the fork/join events
      // (kept until the process ends).  Every stream of the lent batch is
      // the null stream: one molecule gains little from stream concurrency,
      // and streams of their own would take hardware queues from the
      // caller's other work.
      void* ptr = nullptr;
      HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      if (!ptr) throw std::runtime_error("scratch area has no device address");
      nd.loan.base = (char*)ptr;
      nd.loan.bytes = kScratchBytes;
      for (int k = 0; k < 4; ++k) nd.loan.streams[k] = nullptr;
      for (int k = 0; k < 4; ++k) HIPCHECK(hipEventCreateWithFlags(&nd.loan.events[k], hipEventDisableTiming));
      nd.loan.host_words = nd.words;
      nd.ready = true;
    }
    std::vector<xtbamd::MoleculeInput> mols(1);
    mols[0].numbers.assign(numbers, numbers + natoms);
    mols[0].xyz.assign(positions, positions + 3 * (size_t)natoms);
    mols[0].charge = charge;
    mols[0].uhf = uhf;
    xtbamd::Options o;
    o.accuracy = accuracy;
    o.etemp = etemp;
    o.maxiter = maxiter;
    o.guess = 0;
    xtbamd::Batch b(mols, o, nd.loan);
    b.run();
    copy_result(b.result(0, out->coeff || out->density), out);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_eigh(int nprob, const int* n, const double* a, double* evals, double* evecs, double* tri) {
  std::vector<void*> owned;
  struct HipMem : xtbamd::EigMem {
    std::vector<void*>* owned;
    void* get(size_t bytes) override {
      void* p = nullptr;
      HIPCHECK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
      owned->push_back(p);
      return p;
    }
    void put(void* p) override {
      auto it = std::find(owned->begin(), owned->end(), p);
      if (it != owned->end()) owned->erase(it);
      (void)hipFree(p);
    }
  } mem;
  mem.owned = &owned;
  auto release = [&]() {
    (void)hipDeviceSynchronize();
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
  };
  try {
    if (nprob <= 0 || !n || !a || !evals || !evecs) throw std::runtime_error("invalid eigenproblem arguments");
    std::vector<xtbamd::JacobiProb> jp(nprob);
    std::vector<int> dims(nprob);
    long long off = 0, voff = 0;
    for (int k = 0; k < nprob; ++k) {
      if (n[k] < 2) throw std::runtime_error("eigenproblems of dimension < 2 are not supported");
      std::memset(&jp[k], 0, sizeof(jp[k]));
      jp[k].dim = dims[k] = n[k];
      jp[k].ld = n[k];
      jp[k].a_off = jp[k].w_off = off;
      off += (long long)n[k] * n[k];
    }
    double* dA = (double*)mem.get(off * sizeof(double));
    double* dW = (double*)mem.get(off * sizeof(double));
    int* dst = (int*)mem.get((size_t)nprob * xtbamd::kJacState * sizeof(int));
    xtbamd::JacobiProb* djp = (xtbamd::JacobiProb*)mem.get(nprob * sizeof(xtbamd::JacobiProb));
    HIPCHECK(hipMemcpy(dA, a, off * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(djp, jp.data(), nprob * sizeof(xtbamd::JacobiProb), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(dst, 0, (size_t)nprob * xtbamd::kJacState * sizeof(int)));
    xtbamd::EigSet es;
    for (int k = 0; k < nprob; ++k) es.mols.push_back(k);
    xtbamd::eig_set_build(es, dims, mem);
    xtbamd::eig_set_run(es, djp, dst, nullptr, dA, dW, nullptr);
    HIPCHECK(hipDeviceSynchronize());
    std::vector<double> h((size_t)off);
    HIPCHECK(hipMemcpy(h.data(), dA, off * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(evecs, dW, off * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> st((size_t)nprob * xtbamd::kJacState);
    HIPCHECK(hipMemcpy(st.data(), dst, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    long long eo = 0;
    for (int k = 0; k < nprob; ++k) {
      for (int i = 0; i < n[k]; ++i) evals[eo + i] = h[jp[k].a_off + (long long)i * n[k] + i];
      if (tri) {
        const long long ld = es.ld[k];
        HIPCHECK(hipMemcpy(tri + 2 * eo, es.vec + es.v_off[k] + xtbamd::kVd0 * ld, n[k] * sizeof(double),
                           hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(tri + 2 * eo + n[k], es.vec + es.v_off[k] + xtbamd::kVe0 * ld, n[k] * sizeof(double),
                           hipMemcpyDeviceToHost));
      }
      if (st[(size_t)k * xtbamd::kJacState + 1]) throw std::runtime_error("eigensolver produced non-finite values");
      eo += n[k];
    }
    (void)voff;
    release();
    return 0;
  } catch (const std::exception& e) {
    release();
    return fail(e);
  }
}

int xtbamd_batch_size(xtbamd_batch b) { return b ? b->batch->size() : 0; }

int xtbamd_batch_get(xtbamd_batch b, int m, xtbamd_result* out) {
  try {
    if (!b || !out || m < 0 || m >= b->batch->size() || !b->batch->occupied(m))
      throw std::runtime_error("invalid result request");
    const bool mats = out->coeff || out->density;
    xtbamd::MolResult r = b->batch->result(m, mats);
    copy_result(r, out);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

int xtbamd_batch_charges(xtbamd_batch b, int m, double* q) {
  try {
    if (!b || !q || m < 0 || m >= b->batch->size() || !b->batch->occupied(m))
      throw std::runtime_error("invalid charge request");
    b->batch->charges(m, q);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

void xtbamd_batch_free(xtbamd_batch b) {
  if (!b) return;
  delete b->batch;
  delete b;
}

}  // extern "C"
