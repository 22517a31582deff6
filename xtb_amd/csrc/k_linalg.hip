// Dense FP64 linear algebra of the SCC on MI355X:
//  * batched square GEMM over a (molecule, tile) work list, 64x64 tiles staged
//    through LDS, 4x4 register micro-tiles per thread;
//  * a block-Jacobi symmetric eigensolver.  The matrix is split into 32-wide
//    blocks; each round pairs the blocks by the circle method, diagonalises
//    every 64x64 pivot subproblem entirely in LDS (cyclic parallel Jacobi) and
//    applies the resulting orthogonal factors to A (two-sided) and to the
//    eigenvector matrix W as 64x64 tile products.  It replaces the reference's
//    dsygst/dsyevd/dtrsm sequence (src/mctc/lapack/eigensolve.F90:218-246).
#include "device_util.h"
#include "kernels.h"

namespace xtbamd {

constexpr int kTile = 64, kBK = 16;

// C = op(A) * op(B) [* diag(scale) on the k index] for square nao x nao
// matrices; optional column permutation of B (B(:, perm[j])).
template <bool TA, bool TB, bool SCALE, bool PERM>
__global__ void __launch_bounds__(256) k_gemm(const int4* __restrict__ tiles, const MolDesc* __restrict__ desc,
                                              const double* __restrict__ A, const double* __restrict__ B, double* C,
                                              const double* __restrict__ scale, const int* __restrict__ perm,
                                              double alpha, const MolState* __restrict__ mstate) {
  __shared__ double As[kBK][kTile + 1];
  __shared__ double Bs[kBK][kTile + 1];
  const int4 tl = tiles[blockIdx.x];
  if (mstate && !mstate[tl.x].active) return;
  const MolDesc md = desc[tl.x];
  const int N = md.nao;
  const double* Am = A + md.m0;
  const double* Bm = B + md.m0;
  double* Cm = C + md.m0;
  const double* sc = SCALE ? scale + md.o0 : nullptr;
  const int* pm = PERM ? perm + md.o0 : nullptr;
  const int row0 = tl.y * kTile, col0 = tl.z * kTile;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
  for (int kb = 0; kb < N; kb += kBK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int ii, kk;
      if (TA) { kk = e >> 6; ii = e & 63; } else { ii = e >> 4; kk = e & 15; }
      const int gi = row0 + ii, gk = kb + kk;
      double v = 0.0;
      if (gi < N && gk < N) v = TA ? Am[(long long)gk * N + gi] : Am[(long long)gi * N + gk];
      As[kk][ii] = v;
      int jj, k2;
      if (TB) { jj = e >> 4; k2 = e & 15; } else { k2 = e >> 6; jj = e & 63; }
      const int gj = col0 + jj, gk2 = kb + k2;
      double w = 0.0;
      if (gj < N && gk2 < N) {
        const int cj = PERM ? pm[gj] : gj;
        w = TB ? Bm[(long long)cj * N + gk2] : Bm[(long long)gk2 * N + cj];
        if (SCALE) w *= sc[gk2];
      }
      Bs[k2][jj] = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gi = row0 + ty + 16 * r;
    if (gi >= N) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int gj = col0 + tx + 16 * c;
      if (gj < N) Cm[(long long)gi * N + gj] = alpha * acc[r][c];
    }
  }
}

// instantiations used by the engine
template __global__ void k_gemm<false, false, false, false>(const int4*, const MolDesc*, const double*, const double*,
                                                            double*, const double*, const int*, double, const MolState*);
template __global__ void k_gemm<true, false, false, false>(const int4*, const MolDesc*, const double*, const double*,
                                                           double*, const double*, const int*, double, const MolState*);
template __global__ void k_gemm<false, true, true, false>(const int4*, const MolDesc*, const double*, const double*,
                                                          double*, const double*, const int*, double, const MolState*);
template __global__ void k_gemm<false, false, false, true>(const int4*, const MolDesc*, const double*, const double*,
                                                           double*, const double*, const int*, double, const MolState*);

// ---------------------------------------------------------------------------
// Cyclic parallel Jacobi on a 64x64 symmetric matrix held in LDS.
// a, q: 64 x 65 (padded rows).  Returns the number of rotations applied.
constexpr int kLd = 65;

__device__ int jacobi64(double* a, double* q, double tol, int maxsweep, double* cs, int* pq, int* flags) {
  const int tid = threadIdx.x;
  int total = 0;
  for (int sweep = 0; sweep < maxsweep; ++sweep) {
    if (tid == 0) flags[1] = 0;
    __syncthreads();
    for (int r = 0; r < 63; ++r) {
      if (tid < 64) {
        bool rot = false;
        if (tid < 32) {
          const int k = tid;
          int p, qq;
          if (k == 0) { p = 0; qq = 1 + r; }
          else { p = 1 + (r + k) % 63; qq = 1 + (r - k + 63) % 63; }
          if (p > qq) { const int t = p; p = qq; qq = t; }
          const double apq = a[p * kLd + qq];
          double c = 1.0, s = 0.0, t = 0.0;
          if (fabs(apq) > tol) {
            const double app = a[p * kLd + p], aqq = a[qq * kLd + qq];
            const double theta = (aqq - app) / (2.0 * apq);
            t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
            rot = true;
          }
          pq[2 * k] = p;
          pq[2 * k + 1] = qq;
          cs[3 * k] = c;
          cs[3 * k + 1] = s;
          cs[3 * k + 2] = t;
        }
        const unsigned long long any = __ballot(rot);
        if (tid == 0) {
          flags[0] = any ? 1 : 0;
          if (any) flags[1] += 1;
        }
      }
      __syncthreads();
      if (flags[0]) {
        for (int b = tid; b < 528; b += blockDim.x) {
          // triangular (k1 <= k2) index of the 32 x 32 pair blocks
          int k2 = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
          while (k2 * (k2 + 1) / 2 > b) --k2;
          while ((k2 + 1) * (k2 + 2) / 2 <= b) ++k2;
          const int k1 = b - k2 * (k2 + 1) / 2;
          const int p1 = pq[2 * k1], q1 = pq[2 * k1 + 1], p2 = pq[2 * k2], q2 = pq[2 * k2 + 1];
          const double c1 = cs[3 * k1], s1 = cs[3 * k1 + 1], c2 = cs[3 * k2], s2 = cs[3 * k2 + 1];
          if (k1 == k2) {
            if (s1 != 0.0) {
              const double t1 = cs[3 * k1 + 2];
              const double apq = a[p1 * kLd + q1];
              a[p1 * kLd + p1] = a[p1 * kLd + p1] - t1 * apq;
              a[q1 * kLd + q1] = a[q1 * kLd + q1] + t1 * apq;
              a[p1 * kLd + q1] = 0.0;
              a[q1 * kLd + p1] = 0.0;
            }
          } else if (s1 != 0.0 || s2 != 0.0) {
            const double apr = a[p1 * kLd + p2], aps = a[p1 * kLd + q2];
            const double aqr = a[q1 * kLd + p2], aqs = a[q1 * kLd + q2];
            // columns (pair 2), then rows (pair 1)
            const double bpr = c2 * apr - s2 * aps, bps = s2 * apr + c2 * aps;
            const double bqr = c2 * aqr - s2 * aqs, bqs = s2 * aqr + c2 * aqs;
            const double npr = c1 * bpr - s1 * bqr, nqr = s1 * bpr + c1 * bqr;
            const double nps = c1 * bps - s1 * bqs, nqs = s1 * bps + c1 * bqs;
            a[p1 * kLd + p2] = npr; a[p2 * kLd + p1] = npr;
            a[p1 * kLd + q2] = nps; a[q2 * kLd + p1] = nps;
            a[q1 * kLd + p2] = nqr; a[p2 * kLd + q1] = nqr;
            a[q1 * kLd + q2] = nqs; a[q2 * kLd + q1] = nqs;
          }
        }
        for (int e = tid; e < 64 * 32; e += blockDim.x) {
          const int u = e >> 5, k = e & 31;
          const double s = cs[3 * k + 1];
          if (s == 0.0) continue;
          const double c = cs[3 * k];
          const int p = pq[2 * k], qq = pq[2 * k + 1];
          const double up = q[u * kLd + p], uq = q[u * kLd + qq];
          q[u * kLd + p] = c * up - s * uq;
          q[u * kLd + qq] = s * up + c * uq;
        }
      }
      __syncthreads();
    }
    const int nrot = flags[1];
    __syncthreads();
    total += nrot;
    if (nrot == 0) break;
  }
  return total;
}

__device__ __forceinline__ void circle_pair(int nbp, int r, int k, int& bp, int& bq) {
  const int m = nbp - 1;
  if (k == 0) { bp = 0; bq = 1 + (r % m); }
  else { bp = 1 + (r + k) % m; bq = 1 + ((r - k) % m + m) % m; }
  if (bp > bq) { const int t = bp; bp = bq; bq = t; }
}

// state per problem: [0] converged, [1] failed, [2..4] rotation counts of sweeps mod 3, [5] sweeps
constexpr int kJacState = 8;
constexpr int kMaxSweeps = 40;

// problems of dimension <= 64: one workgroup solves to convergence in LDS
__global__ void __launch_bounds__(256) k_jac_small(const JacobiProb* __restrict__ probs, const int* __restrict__ list,
                                                   double* A, double* W, const double* __restrict__ scale_in,
                                                   int* state) {
  __shared__ double a[64 * kLd];
  __shared__ double q[64 * kLd];
  __shared__ double cs[96];
  __shared__ int pq[64];
  __shared__ int flags[2];
  const int pidx = list[blockIdx.x];
  const JacobiProb pr = probs[pidx];
  if (state[pidx * kJacState]) return;
  const int n = pr.dim;
  double* Ap = A + pr.a_off;
  double* Wp = W + pr.w_off;
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    a[u * kLd + v] = (u < n && v < n) ? Ap[(long long)u * n + v] : 0.0;
    q[u * kLd + v] = (u == v) ? 1.0 : 0.0;
  }
  __syncthreads();
  const double tol = 1e-12 * scale_in[pidx];
  const int nrot = jacobi64(a, q, tol, kMaxSweeps, cs, pq, flags);
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int u = e / n, v = e % n;
    Ap[(long long)u * n + v] = a[u * kLd + v];
    Wp[(long long)u * n + v] = q[u * kLd + v];
  }
  if (threadIdx.x == 0) {
    state[pidx * kJacState + 0] = 1;
    state[pidx * kJacState + 1] = 0;
    state[pidx * kJacState + 5] = nrot;
  }
}

// global index of the u-th row (0..63) of the pivot subproblem (bp, bq)
__device__ __forceinline__ int sub_index(int bp, int bq, int u) { return u < 32 ? 32 * bp + u : 32 * bq + (u - 32); }

__device__ __forceinline__ bool jac_skip(const JacobiProb& pr, int* st, int r, int k, bool bookkeeping,
                                         int* nactive) {
  if (st[0]) return true;
  const int s = r / pr.rounds, lr = r % pr.rounds;
  if (s >= 1 && lr == 0 && st[2 + (s - 1) % 3] == 0) {
    if (bookkeeping && k == 0) {
      st[0] = 1;
      st[5] = s;
      atomicSub(nactive, 1);
    }
    return true;
  }
  if (s >= 1 && st[2 + (s - 1) % 3] == 0) return true;
  if (s >= kMaxSweeps) {
    if (bookkeeping && k == 0 && lr == 0) {
      st[0] = 1;
      st[1] = 1;
      atomicSub(nactive, 1);
    }
    return true;
  }
  return false;
}

// one workgroup per (problem, pair): diagonalise the 64x64 pivot subproblem
__global__ void __launch_bounds__(256) k_jac_pairs(const JacobiProb* __restrict__ probs, const int2* __restrict__ list,
                                                   int r, const double* __restrict__ A, double* Qbuf, int* rotflag,
                                                   const double* __restrict__ scale_in, int* state, int* nactive) {
  __shared__ double a[64 * kLd];
  __shared__ double q[64 * kLd];
  __shared__ double cs[96];
  __shared__ int pq[64];
  __shared__ int flags[2];
  __shared__ int skip;
  const int2 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  int* st = state + it.x * kJacState;
  if (threadIdx.x == 0) {
    skip = jac_skip(pr, st, r, it.y, true, nactive) ? 1 : 0;
    const int s = r / pr.rounds, lr = r % pr.rounds;
    if (!skip && it.y == 0 && lr == 0) st[2 + (s + 1) % 3] = 0;
  }
  __syncthreads();
  if (skip) return;
  const int n = pr.dim;
  const int lr = r % pr.rounds, s = r / pr.rounds;
  int bp, bq;
  circle_pair(pr.nbp, lr, it.y, bp, bq);
  const double* Ap = A + pr.a_off;
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    const int gu = sub_index(bp, bq, u), gv = sub_index(bp, bq, v);
    a[u * kLd + v] = (gu < n && gv < n) ? Ap[(long long)gu * n + gv] : 0.0;
    q[u * kLd + v] = (u == v) ? 1.0 : 0.0;
  }
  __syncthreads();
  const double tol = 1e-12 * scale_in[it.x];
  const int nrot = jacobi64(a, q, tol, kMaxSweeps, cs, pq, flags);
  double* Qp = Qbuf + pr.q_off + (long long)it.y * 64 * 64;
  if (nrot > 0)
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) Qp[e] = q[(e >> 6) * kLd + (e & 63)];
  if (threadIdx.x == 0) {
    rotflag[(int)(pr.q_off / 4096) + it.y] = nrot > 0 ? 1 : 0;
    if (nrot > 0) atomicAdd(&st[2 + s % 3], 1);
  }
}

// A[I_k, I_l] <- Q_k^T A[I_k, I_l] Q_l for pair tiles k <= l
__global__ void __launch_bounds__(256) k_jac_apply_a(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                     int r, double* A, const double* __restrict__ Qbuf,
                                                     const int* __restrict__ rotflag, int* state) {
  __shared__ double t[64 * kLd];
  __shared__ double qk[64 * kLd];
  __shared__ double ql[64 * kLd];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  int* st = state + it.x * kJacState;
  if (jac_skip(pr, st, r, -1, false, nullptr)) return;
  const int k = it.y, l = it.z;
  const int qb = (int)(pr.q_off / 4096);
  const int rk = rotflag[qb + k], rl = rotflag[qb + l];
  if (!rk && !rl) return;
  const int n = pr.dim;
  const int lr = r % pr.rounds;
  int kp, kq, lp, lq;
  circle_pair(pr.nbp, lr, k, kp, kq);
  circle_pair(pr.nbp, lr, l, lp, lq);
  double* Ap = A + pr.a_off;
  const double* Qk = Qbuf + pr.q_off + (long long)k * 4096;
  const double* Ql = Qbuf + pr.q_off + (long long)l * 4096;
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    const int gu = sub_index(kp, kq, u), gv = sub_index(lp, lq, v);
    t[u * kLd + v] = (gu < n && gv < n) ? Ap[(long long)gu * n + gv] : 0.0;
    qk[u * kLd + v] = rk ? Qk[e] : (u == v ? 1.0 : 0.0);
    ql[u * kLd + v] = rl ? Ql[e] : (u == v ? 1.0 : 0.0);
  }
  __syncthreads();
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4];
  // M = A Q_l
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int w = 0; w < 64; ++w) {
    double x[4], y[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = t[(ty + 16 * i) * kLd + w];
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] = ql[w * kLd + tx + 16 * j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fma(x[i], y[j], acc[i][j]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) t[(ty + 16 * i) * kLd + tx + 16 * j] = acc[i][j];
  __syncthreads();
  // Q_k^T M
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int w = 0; w < 64; ++w) {
    double x[4], y[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = qk[w * kLd + ty + 16 * i];
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] = t[w * kLd + tx + 16 * j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fma(x[i], y[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int u = ty + 16 * i;
    const int gu = sub_index(kp, kq, u);
    if (gu >= n) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = tx + 16 * j;
      const int gv = sub_index(lp, lq, v);
      if (gv >= n) continue;
      if (k == l) {
        if (u <= v) { Ap[(long long)gu * n + gv] = acc[i][j]; Ap[(long long)gv * n + gu] = acc[i][j]; }
      } else {
        Ap[(long long)gu * n + gv] = acc[i][j];
        Ap[(long long)gv * n + gu] = acc[i][j];
      }
    }
  }
}

// W[rows, I_l] <- W[rows, I_l] Q_l  for 64-row blocks
__global__ void __launch_bounds__(256) k_jac_apply_w(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                     int r, double* W, const double* __restrict__ Qbuf,
                                                     const int* __restrict__ rotflag, int* state) {
  __shared__ double t[64 * kLd];
  __shared__ double ql[64 * kLd];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  int* st = state + it.x * kJacState;
  if (jac_skip(pr, st, r, -1, false, nullptr)) return;
  const int l = it.z, rb = it.y;
  const int qb = (int)(pr.q_off / 4096);
  if (!rotflag[qb + l]) return;
  const int n = pr.dim;
  const int lr = r % pr.rounds;
  int lp, lq;
  circle_pair(pr.nbp, lr, l, lp, lq);
  double* Wp = W + pr.w_off;
  const double* Ql = Qbuf + pr.q_off + (long long)l * 4096;
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    const int gu = rb * 64 + u, gv = sub_index(lp, lq, v);
    t[u * kLd + v] = (gu < n && gv < n) ? Wp[(long long)gu * n + gv] : 0.0;
    ql[u * kLd + v] = Ql[e];
  }
  __syncthreads();
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int w = 0; w < 64; ++w) {
    double x[4], y[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = t[(ty + 16 * i) * kLd + w];
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] = ql[w * kLd + tx + 16 * j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fma(x[i], y[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gu = rb * 64 + ty + 16 * i;
    if (gu >= n) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gv = sub_index(lp, lq, tx + 16 * j);
      if (gv < n) Wp[(long long)gu * n + gv] = acc[i][j];
    }
  }
}

// initialise W = I and the per-problem scale max|A_ii| and state; molecules
// that finished their SCC are marked converged and skipped
__global__ void __launch_bounds__(256) k_jac_init(const JacobiProb* __restrict__ probs, const double* __restrict__ A,
                                                  double* W, double* scale, int* state,
                                                  const MolState* __restrict__ mstate, int* nactive) {
  __shared__ double sh[4];
  const JacobiProb pr = probs[blockIdx.x];
  if (mstate && !mstate[blockIdx.x].active) {
    if (threadIdx.x == 0) {
      int* st = state + blockIdx.x * kJacState;
      for (int i = 0; i < kJacState; ++i) st[i] = 0;
      st[0] = 1;
    }
    return;
  }
  const int n = pr.dim;
  double* Wp = W + pr.w_off;
  const double* Ap = A + pr.a_off;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(Ap[(long long)i * n + i]));
  for (long long e = threadIdx.x; e < (long long)n * n; e += blockDim.x) Wp[e] = (e / n == e % n) ? 1.0 : 0.0;
  m = block_max<256>(m, sh);
  if (threadIdx.x == 0) {
    scale[blockIdx.x] = fmax(m, 1.0);
    int* st = state + blockIdx.x * kJacState;
    for (int i = 0; i < kJacState; ++i) st[i] = 0;
    if (n > 64) atomicAdd(nactive, 1);
  }
}

// eigenvalues ascending: ev[k] = A[perm[k]][perm[k]] (stable rank sort)
__global__ void k_jac_sort(const JacobiProb* __restrict__ probs, const double* __restrict__ A, double* ev, int* perm,
                           const MolState* __restrict__ mstate) {
  if (mstate && !mstate[blockIdx.x].active) return;
  const JacobiProb pr = probs[blockIdx.x];
  const int n = pr.dim;
  const double* Ap = A + pr.a_off;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double di = Ap[(long long)i * n + i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = Ap[(long long)j * n + j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    ev[pr.ev_off + rank] = di;
    perm[pr.ev_off + rank] = i;
  }
}

}  // namespace xtbamd

namespace xtbamd {
__global__ void k_count_active(const MolState* __restrict__ st, int nmol, int* out) {
  __shared__ double sh[4];
  double c = 0.0;
  for (int m = threadIdx.x; m < nmol; m += blockDim.x) c += st[m].active ? 1.0 : 0.0;
  c = block_sum<256>(c, sh);
  if (threadIdx.x == 0) out[0] = (int)c;
}

// C <- T for the molecules still iterating
__global__ void k_copy_active(const MolDesc* __restrict__ desc, const MolState* __restrict__ st,
                              const double* __restrict__ T, double* C) {
  const MolDesc md = desc[blockIdx.y];
  if (!st[blockIdx.y].active) return;
  const long long n2 = (long long)md.nao * md.nao;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n2) C[md.m0 + t] = T[md.m0 + t];
}
}  // namespace xtbamd
