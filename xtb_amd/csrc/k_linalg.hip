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

// C = alpha op(A) * op(B) [* diag(scale) on the k index] for square nao x nao
// matrices; optional column permutation of B (B(:, perm[j])).  Molecules with
// mask[m * mstride] == 0 are skipped (mask may be null).
template <bool TA, bool TB, bool SCALE, bool PERM>
__global__ void __launch_bounds__(256) k_gemm(const int4* __restrict__ tiles, const MolDesc* __restrict__ desc,
                                              const double* __restrict__ A, const double* __restrict__ B, double* C,
                                              const double* __restrict__ scale, const int* __restrict__ perm,
                                              double alpha, const int* __restrict__ mask, int mstride) {
  __shared__ double As[kBK][kTile + 1];
  __shared__ double Bs[kBK][kTile + 1];
  const int4 tl = tiles[blockIdx.x];
  if (mask && !mask[(long long)tl.x * mstride]) return;
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

// The same batched product on the FP64 matrix cores: each wave owns a 32x32
// quarter of the 64x64 tile as 2x2 blocks of v_mfma_f64_16x16x4f64.
// Operand lanes: A(16x4) row = lane%16, k = lane/16; B(4x16) k = lane/16,
// column = lane%16; result register i of a lane holds row lane/16 + 4 i,
// column lane%16.
typedef double v4d __attribute__((ext_vector_type(4)));

template <bool TA, bool TB, bool SCALE, bool PERM>
__global__ void __launch_bounds__(256) k_gemm_mfma(const int4* __restrict__ tiles, const MolDesc* __restrict__ desc,
                                                   const double* __restrict__ A, const double* __restrict__ B,
                                                   double* C, const double* __restrict__ scale,
                                                   const int* __restrict__ perm, double alpha,
                                                   const int* __restrict__ mask, int mstride) {
  __shared__ double As[kBK][kTile + 1];
  __shared__ double Bs[kBK][kTile + 1];
  const int4 tl = tiles[blockIdx.x];
  if (mask && !mask[(long long)tl.x * mstride]) return;
  const MolDesc md = desc[tl.x];
  const int N = md.nao;
  const double* Am = A + md.m0;
  const double* Bm = B + md.m0;
  double* Cm = C + md.m0;
  const double* sc = SCALE ? scale + md.o0 : nullptr;
  const int* pm = PERM ? perm + md.o0 : nullptr;
  const int row0 = tl.y * kTile, col0 = tl.z * kTile;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, i16 = lane & 15, k4 = lane >> 4;
  v4d acc[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[r][c] = v4d{0.0, 0.0, 0.0, 0.0};
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
      double x = 0.0;
      if (gj < N && gk2 < N) {
        const int cj = PERM ? pm[gj] : gj;
        x = TB ? Bm[(long long)cj * N + gk2] : Bm[(long long)gk2 * N + cj];
        if (SCALE) x *= sc[gk2];
      }
      Bs[k2][jj] = x;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) a[r] = As[kk + k4][32 * wr + 16 * r + i16];
#pragma unroll
      for (int c = 0; c < 2; ++c) b[c] = Bs[kk + k4][32 * wc + 16 * c + i16];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gj = col0 + 32 * wc + 16 * c + i16;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gi = row0 + 32 * wr + 16 * r + k4 + 4 * i;
        if (gi < N && gj < N) Cm[(long long)gi * N + gj] = alpha * acc[r][c][i];
      }
    }
}

// instantiations used by the engine
template __global__ void k_gemm<false, false, false, false>(const int4*, const MolDesc*, const double*, const double*,
                                                            double*, const double*, const int*, double, const int*, int);
template __global__ void k_gemm<true, false, false, false>(const int4*, const MolDesc*, const double*, const double*,
                                                           double*, const double*, const int*, double, const int*, int);
template __global__ void k_gemm<false, true, true, false>(const int4*, const MolDesc*, const double*, const double*,
                                                          double*, const double*, const int*, double, const int*, int);
template __global__ void k_gemm<false, false, false, true>(const int4*, const MolDesc*, const double*, const double*,
                                                           double*, const double*, const int*, double, const int*, int);
template __global__ void k_gemm_mfma<false, false, false, false>(const int4*, const MolDesc*, const double*,
                                                                 const double*, double*, const double*, const int*,
                                                                 double, const int*, int);
template __global__ void k_gemm_mfma<true, false, false, false>(const int4*, const MolDesc*, const double*,
                                                                const double*, double*, const double*, const int*,
                                                                double, const int*, int);
template __global__ void k_gemm_mfma<false, true, true, false>(const int4*, const MolDesc*, const double*,
                                                               const double*, double*, const double*, const int*,
                                                               double, const int*, int);
template __global__ void k_gemm_mfma<false, false, false, true>(const int4*, const MolDesc*, const double*,
                                                                const double*, double*, const double*, const int*,
                                                                double, const int*, int);

// ---------------------------------------------------------------------------
// Cyclic parallel Jacobi on a 64x64 symmetric matrix held in LDS.
// a, q: 64 x 65 (padded rows).  Each sweep starts with the largest
// off-diagonal magnitude; the solve stops when it is <= tol or after maxsweep
// sweeps.  Returns the number of rounds that applied a rotation.
constexpr int kLd = 65;

__device__ int jacobi64(double* a, double* q, double tol, int maxsweep, double* cs, int* pq, int* flags,
                        double* red) {
  const int tid = threadIdx.x;
  // this thread's 2x2 pair blocks (k1 <= k2) of the 32 x 32 pair grid: fixed for all rounds
  int kb1[3], kb2[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int b = tid + 256 * j;
    int k2 = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (k2 * (k2 + 1) / 2 > b) --k2;
    while ((k2 + 1) * (k2 + 2) / 2 <= b) ++k2;
    kb2[j] = b < 528 ? k2 : -1;
    kb1[j] = b - k2 * (k2 + 1) / 2;
  }
  int total = 0;
  for (int sweep = 0; sweep < maxsweep; ++sweep) {
    double m = 0.0;
    for (int e = tid; e < 4096; e += blockDim.x) {
      const int u = e >> 6, v = e & 63;
      if (u < v) m = fmax(m, fabs(a[u * kLd + v]));
    }
    m = block_max<256>(m, red);
    if (m <= tol) break;
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
        if (tid == 0) flags[0] = any ? 1 : 0;
      }
      __syncthreads();
      if (flags[0]) {
        total += 1;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int k1 = kb1[j], k2 = kb2[j];
          if (k2 < 0) continue;
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
        {
          // Q <- Q J: thread owns pair k = tid & 31 and rows tid/32 + 8j (blockDim 256)
          const int k = tid & 31;
          const double s = cs[3 * k + 1];
          if (s != 0.0) {
            const double c = cs[3 * k];
            const int p = pq[2 * k], qq = pq[2 * k + 1];
            const int u0 = tid >> 5;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const int u = u0 + 8 * j;
              const double up = q[u * kLd + p], uq = q[u * kLd + qq];
              q[u * kLd + p] = c * up - s * uq;
              q[u * kLd + qq] = s * up + c * uq;
            }
          }
        }
      }
      __syncthreads();
    }
  }
  return total;
}

__device__ __forceinline__ void circle_pair(int nbp, int r, int k, int& bp, int& bq) {
  const int m = nbp - 1;
  if (k == 0) { bp = 0; bq = 1 + (r % m); }
  else { bp = 1 + (r + k) % m; bq = 1 + ((r - k) % m + m) % m; }
  if (bp > bq) { const int t = bp; bp = bq; bq = t; }
}

// state per problem: [0] done, [1] failed, [5] outer sweeps (block path) or
// inner rounds with rotations (dim <= 64)
constexpr int kJacState = 8;
constexpr int kMaxSweeps = 40;

// problems of dimension <= 64: one workgroup solves to convergence in LDS
__global__ void __launch_bounds__(256) k_jac_small(const JacobiProb* __restrict__ probs, const int* __restrict__ list,
                                                   double* A, double* W, const double* __restrict__ scale_in,
                                                   int* state, int keep_w) {
  __shared__ double a[64 * kLd];
  __shared__ double q[64 * kLd];
  __shared__ double cs[96];
  __shared__ double red[4];
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
  const int nrot = jacobi64(a, q, tol, kMaxSweeps, cs, pq, flags, red);
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int u = e / n, v = e % n;
    Ap[(long long)u * n + v] = a[u * kLd + v];
    if (!keep_w) Wp[(long long)u * n + v] = q[u * kLd + v];
  }
  if (keep_w && nrot > 0) {
    // W <- W Q, with the old W staged in LDS
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
      const int u = e >> 6, v = e & 63;
      a[u * kLd + v] = (u < n && v < n) ? Wp[(long long)u * n + v] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      const int u = e / n, v = e % n;
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc = fma(a[u * kLd + k], q[k * kLd + v], acc);
      Wp[(long long)u * n + v] = acc;
    }
  }
  if (threadIdx.x == 0) {
    state[pidx * kJacState + 0] = 1;
    state[pidx * kJacState + 1] = 0;
    state[pidx * kJacState + 5] = nrot;
  }
}

// global index of the u-th row (0..63) of the pivot subproblem (bp, bq)
__device__ __forceinline__ int sub_index(int bp, int bq, int u) { return u < 32 ? 32 * bp + u : 32 * bq + (u - 32); }

// one workgroup per (problem, pair): diagonalise the 64x64 pivot subproblem
// (at most inner_sweeps sweeps)
__global__ void __launch_bounds__(256) k_jac_pairs(const JacobiProb* __restrict__ probs, const int2* __restrict__ list,
                                                   int r, const double* __restrict__ A, double* Qbuf, int* rotflag,
                                                   const double* __restrict__ scale_in, const int* __restrict__ state,
                                                   int inner_sweeps, unsigned long long* work) {
  __shared__ double a[64 * kLd];
  __shared__ double q[64 * kLd];
  __shared__ double cs[96];
  __shared__ double red[4];
  __shared__ int pq[64];
  __shared__ int flags[2];
  const int2 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int n = pr.dim;
  const int lr = r % pr.rounds;
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
  const int nrot = jacobi64(a, q, tol, inner_sweeps, cs, pq, flags, red);
  __syncthreads();
  double* Qp = Qbuf + pr.q_off + (long long)it.y * 64 * 64;
  if (nrot > 0)
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) Qp[e] = q[(e >> 6) * kLd + (e & 63)];
  if (threadIdx.x == 0) {
    rotflag[(int)(pr.q_off / 4096) + it.y] = nrot > 0 ? 1 : 0;
    if (work && nrot > 0) atomicAdd(work, 36864ull * (unsigned long long)nrot);
  }
}

// A[I_k, I_l] <- Q_k^T A[I_k, I_l] Q_l for pair tiles k <= l
__global__ void __launch_bounds__(256) k_jac_apply_a(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                     int r, double* A, const double* __restrict__ Qbuf,
                                                     const int* __restrict__ rotflag, const int* __restrict__ state) {
  __shared__ double t[64 * kLd];
  __shared__ double qk[64 * kLd];
  __shared__ double ql[64 * kLd];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
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
                                                     const int* __restrict__ rotflag, const int* __restrict__ state) {
  __shared__ double t[64 * kLd];
  __shared__ double ql[64 * kLd];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
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
// that finished their SCC are marked done and skipped
__global__ void __launch_bounds__(256) k_jac_init(const JacobiProb* __restrict__ probs, const double* __restrict__ A,
                                                  double* W, double* scale, int* state,
                                                  const MolState* __restrict__ mstate, unsigned long long* offmax,
                                                  int keep_w) {
  __shared__ double sh[4];
  const JacobiProb pr = probs[blockIdx.x];
  int* st = state + blockIdx.x * kJacState;
  if (mstate && !mstate[blockIdx.x].active) {
    if (threadIdx.x == 0) {
      for (int i = 0; i < kJacState; ++i) st[i] = 0;
      st[0] = 1;
      offmax[blockIdx.x] = 0ull;
    }
    return;
  }
  const int n = pr.dim;
  double* Wp = W + pr.w_off;
  const double* Ap = A + pr.a_off;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(Ap[(long long)i * n + i]));
  if (!keep_w)
    for (long long e = threadIdx.x; e < (long long)n * n; e += blockDim.x) Wp[e] = (e / n == e % n) ? 1.0 : 0.0;
  m = block_max<256>(m, sh);
  if (threadIdx.x == 0) {
    scale[blockIdx.x] = fmax(m, 1.0);
    for (int i = 0; i < kJacState; ++i) st[i] = 0;
    offmax[blockIdx.x] = 0ull;
  }
}

// largest off-diagonal magnitude of every unfinished block-path problem, one
// workgroup per upper-triangular 64 x 64 tile; max over non-negative doubles
// is taken on their bit patterns, which order the same way (order independent)
__global__ void __launch_bounds__(256) k_jac_offmax(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                    const double* __restrict__ A, const int* __restrict__ state,
                                                    unsigned long long* offmax) {
  __shared__ double sh[4];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int n = pr.dim;
  const double* Ap = A + pr.a_off;
  double m = 0.0;
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int gu = it.y * 64 + (e >> 6), gv = it.z * 64 + (e & 63);
    if (gu < gv && gv < n) m = fmax(m, fabs(Ap[(long long)gu * n + gv]));
  }
  m = block_max<256>(m, sh);
  if (threadIdx.x == 0 && m > 0.0) atomicMax(&offmax[it.x], (unsigned long long)__double_as_longlong(m));
}

// convergence test at the start of an outer sweep: done when every
// off-diagonal element is <= tol, failed after kMaxSweeps sweeps; counts the
// problems still running into nactive and resets offmax
__global__ void __launch_bounds__(256) k_jac_mark(const JacobiProb* __restrict__ probs, int nprob,
                                                  const double* __restrict__ scale, int* state,
                                                  unsigned long long* offmax, int sweep, int* nactive) {
  __shared__ double sh[4];
  double c = 0.0;
  for (int p = threadIdx.x; p < nprob; p += blockDim.x) {
    int* st = state + p * kJacState;
    if (st[0]) continue;
    if (probs[p].dim <= 64) continue;   // solved by k_jac_small
    const double m = __longlong_as_double((long long)offmax[p]);
    offmax[p] = 0ull;
    if (m <= 1e-12 * scale[p]) {
      st[0] = 1;
      st[5] = sweep;
    } else if (sweep >= kMaxSweeps) {
      st[0] = 1;
      st[1] = 1;
      st[5] = sweep;
    } else {
      c += 1.0;
    }
  }
  c = block_sum<256>(c, sh);
  if (threadIdx.x == 0) nactive[0] = (int)c;
}

// ---------------------------------------------------------------------------
// Iterative refinement of an approximate eigenvector matrix X of a symmetric A
// (Ogita & Aishima, "Iterative refinement for symmetric eigenvalue
// decomposition", Japan J. Indust. Appl. Math. 35 (2018) 1007-1035, Alg. 1):
//   R = I - X^T X, S = X^T A X, l_i = s_ii / (1 - r_ii),
//   e_ij = (s_ij + l_j r_ij) / (l_j - l_i)   (well separated pair),
//   e_ij = r_ij / 2                         (pair left to the Jacobi sweeps),
//   X <- X + X E.
// A pair counts as well separated when |s_ij + l_j r_ij| <= kOaSep |l_j - l_i|.
constexpr double kOaSep = 0.25;

// E from S (and R = I - XtX, or none when X = I); per-molecule maxima of the
// off-diagonal |s_ij|, of |r_ij| and of the applied |e_ij|, and max |s_ii|
__global__ void __launch_bounds__(256) k_oa_e(const MolDesc* __restrict__ desc, const int* __restrict__ act,
                                              const double* __restrict__ Sm, const double* __restrict__ XtX,
                                              double* E, unsigned long long* stat) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  if (!act[md.mol]) return;
  const long long N = md.nao;
  if ((long long)blockIdx.x * blockDim.x >= N * N) return;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double soff = 0.0, rmax = 0.0, emax = 0.0, dmax = 0.0;
  if (t < N * N) {
    const long long i = t / N, j = t % N;
    const double* S = Sm + md.m0;
    const double sij = S[t];
    double rij = 0.0, rii = 0.0, rjj = 0.0;
    if (XtX) {
      const double* G = XtX + md.m0;
      rij = (i == j ? 1.0 : 0.0) - G[t];
      rii = 1.0 - G[i * N + i];
      rjj = 1.0 - G[j * N + j];
    }
    double e;
    if (i == j) {
      e = 0.5 * rij;
      dmax = fabs(sij);
    } else {
      const double li = S[i * N + i] / (1.0 - rii), lj = S[j * N + j] / (1.0 - rjj);
      const double num = sij + lj * rij, den = lj - li;
      if (den != 0.0 && fabs(num) <= kOaSep * fabs(den)) {
        e = num / den;
        emax = fabs(e);
      } else {
        e = 0.5 * rij;
      }
      soff = fabs(sij);
    }
    rmax = fabs(rij);
    E[md.m0 + t] = e;
  }
  soff = block_max<256>(soff, sh);
  rmax = block_max<256>(rmax, sh);
  emax = block_max<256>(emax, sh);
  dmax = block_max<256>(dmax, sh);
  if (threadIdx.x == 0) {
    unsigned long long* st = stat + 4LL * md.mol;
    if (soff > 0.0) atomicMax(&st[0], (unsigned long long)__double_as_longlong(soff));
    if (rmax > 0.0) atomicMax(&st[1], (unsigned long long)__double_as_longlong(rmax));
    if (emax > 0.0) atomicMax(&st[2], (unsigned long long)__double_as_longlong(emax));
    if (dmax > 0.0) atomicMax(&st[3], (unsigned long long)__double_as_longlong(dmax));
  }
}

// stop refining a molecule when its off-diagonal is at the Jacobi tolerance
// and X is orthonormal (to 5e-14 N, the rounding level of X^T X), or when no
// well-separated pair is left to correct (|e_ij| <= 1e-12); the result is then
// accepted (ok = 1).  After maxstep steps it is rejected: the Jacobi sweeps
// start from the unrefined matrix.  act = 1 keeps it for one more update.
__global__ void __launch_bounds__(256) k_oa_mark(int nmol, const MolDesc* __restrict__ desc, int* act, int* ok,
                                                 unsigned long long* stat, int step, int maxstep, int* nactive) {
  __shared__ double sh[4];
  double c = 0.0;
  for (int m = threadIdx.x; m < nmol; m += blockDim.x) {
    unsigned long long* st = stat + 4LL * m;
    if (act[m]) {
      const double soff = __longlong_as_double((long long)st[0]);
      const double rmax = __longlong_as_double((long long)st[1]);
      const double emax = __longlong_as_double((long long)st[2]);
      const double scale = fmax(__longlong_as_double((long long)st[3]), 1.0);
      const double rtol = 5e-14 * desc[m].nao;
      const bool conv = rmax <= rtol && (soff <= 1e-12 * scale || emax <= 1e-12);
      if (conv) {
        act[m] = 0;
        ok[m] = 1;
      } else if (step >= maxstep) {
        act[m] = 0;
        ok[m] = 0;
      }
    }
    for (int k = 0; k < 4; ++k) st[k] = 0ull;
    c += act[m];
  }
  c = block_sum<256>(c, sh);
  if (threadIdx.x == 0) nactive[0] = (int)c;
}

// accepted refinement: A <- X^T A X (kept in Sm); rejected: W <- I
__global__ void k_oa_finish(const MolDesc* __restrict__ desc, const MolState* __restrict__ mstate,
                            const int* __restrict__ ok, const double* __restrict__ Sm, double* A, double* W) {
  const MolDesc md = desc[blockIdx.y];
  if (!mstate[md.mol].active) return;
  const long long N = md.nao;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * N) return;
  if (ok[md.mol]) A[md.m0 + t] = Sm[md.m0 + t];
  else W[md.m0 + t] = (t / N == t % N) ? 1.0 : 0.0;
}

// X <- X + D (D = X E) for refining molecules; with identity = 1, X <- I + E
__global__ void k_oa_update(const MolDesc* __restrict__ desc, const int* __restrict__ act,
                            const double* __restrict__ D, double* X, int identity) {
  const MolDesc md = desc[blockIdx.y];
  if (!act[md.mol]) return;
  const long long N = md.nao;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * N) return;
  const double x = identity ? ((t / N == t % N) ? 1.0 : 0.0) : X[md.m0 + t];
  X[md.m0 + t] = x + D[md.m0 + t];
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
