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
// mask[m * mstride] == 0 are skipped (mask may be null).  A tile entry
// (m, a, b, sym) with sym = 1 is one of a product known to be symmetric: the
// list holds only tiles b >= a, each tile writes its elements and their
// mirrors (a diagonal tile its upper triangle), so C comes out exactly
// symmetric from half the work.
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
  const bool sym = tl.w != 0, dtile = tl.y == tl.z;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gi = row0 + ty + 16 * r;
    if (gi >= N) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int gj = col0 + tx + 16 * c;
      if (gj >= N) continue;
      if (!sym) {
        Cm[(long long)gi * N + gj] = alpha * acc[r][c];
      } else if (!dtile || gi <= gj) {
        Cm[(long long)gi * N + gj] = alpha * acc[r][c];
        Cm[(long long)gj * N + gi] = alpha * acc[r][c];
      }
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
  // two LDS stages: the global loads of k-block kb + kBK are in flight while
  // the matrix cores work on k-block kb, one barrier per k-block
  __shared__ double As[2][kBK][kTile + 1];
  __shared__ double Bs[2][kBK][kTile + 1];
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
  // this thread's staging slots: A (ii, kk) and B (k2, jj) for q = 0..3
  int sa_i[4], sa_k[4], sb_k[4], sb_j[4], cjv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = tid + 256 * q;
    if (TA) { sa_k[q] = e >> 6; sa_i[q] = e & 63; } else { sa_i[q] = e >> 4; sa_k[q] = e & 15; }
    if (TB) { sb_j[q] = e >> 4; sb_k[q] = e & 15; } else { sb_k[q] = e >> 6; sb_j[q] = e & 63; }
    const int gj = col0 + sb_j[q];
    cjv[q] = (gj < N) ? (PERM ? pm[gj] : gj) : -1;
  }
  double ra[4], rb[4];
  auto fetch = [&](int kb) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gi = row0 + sa_i[q], gk = kb + sa_k[q];
      ra[q] = (gi < N && gk < N) ? (TA ? Am[(long long)gk * N + gi] : Am[(long long)gi * N + gk]) : 0.0;
      const int cj = cjv[q], gk2 = kb + sb_k[q];
      double x = 0.0;
      if (cj >= 0 && gk2 < N) {
        x = TB ? Bm[(long long)cj * N + gk2] : Bm[(long long)gk2 * N + cj];
        if (SCALE) x *= sc[gk2];
      }
      rb[q] = x;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[buf][sa_k[q]][sa_i[q]] = ra[q];
      Bs[buf][sb_k[q]][sb_j[q]] = rb[q];
    }
  };
  v4d acc[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[r][c] = v4d{0.0, 0.0, 0.0, 0.0};
  fetch(0);
  stage(0);
  __syncthreads();
  int buf = 0;
  for (int kb = 0; kb < N; kb += kBK) {
    const bool more = kb + kBK < N;
    if (more) fetch(kb + kBK);
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) a[r] = As[buf][kk + k4][32 * wr + 16 * r + i16];
#pragma unroll
      for (int c = 0; c < 2; ++c) b[c] = Bs[buf][kk + k4][32 * wc + 16 * c + i16];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
    }
    // the other stage was last read before the previous barrier
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  const bool sym = tl.w != 0, dtile = tl.y == tl.z;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gj = col0 + 32 * wc + 16 * c + i16;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gi = row0 + 32 * wr + 16 * r + k4 + 4 * i;
        if (gi >= N || gj >= N) continue;
        if (!sym) {
          Cm[(long long)gi * N + gj] = alpha * acc[r][c][i];
        } else if (!dtile || gi <= gj) {
          Cm[(long long)gi * N + gj] = alpha * acc[r][c][i];
          Cm[(long long)gj * N + gi] = alpha * acc[r][c][i];
        }
      }
    }
}

// The same product on the FP32 matrix cores with an FP64 residual correction
// (the mixed-precision option): each FP64 operand is split into hi = fl32(x)
// and lo = fl32(x - hi), every 4-deep step forms hi*hi + hi*lo + lo*hi with
// v_mfma_f32_16x16x4f32 (FP32 accumulation of 12 products) and adds that
// partial to FP64 accumulators, so the rounding of FP32 accumulation is
// confined to one step.  The f32 result layout: register i of a lane holds
// row 4 (lane / 16) + i, column lane % 16.
typedef float v4f __attribute__((ext_vector_type(4)));

template <bool TA, bool TB, bool SCALE>
__global__ void __launch_bounds__(256) k_gemm_mix(const int4* __restrict__ tiles, const MolDesc* __restrict__ desc,
                                                  const double* __restrict__ A, const double* __restrict__ B,
                                                  double* C, const double* __restrict__ scale, double alpha,
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
  const int row0 = tl.y * kTile, col0 = tl.z * kTile;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, i16 = lane & 15, k4 = lane >> 4;
  double acc[2][2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[r][c][i] = 0.0;
  for (int kb = 0; kb < N; kb += kBK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int ii, kk, jj, k2;
      if (TA) { kk = e >> 6; ii = e & 63; } else { ii = e >> 4; kk = e & 15; }
      if (TB) { jj = e >> 4; k2 = e & 15; } else { k2 = e >> 6; jj = e & 63; }
      const int gi = row0 + ii, gk = kb + kk, gj = col0 + jj, gk2 = kb + k2;
      As[kk][ii] = (gi < N && gk < N) ? (TA ? Am[(long long)gk * N + gi] : Am[(long long)gi * N + gk]) : 0.0;
      double x = 0.0;
      if (gj < N && gk2 < N) {
        x = TB ? Bm[(long long)gj * N + gk2] : Bm[(long long)gk2 * N + gj];
        if (SCALE) x *= sc[gk2];
      }
      Bs[k2][jj] = x;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      float ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const double a = As[kk + k4][32 * wr + 16 * r + i16];
        ah[r] = (float)a;
        al[r] = (float)(a - (double)ah[r]);
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double b = Bs[kk + k4][32 * wc + 16 * c + i16];
        bh[c] = (float)b;
        bl[c] = (float)(b - (double)bh[c]);
      }
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          v4f t = v4f{0.f, 0.f, 0.f, 0.f};
          t = __builtin_amdgcn_mfma_f32_16x16x4f32(al[r], bh[c], t, 0, 0, 0);
          t = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[r], bl[c], t, 0, 0, 0);
          t = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[r], bh[c], t, 0, 0, 0);
#pragma unroll
          for (int i = 0; i < 4; ++i) acc[r][c][i] += (double)t[i];
        }
    }
    __syncthreads();
  }
  const bool sym = tl.w != 0, dtile = tl.y == tl.z;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gj = col0 + 32 * wc + 16 * c + i16;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gi = row0 + 32 * wr + 16 * r + 4 * k4 + i;
        if (gi >= N || gj >= N) continue;
        const double v = alpha * acc[r][c][i];
        if (!sym) {
          Cm[(long long)gi * N + gj] = v;
        } else if (!dtile || gi <= gj) {
          Cm[(long long)gi * N + gj] = v;
          Cm[(long long)gj * N + gi] = v;
        }
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
template __global__ void k_gemm<false, true, false, true>(const int4*, const MolDesc*, const double*, const double*,
                                                          double*, const double*, const int*, double, const int*, int);
template __global__ void k_gemm_mfma<false, true, false, true>(const int4*, const MolDesc*, const double*,
                                                               const double*, double*, const double*, const int*,
                                                               double, const int*, int);
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

// Round numbers from kBanded on select the banded schedule: even rounds pair
// blocks (2k, 2k+1), odd rounds (2k+1, 2k+2) and the two end blocks; it
// reaches every coupling between neighbouring blocks in two rounds.
constexpr int kBanded = 1 << 28;

__device__ __forceinline__ void jac_pair(int nbp, int rounds, int r, int k, int& bp, int& bq) {
  if (r < kBanded) {
    circle_pair(nbp, r % rounds, k, bp, bq);
    return;
  }
  if (((r - kBanded) & 1) == 0) {
    bp = 2 * k;
    bq = 2 * k + 1;
  } else if (2 * k + 2 < nbp) {
    bp = 2 * k + 1;
    bq = 2 * k + 2;
  } else {
    bp = 0;
    bq = nbp - 1;
  }
}

// a problem tests convergence when it starts a sweep (circle schedule) or a
// pair of rounds (banded schedule)
__device__ __forceinline__ bool jac_boundary(int rounds, int r) {
  return r < kBanded ? (r % rounds == 0) : (((r - kBanded) & 1) == 0);
}

// state per problem: [0] done, [1] failed, [5] outer sweeps (block path) or
// inner rounds with rotations (dim <= 64)
constexpr int kJacState = 8;
// XCDs of the gfx950 part (MI355X): workgroup b of a launch runs on XCD b mod kXcds
constexpr int kXcds = 8;

// problems of dimension <= 64: one workgroup solves to convergence in LDS
__global__ void __launch_bounds__(256) k_jac_small(const JacobiProb* __restrict__ probs, const int* __restrict__ list,
                                                   double* A, double* W, const double* __restrict__ scale_in,
                                                   int* state, int keep_w, int rows_out, double tolrel) {
  __shared__ double a[64 * kLd];
  __shared__ double q[64 * kLd];
  __shared__ double cs[96];
  __shared__ double red[4];
  __shared__ int pq[64];
  __shared__ int flags[2];
  const int pidx = list[blockIdx.x];
  const JacobiProb pr = probs[pidx];
  if (state[pidx * kJacState]) return;
  const int n = pr.dim, ld = pr.ld;
  double* Ap = A + pr.a_off;
  double* Wp = W + pr.w_off;
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    a[u * kLd + v] = (u < n && v < n) ? Ap[(long long)u * ld + v] : 0.0;
    q[u * kLd + v] = (u == v) ? 1.0 : 0.0;
  }
  __syncthreads();
  const double tol = tolrel * scale_in[pidx];
  const int nrot = jacobi64(a, q, tol, pr.maxsw, cs, pq, flags, red);
  __syncthreads();
  // the sweeps stop at convergence or after maxsw of them: given up when the
  // largest off-diagonal magnitude is still above the tolerance
  double offm = 0.0;
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    if (u < v) offm = fmax(offm, fabs(a[u * kLd + v]));
  }
  offm = block_max<256>(offm, red);
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int u = e / n, v = e % n;
    Ap[(long long)u * ld + v] = a[u * kLd + v];
    if (!keep_w) Wp[(long long)u * ld + v] = rows_out ? q[v * kLd + u] : q[u * kLd + v];
  }
  if (keep_w && nrot > 0) {
    // W <- W Q, with the old W staged in LDS
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
      const int u = e >> 6, v = e & 63;
      a[u * kLd + v] = (u < n && v < n) ? Wp[(long long)u * ld + v] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      const int u = e / n, v = e % n;
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc = fma(a[u * kLd + k], q[k * kLd + v], acc);
      Wp[(long long)u * ld + v] = acc;
    }
  }
  if (threadIdx.x == 0) {
    state[pidx * kJacState + 0] = 1;
    state[pidx * kJacState + 1] = offm > tol ? 1 : 0;
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
  int bp, bq;
  jac_pair(pr.nbp, pr.rounds, r, it.y, bp, bq);
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
    if (work && nrot > 0) prof_add(work, 36864ull * (unsigned long long)nrot);
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
  int kp, kq, lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, k, kp, kq);
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
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
  int lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
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

// ---- the two apply kernels on the FP64 matrix cores: 64x64x64 products
// staged in LDS (two 64x65 tiles, two workgroups per CU); each wave owns a
// 32x32 quadrant as 2x2 v_mfma_f64_16x16x4f64 blocks (lane layouts as in
// k_gemm_mfma).  out(32x32 quadrant) = sum_k L[row][k] R[k][col] with the
// left operand read as L[row][k] = lt ? Lm[k][row] : Lm[row][k].
__device__ __forceinline__ void mfma_tile64(const double* Lm, bool lt, const double* Rm, v4d acc[2][2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w >> 1, wc = w & 1, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[r][c] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int kk = 0; kk < 64; kk += 4) {
    double a[2], b[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int row = 32 * wr + 16 * r + i16, k = kk + k4;
      a[r] = lt ? Lm[k * kLd + row] : Lm[row * kLd + k];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) b[c] = Rm[(kk + k4) * kLd + 32 * wc + 16 * c + i16];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
  }
}

__device__ __forceinline__ void store_quadrants(double* Tm, const v4d acc[2][2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w >> 1, wc = w & 1, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) Tm[(32 * wr + 16 * r + k4 + 4 * i) * kLd + 32 * wc + 16 * c + i16] = acc[r][c][i];
}

// A[I_k, I_l] <- Q_k^T A[I_k, I_l] Q_l for pair tiles k <= l (matrix cores)
__global__ void __launch_bounds__(256) k_jac_apply_a_mfma(const JacobiProb* __restrict__ probs,
                                                          const int4* __restrict__ list, int r, double* A,
                                                          const double* __restrict__ Qbuf,
                                                          const int* __restrict__ rotflag,
                                                          const int* __restrict__ state) {
  __shared__ double t[64 * kLd];
  __shared__ double qm[64 * kLd];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int k = it.y, l = it.z;
  const int qb = (int)(pr.q_off / 4096);
  const int rk = rotflag[qb + k], rl = rotflag[qb + l];
  if (!rk && !rl) return;
  const int n = pr.dim;
  int kp, kq, lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, k, kp, kq);
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
  double* Ap = A + pr.a_off;
  const double* Qk = Qbuf + pr.q_off + (long long)k * 4096;
  const double* Ql = Qbuf + pr.q_off + (long long)l * 4096;
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    const int gu = sub_index(kp, kq, u), gv = sub_index(lp, lq, v);
    t[u * kLd + v] = (gu < n && gv < n) ? Ap[(long long)gu * n + gv] : 0.0;
    qm[u * kLd + v] = rl ? Ql[e] : (u == v ? 1.0 : 0.0);
  }
  __syncthreads();
  v4d acc[2][2];
  mfma_tile64(t, false, qm, acc);   // M = T Q_l
  __syncthreads();
  store_quadrants(t, acc);
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    qm[u * kLd + v] = rk ? Qk[e] : (u == v ? 1.0 : 0.0);
  }
  __syncthreads();
  mfma_tile64(qm, true, t, acc);    // R = Q_k^T M
  __syncthreads();
  store_quadrants(t, acc);
  __syncthreads();
  // R to [I_k, I_l] (row-coalesced) and R^T to [I_l, I_k] (column reads of
  // the LDS tile, row-coalesced writes); a diagonal tile is written from its
  // upper half to both positions, which keeps A exactly symmetric
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    const int gu = sub_index(kp, kq, u), gv = sub_index(lp, lq, v);
    if (gu < n && gv < n) {
      const double val = (k == l && u > v) ? t[v * kLd + u] : t[u * kLd + v];
      Ap[(long long)gu * n + gv] = val;
    }
  }
  if (k != l)
    for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
      const int v = e >> 6, u = e & 63;
      const int gu = sub_index(kp, kq, u), gv = sub_index(lp, lq, v);
      if (gu < n && gv < n) Ap[(long long)gv * n + gu] = t[u * kLd + v];
    }
}

// W[rows, I_l] <- W[rows, I_l] Q_l for 64-row blocks (matrix cores)
__global__ void __launch_bounds__(256) k_jac_apply_w_mfma(const JacobiProb* __restrict__ probs,
                                                          const int4* __restrict__ list, int r, double* W,
                                                          const double* __restrict__ Qbuf,
                                                          const int* __restrict__ rotflag,
                                                          const int* __restrict__ state) {
  __shared__ double t[64 * kLd];
  __shared__ double qm[64 * kLd];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int l = it.z, rb = it.y;
  const int qb = (int)(pr.q_off / 4096);
  if (!rotflag[qb + l]) return;
  const int n = pr.dim;
  int lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
  double* Wp = W + pr.w_off;
  const double* Ql = Qbuf + pr.q_off + (long long)l * 4096;
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    const int gu = rb * 64 + u, gv = sub_index(lp, lq, v);
    t[u * kLd + v] = (gu < n && gv < n) ? Wp[(long long)gu * n + gv] : 0.0;
    qm[u * kLd + v] = Ql[e];
  }
  __syncthreads();
  v4d acc[2][2];
  mfma_tile64(t, false, qm, acc);
  __syncthreads();
  store_quadrants(t, acc);
  __syncthreads();
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int u = e >> 6, v = e & 63;
    const int gu = rb * 64 + u, gv = sub_index(lp, lq, v);
    if (gu < n && gv < n) Wp[(long long)gu * n + gv] = t[u * kLd + v];
  }
}

// ---------------------------------------------------------------------------
// Block-Jacobi kernels for a pivot size P (P/2-wide blocks): P = 64 with 256
// threads (four waves) or P = 32 with 64 threads (one wave).  Same scheme as
// above: cyclic parallel Jacobi of each P x P pivot problem in LDS, then
// two-sided tile products on the matrix cores.
template <int P, int NT>
__device__ int jacobiP(double* a, double* q, double tol, int maxsweep, double* cs, int* pq, int* flags, double* red) {
  constexpr int L = P + 1, H = P / 2, NB2 = H * (H + 1) / 2, NPB = (NB2 + NT - 1) / NT;
  constexpr int RS = NT / H, RPT = P / RS;   // Q update: row stride and rows per thread
  const int tid = threadIdx.x;
  int kb1[NPB], kb2[NPB];
#pragma unroll
  for (int j = 0; j < NPB; ++j) {
    const int b = tid + NT * j;
    int k2 = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (k2 * (k2 + 1) / 2 > b) --k2;
    while ((k2 + 1) * (k2 + 2) / 2 <= b) ++k2;
    kb2[j] = b < NB2 ? k2 : -1;
    kb1[j] = b - k2 * (k2 + 1) / 2;
  }
  int total = 0;
  for (int sweep = 0; sweep < maxsweep; ++sweep) {
    double m = 0.0;
    for (int e = tid; e < P * P; e += NT) {
      const int u = e / P, v = e % P;
      if (u < v) m = fmax(m, fabs(a[u * L + v]));
    }
    m = block_max<NT>(m, red);
    if (m <= tol) break;
    for (int r = 0; r < P - 1; ++r) {
      if (tid < 64) {
        bool rot = false;
        if (tid < H) {
          const int k = tid;
          int p, qq;
          if (k == 0) { p = 0; qq = 1 + r; }
          else { p = 1 + (r + k) % (P - 1); qq = 1 + (r - k + (P - 1)) % (P - 1); }
          if (p > qq) { const int t = p; p = qq; qq = t; }
          const double apq = a[p * L + qq];
          double c = 1.0, sn = 0.0, t = 0.0;
          if (fabs(apq) > tol) {
            const double app = a[p * L + p], aqq = a[qq * L + qq];
            const double theta = (aqq - app) / (2.0 * apq);
            t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            sn = t * c;
            rot = true;
          }
          pq[2 * k] = p;
          pq[2 * k + 1] = qq;
          cs[3 * k] = c;
          cs[3 * k + 1] = sn;
          cs[3 * k + 2] = t;
        }
        const unsigned long long any = __ballot(rot);
        if (tid == 0) flags[0] = any ? 1 : 0;
      }
      __syncthreads();
      if (flags[0]) {
        total += 1;
#pragma unroll
        for (int j = 0; j < NPB; ++j) {
          const int k1 = kb1[j], k2 = kb2[j];
          if (k2 < 0) continue;
          const int p1 = pq[2 * k1], q1 = pq[2 * k1 + 1], p2 = pq[2 * k2], q2 = pq[2 * k2 + 1];
          const double c1 = cs[3 * k1], s1 = cs[3 * k1 + 1], c2 = cs[3 * k2], s2 = cs[3 * k2 + 1];
          if (k1 == k2) {
            if (s1 != 0.0) {
              const double t1 = cs[3 * k1 + 2];
              const double apq = a[p1 * L + q1];
              a[p1 * L + p1] = a[p1 * L + p1] - t1 * apq;
              a[q1 * L + q1] = a[q1 * L + q1] + t1 * apq;
              a[p1 * L + q1] = 0.0;
              a[q1 * L + p1] = 0.0;
            }
          } else if (s1 != 0.0 || s2 != 0.0) {
            const double apr = a[p1 * L + p2], aps = a[p1 * L + q2];
            const double aqr = a[q1 * L + p2], aqs = a[q1 * L + q2];
            const double bpr = c2 * apr - s2 * aps, bps = s2 * apr + c2 * aps;
            const double bqr = c2 * aqr - s2 * aqs, bqs = s2 * aqr + c2 * aqs;
            const double npr = c1 * bpr - s1 * bqr, nqr = s1 * bpr + c1 * bqr;
            const double nps = c1 * bps - s1 * bqs, nqs = s1 * bps + c1 * bqs;
            a[p1 * L + p2] = npr; a[p2 * L + p1] = npr;
            a[p1 * L + q2] = nps; a[q2 * L + p1] = nps;
            a[q1 * L + p2] = nqr; a[p2 * L + q1] = nqr;
            a[q1 * L + q2] = nqs; a[q2 * L + q1] = nqs;
          }
        }
        {
          const int k = tid % H;
          const double sn = cs[3 * k + 1];
          if (sn != 0.0) {
            const double c = cs[3 * k];
            const int p = pq[2 * k], qq = pq[2 * k + 1];
            const int u0 = tid / H;
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
              const int u = u0 + RS * j;
              const double up = q[u * L + p], uq = q[u * L + qq];
              q[u * L + p] = c * up - sn * uq;
              q[u * L + qq] = sn * up + c * uq;
            }
          }
        }
      }
      __syncthreads();
    }
  }
  return total;
}

template <int P>
__device__ __forceinline__ int sub_indexP(int bp, int bq, int u) {
  return u < P / 2 ? (P / 2) * bp + u : (P / 2) * bq + (u - P / 2);
}

template <int P, int NT>
__global__ void __launch_bounds__(NT) k_jac_pairsP(const JacobiProb* __restrict__ probs, const int2* __restrict__ list,
                                                   int r, const double* __restrict__ A, double* Qbuf, int* rotflag,
                                                   const double* __restrict__ scale_in, const int* __restrict__ state,
                                                   int inner_sweeps, unsigned long long* work) {
  constexpr int L = P + 1;
  __shared__ double a[P * L];
  __shared__ double q[P * L];
  __shared__ double cs[3 * (P / 2)];
  __shared__ double red[4];
  __shared__ int pq[P];
  __shared__ int flags[2];
  const int2 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int n = pr.dim;
  int bp, bq;
  jac_pair(pr.nbp, pr.rounds, r, it.y, bp, bq);
  const double* Ap = A + pr.a_off;
  for (int e = threadIdx.x; e < P * P; e += NT) {
    const int u = e / P, v = e % P;
    const int gu = sub_indexP<P>(bp, bq, u), gv = sub_indexP<P>(bp, bq, v);
    a[u * L + v] = (gu < n && gv < n) ? Ap[(long long)gu * n + gv] : 0.0;
    q[u * L + v] = (u == v) ? 1.0 : 0.0;
  }
  __syncthreads();
  const double tol = 1e-12 * scale_in[it.x];
  const int nrot = jacobiP<P, NT>(a, q, tol, inner_sweeps, cs, pq, flags, red);
  __syncthreads();
  double* Qp = Qbuf + pr.q_off + (long long)it.y * P * P;
  if (nrot > 0)
    for (int e = threadIdx.x; e < P * P; e += NT) Qp[e] = q[(e / P) * L + (e % P)];
  if (threadIdx.x == 0) {
    rotflag[(int)(pr.q_off / (P * P)) + it.y] = nrot > 0 ? 1 : 0;
    if (work && nrot > 0) prof_add(work, (unsigned long long)(9 * P * P) * (unsigned long long)nrot);
  }
}

// out(P x P) = L R with L read transposed when lt; one 32x32 quadrant per wave
template <int P>
__device__ __forceinline__ void mfma_tileP(const double* Lm, bool lt, const double* Rm, v4d acc[2][2]) {
  constexpr int L = P + 1, QW = P / 32;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w / QW, wc = w % QW, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[r][c] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int kk = 0; kk < P; kk += 4) {
    double a[2], b[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int row = 32 * wr + 16 * r + i16, k = kk + k4;
      a[r] = lt ? Lm[k * L + row] : Lm[row * L + k];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) b[c] = Rm[(kk + k4) * L + 32 * wc + 16 * c + i16];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
  }
}

template <int P>
__device__ __forceinline__ void store_quadrantsP(double* Tm, const v4d acc[2][2]) {
  constexpr int L = P + 1, QW = P / 32;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w / QW, wc = w % QW, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) Tm[(32 * wr + 16 * r + k4 + 4 * i) * L + 32 * wc + 16 * c + i16] = acc[r][c][i];
}

// A[I_k, I_l] <- Q_k^T A[I_k, I_l] Q_l for pair tiles k <= l
template <int P>
__global__ void __launch_bounds__((P / 32) * (P / 32) * 64) k_jac_apply_aP(
    const JacobiProb* __restrict__ probs, const int4* __restrict__ list, int r, double* A,
    const double* __restrict__ Qbuf, const int* __restrict__ rotflag, const int* __restrict__ state) {
  constexpr int L = P + 1, NT = (P / 32) * (P / 32) * 64;
  __shared__ double t[P * L];
  __shared__ double qm[P * L];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int k = it.y, l = it.z;
  const int qb = (int)(pr.q_off / (P * P));
  const int rk = rotflag[qb + k], rl = rotflag[qb + l];
  if (!rk && !rl) return;
  const int n = pr.dim;
  int kp, kq, lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, k, kp, kq);
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
  double* Ap = A + pr.a_off;
  const double* Qk = Qbuf + pr.q_off + (long long)k * P * P;
  const double* Ql = Qbuf + pr.q_off + (long long)l * P * P;
  for (int e = threadIdx.x; e < P * P; e += NT) {
    const int u = e / P, v = e % P;
    const int gu = sub_indexP<P>(kp, kq, u), gv = sub_indexP<P>(lp, lq, v);
    t[u * L + v] = (gu < n && gv < n) ? Ap[(long long)gu * n + gv] : 0.0;
    qm[u * L + v] = rl ? Ql[e] : (u == v ? 1.0 : 0.0);
  }
  __syncthreads();
  v4d acc[2][2];
  mfma_tileP<P>(t, false, qm, acc);
  __syncthreads();
  store_quadrantsP<P>(t, acc);
  for (int e = threadIdx.x; e < P * P; e += NT) {
    const int u = e / P, v = e % P;
    qm[u * L + v] = rk ? Qk[e] : (u == v ? 1.0 : 0.0);
  }
  __syncthreads();
  mfma_tileP<P>(qm, true, t, acc);
  __syncthreads();
  store_quadrantsP<P>(t, acc);
  __syncthreads();
  for (int e = threadIdx.x; e < P * P; e += NT) {
    const int u = e / P, v = e % P;
    const int gu = sub_indexP<P>(kp, kq, u), gv = sub_indexP<P>(lp, lq, v);
    if (gu < n && gv < n) Ap[(long long)gu * n + gv] = (k == l && u > v) ? t[v * L + u] : t[u * L + v];
  }
  if (k != l)
    for (int e = threadIdx.x; e < P * P; e += NT) {
      const int v = e / P, u = e % P;
      const int gu = sub_indexP<P>(kp, kq, u), gv = sub_indexP<P>(lp, lq, v);
      if (gu < n && gv < n) Ap[(long long)gv * n + gu] = t[u * L + v];
    }
}

// W[P-row block, I_l] <- W[..., I_l] Q_l
template <int P>
__global__ void __launch_bounds__((P / 32) * (P / 32) * 64) k_jac_apply_wP(
    const JacobiProb* __restrict__ probs, const int4* __restrict__ list, int r, double* W,
    const double* __restrict__ Qbuf, const int* __restrict__ rotflag, const int* __restrict__ state) {
  constexpr int L = P + 1, NT = (P / 32) * (P / 32) * 64;
  __shared__ double t[P * L];
  __shared__ double qm[P * L];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int l = it.z, rb = it.y;
  const int qb = (int)(pr.q_off / (P * P));
  if (!rotflag[qb + l]) return;
  const int n = pr.dim;
  int lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
  double* Wp = W + pr.w_off;
  const double* Ql = Qbuf + pr.q_off + (long long)l * P * P;
  for (int e = threadIdx.x; e < P * P; e += NT) {
    const int u = e / P, v = e % P;
    const int gu = rb * P + u, gv = sub_indexP<P>(lp, lq, v);
    t[u * L + v] = (gu < n && gv < n) ? Wp[(long long)gu * n + gv] : 0.0;
    qm[u * L + v] = Ql[e];
  }
  __syncthreads();
  v4d acc[2][2];
  mfma_tileP<P>(t, false, qm, acc);
  __syncthreads();
  store_quadrantsP<P>(t, acc);
  __syncthreads();
  for (int e = threadIdx.x; e < P * P; e += NT) {
    const int u = e / P, v = e % P;
    const int gu = rb * P + u, gv = sub_indexP<P>(lp, lq, v);
    if (gu < n && gv < n) Wp[(long long)gu * n + gv] = t[u * L + v];
  }
}

template __global__ void k_jac_pairsP<32, 64>(const JacobiProb*, const int2*, int, const double*, double*, int*,
                                              const double*, const int*, int, unsigned long long*);
template __global__ void k_jac_apply_aP<32>(const JacobiProb*, const int4*, int, double*, const double*, const int*,
                                            const int*);
template __global__ void k_jac_apply_wP<32>(const JacobiProb*, const int4*, int, double*, const double*, const int*,
                                            const int*);

// ---------------------------------------------------------------------------
// Register-resident Jacobi sweep of a 32 x 32 pivot problem, one wave.
//
// The 32 indices sit in 16 slot pairs k = 0..15 (a top and a bottom slot;
// initially index k and index 16 + k).  Lane L holds column pair c = L & 15 of
// the row pairs 4g..4g+3 (g = L >> 4) of A, 2 x 2 doubles per pair block, and
// the same row pairs of Q^T (rows are slots, columns the indices c and
// c + 16).  An inner round rotates every pair (top, bottom) at once, the lane
// owning the diagonal block of a pair computing its rotation, then moves the
// slots by the circle method: the top of pair 0 stays, the bottom of pair 0
// becomes the top of pair 1, tops move one pair up, the top of pair 15
// becomes its bottom and bottoms move one pair down.  31 rounds meet every
// pair of indices once and restore the slot order, so Q comes out with its
// columns in index order.  Column moves of A are DPP shifts within a 16-lane
// row.  Row moves of A and Q^T stay in a lane except across the 16-lane
// groups (one shuffle each way); the in-lane ones are register renamings:
// at phase ph = round mod 4 the top row of pair 4g + i sits in T[(i - ph) & 3]
// and its bottom row in B[(i + ph) & 3], so rounds are unrolled by four and
// only the moves across groups and the fixed pair 0 / pair 15 rows cost
// instructions.  Only the rotation cosines and sines go through LDS.  The
// rotation is the one of the LDS kernels (Rutishauser's formulas, exact
// update of the pivot's diagonal pair).

// src moved one lane within each row of 16: CTRL 0x111 reads lane - 1
// (row_shr:1), 0x101 reads lane + 1 (row_shl:1); a lane without a source
// (the first, resp. last, of its row) returns old
template <int CTRL>
__device__ __forceinline__ double dpp_row_f64(double old, double src) {
  const long long o = __double_as_longlong(old), b = __double_as_longlong(src);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned int)(o & 0xffffffffll), (int)(unsigned int)(b & 0xffffffffll),
                                             CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// circle-method move of a (top, bottom) column pair held by lane c of a row:
// top(c) <- top(c-1) for c >= 2, bottom(0) for c = 1, itself for c = 0;
// bottom(c) <- bottom(c+1) for c <= 14, top(15) for c = 15
__device__ __forceinline__ void circle_cols(double& top, double& bot, int c) {
  const double send = (c == 0) ? bot : top;
  const double nt = dpp_row_f64<0x111>(top, send);
  const double nb = dpp_row_f64<0x101>(top, bot);
  top = nt;
  bot = nb;
}

// circle-method row move of one lane's rows at phase PH (top row of pair
// 4g + i in T[(i - PH) & 3], bottom row in B[(i + PH) & 3]), leaving phase
// PH + 1: the new top of pair 4g comes from the previous group (pair 0: its
// own top), the new top of pair 4g + 1 of group 0 is the old bottom of pair 0,
// the new bottom of pair 4g + 3 comes from the next group (pair 15: its own
// top); every other row keeps its register
template <int PH>
__device__ __forceinline__ void circle_rows(double (&T)[4][2], double (&B)[4][2], int lane, int g) {
  const int up = (lane + 48) & 63, dn = (lane + 16) & 63;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const double t0 = T[(0 - PH) & 3][h], t3 = T[(3 - PH) & 3][h], b0 = B[PH & 3][h];
    const double x = __shfl(t3, up, 64), y = __shfl(b0, dn, 64);
    T[(3 - PH) & 3][h] = (g == 0) ? t0 : x;
    T[(0 - PH) & 3][h] = (g == 0) ? b0 : t0;
    B[PH & 3][h] = (g == 3) ? t3 : y;
  }
}

// Cross rounds rotate only the pairs (index k, index 16 + (k + r) mod 16),
// r = 0..15: the tops stay in place (top row of pair 4g + i in T[i]) and the
// bottoms move one pair down cyclically (bottom row of pair 4g + i at phase PH
// in B[(i + PH) & 3]); 16 rounds meet every pair across the two halves once
// and restore the slot order.  Row move: only the bottom of pair 4g + 3 is new
// to its register, taken from the next group's pair 4g + 4 (pair 15 from
// pair 0).
template <int PH>
__device__ __forceinline__ void cross_rows(double (&B)[4][2], int lane) {
  const int dn = (lane + 16) & 63;
#pragma unroll
  for (int h = 0; h < 2; ++h) B[PH & 3][h] = __shfl(B[PH & 3][h], dn, 64);
}

// column move of a cross round: bottom(c) <- bottom((c + 1) mod 16), the
// row_ror:15 rotation within each row of 16 lanes (every lane has a source,
// so the old value passed is the source itself: no register to clear)
__device__ __forceinline__ double cross_col(double bot) {
  const long long b = __double_as_longlong(bot);
  const int blo = (int)(unsigned int)(b & 0xffffffffll), bhi = (int)(b >> 32);
  const int lo = __builtin_amdgcn_update_dpp(blo, blo, 0x12F, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(bhi, bhi, 0x12F, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// 1/x and 1/sqrt(x) for normal positive x: the hardware approximations
// refined by two Newton steps each (relative error near 1 ulp)
__device__ __forceinline__ double rcp_refined(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

__device__ __forceinline__ double rsq_refined(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  return fma(y, e, y);
}

template <int PH, bool CROSS = false>
__device__ __forceinline__ void jac_roundR(double (&T)[4][2], double (&B)[4][2], double (&QT)[4][2],
                                           double (&QB)[4][2], double2* rcs, int lane, int g, int c, int id, bool diag,
                                           double tol, int& total) {
  double app = 0.0, apq = 0.0, aqq = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (id == i) {
      app = T[CROSS ? i : (i - PH) & 3][0];
      apq = T[CROSS ? i : (i - PH) & 3][1];
      aqq = B[(i + PH) & 3][1];
    }
  double cr = 1.0, sr = 0.0, tr = 0.0;
  bool rot = false;
  if (diag && fabs(apq) > tol) {
    // t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)), theta = d / e, written as
    // sgn(d) e / (|d| + sqrt(d^2 + e^2)); reciprocals from the hardware
    // approximations with two Newton steps (a shorter dependent chain than
    // three IEEE divisions and two square roots)
    const double d = aqq - app, e = apq + apq;
    const double x = fma(d, d, e * e);
    const double t = (d >= 0.0 ? e : -e) * rcp_refined(fabs(d) + x * rsq_refined(x));
    tr = t;
    cr = rsq_refined(fma(t, t, 1.0));
    sr = t * cr;
    rot = true;
  }
  if (__ballot(rot)) {
    total += 1;
    if (diag) rcs[c] = make_double2(cr, sr);
    __syncthreads();
    const double2 col = rcs[c];
    double2 row[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) row[i] = rcs[4 * g + i];
    __syncthreads();
    const double c2 = col.x, s2 = col.y;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double* t = T[CROSS ? i : (i - PH) & 3];
      double* b = B[(i + PH) & 3];
      const double c1 = row[i].x, s1 = row[i].y;
      const double tt = t[0], tb = t[1], bt = b[0], bb = b[1];
      const double btt = c2 * tt - s2 * tb, btb = s2 * tt + c2 * tb;
      const double bbt = c2 * bt - s2 * bb, bbb = s2 * bt + c2 * bb;
      t[0] = c1 * btt - s1 * bbt;
      b[0] = s1 * btt + c1 * bbt;
      t[1] = c1 * btb - s1 * bbb;
      b[1] = s1 * btb + c1 * bbb;
      if (rot && i == id) {
        t[0] = app - tr * apq;
        b[1] = aqq + tr * apq;
        t[1] = 0.0;
        b[0] = 0.0;
      }
      double* qt = QT[CROSS ? i : (i - PH) & 3];
      double* qb = QB[(i + PH) & 3];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double x = qt[h], y = qb[h];
        qt[h] = c1 * x - s1 * y;
        qb[h] = s1 * x + c1 * y;
      }
    }
  }
  if (CROSS) {
    cross_rows<PH>(B, lane);
    cross_rows<PH>(QB, lane);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T[i][1] = cross_col(T[i][1]);
      B[i][1] = cross_col(B[i][1]);
    }
  } else {
    circle_rows<PH>(T, B, lane, g);
    circle_rows<PH>(QT, QB, lane, g);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      circle_cols(T[i][0], T[i][1], c);
      circle_cols(B[i][0], B[i][1], c);
    }
  }
}

// With offmax given, a problem at a sweep boundary is first tested here (the
// work of a separate marking pass): its waves read the off-diagonal maximum
// the tile updates measured at the end of the previous sweep (buffer of that
// sweep's parity); the wave of pair 0 clears the buffer the current sweep
// will fill and, for a finished problem, records the outcome in state, which
// every wave of the problem decides alike (done when max <= tol, failed
// after the problem's maxsw sweeps).
constexpr int kCrossPeriod = 8;

template <bool CROSSK>
__device__ __forceinline__ void jac_pivot_body(const JacobiProb* __restrict__ probs, const int2* __restrict__ list,
                                               int idx, int r, const double* __restrict__ A, double* Qbuf, int* rotflag,
                                               const double* __restrict__ scale_in, int* state, int inner_sweeps,
                                               unsigned long long* work, double tolrel, unsigned long long* offmax,
                                               int stride) {
  __shared__ double2 rcs[16];
  __shared__ double red[2];
  const int2 it = list[idx];
  if (it.x < 0) return;   // skip entry of an XCD-ordered list
  const JacobiProb pr = probs[it.x];
  int* st = state + it.x * kJacState;
  if (st[0]) return;
  if (offmax && r < kBanded && r % pr.rounds == 0) {
    const int sweep = r / pr.rounds;
    const double m = __longlong_as_double((long long)offmax[((sweep + 1) & 1) * stride + it.x]);
    const bool conv = m <= tolrel * scale_in[it.x];
    const bool fail = !conv && sweep >= pr.maxsw;
    if (it.y == 0 && threadIdx.x == 0) {
      offmax[(sweep & 1) * stride + it.x] = 0ull;
      if (conv || fail) {
        st[5] = sweep;
        st[1] = fail ? 1 : 0;
        st[0] = 1;
      }
    }
    if (conv || fail) return;
  }
  const int n = pr.dim, ld = pr.ld;
  int bp, bq;
  jac_pair(pr.nbp, pr.rounds, r, it.y, bp, bq);
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const double* Ap = A + pr.a_off;
  double T[4][2], B[4][2];   // rows (top, bottom) of pair 4g + i, columns (top, bottom) of pair c
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int hc = 0; hc < 2; ++hc) {
      const int gt = 16 * bp + 4 * g + i, gb = 16 * bq + 4 * g + i, gv = (hc ? 16 * bq : 16 * bp) + c;
      T[i][hc] = (gt < n && gv < n) ? Ap[(long long)gt * ld + gv] : 0.0;
      B[i][hc] = (gb < n && gv < n) ? Ap[(long long)gb * ld + gv] : 0.0;
    }
  double QT[4][2], QB[4][2];   // Q^T rows (top, bottom slot of pair 4g + i), columns (c, c + 16)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double d = (4 * g + i == c) ? 1.0 : 0.0;
    QT[i][0] = d;
    QT[i][1] = 0.0;
    QB[i][0] = 0.0;
    QB[i][1] = d;
  }
  const double tol = tolrel * scale_in[it.x];
  const bool diag = (c >> 2) == g;   // this lane owns the diagonal block of pair c
  const int id = c & 3;
  int total = 0;
  // cross mode (CROSSK, a separate instantiation: its second loop costs
  // registers): outer rounds other than the first of every kCrossPeriod
  // rotate only the pairs across the two blocks.  The period is the same for
  // every problem so that the waves of a launch take equally long; the pairs
  // inside a block are met at least every kCrossPeriod rounds.
  const bool cross = CROSSK && r < kBanded && r % kCrossPeriod != 0;
  for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool on = diag && i == id;
      if (!on) m = fmax(m, fabs(T[i][0]));
      m = fmax(m, fabs(T[i][1]));
      m = fmax(m, fabs(B[i][0]));
      if (!on) m = fmax(m, fabs(B[i][1]));
    }
    m = block_max<64>(m, red);
    if (m <= tol) break;
    if (cross) {
#pragma unroll 1
      for (int rr = 0; rr < 16; rr += 4) {
        jac_roundR<0, true>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
        jac_roundR<1, true>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
        jac_roundR<2, true>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
        jac_roundR<3, true>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
      }
      continue;
    }
    for (int rr = 0; rr < 28; rr += 4) {
      jac_roundR<0>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
      jac_roundR<1>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
      jac_roundR<2>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
      jac_roundR<3>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
    }
    jac_roundR<0>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
    jac_roundR<1>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
    jac_roundR<2>(T, B, QT, QB, rcs, lane, g, c, id, diag, tol, total);
    // phase 3 -> 0: top row i from T[(i + 1) & 3], bottom row i from B[(i + 3) & 3]
    double t0[4][2], b0[4][2], qt0[4][2], qb0[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        t0[i][h] = T[(i + 1) & 3][h];
        b0[i][h] = B[(i + 3) & 3][h];
        qt0[i][h] = QT[(i + 1) & 3][h];
        qb0[i][h] = QB[(i + 3) & 3][h];
      }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        T[i][h] = t0[i][h];
        B[i][h] = b0[i][h];
        QT[i][h] = qt0[i][h];
        QB[i][h] = qb0[i][h];
      }
  }
  double* Qp = Qbuf + pr.q_off + (long long)it.y * 32 * 32;
  if (total > 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = 4 * g + i;
      Qp[c * 32 + k] = QT[i][0];
      Qp[(c + 16) * 32 + k] = QT[i][1];
      Qp[c * 32 + 16 + k] = QB[i][0];
      Qp[(c + 16) * 32 + 16 + k] = QB[i][1];
    }
  if (threadIdx.x == 0) {
    rotflag[(int)(pr.q_off / (32 * 32)) + it.y] = total > 0 ? 1 : 0;
    if (work && total > 0) prof_add(work, (unsigned long long)(9 * 32 * 32) * (unsigned long long)total);
  }
}

// sampled launches (ts != nullptr) record the device wall clock at the start
// of their earliest wave and the end of their latest: the execution span.
// The occupancy target sets the register budget: the cross-round kernel needs
// about 180 VGPRs, so a target of 3 waves per SIMD (168 VGPRs) spills the
// rotation buffer's address and reloads it from scratch in every inner round
// (26.6 us per launch against 23.0 us with 2 waves per SIMD on the
// 512-molecule batch, which has about 2 pivot waves per SIMD per round).
#ifndef XTBAMD_PIVOT_WAVES
#define XTBAMD_PIVOT_WAVES 2
#endif
template <bool CROSSK>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(XTBAMD_PIVOT_WAVES)))
k_jac_pairsR(const JacobiProb* __restrict__ probs, const int2* __restrict__ list, int r, const double* __restrict__ A,
             double* Qbuf, int* rotflag, const double* __restrict__ scale_in, int* state, int inner_sweeps,
             unsigned long long* work, double tolrel, unsigned long long* offmax, int stride, unsigned long long* ts) {
  prof_clock_start(ts);
  jac_pivot_body<CROSSK>(probs, list, blockIdx.x, r, A, Qbuf, rotflag, scale_in, state, inner_sweeps, work, tolrel,
                         offmax, stride);
  prof_clock_end(ts);
}

template __global__ void k_jac_pairsR<false>(const JacobiProb*, const int2*, int, const double*, double*, int*,
                                             const double*, int*, int, unsigned long long*, double,
                                             unsigned long long*, int, unsigned long long*);
template __global__ void k_jac_pairsR<true>(const JacobiProb*, const int2*, int, const double*, double*, int*,
                                            const double*, int*, int, unsigned long long*, double,
                                            unsigned long long*, int, unsigned long long*);

// ---------------------------------------------------------------------------
// Tile updates of the block-Jacobi rounds with 32 x 32 pivots, one wave per
// tile, without LDS: operands go from global memory straight into the lane
// layout of v_mfma_f64_16x16x4f64 (lane = 16 k4 + i16; A operand (row i16,
// k k4), B operand (k k4, col i16), accumulator register i (row k4 + 4i,
// col i16)) and results are stored from it.  A 16 x 16 accumulator block is
// the B operand of a following product for its k-steps 4i..4i+3, so the
// two-sided update chains two products in registers.  W is kept as rows
// (row j = eigenvector j), which makes every load and store coalesced.

// global index of position u (0..31) of pivot pair (bp, bq)
__device__ __forceinline__ int pair_index(int bp, int bq, int u) { return u < 16 ? 16 * bp + u : 16 * bq + u - 16; }

// A[I_k, I_l] <- Q_k^T A[I_k, I_l] Q_l for pair tiles k <= l (A symmetric).
// In the last round of a problem's sweep every tile also folds the largest
// off-diagonal magnitude it leaves (strict upper triangle; the tiles k <= l
// hold every index pair once) into offmax, the convergence test of the next
// sweep boundary (order independent: max on the bit patterns).
__device__ __forceinline__ void jac_tile_a(const JacobiProb* __restrict__ probs, const int4 it, int r, double* A,
                                           const double* __restrict__ Qbuf, const int* __restrict__ rotflag,
                                           const int* __restrict__ state, unsigned long long* offmax, int stride) {
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int k = it.y, l = it.z;
  const int qb = (int)(pr.q_off / (32 * 32));
  const int rk = rotflag[qb + k], rl = rotflag[qb + l];
  const bool measure = offmax && r < kBanded && (r + 1) % pr.rounds == 0;
  if (measure) offmax += ((r / pr.rounds) & 1) * stride;
  if (!rk && !rl && !measure) return;
  const int n = pr.dim, ld = pr.ld;
  int kp, kq, lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, k, kp, kq);
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
  double* Ap = A + pr.a_off;
  const double* Qk = Qbuf + pr.q_off + (long long)k * 32 * 32;
  const double* Ql = Qbuf + pr.q_off + (long long)l * 32 * 32;
  const int lane = threadIdx.x, i16 = lane & 15, k4 = lane >> 4;
  // Q_k is loaded with the first product's operands, so that its latency
  // overlaps them instead of following the first product
  double qv[8][2];
  if (rk)
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int h = 0; h < 2; ++h) qv[s][h] = Qk[(4 * s + k4) * 32 + 16 * h + i16];
  v4d t[2][2];
  if (rl) {
    // T = A_kl Q_l; A_kl(u, w) is read as A(I_l[w], I_k[u]) (coalesced over u)
    double av[8][2], bv[8][2];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int w = 4 * s + k4, gw = pair_index(lp, lq, w);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int gu = pair_index(kp, kq, 16 * h + i16);
        av[s][h] = (gw < n && gu < n) ? Ap[(long long)gw * ld + gu] : 0.0;
        bv[s][h] = Ql[w * 32 + 16 * h + i16];
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) t[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) t[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s][a], bv[s][b], t[a][b], 0, 0, 0);
  } else {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int gu = pair_index(kp, kq, 16 * a + k4 + 4 * i), gv = pair_index(lp, lq, 16 * b + i16);
          t[a][b][i] = (gu < n && gv < n) ? Ap[(long long)gu * ld + gv] : 0.0;
        }
  }
  v4d o[2][2];
  if (rk) {
    // out = Q_k^T T; the A operand Q_k^T(u, w) = Q_k(w, u), the B operand is T
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) o[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          o[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(qv[s][a], t[s >> 2][b][s & 3], o[a][b], 0, 0, 0);
  } else {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) o[a][b] = t[a][b];
  }
  // store the tile and its mirror; a diagonal tile stores its upper triangle
  // to both places, so A stays exactly symmetric
  if (rk || rl)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * a + k4 + 4 * i, v = 16 * b + i16;
          const int gu = pair_index(kp, kq, u), gv = pair_index(lp, lq, v);
          if (gu >= n || gv >= n || (k == l && u > v)) continue;
          Ap[(long long)gu * ld + gv] = o[a][b][i];
          Ap[(long long)gv * ld + gu] = o[a][b][i];
        }
  if (measure) {
    double m = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int u = 16 * a + k4 + 4 * i, v = 16 * b + i16;
          const int gu = pair_index(kp, kq, u), gv = pair_index(lp, lq, v);
          if (gu < n && gv < n && (k < l || u < v)) m = fmax(m, fabs(o[a][b][i]));
        }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (lane == 0 && m > 0.0) atomicMax(&offmax[it.x], (unsigned long long)__double_as_longlong(m));
  }
}

// W rows I_l, columns of row block rb: Wr[I_l, ...] <- Q_l^T Wr[I_l, ...]
__device__ __forceinline__ void jac_tile_w(const JacobiProb* __restrict__ probs, const int4 it, int r, double* W,
                                           const double* __restrict__ Qbuf, const int* __restrict__ rotflag,
                                           const int* __restrict__ state) {
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return;
  const int l = it.z, rb = it.y;
  const int qb = (int)(pr.q_off / (32 * 32));
  if (!rotflag[qb + l]) return;
  const int n = pr.dim, ld = pr.ld;
  int lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, l, lp, lq);
  double* Wp = W + pr.w_off;
  const double* Ql = Qbuf + pr.q_off + (long long)l * 32 * 32;
  const int lane = threadIdx.x, i16 = lane & 15, k4 = lane >> 4;
  double qv[8][2], wv[8][2];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int w = 4 * s + k4, gw = pair_index(lp, lq, w);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int x = 32 * rb + 16 * h + i16;
      qv[s][h] = Ql[w * 32 + 16 * h + i16];
      wv[s][h] = (gw < n && x < n) ? Wp[(long long)gw * ld + x] : 0.0;
    }
  }
  v4d o[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) o[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) o[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(qv[s][a], wv[s][b], o[a][b], 0, 0, 0);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gv = pair_index(lp, lq, 16 * a + k4 + 4 * i), x = 32 * rb + 16 * b + i16;
        if (gv < n && x < n) Wp[(long long)gv * ld + x] = o[a][b][i];
      }
}

__global__ void __launch_bounds__(64) k_jac_apply_aR(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                     int r, double* A, const double* __restrict__ Qbuf,
                                                     const int* __restrict__ rotflag, const int* __restrict__ state) {
  jac_tile_a(probs, list[blockIdx.x], r, A, Qbuf, rotflag, state, nullptr, 0);
}

__global__ void __launch_bounds__(64) k_jac_apply_wR(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                     int r, double* W, const double* __restrict__ Qbuf,
                                                     const int* __restrict__ rotflag, const int* __restrict__ state) {
  jac_tile_w(probs, list[blockIdx.x], r, W, Qbuf, rotflag, state);
}

// both tile updates of a round in one launch, one list entry (m, y, z, kind)
// per workgroup: kind 0 updates A tile (y, z) = (k, l) (and measures the
// off-diagonal maximum into the buffer of the sweep's parity, offmax[(sweep &
// 1) * stride + problem]), kind 1 W tile (rb, l) = (y, z), kind 2 is a skip
// algorithmic HBM bytes of one tile update: every element of A and W that a
// rotation touches is read once and written once per round, A through one of
// its two mirrored tiles (an off-diagonal A tile: 8 B read, 16 B written per
// element; a diagonal one 8 + 8; a W tile 8 + 8); untouched tiles count zero
__device__ __forceinline__ unsigned long long jac_tile_bytes(const JacobiProb* __restrict__ probs, const int4 it,
                                                             int r, const int* __restrict__ rotflag,
                                                             const int* __restrict__ state) {
  if (it.w > 1) return 0ull;
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState]) return 0ull;
  const int qb = (int)(pr.q_off / (32 * 32)), n = pr.dim;
  auto valid = [&](int b, int w) { return (long long)max(0, min(w, n - w * b)); };
  int lp, lq;
  jac_pair(pr.nbp, pr.rounds, r, it.z, lp, lq);
  const long long cl = valid(lp, 16) + valid(lq, 16);
  if (it.w == 1) return rotflag[qb + it.z] ? 16ull * cl * valid(it.y, 32) : 0ull;
  const bool measure = r < kBanded && (r + 1) % pr.rounds == 0;
  if (!rotflag[qb + it.y] && !rotflag[qb + it.z] && !measure) return 0ull;
  int kp, kq;
  jac_pair(pr.nbp, pr.rounds, r, it.y, kp, kq);
  const long long ck = valid(kp, 16) + valid(kq, 16);
  return (it.y == it.z ? 16ull : 24ull) * ck * cl;
}

__global__ void __launch_bounds__(64) k_jac_apply_R(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                    int r, double* A, double* W, const double* __restrict__ Qbuf,
                                                    const int* __restrict__ rotflag, const int* __restrict__ state,
                                                    unsigned long long* offmax, int stride,
                                                    unsigned long long* work, unsigned long long* ts) {
  // sampled launches record the device wall clock at the start of their
  // earliest wave (ts[0], minimum) and the end of their latest (ts[1],
  // maximum): the launch's execution span
  prof_clock_start(ts);
  const int4 it = list[blockIdx.x];
  // sampled launches count their algorithmic bytes (one atomic per tile)
  if (work && threadIdx.x == 0) {
    const unsigned long long b = jac_tile_bytes(probs, it, r, rotflag, state);
    if (b) prof_add(work, b);
  }
  if (it.w == 0)
    jac_tile_a(probs, it, r, A, Qbuf, rotflag, state, offmax, stride);
  else if (it.w == 1)
    jac_tile_w(probs, it, r, W, Qbuf, rotflag, state);
  prof_clock_end(ts);
}


// initialise W = I and the per-problem scale max|A_ii| and state, clear
// both off-diagonal maximum buffers (offmax[0, 2 * nprob)); molecules that
// finished their SCC are marked done and skipped
__global__ void __launch_bounds__(256) k_jac_init(const JacobiProb* __restrict__ probs, const double* __restrict__ A,
                                                  double* W, double* scale, int* state,
                                                  const MolState* __restrict__ mstate, unsigned long long* offmax,
                                                  int keep_w, double loose, double qrel, double tol) {
  __shared__ double sh[4];
  const JacobiProb pr = probs[blockIdx.x];
  int* st = state + blockIdx.x * kJacState;
  if (mstate && !mstate[blockIdx.x].active) {
    if (threadIdx.x == 0) {
      for (int i = 0; i < kJacState; ++i) st[i] = 0;
      st[0] = 1;
      offmax[blockIdx.x] = 0ull;
      offmax[gridDim.x + blockIdx.x] = 0ull;
    }
    return;
  }
  const int n = pr.dim, ld = pr.ld;
  double* Wp = W + pr.w_off;
  const double* Ap = A + pr.a_off;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(Ap[(long long)i * ld + i]));
  if (!keep_w)
    for (long long e = threadIdx.x; e < (long long)n * n; e += blockDim.x)
      Wp[(e / n) * ld + e % n] = (e / n == e % n) ? 1.0 : 0.0;
  m = block_max<256>(m, sh);
  if (threadIdx.x == 0) {
    // a molecule before its final diagonalization converges to a looser
    // tolerance tied to its SCC progress: off-diagonal residuals of qrel x
    // the last iteration's RMS charge change (between tol and loose x tol;
    // the first iteration takes the loosest), so that the eigenvector error
    // stays far below the change the next mixing step makes anyway
    double f = 1.0;
    if (mstate && !mstate[blockIdx.x].lastdiag) {
      const MolState& ms = mstate[blockIdx.x];
      f = ms.iter > 0 ? fmin(loose, fmax(1.0, qrel * ms.rmsq / tol)) : loose;
    }
    scale[blockIdx.x] = fmax(m, 1.0) * f;
    for (int i = 0; i < kJacState; ++i) st[i] = 0;
    offmax[blockIdx.x] = 0ull;
    offmax[gridDim.x + blockIdx.x] = 0ull;
  }
}

// copies between a problem's matrix in the caller's packed layout (row stride
// dim at src_off) and its padded working copy (row stride ld at a_off / w_off):
// in = 1: src packed -> dst padded (A); in = 0: src padded -> dst packed (W)
__global__ void __launch_bounds__(256) k_jac_pad(const JacobiProb* __restrict__ probs, const double* __restrict__ src,
                                                 double* dst, const MolState* __restrict__ mstate, int in) {
  const int p = blockIdx.y;
  if (mstate && !mstate[p].active) return;
  const JacobiProb pr = probs[p];
  const long long n = pr.dim;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (long long)gridDim.x * blockDim.x) {
    const long long u = e / n, v = e % n;
    if (in)
      dst[pr.a_off + u * pr.ld + v] = src[pr.src_off + e];
    else
      dst[pr.src_off + e] = src[pr.w_off + u * pr.ld + v];
  }
}

// largest off-diagonal magnitude of the unfinished block-path problems that
// start a sweep at global round r (r a multiple of the problem's rounds per
// sweep), one workgroup per upper-triangular 64 x 64 tile; max over
// non-negative doubles is taken on their bit patterns, which order the same
// way (order independent).
__global__ void __launch_bounds__(256) k_jac_offmax(const JacobiProb* __restrict__ probs, const int4* __restrict__ list,
                                                    const double* __restrict__ A, const int* __restrict__ state,
                                                    unsigned long long* offmax, int r) {
  __shared__ double sh[4];
  const int4 it = list[blockIdx.x];
  const JacobiProb pr = probs[it.x];
  if (state[it.x * kJacState] || !jac_boundary(pr.rounds, r)) return;
  const int n = pr.dim, ld = pr.ld;
  const double* Ap = A + pr.a_off;
  double m = 0.0;
  for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
    const int gu = it.y * 64 + (e >> 6), gv = it.z * 64 + (e & 63);
    if (gu < gv && gv < n) m = fmax(m, fabs(Ap[(long long)gu * ld + gv]));
  }
  m = block_max<256>(m, sh);
  if (threadIdx.x == 0 && m > 0.0) atomicMax(&offmax[it.x], (unsigned long long)__double_as_longlong(m));
}

// convergence test of the problems at a sweep boundary (per problem, so a
// problem's rounds never depend on the other problems of the batch): done
// when every off-diagonal element is <= tol, failed after maxsw sweeps;
// counts the problems still running into nactive.  group >= 0 restricts the
// test to the problems of that round lane.
__global__ void __launch_bounds__(256) k_jac_mark(const JacobiProb* __restrict__ probs, int nprob,
                                                  const double* __restrict__ scale, int* state,
                                                  unsigned long long* offmax, int r, int* nactive, double tolrel,
                                                  int group) {
  __shared__ double sh[4];
  double c = 0.0;
  for (int p = threadIdx.x; p < nprob; p += blockDim.x) {
    const JacobiProb pr = probs[p];
    if (pr.dim <= 64) continue;   // solved by k_jac_small
    if (group >= 0 && pr.group != group) continue;
    int* st = state + p * kJacState;
    if (st[0]) continue;
    if (jac_boundary(pr.rounds, r)) {
      const int sweep = r < kBanded ? r / pr.rounds : 0;
      const double m = __longlong_as_double((long long)offmax[p]);
      offmax[p] = 0ull;
      if (m <= tolrel * scale[p]) {
        st[0] = 1;
        st[5] = sweep;
        continue;
      }
      if (sweep >= pr.maxsw) {
        st[0] = 1;
        st[1] = 1;
        st[5] = sweep;
        continue;
      }
    }
    c += 1.0;
  }
  c = block_sum<256>(c, sh);
  if (threadIdx.x == 0) nactive[0] = (int)c;
}

// problems of a lane (group) still running on the block path
__global__ void __launch_bounds__(256) k_jac_count(const JacobiProb* __restrict__ probs, int nprob,
                                                   const int* __restrict__ state, int group, int* nactive) {
  __shared__ double sh[4];
  double c = 0.0;
  for (int p = threadIdx.x; p < nprob; p += blockDim.x) {
    const JacobiProb pr = probs[p];
    if (pr.dim <= 64 || pr.group != group) continue;
    if (!state[p * kJacState]) c += 1.0;
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

// ---------------------------------------------------------------------------
// Dense symmetric eigensolver for one problem per workgroup: Householder
// tridiagonalisation (A = Q T Q^T, vectors kept in the rows of A), explicit
// Q^T, implicit-shift QL on T with the rotations applied to the rows of
// W^T = Q^T (Golub & Van Loan, "Matrix Computations", 4th ed., 8.3.1 and
// 8.3.3; the same sequence as EISPACK tred2/tql2).  Row i of W (stored at
// w_off) is the eigenvector of eigenvalue d_i; ev/perm receive the ascending
// order.  Dynamic LDS: 5 * nmax doubles (d, e, tau and two work vectors).
__global__ void __launch_bounds__(256) k_tred_ql(const JacobiProb* __restrict__ probs, const int* __restrict__ list,
                                                 double* A, double* W, double* scratch, double* ev, int* perm,
                                                 const MolState* __restrict__ mstate, int lds_cap, int* failed,
                                                 unsigned long long* work) {
  extern __shared__ double lds[];
  __shared__ double red[4];
  __shared__ double sc[4];
  __shared__ int ctl[4];
  const int pidx = list[blockIdx.x];
  if (mstate && !mstate[pidx].active) return;
  const JacobiProb pr = probs[pidx];
  const int N = pr.dim;
  // vectors in LDS when they fit, else in the problem's scratch matrix
  double* base = (N <= lds_cap) ? lds : scratch + pr.a_off;
  double* d = base;
  double* e = base + N;
  double* tau = base + 2 * N;
  double* v = base + 3 * N;   // Householder vector, then rotation cosines
  double* w = base + 4 * N;   // p / w vectors, then rotation sines
  const int tid = threadIdx.x, nt = blockDim.x;
  double* M = A + pr.a_off;
  double* Wt = W + pr.w_off;
  // ---- tridiagonalisation
  for (int k = 0; k + 2 < N; ++k) {
    double s2 = 0.0;
    for (int j = k + 2 + tid; j < N; j += nt) s2 += M[(long long)k * N + j] * M[(long long)k * N + j];
    s2 = block_sum<256>(s2, red);
    if (tid == 0) {
      const double alpha = M[(long long)k * N + k + 1];
      d[k] = M[(long long)k * N + k];
      if (s2 == 0.0) {
        tau[k] = 0.0;
        e[k] = alpha;
        sc[0] = 0.0;
      } else {
        const double beta = -copysign(sqrt(alpha * alpha + s2), alpha);
        tau[k] = (beta - alpha) / beta;
        e[k] = beta;
        sc[0] = 1.0 / (alpha - beta);
      }
    }
    __syncthreads();
    const double t = tau[k];
    if (t == 0.0) continue;   // uniform
    const double scale = sc[0];
    for (int j = k + 1 + tid; j < N; j += nt) {
      const double vj = (j == k + 1) ? 1.0 : M[(long long)k * N + j] * scale;
      v[j] = vj;
      M[(long long)k * N + j] = vj;   // kept for Q
    }
    __syncthreads();
    // p = t A22 v (columns by thread, rows in the loop: coalesced)
    for (int i = k + 1 + tid; i < N; i += nt) {
      double acc = 0.0;
      for (int j = k + 1; j < N; ++j) acc = fma(M[(long long)j * N + i], v[j], acc);
      w[i] = t * acc;
    }
    __syncthreads();
    double pv = 0.0;
    for (int i = k + 1 + tid; i < N; i += nt) pv += w[i] * v[i];
    pv = block_sum<256>(pv, red);
    const double K = 0.5 * t * pv;
    for (int i = k + 1 + tid; i < N; i += nt) w[i] -= K * v[i];
    __syncthreads();
    // A22 -= v w^T + w v^T
    for (int i = k + 1 + tid; i < N; i += nt) {
      const double vi = v[i], wi = w[i];
      for (int j = k + 1; j < N; ++j) M[(long long)j * N + i] -= v[j] * wi + w[j] * vi;
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (N >= 2) {
      d[N - 2] = M[(long long)(N - 2) * N + N - 2];
      e[N - 2] = M[(long long)(N - 2) * N + N - 1];
    }
    d[N - 1] = M[(long long)(N - 1) * N + N - 1];
    e[N - 1] = 0.0;
  }
  // ---- W^T = Q^T = H_{N-3} ... H_0, built as W^T <- H_k W^T
  for (long long x = tid; x < (long long)N * N; x += nt) Wt[x] = (x / N == x % N) ? 1.0 : 0.0;
  __syncthreads();
  for (int k = 0; k + 2 < N; ++k) {
    const double t = tau[k];
    if (t == 0.0) continue;
    for (int j = k + 1 + tid; j < N; j += nt) v[j] = M[(long long)k * N + j];
    __syncthreads();
    for (int c = tid; c < N; c += nt) {
      double u = 0.0;
      for (int r = k + 1; r < N; ++r) u = fma(v[r], Wt[(long long)r * N + c], u);
      u *= t;
      for (int r = k + 1; r < N; ++r) Wt[(long long)r * N + c] -= v[r] * u;
    }
    __syncthreads();
  }
  // ---- implicit QL; thread 0 runs the recurrence of a chase, all threads
  // then rotate their columns of W^T by its (c, s) sequence
  unsigned long long nrot = 0;
  for (int l = 0; l < N; ++l) {
    int iter = 0;
    for (;;) {
      if (tid == 0) {
        int m = l;
        for (; m < N - 1; ++m) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
        }
        ctl[0] = (m == l || iter >= 60) ? 1 : 0;
        if (iter >= 60 && m != l) failed[pidx] = 1;
        ctl[1] = m;
        ctl[2] = l;   // lowest rotated index
        if (m != l && iter < 60) {
          double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
          double r = sqrt(g * g + 1.0);
          g = d[m] - d[l] + e[l] / (g + copysign(r, g));
          double s = 1.0, c = 1.0, p = 0.0;
          int i = m - 1;
          bool early = false;
          for (; i >= l; --i) {
            const double f = s * e[i], b = c * e[i];
            r = sqrt(f * f + g * g);
            e[i + 1] = r;
            if (r == 0.0) {
              d[i + 1] -= p;
              e[m] = 0.0;
              early = true;
              break;
            }
            s = f / r;
            c = g / r;
            g = d[i + 1] - p;
            r = (d[i] - g) * s + 2.0 * c * b;
            p = s * r;
            d[i + 1] = g + p;
            g = c * r - b;
            v[i] = c;
            w[i] = s;
          }
          ctl[2] = early ? i + 1 : l;
          if (!early) {
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
          }
        }
      }
      __syncthreads();
      if (ctl[0]) break;   // uniform
      const int m = ctl[1], lo = ctl[2];
      nrot += (unsigned long long)(m - lo);
      for (int col = tid; col < N; col += nt) {
        for (int i = m - 1; i >= lo; --i) {
          const double c = v[i], s = w[i];
          const double zi = Wt[(long long)i * N + col], zj = Wt[(long long)(i + 1) * N + col];
          Wt[(long long)(i + 1) * N + col] = s * zi + c * zj;
          Wt[(long long)i * N + col] = c * zi - s * zj;
        }
      }
      ++iter;
      __syncthreads();
    }
    __syncthreads();
  }
  // ---- ascending order (stable rank sort)
  for (int i = tid; i < N; i += nt) {
    const double di = d[i];
    int rank = 0;
    for (int j = 0; j < N; ++j) {
      const double dj = d[j];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    ev[pr.ev_off + rank] = di;
    perm[pr.ev_off + rank] = i;
  }
  // algorithmic flops: tridiagonalisation 4/3 N^3 and Q 4/3 N^3 (the
  // LAPACK dsytrd/dorgtr counts), 6 N per applied rotation
  if (work && tid == 0) {
    const unsigned long long n3 = (unsigned long long)N * N * N;
    prof_add(work, (8ull * n3) / 3ull + 6ull * (unsigned long long)N * nrot);
  }
}

// eigenvalues ascending: ev[k] = A[perm[k]][perm[k]] (stable rank sort)
// eigenvalues A_ii in ascending order (ties by index) and their permutation;
// grid (chunks of 256 eigenvalues, problems), the diagonal staged through
// LDS 2048 values at a time
__global__ void __launch_bounds__(256) k_jac_sort(const JacobiProb* __restrict__ probs, const double* __restrict__ A,
                                                  double* ev, int* perm, const MolState* __restrict__ mstate,
                                                  int max_dim) {
  __shared__ double dg[2048];
  const int p = blockIdx.y;
  if (mstate && !mstate[p].active) return;
  const JacobiProb pr = probs[p];
  if (pr.dim > max_dim) return;
  const int n = pr.dim, ld = pr.ld;
  if (blockIdx.x * blockDim.x >= (unsigned)n) return;
  const double* Ap = A + pr.a_off;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double di = i < n ? Ap[(long long)i * ld + i] : 0.0;
  int rank = 0;
  for (int j0 = 0; j0 < n; j0 += 2048) {
    const int m = min(2048, n - j0);
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += blockDim.x) dg[j] = Ap[(long long)(j0 + j) * ld + j0 + j];
    __syncthreads();
    if (i < n)
      for (int j = 0; j < m; ++j) {
        const double dj = dg[j];
        rank += (dj < di || (dj == di && j0 + j < i)) ? 1 : 0;
      }
  }
  if (i < n) {
    ev[pr.ev_off + rank] = di;
    perm[pr.ev_off + rank] = i;
  }
}

// W = I (row stride ld) of the problems still to solve; grid (chunks, problems)
__global__ void __launch_bounds__(256) k_jac_ident(const JacobiProb* __restrict__ probs, double* W,
                                                   const MolState* __restrict__ mstate) {
  const int p = blockIdx.y;
  if (mstate && !mstate[p].active) return;
  const JacobiProb pr = probs[p];
  const long long n = pr.dim;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (long long)gridDim.x * blockDim.x) {
    const long long u = e / n, v = e % n;
    W[pr.w_off + u * pr.ld + v] = (u == v) ? 1.0 : 0.0;
  }
}

}  // namespace xtbamd

namespace xtbamd {
// st[desc[k].mol] <- stage[k]: initial SCC states of a sub-batch, uploaded in one copy
__global__ void k_state_scatter(const MolDesc* __restrict__ desc, const MolState* __restrict__ stage, MolState* st,
                                int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) st[desc[k].mol] = stage[k];
}

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
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (long long)gridDim.x * blockDim.x) C[md.m0 + t] = T[md.m0 + t];
}
}  // namespace xtbamd
