// Blocked FP64 symmetric eigensolver for the large problems of a batch
// (replaces dsytrd + dstedc + dormtr inside the reference's dsyevd,
// src/mctc/lapack/eigensolve.F90:218-246, called from src/scc_core.f90:475):
//
//  1. Householder tridiagonalisation in panels of kEigNb columns.  Every
//     column is one launch of k_trd_col (the column's reflector, formed by
//     every workgroup from the partial sums the launch before left, and
//     w = A22 v with its panel corrections: one wave per trailing row,
//     coalesced row reads) and one of k_trd_row (w completed, the next
//     column's corrected row and its partial sums, 256 rows per workgroup);
//     a panel ends with the rank-2 nb update A22 -= X^T Y of the trailing
//     matrix on the matrix cores (k_bgemm).
//  2. Divide and conquer on the tridiagonal matrix: leaves of at most 64
//     solved whole in LDS, merges level by level with deflation (one
//     workgroup per merge), the secular equation's roots by safeguarded
//     Newton steps (one wave per root), eigenvectors from the recomputed
//     z of Gu and Eisenstat (SIAM J. Matrix Anal. Appl. 16 (1995) 172) and
//     one matrix-core product per merge.
//  3. Back-transformation of the eigenvectors with the compact WY form
//     I - V T V^T of blocks of kBtNb reflectors (the Gram matrices of all
//     blocks in one product launch, T by the dlarft recurrence, then three
//     matrix-core products per block).
//
// Layouts: every n x n array of a problem is row-major with row stride ld;
// the Householder vector of column j is row j of Vh (entries j+1..n-1, a 1 at
// j+1); eigenvectors are rows (row j = eigenvector j) throughout the divide
// and conquer and the back-transformation.
#pragma once
#include <hip/hip_runtime.h>

#include "device_util.h"
#include "engine.h"

namespace xtbamd {

constexpr int kEigNb = 32;      // panel width of the tridiagonalisation
constexpr int kDcLeaf = 64;     // largest leaf of the divide and conquer tree
constexpr int kEigVecs = 18;    // vectors of length ld per problem in the vector workspace
// slots of the vector workspace (doubles)
constexpr int kVd = 0, kVe = 1, kVtau = 2, kVw = 3, kVlam = 4, kVlam2 = 5, kVdn = 6, kVz = 7, kVdd = 8,
              kVzz = 9, kVtr = 10, kVzh = 11, kVrc = 12, kVrs = 13, kVy = 14, kVev = 15, kVd0 = 16, kVe0 = 17;
// slots of the integer workspace (ints, length ld each)
constexpr int kIperm = 0, kIkeep = 1, kIpos = 2, kIorg = 3, kIrot = 4, kIrot2 = 5, kIrow = 6, kIcnt = 7;
constexpr int kEigInts = 8;
constexpr int kBtChunk = 512;   // K chunk of the back-transformation's product Z V
constexpr int kBtNb = 128;      // reflectors per block of the back-transformation (four panels)

// One batched product job: C = beta C + alpha op(A) op(B), op(A) M x K,
// op(B) K x N, row-major with the given row strides.  TA: A is stored K x M;
// TB: B is stored N x K.  bk (optional) gathers the K index of B's rows
// (B row bk[k] for op(B) row k, !TB only); crow (optional) scatters C's rows;
// kdev (optional) reads M = K = *kdev on the device (at most the tile grid's M).
struct GemmJob {
  int M, N, K;
  int tm, tn;             // tiles along M and N (64 each)
  long long tile0;        // first tile of this job in the launch
  const double* A;
  const double* B;
  double* C;
  long long lda, ldb, ldc;
  const int* bk;
  const int* crow;
  const int* kdev;
  const int* act;         // (optional) the job is skipped when *act == 0
  double alpha, beta;
};

// One problem of the blocked eigensolver.
struct EigProb {
  int n, ld;         // dimension, row stride of the solver's n x n arrays
  long long m_off;   // this problem's n x ld blocks in the solver's A, Vh, Z0, Z1, U
  long long v_off;   // this problem's kEigVecs x ld vector block
  long long i_off;   // this problem's kEigInts x ld integer block
  long long x_off;   // this problem's panel buffers X, Y (2 kEigNb x ld each)
  long long p_off;   // this problem's product buffers of the back-transformation
  long long p_off2;  // this problem's k_trd_row partial sums (ceil(n / 256) x kTrdPart)
  long long t_off;   // this problem's T factors (nblk x kBtNb x nchunk kBtNb: T repeated nchunk times)
  long long g_off;   // this problem's Gram matrices of the reflector blocks (nblk x kBtNb x kBtNb)
  int nchunk;        // K chunks of the back-transformation's first product
  int pad2_;
  int mol;           // batch slot (index of its Jacobi problem)
  int depth;         // levels of the divide and conquer tree (leaves at this depth)
};

// A merge of the divide and conquer tree: rows [start, start + n1 + n2) of
// problem prob, split after n1 rows.
struct DcNode {
  int prob, start, n1, n2;
};

// ---------------------------------------------------------------------------
// batched FP64 product on the matrix cores: 64 x 64 tiles, four waves of 2 x 2
// v_mfma_f64_16x16x4f64 blocks, operands staged through LDS in k-blocks of 16
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_bgemm(const GemmJob* __restrict__ jobs, int njobs,
                                               const int2* __restrict__ tiles) {
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int2 tl = tiles[blockIdx.x];
  const GemmJob jb = jobs[tl.x];
  if (jb.act && !*jb.act) return;
  int M = jb.M, K = jb.K;
  if (jb.kdev) {
    M = *jb.kdev;
    K = M;
  }
  const int N = jb.N;
  const int row0 = (tl.y / jb.tn) * 64, col0 = (tl.y % jb.tn) * 64;
  if (row0 >= M || col0 >= N) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, i16 = lane & 15, k4 = lane >> 4;
  typedef double v4d_ __attribute__((ext_vector_type(4)));
  v4d_ acc[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[r][c] = v4d_{0.0, 0.0, 0.0, 0.0};
  for (int kb = 0; kb < K; kb += 16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int ii, kk;
      if (TA) { kk = e >> 6; ii = e & 63; } else { ii = e >> 4; kk = e & 15; }
      const int gi = row0 + ii, gk = kb + kk;
      double x = 0.0;
      if (gi < M && gk < K) x = TA ? jb.A[(long long)gk * jb.lda + gi] : jb.A[(long long)gi * jb.lda + gk];
      As[kk][ii] = x;
      int jj, k2;
      if (TB) { jj = e >> 4; k2 = e & 15; } else { k2 = e >> 6; jj = e & 63; }
      const int gj = col0 + jj, gk2 = kb + k2;
      double y = 0.0;
      if (gj < N && gk2 < K) {
        if (TB) y = jb.B[(long long)gj * jb.ldb + gk2];
        else y = jb.B[(long long)(jb.bk ? jb.bk[gk2] : gk2) * jb.ldb + gj];
      }
      Bs[k2][jj] = y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
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
        if (gi >= M || gj >= N) continue;
        const long long row = jb.crow ? jb.crow[gi] : gi;
        double* cp = jb.C + row * jb.ldc + gj;
        *cp = jb.beta == 0.0 ? jb.alpha * acc[r][c][i] : jb.beta * *cp + jb.alpha * acc[r][c][i];
      }
    }
}

// ---------------------------------------------------------------------------
// 1. tridiagonalisation

// Two wide launches per column j (panel start j0, i = j - j0 earlier
// columns in the panel).  k_trd_row has left the corrected row a = A(j, j:n)
// of column j (with the panel's earlier vector pairs applied) and, per
// workgroup, the partial sums that column j's reflector needs: the squared
// norm of a(j+2:n) and the products of a(j+2:n) with the panel's v and w.
// k_trd_col forms the reflector from them (every workgroup for itself),
// w = A22 v over the trailing rows (one wave per row, coalesced row reads,
// v formed on the fly from a), the panel corrections and the scaling by tau,
// and the partial sums of w . v.  The next k_trd_row completes
// w += alpha v, stores the pair in the panel buffers and forms the next
// column's corrected row and partial sums.
// X rows [0, nb): v of the panel's columns, rows [nb, 2nb): their w; Y the
// other way round, so that the trailing update is A22 -= X^T Y.
constexpr int kTrdPart = 1 + 2 * kEigNb;   // partial sums per k_trd_row workgroup

__device__ __forceinline__ double sum256(double v, double* sh) { return block_sum<256>(v, sh); }

// column j's reflector from the corrected row a and k_trd_row's partials:
// beta, tau, scale (v = scale a beyond j + 1, v[j+1] = 1) and the panel
// products y1 = V^T v, y2 = W^T v (t < i) into LDS
__device__ __forceinline__ void trd_reflector(const double* __restrict__ a, const double* __restrict__ part,
                                              const double* __restrict__ Xp, int ld, int n, int j, int i,
                                              double& beta, double& tau, double& scale, double* y) {
  const int nw = (n - j + 255) / 256;   // k_trd_row workgroups over rows j..n-1
  // the partial sums with four loads in flight (a loop that adds each load
  // before the next is issued waits nw memory latencies)
  auto psum = [&](int k) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int w = 0;
    for (; w + 3 < nw; w += 4) {
      s0 += part[w * kTrdPart + k];
      s1 += part[(w + 1) * kTrdPart + k];
      s2 += part[(w + 2) * kTrdPart + k];
      s3 += part[(w + 3) * kTrdPart + k];
    }
    for (; w < nw; ++w) s0 += part[w * kTrdPart + k];
    return (s0 + s1) + (s2 + s3);
  };
  const double ss = psum(0);
  const double alpha = a[j + 1];
  beta = alpha;
  tau = 0.0;
  scale = 0.0;
  if (ss > 0.0) {
    beta = -copysign(sqrt(alpha * alpha + ss), alpha);
    tau = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  for (int t = threadIdx.x; t < 2 * i; t += blockDim.x) {
    const int tt = t < i ? t : t - i;
    const int row = t < i ? tt : kEigNb + tt;   // v rows, then w rows
    y[t] = Xp[(long long)row * ld + j + 1] + scale * psum(1 + (t < i ? 0 : kEigNb) + tt);
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256) k_trd_col(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                 const double* __restrict__ A, double* Vh, double* X, double* vec,
                                                 const double* __restrict__ part2, int j, int j0) {
  __shared__ double y[2 * kEigNb];
  __shared__ double sh[4];
  if (!act[blockIdx.y]) return;
  const EigProb p = probs[blockIdx.y];
  const int n = p.n, ld = p.ld, i = j - j0;
  if (j >= n - 1) return;
  const int nrows = n - 1 - j;
  if ((int)blockIdx.x * 4 >= nrows) return;
  double* V = vec + p.v_off;
  const double* a = V + (long long)kVy * ld;
  double* Xp = X + p.x_off;
  double* Yp = Xp + (long long)2 * kEigNb * ld;
  double beta, tau, scale;
  trd_reflector(a, part2 + p.p_off2, Xp, ld, n, j, i, beta, tau, scale, y);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (blockIdx.x == 0) {
    // publish column j: d, e, tau and v (Vh row j, panel buffers)
    if (tid == 0) {
      V[(long long)kVd * ld + j] = a[j];
      V[(long long)kVe * ld + j] = beta;
      V[(long long)kVtau * ld + j] = tau;
    }
    double* vr = Vh + p.m_off + (long long)j * ld;
    for (int r = tid; r < n; r += 256) {
      const double x = r <= j ? 0.0 : (r == j + 1 ? 1.0 : a[r] * scale);
      vr[r] = x;
      Xp[(long long)i * ld + r] = x;
      Yp[(long long)(kEigNb + i) * ld + r] = x;
    }
  }
  // w = tau (A22 v - W y1 - V y2) for this wave's row: the row product with
  // four independent partial sums per lane, the 2i panel corrections one per
  // lane, both reduced across the wave
  const int r = j + 1 + blockIdx.x * 4 + wv;
  double wvdot = 0.0;
  if (r < n) {
    const double* Ar = A + p.m_off + (long long)r * ld;
    // eight independent partial sums per lane: eight loads of the row in
    // flight per lane (the product is bound by memory latency, not issue)
    double sp[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sp[q] = 0.0;
    int c = j + 2 + lane;
#pragma unroll 2
    for (; c + 448 < n; c += 512)
#pragma unroll
      for (int q = 0; q < 8; ++q) sp[q] = fma(Ar[c + 64 * q], a[c + 64 * q], sp[q]);
    for (; c < n; c += 64) sp[0] = fma(Ar[c], a[c], sp[0]);
    double s = scale * (((sp[0] + sp[1]) + (sp[2] + sp[3])) + ((sp[4] + sp[5]) + (sp[6] + sp[7])));
    if (lane == 0) s += Ar[j + 1];   // v[j+1] = 1
    // corrections: lane t < i takes W(r, t) y1[t], lane i + t takes V(r, t) y2[t]
    if (lane < 2 * i) {
      const int t = lane < i ? lane : lane - i;
      const double xv = lane < i ? Xp[(long long)(kEigNb + t) * ld + r] : Xp[(long long)t * ld + r];
      s -= xv * y[lane];
    }
    s = tau * wave_sum(s);
    if (lane == 0) {
      V[(long long)kVw * ld + r] = s;
      const double vr = r == j + 1 ? 1.0 : a[r] * scale;
      wvdot = s * vr;
    }
  }
  wvdot = sum256(wvdot, sh);
  if (tid == 0) V[(long long)kVrc * ld + blockIdx.x] = wvdot;
}

// finish >= 0: w of column finish gets alpha v (alpha = -tau (w . v) / 2 from
// k_trd_col's partials) and goes to the panel buffers.  next >= 0: the
// corrected row a = A(next, next:n) minus the panel's first next - j0 vector
// pairs, and this workgroup's partial sums for next's reflector; at
// next = n - 1 only the last diagonal entry.  One thread per row.
__global__ void __launch_bounds__(256) k_trd_row(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                 const double* __restrict__ A, double* X, double* vec, double* part2,
                                                 int finish, int next, int j0) {
  __shared__ double sh[4];
  __shared__ double cw[2 * kEigNb + 2];
  __shared__ double ra[256];   // this workgroup's rows of the corrected row a (0 outside the tail)
  __shared__ double rw[256];   // and of the just finished w (t = i)
  if (!act[blockIdx.y]) return;
  const EigProb p = probs[blockIdx.y];
  const int n = p.n, ld = p.ld, tid = threadIdx.x;
  const int first = next >= 0 ? next : finish + 1;   // rows [first, n)
  if (first >= n || (int)blockIdx.x * 256 >= n - first) return;
  double* V = vec + p.v_off;
  double* Xp = X + p.x_off;
  double* Yp = Xp + (long long)2 * kEigNb * ld;
  const int r0 = first + blockIdx.x * 256;
  const int r = r0 + tid;
  double alpha = 0.0, wi = 0.0;
  int i = -1;
  if (finish >= 0 && finish < n - 1) {
    const int j = finish;
    i = j - j0;
    const int nw = (n - 1 - j + 3) / 4;   // k_trd_col workgroups
    const double* pw = V + (long long)kVrc * ld;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int w = tid;
    for (; w + 768 < nw; w += 1024) {
      s0 += pw[w];
      s1 += pw[w + 256];
      s2 += pw[w + 512];
      s3 += pw[w + 768];
    }
    for (; w < nw; w += 256) s0 += pw[w];
    const double s = sum256((s0 + s1) + (s2 + s3), sh);
    alpha = -0.5 * V[(long long)kVtau * ld + j] * s;
    if (r < n && r > j) {
      wi = fma(alpha, Xp[(long long)i * ld + r], V[(long long)kVw * ld + r]);
      Xp[(long long)(kEigNb + i) * ld + r] = wi;
      Yp[(long long)i * ld + r] = wi;
    }
  }
  if (next < 0 || next >= n) return;
  const int jn = next, ni = jn - j0;   // pairs 0..ni-1 correct the row (ni = i + 1 after a finish)
  // all kEigNb pairs take part with fixed trip counts, so that every load is
  // issued before the sums need it: the panel rows t >= ni are still zero
  // (cleared at the panel start) and so are their coefficients here
  if (tid < kEigNb) {
    const bool on = tid < ni;
    cw[tid] = on ? Xp[(long long)tid * ld + jn] : 0.0;   // v_t[jn]
    // w_t[jn]; the pair just finished is completed here, not yet visible in X
    cw[kEigNb + tid] = !on ? 0.0
                       : tid == i ? fma(alpha, Xp[(long long)tid * ld + jn], V[(long long)kVw * ld + jn])
                                  : Xp[(long long)(kEigNb + tid) * ld + jn];
  }
  __syncthreads();
  double ar = 0.0;
  if (r < n) {
    ar = A[p.m_off + (long long)jn * ld + r];
    double vt[kEigNb], wt[kEigNb];
#pragma unroll
    for (int t = 0; t < kEigNb; ++t) {
      vt[t] = Xp[(long long)t * ld + r];
      wt[t] = t == i ? wi : Xp[(long long)(kEigNb + t) * ld + r];
    }
#pragma unroll
    for (int t = 0; t < kEigNb; ++t) ar -= cw[t] * wt[t] + cw[kEigNb + t] * vt[t];
    V[(long long)kVy * ld + r] = ar;
    if (jn == n - 1) V[(long long)kVd * ld + jn] = ar;
  }
  // partial sums over rows jn + 2..n-1: |a|^2, v_t . a, w_t . a (products
  // k = 0, 1 + t, 1 + kEigNb + t; zero for t >= ni); wave wv takes the
  // products k = wv (mod 4), each lane four of the workgroup's 256 rows, all
  // loads of a wave issued before its sums
  const bool tail = r < n && r >= jn + 2;
  ra[tid] = tail ? ar : 0.0;
  rw[tid] = wi;
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  const int rows = min(256, n - r0);
  double* pr = part2 + p.p_off2 + (long long)blockIdx.x * kTrdPart;
  constexpr int kPerWave = (kTrdPart + 3) / 4;
  double au[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) au[q] = ra[lane + 64 * q];
  double acc[kPerWave];
#pragma unroll
  for (int c = 0; c < kPerWave; ++c) {
    const int k = wv + 4 * c;
    double x = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int u = lane + 64 * q;
      double o = 0.0;
      if (k < kTrdPart && u < rows) {
        if (k == 0) o = au[q];
        else if (k < 1 + kEigNb) o = Xp[(long long)(k - 1) * ld + r0 + u];
        else o = (k - 1 - kEigNb) == i ? rw[u] : Xp[(long long)(k - 1) * ld + r0 + u];
      }
      x = fma(o, au[q], x);
    }
    acc[c] = x;
  }
#pragma unroll
  for (int c = 0; c < kPerWave; ++c) {
    const int k = wv + 4 * c;
    const double x = wave_sum(acc[c]);
    if (lane == 0 && k < kTrdPart) pr[k] = x;
  }
}

// the tridiagonal form as it leaves the tridiagonalisation (kept beside the
// working copy that the divide and conquer splits)
__global__ void k_trd_save(const EigProb* __restrict__ probs, const int* __restrict__ act, double* vec) {
  if (!act[blockIdx.y]) return;
  const EigProb p = probs[blockIdx.y];
  double* V = vec + p.v_off;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < p.n; r += gridDim.x * blockDim.x) {
    V[kVd0 * p.ld + r] = V[kVd * p.ld + r];
    V[kVe0 * p.ld + r] = r + 1 < p.n ? V[kVe * p.ld + r] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// 2. divide and conquer

// Cuppen's splitting: at every merge the coupling beta = |e| is taken off the
// two diagonal entries next to the split (T = diag(T1, T2) + beta u u^T)
__global__ void k_dc_split(const EigProb* __restrict__ probs, const int* __restrict__ act,
                           const DcNode* __restrict__ nodes, int nn, double* vec) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const DcNode nd = nodes[k];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  double* d = vec + p.v_off + (long long)kVd * p.ld;
  const double* e = vec + p.v_off + (long long)kVe * p.ld;
  const int m = nd.start + nd.n1;
  const double b = fabs(e[m - 1]);
  d[m - 1] -= b;
  d[m] -= b;
}

// leaves: the dense tridiagonal block in LDS, cyclic Jacobi to convergence,
// eigenvalues ascending into lam, eigenvectors as rows of Z (zero outside
// the leaf's columns)
__global__ void __launch_bounds__(256) k_dc_leaf(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                 const int4* __restrict__ leaves, double* vec, double* Z) {
  __shared__ double a[64 * kLd];
  __shared__ double q[64 * kLd];
  __shared__ double cs[96];
  __shared__ double red[4];
  __shared__ int pq[64];
  __shared__ int flags[2];
  __shared__ int rank[64];
  const int4 lf = leaves[blockIdx.x];   // (prob, start, size, -)
  if (!act[lf.x]) return;
  const EigProb p = probs[lf.x];
  const int s0 = lf.y, m = lf.z, ld = p.ld;
  const double* d = vec + p.v_off + (long long)kVd * ld;
  const double* e = vec + p.v_off + (long long)kVe * ld;
  double* lam = vec + p.v_off + (long long)kVlam * ld;
  double scl = 0.0;
  for (int u = threadIdx.x; u < m; u += blockDim.x)
    scl = fmax(scl, fabs(d[s0 + u]) + (u + 1 < m ? fabs(e[s0 + u]) : 0.0));
  scl = block_max<256>(scl, red);
  for (int x = threadIdx.x; x < 64 * 64; x += blockDim.x) {
    const int u = x >> 6, v = x & 63;
    double val = 0.0;
    if (u < m && v < m) {
      if (u == v) val = d[s0 + u];
      else if (v == u + 1) val = e[s0 + u];
      else if (u == v + 1) val = e[s0 + v];
    }
    a[u * kLd + v] = val;
    q[u * kLd + v] = (u == v) ? 1.0 : 0.0;
  }
  __syncthreads();
  jacobi64(a, q, 1e-15 * fmax(scl, 1e-300), 40, cs, pq, flags, red);
  __syncthreads();
  // ascending order of the leaf's eigenvalues (ties by index)
  for (int u = threadIdx.x; u < m; u += blockDim.x) {
    const double x = a[u * kLd + u];
    int r = 0;
    for (int v = 0; v < m; ++v) {
      const double y = a[v * kLd + v];
      r += (y < x || (y == x && v < u)) ? 1 : 0;
    }
    rank[u] = r;
  }
  __syncthreads();
  for (int x = threadIdx.x; x < m * m; x += blockDim.x) {
    const int u = x / m, r = x % m;   // eigenvector u (column u of q), component r
    Z[p.m_off + (long long)(s0 + rank[u]) * ld + s0 + r] = q[r * kLd + u];
  }
  for (int u = threadIdx.x; u < m; u += blockDim.x) lam[s0 + rank[u]] = a[u * kLd + u];
}

// merge, part 1 (one workgroup of 256 per node): z from the children's
// boundary components, the merged ascending order, normalisation, deflation
// (small z components; close pairs rotated so that one z component vanishes,
// the rotations applied to the eigenvector rows), and the list of the k
// non-deflated entries for the secular equation.
__global__ void __launch_bounds__(256) k_dc_deflate(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                    const DcNode* __restrict__ nodes, double* vec, int* ivec,
                                                    double* Z) {
  __shared__ double red[4];
  __shared__ double bz_s;
  const DcNode nd = nodes[blockIdx.x];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  const int ld = p.ld, s0 = nd.start, n1 = nd.n1, m = nd.n1 + nd.n2;
  double* V = vec + p.v_off;
  int* I = ivec + p.i_off;
  const double* lam = V + kVlam * ld;
  const double* e = V + kVe * ld;
  double* dn = V + kVdn * ld;     // sorted (then deflation-modified) poles, node-local index
  double* z = V + kVz * ld;       // sorted normalised z
  double* rc = V + kVrc * ld;     // rotation cosines / sines by step
  double* rs = V + kVrs * ld;
  int* perm = I + kIperm * ld;    // sorted index -> child row (node-local)
  int* keep = I + kIkeep * ld;    // non-deflated sorted indices, ascending pole
  int* rot = I + kIrot * ld;      // rotation pairs (prev, cur) by step
  int* rot2 = I + kIrot2 * ld;
  int* cnt = I + kIcnt * ld;      // [s0]: k, [s0 + 1]: rotations
  double* zbuf = V + kVzz * ld;
  Z += p.m_off;
  const double ecoup = e[s0 + n1 - 1];
  const double beta = fabs(ecoup), sg = ecoup >= 0.0 ? 1.0 : -1.0;
  // z in child order: the last component of the left child's vectors, the
  // first component (times the sign of the coupling) of the right child's
  for (int u = threadIdx.x; u < m; u += blockDim.x)
    zbuf[s0 + u] = u < n1 ? Z[(long long)(s0 + u) * ld + s0 + n1 - 1] : sg * Z[(long long)(s0 + u) * ld + s0 + n1];
  __syncthreads();
  // merged ascending order of the two sorted children (ties: left first)
  for (int u = threadIdx.x; u < m; u += blockDim.x) {
    const double x = lam[s0 + u];
    int lo, hi, base;
    if (u < n1) { lo = n1; hi = m; base = u; } else { lo = 0; hi = n1; base = u - n1; }
    // number of entries of the other child before x
    int a0 = lo, a1 = hi;
    while (a0 < a1) {
      const int mid = (a0 + a1) >> 1;
      const double y = lam[s0 + mid];
      const bool before = u < n1 ? (y < x) : (y <= x);
      if (before) a0 = mid + 1; else a1 = mid;
    }
    const int r = base + (a0 - lo);
    perm[s0 + r] = u;
  }
  __syncthreads();
  double zn = 0.0, dmax = 0.0;
  for (int r = threadIdx.x; r < m; r += blockDim.x) {
    const int u = perm[s0 + r];
    dn[s0 + r] = lam[s0 + u];
    const double zz = zbuf[s0 + u];
    z[s0 + r] = zz;
    zn = fma(zz, zz, zn);
    dmax = fmax(dmax, fabs(lam[s0 + u]));
  }
  zn = block_sum<256>(zn, red);
  dmax = block_max<256>(dmax, red);
  const double rn = 1.0 / sqrt(zn);
  for (int r = threadIdx.x; r < m; r += blockDim.x) z[s0 + r] *= rn;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double bz = beta * zn;
    bz_s = bz;
    const double tol = 8.0 * 2.220446049250313e-16 * fmax(dmax, bz);
    int k = 0, nr = 0, prev = -1;
    for (int r = 0; r < m; ++r) {
      if (bz * fabs(z[s0 + r]) <= tol) {   // negligible coupling: d is an eigenvalue
        z[s0 + r] = 0.0;
        continue;
      }
      if (prev < 0) {
        prev = r;
        continue;
      }
      double c = z[s0 + r], s = z[s0 + prev];
      const double h = sqrt(c * c + s * s);
      c /= h;
      s = -s / h;
      const double t = dn[s0 + r] - dn[s0 + prev];
      if (fabs(t * c * s) <= tol) {
        // close pair: rotate so that z[prev] vanishes; prev is deflated
        const double dp = dn[s0 + prev], dr = dn[s0 + r];
        dn[s0 + prev] = dp * c * c + dr * s * s;
        dn[s0 + r] = dp * s * s + dr * c * c;
        z[s0 + r] = h;
        z[s0 + prev] = 0.0;
        rot[s0 + nr] = prev;
        rot2[s0 + nr] = r;
        rc[s0 + nr] = c;
        rs[s0 + nr] = s;
        ++nr;
      } else {
        keep[s0 + k++] = prev;
      }
      prev = r;
    }
    if (prev >= 0) keep[s0 + k++] = prev;
    cnt[s0] = k;
    cnt[s0 + 1] = nr;
  }
  __syncthreads();
  // the rotations, in order, on the eigenvector rows (node columns only)
  const int nr = cnt[s0 + 1];
  for (int t = 0; t < nr; ++t) {
    const long long ra = s0 + perm[s0 + rot[s0 + t]], rb = s0 + perm[s0 + rot2[s0 + t]];
    const double c = rc[s0 + t], s = rs[s0 + t];
    for (int col = threadIdx.x; col < m; col += blockDim.x) {
      const double xa = Z[ra * ld + s0 + col], xb = Z[rb * ld + s0 + col];
      Z[ra * ld + s0 + col] = c * xa + s * xb;
      Z[rb * ld + s0 + col] = -s * xa + c * xb;
    }
    __syncthreads();
  }
  // poles and z of the secular equation (ascending poles), beta |z|^2 in vec
  const int k = cnt[s0];
  double* dd = V + kVdd * ld;
  double* zz = V + kVzz * ld;
  for (int t = threadIdx.x; t < k; t += blockDim.x) {
    dd[s0 + t] = dn[s0 + keep[s0 + t]];
    zz[s0 + t] = z[s0 + keep[s0 + t]];
  }
  if (threadIdx.x == 0) V[kVy * ld + s0] = bz_s;
}

// secular equation 1 + bz sum_t zz_t^2 / (dd_t - lambda) = 0: root t lies in
// (dd_t, dd_{t+1}) (the last in (dd_{k-1}, dd_{k-1} + bz)); it is kept as
// its pole of origin (the nearer end of the interval) and the offset tau from
// it, so that every difference dd_i - lambda is formed without cancellation.
// Safeguarded Newton steps inside a shrinking bracket.
__global__ void __launch_bounds__(256) k_dc_secular(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                    const DcNode* __restrict__ nodes, double* vec, int* ivec) {
  const DcNode nd = nodes[blockIdx.y];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  const int ld = p.ld, s0 = nd.start;
  double* V = vec + p.v_off;
  int* I = ivec + p.i_off;
  const int k = I[kIcnt * ld + s0];
  // one wave per root, the lanes share the sums over the poles
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (t >= k) return;
  const double* dd = V + kVdd * ld + s0;
  const double* zz = V + kVzz * ld + s0;
  const double bz = V[kVy * ld + s0];
  int org;
  double a, b;
  if (t < k - 1) {
    const double gap = dd[t + 1] - dd[t];
    const double mid = 0.5 * gap;
    double f = 0.0;
    for (int i = lane; i < k; i += 64) f += bz * zz[i] * zz[i] / ((dd[i] - dd[t]) - mid);
    f = 1.0 + wave_sum(f);
    if (f >= 0.0) { org = t; a = 0.0; b = mid; }
    else { org = t + 1; a = -mid; b = 0.0; }
  } else {
    org = k - 1;
    a = 0.0;
    double s = 0.0;
    for (int i = lane; i < k; i += 64) s += zz[i] * zz[i];
    b = bz * wave_sum(s);
  }
  const double dorg = dd[org];
  double x = 0.5 * (a + b);
  for (int it = 0; it < 200; ++it) {
    double f = 0.0, fp = 0.0;
    for (int i = lane; i < k; i += 64) {
      const double del = (dd[i] - dorg) - x;
      const double q = zz[i] / del;
      f = fma(bz * zz[i], q, f);
      fp = fma(bz * q, q, fp);
    }
    f = 1.0 + wave_sum(f);
    fp = wave_sum(fp);
    if (f == 0.0) { a = b = x; break; }
    if (f > 0.0) b = x; else a = x;
    // Newton step (f is increasing in lambda), kept inside the bracket
    double xn = x - f / fp;
    if (!(xn > a && xn < b)) xn = 0.5 * (a + b);
    if (xn == x || b - a <= 2.220446049250313e-16 * fmax(fabs(a), fabs(b))) {
      x = xn;
      break;
    }
    x = xn;
  }
  if (lane == 0) {
    I[kIorg * ld + s0 + t] = org;
    V[kVtr * ld + s0 + t] = x;
  }
}

// recomputed z (Gu and Eisenstat): zh_i^2 = prod_j (lambda_j - d_i) /
// (bz prod_{j != i} (d_j - d_i)), with the sign of z_i; one thread per i
__global__ void k_dc_zhat(const EigProb* __restrict__ probs, const int* __restrict__ act,
                          const DcNode* __restrict__ nodes, double* vec, const int* __restrict__ ivec) {
  const DcNode nd = nodes[blockIdx.y];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  const int ld = p.ld, s0 = nd.start;
  double* V = vec + p.v_off;
  const int* I = ivec + p.i_off;
  const int k = I[kIcnt * ld + s0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const double* dd = V + kVdd * ld + s0;
  const double* zz = V + kVzz * ld + s0;
  const double* tr = V + kVtr * ld + s0;
  const int* org = I + kIorg * ld + s0;
  const double bz = V[kVy * ld + s0];
  const double di = dd[i];
  // lambda_j - d_i = (d_org(j) - d_i) + tau_j
  double val = ((dd[org[k - 1]] - di) + tr[k - 1]) / bz;
  for (int j = 0; j < k - 1; ++j) {
    const double lj = (dd[org[j]] - di) + tr[j];
    val *= j < i ? lj / (dd[j] - di) : lj / (dd[j + 1] - di);
  }
  V[kVzh * ld + s0 + i] = copysign(sqrt(fmax(val, 0.0)), zz[i]);
}

// merge, part 2: every eigenvalue of the node (roots and deflated poles) gets
// its ascending position (one thread per entry, grid over the node); the
// product's row gather (child rows of the non-deflated entries) and row
// scatter (output rows of the roots)
__global__ void __launch_bounds__(256) k_dc_place(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                  const DcNode* __restrict__ nodes, double* vec, int* ivec) {
  const DcNode nd = nodes[blockIdx.y];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  const int ld = p.ld, s0 = nd.start, m = nd.n1 + nd.n2;
  double* V = vec + p.v_off;
  int* I = ivec + p.i_off;
  const int k = I[kIcnt * ld + s0];
  const double* dn = V + kVdn * ld + s0;
  const double* z = V + kVz * ld + s0;
  const double* dd = V + kVdd * ld + s0;
  const double* tr = V + kVtr * ld + s0;
  const int* org = I + kIorg * ld + s0;
  const int* perm = I + kIperm * ld + s0;
  const int* keep = I + kIkeep * ld + s0;
  int* pos = I + kIpos * ld + s0;
  double* lam2 = V + kVlam2 * ld;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  // eigenvalue of sorted index u: a root for kept entries (z != 0), the pole otherwise;
  // the root of a kept entry: keep[t] == u, found by its position in keep
  auto ev = [&](int u) -> double {
    if (z[u] == 0.0) return dn[u];
    int lo = 0, hi = k - 1;   // keep is ascending
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (keep[mid] < u) lo = mid + 1; else hi = mid;
    }
    return dd[org[lo]] + tr[lo];
  };
  const double x = ev(r);
  int c = 0;
  for (int u = 0; u < m; ++u) {
    const double y = ev(u);
    c += (y < x || (y == x && u < r)) ? 1 : 0;
  }
  pos[r] = c;
  lam2[s0 + c] = x;
  (void)perm;
}

// the product's row lists (gather: child row of kept entry t; scatter: its
// output row) and the deflated rows copied whole to their positions (one
// workgroup per node row)
__global__ void __launch_bounds__(256) k_dc_rows(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                 const DcNode* __restrict__ nodes, const double* __restrict__ vec,
                                                 int* ivec, const double* __restrict__ Zin, double* Zout) {
  const DcNode nd = nodes[blockIdx.y];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  const int ld = p.ld, s0 = nd.start, m = nd.n1 + nd.n2;
  const int r = blockIdx.x;
  if (r >= m) return;
  const double* V = vec + p.v_off;
  int* I = ivec + p.i_off;
  const int k = I[kIcnt * ld + s0];
  const double* z = V + kVz * ld + s0;
  const int* perm = I + kIperm * ld + s0;
  const int* keep = I + kIkeep * ld + s0;
  const int* pos = I + kIpos * ld + s0;
  if (r < k && threadIdx.x == 0) {
    I[kIrow * ld + s0 + r] = s0 + perm[keep[r]];
    I[kIrot * ld + s0 + r] = s0 + pos[keep[r]];   // output row of root r (rotation list no longer needed)
  }
  if (z[r] != 0.0) return;
  const long long src = s0 + perm[r], dst = s0 + pos[r];
  const double* Zi = Zin + p.m_off;
  double* Zo = Zout + p.m_off;
  for (int col = threadIdx.x; col < m; col += blockDim.x) Zo[dst * ld + s0 + col] = Zi[src * ld + s0 + col];
}

// the eigenvectors of D + bz z z^T in the recomputed z: row t of U (t < k)
// holds u_t(i) = zh_i / (d_i - lambda_t), normalised; one wave per row
__global__ void __launch_bounds__(256) k_dc_vectors(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                    const DcNode* __restrict__ nodes, double* vec,
                                                    const int* __restrict__ ivec, double* U) {
  const DcNode nd = nodes[blockIdx.y];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  const int ld = p.ld, s0 = nd.start;
  const double* V = vec + p.v_off;
  const int* I = ivec + p.i_off;
  const int k = I[kIcnt * ld + s0];
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (t >= k) return;
  const double* dd = V + kVdd * ld + s0;
  const double* zh = V + kVzh * ld + s0;
  const int og = I[kIorg * ld + s0 + t];
  const double tt = V[kVtr * ld + s0 + t];
  const double dorg = dd[og];
  double* Ur = U + p.m_off + (long long)(s0 + t) * ld;
  double s = 0.0;
  for (int i = lane; i < k; i += 64) {
    const double x = zh[i] / ((dd[i] - dorg) - tt);
    Ur[i] = x;
    s = fma(x, x, s);
  }
  s = wave_sum(s);
  const double rn = 1.0 / sqrt(s);
  for (int i = lane; i < k; i += 64) Ur[i] *= rn;
}

// the node's eigenvalues in ascending order back into lam
__global__ void k_dc_lam(const EigProb* __restrict__ probs, const int* __restrict__ act,
                         const DcNode* __restrict__ nodes, double* vec) {
  const DcNode nd = nodes[blockIdx.y];
  if (!act[nd.prob]) return;
  const EigProb p = probs[nd.prob];
  const int m = nd.n1 + nd.n2;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double* V = vec + p.v_off;
  V[kVlam * p.ld + nd.start + r] = V[kVlam2 * p.ld + nd.start + r];
}

// ---------------------------------------------------------------------------
// 3. back-transformation

// T factor of each panel's compact WY form (I - V T V^T = H_j0 ... H_j0+nb-1,
// forward, columnwise): T_ii = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) (V^T v_i)
// compact WY factor of the reflectors [j0, j0 + jb) of one back-transformation
// block (j0 = blk kBtNb, jb <= kBtNb): T upper triangular with T(i, i) = tau_i
// and T(a, i) = -tau_i sum_{a <= c < i} T(a, c) G(c, i) (the recurrence of
// LAPACK's dlarft, forward and columnwise), G = V V^T the block's Gram matrix
// (a product launch before this one).  Thread a owns row a of T, so each
// thread reads back only its own writes; column i of G goes through LDS.  T
// is written once per K chunk of the first product (the chunk products are
// summed by the product with this wide T).
__global__ void __launch_bounds__(256) k_trd_tblock(const EigProb* __restrict__ probs, const int* __restrict__ act,
                                                    const double* __restrict__ G, const double* __restrict__ vec,
                                                    double* Tf) {
  __shared__ double g[kBtNb];
  if (!act[blockIdx.y]) return;
  const int blk = blockIdx.x;
  const EigProb p = probs[blockIdx.y];
  const int n = p.n, ld = p.ld;
  const int j0 = blk * kBtNb;
  if (j0 >= n - 1) return;
  const int jb = min(kBtNb, n - 1 - j0);
  const int tw = p.nchunk * kBtNb;   // row stride of the repeated T
  double* To = Tf + p.t_off + (long long)blk * kBtNb * tw;
  const double* Gp = G + p.g_off + (long long)blk * kBtNb * kBtNb;
  const double* tau = vec + p.v_off + (long long)kVtau * ld + j0;
  const int tid = threadIdx.x;
  if (tid < kBtNb)
    for (int b = 0; b < kBtNb; ++b) To[(long long)tid * tw + b] = 0.0;
  for (int i = 0; i < jb; ++i) {
    if (tid < i) g[tid] = Gp[tid * kBtNb + i];
    __syncthreads();
    if (tid < i) {
      const double* Ta = To + (long long)tid * tw;
      double s0 = 0.0, s1 = 0.0;
      int c = tid;
      for (; c + 1 < i; c += 2) {
        s0 = fma(Ta[c], g[c], s0);
        s1 = fma(Ta[c + 1], g[c + 1], s1);
      }
      if (c < i) s0 = fma(Ta[c], g[c], s0);
      To[(long long)tid * tw + i] = -tau[i] * (s0 + s1);
    } else if (tid == i) {
      To[(long long)i * tw + i] = tau[i];
    }
    __syncthreads();
  }
  if (tid < kBtNb)
    for (int c = 1; c < p.nchunk; ++c)
      for (int b = 0; b < kBtNb; ++b) To[(long long)tid * tw + c * kBtNb + b] = To[(long long)tid * tw + b];
}

// the blocked path's problems are done for the block-Jacobi lanes
__global__ void k_eig_claim(const EigProb* __restrict__ probs, int np, int* jstate) {
  for (int k = threadIdx.x; k < np; k += blockDim.x) jstate[probs[k].mol * kJacState] = 1;
}

// the problems' matrices into the solver's layout (problems of molecules
// that no longer iterate are switched off)
__global__ void __launch_bounds__(256) k_eig_in(const EigProb* __restrict__ probs, const JacobiProb* __restrict__ jp,
                                                const MolState* __restrict__ mstate, int* act,
                                                const double* __restrict__ Asrc, double* Aw) {
  const EigProb p = probs[blockIdx.y];
  const JacobiProb q = jp[p.mol];
  const bool on = !mstate || mstate[p.mol].active;
  if (blockIdx.x == 0 && threadIdx.x == 0) act[blockIdx.y] = on ? 1 : 0;
  if (!on) return;
  const int n = p.n;
  for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < (long long)n * n;
       x += (long long)gridDim.x * blockDim.x) {
    const long long u = x / n, v = x % n;
    Aw[p.m_off + u * p.ld + v] = Asrc[q.a_off + u * q.ld + v];
  }
}

// eigenvalues onto the diagonal of the caller's A, eigenvector rows into its
// W (the Jacobi problem's layout); a non-finite eigenvalue fails the problem
__global__ void __launch_bounds__(256) k_eig_out(const EigProb* __restrict__ probs, const JacobiProb* __restrict__ jp,
                                                 const int* __restrict__ act, const double* __restrict__ vec,
                                                 const double* __restrict__ Z0, const double* __restrict__ Z1,
                                                 const int* __restrict__ zfin, double* A, double* W, int* jstate) {
  if (!act[blockIdx.y]) return;
  const EigProb p = probs[blockIdx.y];
  const JacobiProb q = jp[p.mol];
  const int n = p.n;
  const double* lam = vec + p.v_off + (long long)kVlam * p.ld;
  const double* Zf = (zfin[blockIdx.y] ? Z1 : Z0) + p.m_off;
  for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < (long long)n * n;
       x += (long long)gridDim.x * blockDim.x) {
    const long long u = x / n, v = x % n;
    W[q.w_off + u * q.ld + v] = Zf[u * p.ld + v];
    A[q.a_off + u * q.ld + v] = u == v ? lam[u] : 0.0;
    if (u == v && !isfinite(lam[u])) jstate[p.mol * kJacState + 1] = 1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) jstate[p.mol * kJacState] = 1;
}

}  // namespace xtbamd
