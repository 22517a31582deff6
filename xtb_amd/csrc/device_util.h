// Device helpers shared by the engine kernels: deterministic reductions,
// pair indexing, Gaussian moment integrals.
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"

namespace xtbamd {

// sampled-launch counters and clocks, spread over kProfLines lines (engine.h)
__device__ __forceinline__ unsigned long long* prof_line(unsigned long long* p) {
  return p + (blockIdx.x % kProfLines) * kProfLineU64;
}
__device__ __forceinline__ void prof_add(unsigned long long* work, unsigned long long v) {
  atomicAdd(prof_line(work), v);
}
// ts[0] (per line): earliest start, ts[1]: latest end on the device wall clock
__device__ __forceinline__ void prof_clock_start(unsigned long long* ts) {
  if (ts && threadIdx.x == 0) atomicMin(prof_line(ts), (unsigned long long)__builtin_amdgcn_s_memrealtime());
}
__device__ __forceinline__ void prof_clock_end(unsigned long long* ts) {
  if (ts && threadIdx.x == 0) atomicMax(prof_line(ts) + 1, (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic block sum for blockDim.x = NT (multiple of 64); sh needs NT/64 slots.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r = fmax(r, sh[i]);
  __syncthreads();
  return r;
}

// D4 charge scaling of a reference with charge qref at the effective charge
// qmod (both including the effective nuclear charge) and its qmod derivative
__device__ __forceinline__ double zeta_d(double a, double c, double qref, double qmod) {
  if (qmod < 0.0) return exp(a);
  return exp(a * (1.0 - exp(c * (1.0 - qref / qmod))));
}
__device__ __forceinline__ double dzeta_d(double a, double c, double qref, double qmod) {
  if (qmod < 0.0) return 0.0;
  return -a * c * exp(c * (1.0 - qref / qmod)) * zeta_d(a, c, qref, qmod) * qref / (qmod * qmod);
}

// index t of the strictly lower triangle -> (i, j), i > j
__device__ __forceinline__ void tri_index(long long t, int& i, int& j) {
  double d = sqrt(8.0 * (double)t + 1.0);
  int ii = (int)((1.0 + d) * 0.5);
  while ((long long)ii * (ii - 1) / 2 > t) --ii;
  while ((long long)(ii + 1) * ii / 2 <= t) ++ii;
  i = ii;
  j = (int)(t - (long long)ii * (ii - 1) / 2);
}

__device__ __forceinline__ double ipowd(double x, int l) {
  double r = 1.0;
  for (int i = 0; i < l; ++i) r *= x;
  return r;
}

// Gaussian moment int t^m exp(-g t^2) dt / sqrt(pi/g)   (olapp: src/intgrad.f90:95-111)
__device__ __forceinline__ double gmoment(int m, double g) {
  if (m & 1) return 0.0;
  const int h = m >> 1;
  const double df[5] = {1.0, 1.0, 3.0, 15.0, 105.0};
  return ipowd(0.5 / g, h) * df[h];
}

// cartesian exponent of component c of a shell with angular momentum l in
// direction k (order s; x,y,z; xx,yy,zz,xy,xz,yz)
__device__ __forceinline__ int cart_pow(int l, int c, int k) {
  if (l == 0) return 0;
  if (l == 1) return c == k ? 1 : 0;
  // d
  const int tab[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
  return tab[c][k];
}

// cartesian->spherical d transformation weight T[c][m] (dtrf2: src/intgrad.f90:144-231)
__device__ __forceinline__ double dsph(int c, int m) {
  const double h3 = 0.86602540378443864676;  // sqrt(3)/2
  switch (m) {
    case 0: return c == 0 ? h3 : (c == 1 ? -h3 : 0.0);
    case 1: return c == 0 ? 0.5 : (c == 1 ? 0.5 : (c == 2 ? -1.0 : 0.0));
    default: return c == m + 1 ? 1.0 : 0.0;
  }
}
__device__ __forceinline__ int ncart(int l) { return l == 0 ? 1 : (l == 1 ? 3 : 6); }
// normalisation of the cartesian d functions relative to xx (set_d_function)
__device__ __forceinline__ double cnorm(int l, int c) { return (l == 2 && c >= 3) ? 1.7320508075688772 : 1.0; }

// weight of cartesian c in spherical AO m of a shell with angular momentum l
__device__ __forceinline__ double sph_w(int l, int c, int m) { return l < 2 ? (c == m ? 1.0 : 0.0) : dsph(c, m); }

// 1D moments of (x-A)^a (x-B)^b times 1, (x-C), (x-C)^2 about the product
// centre; v[a][b][k] for a <= LA, b <= LB.
template <int LA, int LB>
__device__ __forceinline__ void oned(double pa, double pb, double pc, double g, double (&v)[LA + 1][LB + 1][3]) {
  double e[LA + LB + 3];
#pragma unroll
  for (int m = 0; m < LA + LB + 3; ++m) e[m] = gmoment(m, g);
#pragma unroll
  for (int a = 0; a <= LA; ++a)
#pragma unroll
    for (int b = 0; b <= LB; ++b) {
      double ca[LA + 1], cb[LB + 1];
#pragma unroll
      for (int k = 0; k <= a; ++k) {
        const double bin = (k == 0 || k == a) ? 1.0 : (double)a;  // a <= 3: binomials 1, a, a(a-1)/2 ...
        double bb = bin;
        if (a == 3 && (k == 1 || k == 2)) bb = 3.0;
        ca[k] = bb * ipowd(pa, a - k);
      }
#pragma unroll
      for (int k = 0; k <= b; ++k) {
        double bb = (k == 0 || k == b) ? 1.0 : (double)b;
        cb[k] = bb * ipowd(pb, b - k);
      }
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
      for (int i = 0; i <= a; ++i)
#pragma unroll
        for (int j = 0; j <= b; ++j) {
          const double c = ca[i] * cb[j];
          const int t = i + j;
          s0 += c * e[t];
          s1 += c * (e[t + 1] + pc * e[t]);
          s2 += c * (e[t + 2] + 2.0 * pc * e[t + 1] + pc * pc * e[t]);
        }
      v[a][b][0] = s0;
      v[a][b][1] = s1;
      v[a][b][2] = s2;
    }
}

}  // namespace xtbamd
