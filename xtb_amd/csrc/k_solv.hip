// Implicit solvation on the device: GBSA (Still kernel) and ALPB (P16 kernel
// with the shape correction), coupled to GFN2 through the Mulliken atomic
// charges (src/scf_module.F90:299-321).
//
// Set-up (once per geometry): Born radii (src/solv/born.f90:35-378), the
// solvent accessible surface on the 230-point Lebedev grid
// (src/solv/sasa.f90:42-205), the H-bond term (src/solv/gbsa.f90:991-1030),
// the ALPB shape descriptor (src/solv/gbsa.f90:661-705) and the interaction
// matrix B (src/solv/kernel.f90:649-760, src/solv/gbsa.f90:445-491).  SCC:
// atomic potential B q (src/solv/gbsa.f90:495-521) and energy q^T B q / 2 +
// surface + shift (:525-548).  Gradient: src/solv/gbsa.f90:600-777,
// src/solv/kernel.f90:171-376.
//
// Every per-atom quantity is gathered by one thread or one workgroup of that
// atom in a fixed order, so results do not depend on the rest of the batch
// (no floating-point atomics).  The surface derivative is gathered as well:
// the quadrature weight of every grid point of every atom is kept (230 per
// atom) and each atom collects the terms in which it is the moving neighbour.
#include "solv_params.h"

namespace xtbamd {

constexpr int kSolvGrid = 230;
constexpr int kSolvNbCap = 2048;                   // neighbours kept in LDS by k_solv_sasa
constexpr double kSolvW = 0.3 * kAAtoAU;           // smoothing width of the surface function
constexpr double kSolvW3 = kSolvW * (kSolvW * kSolvW);
constexpr double kSolvAh0 = 0.5;
constexpr double kSolvAh1 = 3.0 / (4.0 * kSolvW);
constexpr double kSolvAh3 = -1.0 / (4.0 * kSolvW3);
constexpr double kSolvLrcut = 35.0 * kAAtoAU;      // Born radius neighbour cut-off
constexpr double kSolvSrOff = 2.0 * kAAtoAU;       // surface cut-off offset
constexpr double kKcalToAU = 1.0 / 627.50947428;
constexpr double kZetaP16 = 1.028;
constexpr double kSolvTiny = 2.2250738585072014e-308;

// solvent model resolved on the host from one parameter set
struct SolvDev {
  int kernel;   // 1 Still, 2 P16
  double keps, alpbet, c1, boff, gshift;
  double vdwr[XTB_MAX_Z + 1], rho[XTB_MAX_Z + 1], vdwsa[XTB_MAX_Z + 1], wrp[XTB_MAX_Z + 1];
  double trj2a[XTB_MAX_Z + 1], trj2b[XTB_MAX_Z + 1], gam[XTB_MAX_Z + 1], hbmag[XTB_MAX_Z + 1];
  double grid[kSolvGrid][4];
};

// contribution of a sphere of radius rh at distance r to the Born integral of
// an atom of radius rv (compute_psi): far (no overlap) and near forms, value g
// and radial derivative factor dg
__device__ __forceinline__ void psi_far(double r, double rh, double& g, double& dg) {
  const double r1 = 1.0 / r, ap = r + rh, am = r - rh, ab = ap * am;
  const double rhab = rh / ab, lnab = 0.5 * log(am / ap) * r1;
  g = rhab + lnab;
  dg = -2.0 * rhab / ab + (rhab - lnab) * r1 * r1;
}
__device__ __forceinline__ void psi_near(double r, double rh, double rv, double& g, double& dg) {
  const double r1 = 1.0 / r, r12 = 0.5 * r1, r24 = r12 * r12, ap = r + rh, am = r - rh;
  const double rh1 = 1.0 / rv, rhr1 = 1.0 / ap, aprh1 = ap * rh1, lnab = log(aprh1);
  g = rh1 - rhr1 + r12 * (0.5 * am * (rhr1 - rh1 * aprh1) - lnab);
  dg = rhr1 * rhr1 * (1.0 - 0.25 * am * r1 * (1.0 + aprh1 * aprh1)) + rh * r24 * (rhr1 - rh1 * aprh1) +
       r12 * (r1 * lnab - rhr1);
  dg = dg * r1;
}

// contribution of atom o to the Born integral of atom a (a != o), with the
// case structure of compute_psi for the pair (max(a,o), min(a,o)); a_first:
// a has the larger index; false when the pair adds nothing to psi_a
__device__ __forceinline__ bool psi_pair(const SolvDev& sv, int za, int zo, bool a_first, double r, double& g,
                                         double& dg) {
  const double rho_a = sv.rho[za], rho_o = sv.rho[zo], rv_a = sv.vdwr[za], rv_o = sv.vdwr[zo];
  const bool ov_a = r < rv_a + rho_o, ov_o = r < rv_o + rho_a;
  if (!ov_a) {
    // equal reduced radii: the pair's first-atom term is reused for the second
    if (!a_first && !ov_o && fabs(rho_a - rho_o) < 1.e-8) psi_far(r, rho_a, g, dg);
    else psi_far(r, rho_o, g, dg);
    return true;
  }
  if ((r + rho_o) > rv_a) {
    psi_near(r, rho_o, rv_a, g, dg);
    return true;
  }
  return false;
}

// Born radii and the chain factor d brad / d psi (compute_bornr), thread per atom
__global__ void k_solv_born(const MolDesc* __restrict__ desc, const SolvDev* __restrict__ svp,
                            const int* __restrict__ at, const double* __restrict__ xyz, double* brad, double* bdpsi) {
  const MolDesc md = desc[blockIdx.y];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= md.n) return;
  const SolvDev& sv = *svp;
  const int za = at[md.a0 + a];
  const double* xa = xyz + 3 * (md.a0 + a);
  const double lr2 = kSolvLrcut * kSolvLrcut;
  double psi = 0.0;
  for (int j = 0; j < md.n; ++j) {
    if (j == a) continue;
    const double* xj = xyz + 3 * (md.a0 + j);
    const double dx = xa[0] - xj[0], dy = xa[1] - xj[1], dz = xa[2] - xj[2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 < lr2)) continue;
    double g, dg;
    if (psi_pair(sv, za, at[md.a0 + j], a > j, sqrt(d2), g, dg)) psi = psi + g;
  }
  const double alp = 1.0, bet = 0.8, gam = 4.85;
  const double vdwri = sv.vdwr[za], svdwi = vdwri - sv.boff;
  const double s1 = 1.0 / svdwi, v1 = 1.0 / vdwri, s2 = 0.5 * svdwi;
  double br = psi * s2;
  double arg2 = br * (gam * br - bet);
  const double arg = br * (alp + arg2);
  arg2 = 2.0 * arg2 + alp + gam * br * br;
  const double th = tanh(arg), ch = cosh(arg);
  br = 1.0 / (s1 - v1 * th);
  br = sv.c1 * br;
  double dp = ch * (s1 - v1 * th);
  dp = s2 * v1 * arg2 / (dp * dp);
  dp = sv.c1 * dp;
  brad[md.a0 + a] = br;
  bdpsi[md.a0 + a] = dp;
}

// surface cut-off of a molecule: 2 (w + max probe radius) + offset (initBorn)
__device__ double solv_srcut(const MolDesc& md, const SolvDev& sv, const int* at) {
  double m = 0.0;
  for (int i = 0; i < md.n; ++i) m = fmax(m, sv.vdwsa[at[md.a0 + i]]);
  return 2.0 * (kSolvW + m) + kSolvSrOff;
}

// accessibility of one grid point (compute_w_sp) over the neighbour list nb;
// returns the product, 0 when buried
__device__ __forceinline__ double solv_point(const SolvDev& sv, const int* __restrict__ at, const double* __restrict__ xyz,
                                             int a0, const int* nb, int nno, const double xp[3]) {
  double sp = 1.0;
  for (int m = 0; m < nno; ++m) {
    const int o = nb[m];
    const int zo = at[a0 + o];
    const double* xo = xyz + 3 * (a0 + o);
    const double t0 = xp[0] - xo[0], t1 = xp[1] - xo[1], t2 = xp[2] - xo[2];
    const double tj2 = t0 * t0 + t1 * t1 + t2 * t2;
    if (tj2 < sv.trj2b[zo]) {
      if (tj2 <= sv.trj2a[zo]) return 0.0;
      const double uj = sqrt(tj2) - sv.vdwsa[zo], ah3uj2 = kSolvAh3 * uj * uj;
      double sij = kSolvAh0 + (kSolvAh1 + ah3uj2) * uj;
      sij = fmin(sij, 1.0);
      sij = fmax(0.0, sij);
      if (sij < kSolvTiny) return 0.0;
      sp = sp * sij;
    }
  }
  return sp;
}

// derivative factor of a neighbour's surface function at a grid point:
// d ln s / d t times t (the grds of compute_w_sp), zero outside the band
__device__ __forceinline__ bool solv_dfac(const SolvDev& sv, int zo, const double t[3], double g[3]) {
  const double tj2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
  if (!(tj2 < sv.trj2b[zo]) || tj2 <= sv.trj2a[zo]) return false;
  const double sqtj = sqrt(tj2), uj = sqtj - sv.vdwsa[zo], ah3uj2 = kSolvAh3 * uj * uj;
  double ds = kSolvAh1 + 3.0 * ah3uj2;
  double sij = kSolvAh0 + (kSolvAh1 + ah3uj2) * uj;
  sij = fmin(sij, 1.0);
  sij = fmax(0.0, sij);
  ds = ds / (sij * sqtj);
  g[0] = ds * t[0];
  g[1] = ds * t[1];
  g[2] = ds * t[2];
  return true;
}

// neighbour list of atom a within the surface cut-off, ascending, in LDS;
// returns the count (more than kSolvNbCap: the list is not usable)
__device__ int solv_neighbours(const MolDesc& md, const int* at, const double* xyz, int a, double srcut, int* nb) {
  __shared__ int cnt;
  const int lane = threadIdx.x;
  if (lane == 0) cnt = 0;
  __syncthreads();
  const double sr2 = srcut * srcut, lr2 = kSolvLrcut * kSolvLrcut;
  const double* xa = xyz + 3 * (md.a0 + a);
  for (int base = 0; base < md.n; base += 64) {
    const int j = base + lane;
    bool in = false;
    if (j < md.n && j != a) {
      const double* xj = xyz + 3 * (md.a0 + j);
      const double dx = xa[0] - xj[0], dy = xa[1] - xj[1], dz = xa[2] - xj[2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      in = d2 < lr2 && d2 < sr2;
    }
    const unsigned long long mask = __ballot(in);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    const int c0 = cnt;
    if (in && c0 + before < kSolvNbCap) nb[c0 + before] = j;
    __syncthreads();
    if (lane == 0) cnt = c0 + __popcll(mask);
    __syncthreads();
  }
  return cnt;
}

// surface of every atom: one wave per atom, lanes over the grid points;
// stores the quadrature weight of each point (0 when below 1e-6 accessibility)
__global__ void __launch_bounds__(64) k_solv_sasa(const MolDesc* __restrict__ desc, const SolvDev* __restrict__ svp,
                                                  const int* __restrict__ at, const double* __restrict__ xyz,
                                                  double* wpt, double* sasa, int* failed) {
  __shared__ int nb[kSolvNbCap];
  const MolDesc md = desc[blockIdx.y];
  const int a = blockIdx.x;
  if (a >= md.n) return;
  const SolvDev& sv = *svp;
  const double srcut = solv_srcut(md, sv, at);
  const int nno = solv_neighbours(md, at, xyz, a, srcut, nb);
  if (nno > kSolvNbCap) {
    if (threadIdx.x == 0) failed[md.mol] = 1;
    return;
  }
  const int za = at[md.a0 + a];
  const double* xa = xyz + 3 * (md.a0 + a);
  const double rsas = sv.vdwsa[za], wr = sv.wrp[za];
  double part = 0.0;
  for (int ip = threadIdx.x; ip < kSolvGrid; ip += 64) {
    const double xp[3] = {xa[0] + rsas * sv.grid[ip][0], xa[1] + rsas * sv.grid[ip][1], xa[2] + rsas * sv.grid[ip][2]};
    const double sp = solv_point(sv, at, xyz, md.a0, nb, nno, xp);
    const double w = sp > 1.e-6 ? sv.grid[ip][3] * wr * sp : 0.0;
    wpt[(long long)(md.a0 + a) * kSolvGrid + ip] = w;
    part += w;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
  if (threadIdx.x == 0) sasa[md.a0 + a] = part;
}

// per molecule: surface energy, H-bond weights, ALPB shape descriptor
__global__ void __launch_bounds__(256) k_solv_mol(const MolDesc* __restrict__ desc, const SolvDev* __restrict__ svp,
                                                  const int* __restrict__ at, const double* __restrict__ xyz,
                                                  const double* __restrict__ sasa, double* hbw, double* gsasa,
                                                  double* adet) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.x];
  const SolvDev& sv = *svp;
  double g = 0.0;
  for (int i = threadIdx.x; i < md.n; i += blockDim.x) {
    const int z = at[md.a0 + i];
    const double s = sasa[md.a0 + i];
    g += s * sv.gam[z];
    const double smaxd = 1.0 / (sv.vdwsa[z] * sv.vdwsa[z]);
    hbw[md.a0 + i] = sv.hbmag[z] * (s * smaxd);
  }
  g = block_sum<256>(g, sh);
  if (threadIdx.x == 0) {
    gsasa[md.mol] = g;
    double ad = 0.0;
    if (sv.alpbet > 0.0) {
      const double tof = 2.0 / 5.0;
      double tot = 0.0, c[3] = {0.0, 0.0, 0.0}, I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int i = 0; i < md.n; ++i) {
        const double r = sv.vdwr[at[md.a0 + i]], r3 = r * r * r;
        tot += r3;
        for (int k = 0; k < 3; ++k) c[k] += xyz[3 * (md.a0 + i) + k] * r3;
      }
      for (int k = 0; k < 3; ++k) c[k] /= tot;
      for (int i = 0; i < md.n; ++i) {
        const double r = sv.vdwr[at[md.a0 + i]], r2 = r * r, r3 = r2 * r;
        double v[3], d2 = 0.0;
        for (int k = 0; k < 3; ++k) { v[k] = xyz[3 * (md.a0 + i) + k] - c[k]; d2 += v[k] * v[k]; }
        for (int k = 0; k < 3; ++k) I[k][k] += r3 * (d2 + tof * r2);
        for (int cc = 0; cc < 3; ++cc)
          for (int k = 0; k < 3; ++k) I[k][cc] -= r3 * v[cc] * v[k];
      }
      const double det = I[0][0] * (I[1][1] * I[2][2] - I[1][2] * I[2][1]) -
                         I[0][1] * (I[1][0] * I[2][2] - I[1][2] * I[2][0]) +
                         I[0][2] * (I[1][0] * I[2][1] - I[1][1] * I[2][0]);
      ad = sqrt(pow(det, 1.0 / 3.0) / (tof * tot));
    }
    adet[md.mol] = ad;
  }
}

// interaction matrix B (n x n per molecule, at the n^2 offset p0)
__global__ void __launch_bounds__(256) k_solv_bmat(const MolDesc* __restrict__ desc, const SolvDev* __restrict__ svp,
                                                   const double* __restrict__ xyz, const double* __restrict__ brad,
                                                   const double* __restrict__ hbw, const double* __restrict__ adet,
                                                   double* B) {
  const MolDesc md = desc[blockIdx.y];
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)md.n * md.n) return;
  const SolvDev& sv = *svp;
  const int i = (int)(e / md.n), j = (int)(e % md.n);
  double v;
  if (i == j) {
    v = sv.keps * (1.0 / brad[md.a0 + i]);
    v += 2 * hbw[md.a0 + i];
  } else {
    const double* xi = xyz + 3 * (md.a0 + i);
    const double* xj = xyz + 3 * (md.a0 + j);
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const double* xh = xyz + 3 * (md.a0 + hi);
    const double* xl = xyz + 3 * (md.a0 + lo);
    (void)xi; (void)xj;
    const double dx = xh[0] - xl[0], dy = xh[1] - xl[1], dz = xh[2] - xl[2];
    const double r1 = sqrt(dx * dx + dy * dy + dz * dz);
    const double bh = brad[md.a0 + hi], bl = brad[md.a0 + lo];
    double f;
    if (sv.kernel == 1) {
      const double r2 = r1 * r1, aa = bh * bl, dd = 0.25 * r2 / aa, expd = exp(-dd);
      f = 1.0 / sqrt(r2 + aa * expd);
    } else {
      const double ab = sqrt(bh * bl);
      double arg = ab / (ab + (kZetaP16 / 16.0) * r1);
      arg = arg * arg;
      arg = arg * arg;
      arg = arg * arg;
      arg = arg * arg;
      f = 1.0 / (r1 + ab * arg);
    }
    v = sv.keps * f;
  }
  if (sv.alpbet > 0.0) v += sv.keps * sv.alpbet / adet[md.mol];
  B[md.p0 + e] = v;
}

// atomic potential (B q)_a of the molecules still iterating, one wave per row
__global__ void __launch_bounds__(64) k_solv_pot(const MolDesc* __restrict__ desc, const double* __restrict__ B,
                                                 const double* __restrict__ q, double* pot,
                                                 const MolState* __restrict__ mst) {
  const MolDesc md = desc[blockIdx.y];
  if (mst && !mst[blockIdx.y].active) return;
  const int a = blockIdx.x;
  if (a >= md.n) return;
  const double* row = B + md.p0 + (long long)a * md.n;
  double s = 0.0;
  for (int j = threadIdx.x; j < md.n; j += 64) s += row[j] * q[md.a0 + j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) pot[md.a0 + a] = s;
}

// q^T B q / 2 per molecule (getEnergy without the surface and shift terms)
__global__ void __launch_bounds__(256) k_solv_energy(const MolDesc* __restrict__ desc, const double* __restrict__ B,
                                                     const double* __restrict__ q, double* qbq,
                                                     const MolState* __restrict__ mst) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.x];
  if (mst && !mst[blockIdx.x].active) return;
  double e = 0.0;
  for (int a = threadIdx.x; a < md.n; a += blockDim.x) {
    const double* row = B + md.p0 + (long long)a * md.n;
    double s = 0.0;
    for (int j = 0; j < md.n; ++j) s += row[j] * q[md.a0 + j];
    e += q[md.a0 + a] * (0.5 * s);
  }
  e = block_sum<256>(e, sh);
  if (threadIdx.x == 0) qbq[md.mol] = e;
}

// d E / d brad of every atom (grddb of addGradientStill / dEdbr of addGradientP16)
__global__ void k_solv_grddb(const MolDesc* __restrict__ desc, const SolvDev* __restrict__ svp,
                             const double* __restrict__ xyz, const double* __restrict__ q,
                             const double* __restrict__ brad, double* grddb) {
  const MolDesc md = desc[blockIdx.y];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= md.n) return;
  const SolvDev& sv = *svp;
  const double keps = sv.keps, ba = brad[md.a0 + a], qa = q[md.a0 + a];
  const double* xa = xyz + 3 * (md.a0 + a);
  double g = 0.0;
  for (int j = 0; j < md.n; ++j) {
    if (j == a) continue;
    const double* xj = xyz + 3 * (md.a0 + j);
    const double dx = xa[0] - xj[0], dy = xa[1] - xj[1], dz = xa[2] - xj[2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double bj = brad[md.a0 + j], qq = (a > j) ? qa * q[md.a0 + j] : q[md.a0 + j] * qa;
    if (sv.kernel == 1) {
      const double r2 = r * r, aa = (a > j) ? ba * bj : bj * ba, dd = 0.25 * r2 / aa, expd = exp(-dd);
      const double fgb2 = r2 + aa * expd, dfgb2 = 1.0 / fgb2, dfgb = sqrt(dfgb2), dfgb3 = dfgb2 * dfgb * keps;
      const double bp = -0.5 * expd * (1.0 + dd) * dfgb3;
      g = g + (bj * bp) * qq;
    } else {
      const double ab = (a > j) ? sqrt(ba * bj) : sqrt(bj * ba);
      const double arg1 = ab / (ab + (kZetaP16 / 16.0) * r);
      double a16 = arg1 * arg1;
      a16 = a16 * a16;
      a16 = a16 * a16;
      a16 = a16 * a16;
      const double fgb = r + ab * a16, dfgb = 1.0 / fgb, dfgb2 = dfgb * dfgb;
      const double bp = -0.5 * (r * kZetaP16 / ab * arg1 + 1.0) / ab * a16 * dfgb2;
      g = g + bj * bp * keps * qq;
    }
  }
  const double bp = 1.0 / ba, qq = qa * bp;
  g = g + (-0.5 * keps * qq * bp) * qa;
  grddb[md.a0 + a] = g;
}

// gradient of atom a: kernel pair terms, Born radius chain, surface and
// H-bond terms; one wave per atom, accumulated into grad by lane 0
__global__ void __launch_bounds__(64) k_solv_grad(const MolDesc* __restrict__ desc, const SolvDev* __restrict__ svp,
                                                  const int* __restrict__ at, const double* __restrict__ xyz,
                                                  const double* __restrict__ q, const double* __restrict__ brad,
                                                  const double* __restrict__ bdpsi, const double* __restrict__ grddb,
                                                  const double* __restrict__ wpt, double* grad,
                                                  const int* __restrict__ act) {
  __shared__ int nb[kSolvNbCap];
  const MolDesc md = desc[blockIdx.y];
  if (act && !act[md.mol]) return;
  const int a = blockIdx.x;
  if (a >= md.n) return;
  const SolvDev& sv = *svp;
  const int za = at[md.a0 + a];
  const double* xa = xyz + 3 * (md.a0 + a);
  const double keps = sv.keps, ba = brad[md.a0 + a], qa = q[md.a0 + a];
  const double lr2 = kSolvLrcut * kSolvLrcut;
  const double ga = grddb[md.a0 + a] * bdpsi[md.a0 + a];
  double g[3] = {0.0, 0.0, 0.0};
  // pair terms over the other atoms
  for (int j = threadIdx.x; j < md.n; j += 64) {
    if (j == a) continue;
    const int zj = at[md.a0 + j];
    const double* xj = xyz + 3 * (md.a0 + j);
    const double v[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    const double d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2], r = sqrt(d2);
    const double bj = brad[md.a0 + j], qq = qa * q[md.a0 + j];
    double ap;
    if (sv.kernel == 1) {
      const double r2 = r * r, aa = ba * bj, dd = 0.25 * r2 / aa, expd = exp(-dd);
      const double fgb2 = r2 + aa * expd, dfgb2 = 1.0 / fgb2, dfgb = sqrt(dfgb2), dfgb3 = dfgb2 * dfgb * keps;
      ap = (1.0 - 0.25 * expd) * dfgb3;
    } else {
      const double ab = sqrt(ba * bj);
      const double arg1 = ab / (ab + (kZetaP16 / 16.0) * r);
      double a16 = arg1 * arg1;
      a16 = a16 * a16;
      a16 = a16 * a16;
      a16 = a16 * a16;
      const double fgb = r + ab * a16, dfgb = 1.0 / fgb, dfgb2 = dfgb * dfgb;
      ap = (1.0 - kZetaP16 * arg1 * a16) * dfgb2 * keps / r;
    }
    double f = -ap * qq;
    // Born radius chain: psi_a depends on atom j and psi_j on atom a
    if (d2 < lr2) {
      double gg, dg;
      if (psi_pair(sv, za, zj, a > j, r, gg, dg)) f += ga * dg;
      if (psi_pair(sv, zj, za, j > a, r, gg, dg)) f += grddb[md.a0 + j] * bdpsi[md.a0 + j] * dg;
    }
    g[0] += f * v[0];
    g[1] += f * v[1];
    g[2] += f * v[2];
  }
  // surface and H-bond terms: weights gamma_i + dhbdw_i q_i^2 of the atom that owns the grid point
  const double srcut = solv_srcut(md, sv, at);
  const int nno = solv_neighbours(md, at, xyz, a, srcut, nb);
  if (nno <= kSolvNbCap) {
    const double smaxa = 1.0 / (sv.vdwsa[za] * sv.vdwsa[za]);
    const double ca = sv.gam[za] + sv.hbmag[za] * smaxa * qa * qa;
    const double rsas = sv.vdwsa[za];
    // own grid points: every neighbour in its band moves the point's accessibility
    for (int ip = threadIdx.x; ip < kSolvGrid; ip += 64) {
      const double w = wpt[(long long)(md.a0 + a) * kSolvGrid + ip];
      if (w == 0.0) continue;   // excluded point (the grid has negative weights)
      const double xp[3] = {xa[0] + rsas * sv.grid[ip][0], xa[1] + rsas * sv.grid[ip][1], xa[2] + rsas * sv.grid[ip][2]};
      for (int m = 0; m < nno; ++m) {
        const int o = nb[m];
        const double* xo = xyz + 3 * (md.a0 + o);
        const double t[3] = {xp[0] - xo[0], xp[1] - xo[1], xp[2] - xo[2]};
        double d[3];
        if (solv_dfac(sv, at[md.a0 + o], t, d)) {
          const double cw = ca * w;
          g[0] += cw * d[0];
          g[1] += cw * d[1];
          g[2] += cw * d[2];
        }
      }
    }
    // grid points of the neighbours whose accessibility atom a moves
    for (int m = 0; m < nno; ++m) {
      const int i = nb[m], zi = at[md.a0 + i];
      const double* xi = xyz + 3 * (md.a0 + i);
      const double qi = q[md.a0 + i];
      const double smaxi = 1.0 / (sv.vdwsa[zi] * sv.vdwsa[zi]);
      const double ci = sv.gam[zi] + sv.hbmag[zi] * smaxi * qi * qi;
      const double rsi = sv.vdwsa[zi];
      for (int ip = threadIdx.x; ip < kSolvGrid; ip += 64) {
        const double w = wpt[(long long)(md.a0 + i) * kSolvGrid + ip];
        if (w == 0.0) continue;   // excluded point (the grid has negative weights)
        const double t[3] = {xi[0] + rsi * sv.grid[ip][0] - xa[0], xi[1] + rsi * sv.grid[ip][1] - xa[1],
                             xi[2] + rsi * sv.grid[ip][2] - xa[2]};
        double d[3];
        if (solv_dfac(sv, za, t, d)) {
          const double cw = ci * w;
          g[0] -= cw * d[0];
          g[1] -= cw * d[1];
          g[2] -= cw * d[2];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) g[k] += __shfl_xor(g[k], o, 64);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + a) + k] += g[k];
}

// ALPB shape term of the gradient (addADetDeriv), one workgroup per molecule
__global__ void __launch_bounds__(64) k_solv_alpb_grad(const MolDesc* __restrict__ desc,
                                                       const SolvDev* __restrict__ svp, const int* __restrict__ at,
                                                       const double* __restrict__ xyz, const double* __restrict__ q,
                                                       double* grad, const int* __restrict__ act) {
  __shared__ double A[3][3], c[3];
  const MolDesc md = desc[blockIdx.x];
  if (act && !act[md.mol]) return;
  const SolvDev& sv = *svp;
  if (!(sv.alpbet > 0.0)) return;
  if (threadIdx.x == 0) {
    const double tof = 2.0 / 5.0;
    double tot = 0.0, qt = 0.0, cc[3] = {0.0, 0.0, 0.0}, I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < md.n; ++i) {
      const double r = sv.vdwr[at[md.a0 + i]], r3 = r * r * r;
      tot += r3;
      for (int k = 0; k < 3; ++k) cc[k] += xyz[3 * (md.a0 + i) + k] * r3;
      qt += q[md.a0 + i];
    }
    for (int k = 0; k < 3; ++k) cc[k] /= tot;
    for (int i = 0; i < md.n; ++i) {
      const double r = sv.vdwr[at[md.a0 + i]], r2 = r * r, r3 = r2 * r;
      double v[3], d2 = 0.0;
      for (int k = 0; k < 3; ++k) { v[k] = xyz[3 * (md.a0 + i) + k] - cc[k]; d2 += v[k] * v[k]; }
      for (int k = 0; k < 3; ++k) I[k][k] += r3 * (d2 + tof * r2);
      for (int b = 0; b < 3; ++b)
        for (int k = 0; k < 3; ++k) I[k][b] -= r3 * v[b] * v[k];
    }
    const double det = I[0][0] * (I[1][1] * I[2][2] - I[1][2] * I[2][1]) - I[0][1] * (I[1][0] * I[2][2] - I[1][2] * I[2][0]) +
                       I[0][2] * (I[1][0] * I[2][1] - I[1][1] * I[2][0]);
    const double ad = sqrt(pow(det, 1.0 / 3.0) / (tof * tot));
    const double ke = sv.keps * sv.alpbet;
    const double f = (250.0 / (48.0 * tot * tot * tot * pow(ad, 5))) * (-0.5 * ke * qt * qt / (ad * ad));
    A[0][0] = (I[0][0] * (I[1][1] + I[2][2]) - I[0][1] * I[0][1] - I[0][2] * I[0][2]) * f;
    A[1][0] = A[0][1] = (I[0][1] * I[2][2] - I[0][2] * I[1][2]) * f;
    A[2][0] = A[0][2] = (I[0][2] * I[1][1] - I[0][1] * I[2][1]) * f;
    A[1][1] = (I[1][1] * (I[0][0] + I[2][2]) - I[0][1] * I[0][1] - I[1][2] * I[1][2]) * f;
    A[2][1] = A[1][2] = (I[0][0] * I[1][2] - I[0][1] * I[0][2]) * f;
    A[2][2] = (I[2][2] * (I[0][0] + I[1][1]) - I[0][2] * I[0][2] - I[1][2] * I[1][2]) * f;
    for (int k = 0; k < 3; ++k) c[k] = cc[k];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < md.n; i += blockDim.x) {
    const double r = sv.vdwr[at[md.a0 + i]], r3 = r * r * r;
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = xyz[3 * (md.a0 + i) + k] - c[k];
    for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + i) + k] += r3 * (A[k][0] * v[0] + A[k][1] * v[1] + A[k][2] * v[2]);
  }
}

// ---- host: solvent parameter resolution (src/solv/model.f90) -------------

// normalizeSolventName + loadInternalParam: Still kernel -> GFN2/GBSA sets,
// P16 kernel -> GFN2/ALPB sets; nullptr when the solvent is not parametrised
inline const SolvParamSet* solv_param_set(const std::string& name, int kernel) {
  std::string nm;
  for (char ch : name) nm.push_back((char)std::tolower((unsigned char)ch));
  static const char* alias[][2] = {
      {"methylcyanide", "acetonitrile"}, {"methylenechloride", "ch2cl2"}, {"dichloromethane", "ch2cl2"},
      {"dcm", "ch2cl2"}, {"chloroform", "chcl3"}, {"trichloromethane", "chcl3"}, {"tcm", "chcl3"},
      {"carbondisulfide", "cs2"}, {"diethylether", "ether"}, {"dimethylformamide", "dmf"},
      {"dimethylsulfoxide", "dmso"}, {"furan", "furane"}, {"nhexan", "hexane"}, {"n-hexan", "hexane"},
      {"nhexane", "hexane"}, {"n-hexane", "hexane"}, {"tetrahydrofuran", "thf"}, {"h2o", "water"}};
  for (auto& al : alias)
    if (nm == al[0]) { nm = al[1]; break; }
  if (nm == "dichlormethane") nm = "ch2cl2";
  const SolvParamSet* tab = kernel == 1 ? kSolvGbsaGfn2 : kSolvAlpbGfn2;
  const int cnt = kernel == 1 ? kSolvGbsaGfn2Count : kSolvAlpbGfn2Count;
  if (kernel == 1) {
    if (nm == "water") nm = "h2o";
    if (nm == "hexane") nm = "nhexan";
  }
  for (int i = 0; i < cnt; ++i)
    if (nm == tab[i].name) return &tab[i];
  return nullptr;
}

// paramToModel + initBorn: the device model of one solvent.  state: 1 gsolv
// (no shift), 2 reference (1 M solute in the pure solvent), 3 1 bar / 1 M
// (getStateShift, src/solv/state.f90:55-85)
inline SolvDev solv_model(const SolvParamSet& p, int kernel, bool alpb, int state, double temperature) {
  SolvDev s;
  std::memset(&s, 0, sizeof(s));
  s.kernel = kernel;
  const double eps = p.epsv;
  s.gshift = p.gshift * kKcalToAU;
  const double molar_mass = p.smass, density = p.rhos / p.smass * 8.92388e-2;
  const double ideal = 24.79, ambient = 298.15, ref_density = 1.0e-3, ref_mass = 1.0;
  if (state == 2)
    s.gshift = s.gshift + temperature * kBoltzmann *
                              (std::log(ideal * temperature / ambient) + std::log(density / ref_density * ref_mass / molar_mass));
  else if (state == 3)
    s.gshift = s.gshift + temperature * kBoltzmann * std::log(ideal * temperature / ambient);
  s.c1 = p.c1;
  s.boff = p.soset * 0.1 * kAAtoAU;
  const double probe = p.rprobe * kAAtoAU;
  bool anyhb = false;
  for (int k = 0; k < 94; ++k) if (std::fabs(p.hbond[k]) > 1.e-3) anyhb = true;
  s.alpbet = alpb ? 0.571412 / eps : 0.0;
  s.keps = (1.0 / eps - 1.0) / (1.0 + s.alpbet);
  for (int z = 1; z <= XTB_MAX_Z; ++z) {
    const double vdwr = kAAtoAU * kVdwRadD3Aa[z - 1];
    s.vdwr[z] = vdwr;
    s.rho[z] = vdwr * p.sx[z - 1];
    const double sa = vdwr + probe;
    s.vdwsa[z] = sa;
    s.trj2a[z] = (sa - kSolvW) * (sa - kSolvW);
    s.trj2b[z] = (sa + kSolvW) * (sa + kSolvW);
    double r = sa + kSolvW;
    double w = (0.25 / kSolvW + 3.0 * kSolvAh3 * (0.2 * r * r - 0.5 * r * sa + sa * sa / 3.0)) * r * r * r;
    r = sa - kSolvW;
    w = w - (0.25 / kSolvW + 3.0 * kSolvAh3 * (0.2 * r * r - 0.5 * r * sa + sa * sa / 3.0)) * r * r * r;
    s.wrp[z] = w;
    s.gam[z] = p.gamscale[z - 1] * (4.0 * kPi) * 1.0e-5;
    s.hbmag[z] = anyhb ? -(p.hbond[z - 1] * p.hbond[z - 1]) * kKcalToAU : 0.0;
  }
  for (int i = 0; i < kSolvGrid; ++i)
    for (int k = 0; k < 4; ++k) s.grid[i][k] = kLebedev230[i][k];
  return s;
}

}  // namespace xtbamd
