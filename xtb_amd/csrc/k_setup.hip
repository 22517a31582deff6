// Setup-phase kernels of one GFN2 single point (run once per molecule):
// coordination numbers, self energies, overlap/dipole/quadrupole integrals and
// H0, Coulomb matrix, D4 weights and dispersion matrix, EEQ system, AES kernels.
#include "device_util.h"
#include "kernels.h"

namespace xtbamd {

__device__ __forceinline__ double exp_count(double k, double r, double r0) { return 1.0 / (1.0 + exp(-k * (r0 / r - 1.0))); }
__device__ __forceinline__ double erf_count(double k, double r, double r0) { return 0.5 * (1.0 + erf(-k * (r - r0) / r0)); }

// Coordination numbers of one atom (block) against all others (threads):
//  gfn  : ncoordLatP + gfnCount (src/disp/coordinationnumber.F90:351-479,564-599), 40 bohr
//  cov  : ncoordLatP + erfCount with EN scaling (cnType%cov), 40 bohr
//  d4   : ncoord_d4 (src/disp/ncoord.f90:575-617), 1600 bohr^2
//  eeq  : ncoord_erf (src/disp/ncoord.f90:469-506), 1600 bohr^2
__global__ void k_cn(const MolDesc* __restrict__ desc, const Element* __restrict__ el, const int* __restrict__ at,
                     const double* __restrict__ xyz, double* cn_gfn, double* cn_cov, double* cn_d4, double* cn_eeq) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= md.n) return;
  const int* a = at + md.a0;
  const double* x = xyz + 3 * md.a0;
  const Element& ei = el[a[i]];
  const double k4 = 4.10451, k5 = 19.08857, k6 = 2 * 11.28174 * 11.28174;
  double sg = 0, sc = 0, sd = 0, se = 0;
  for (int j = threadIdx.x; j < md.n; j += blockDim.x) {
    if (j == i) continue;
    const Element& ej = el[a[j]];
    const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > 1600.0) continue;
    const double r = sqrt(r2);
    if (r2 >= 1.0e-12) {
      const double rc = ei.rcov_d3 + ej.rcov_d3;
      sg += exp_count(10.0, r, rc) * exp_count(20.0, r, rc + 2);
      const double d = fabs(ei.en - ej.en) + k5;
      sc += k4 * exp(-(d * d) / k6) * erf_count(7.5, r, rc);
    }
    const double rco = ej.rcov_nc + ei.rcov_nc;
    const double ec = erf_count(7.5, r, rco);
    const double dn = fabs(ei.en_nc - ej.en_nc) + k5;
    sd += (k4 * exp(-(dn * dn) / k6)) * ec;
    se += ec;
  }
  sg = block_sum<64>(sg, sh);
  sc = block_sum<64>(sc, sh);
  sd = block_sum<64>(sd, sh);
  se = block_sum<64>(se, sh);
  if (threadIdx.x == 0) {
    cn_gfn[md.a0 + i] = sg;
    cn_cov[md.a0 + i] = sc;
    cn_d4[md.a0 + i] = sd;
    cn_eeq[md.a0 + i] = se;
  }
}

// Self energies and shell data (getSelfEnergy2D: src/xtb/hamiltonian.F90:98-141;
// setzshell: src/scc_core.f90:1509-1542; third order: src/xtb/thirdorder.f90:58-96)
__global__ void k_shell_setup(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                              const int* __restrict__ at, const int* __restrict__ sh_at, const int* __restrict__ sh_m,
                              const double* __restrict__ cn, double* selfe, double* dsedcn, double* shgam, double* zsh) {
  const MolDesc md = desc[blockIdx.y];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= md.nsh) return;
  const int ia = sh_at[md.s0 + s], m = sh_m[md.s0 + s];
  const Element& e = el[at[md.a0 + ia]];
  selfe[md.s0 + s] = e.selfe[m] - e.kcn[m] * cn[md.a0 + ia];
  dsedcn[md.s0 + s] = -e.kcn[m];
  shgam[md.s0 + s] = e.third[m];
  double ntot = -1.0e-6;
  double z = 0.0;
  for (int k = 0; k <= m; ++k) {
    z = e.refocc[k];
    ntot += z;
    if (ntot > e.zval) z = 0.0;
  }
  zsh[md.s0 + s] = z;
}

// ---------------------------------------------------------------------------
// Overlap, dipole and quadrupole integrals of a shell pair, moments about the
// origin, and the H0 element (build_SDQH0: src/xtb/hamiltonian.F90:146-371 with
// get_multiints src/intgrad.f90:879-955, shellPoly/h0scal src/scc_core.f90:644-754)
template <int LI, int LJ>
__device__ void pair_block(const Element& ei, int is, const Element& ej, int js, const double* ri, const double* rj,
                           double intcut, double hav_sp, int aoi, int aoj, int N, long long NN, double* S, double* H0,
                           double* H0n, double* D, double* Q) {
  const double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
  const double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
  constexpr int NSI = 2 * LI + 1, NSJ = 2 * LJ + 1, NCJ = LJ == 0 ? 1 : (LJ == 1 ? 3 : 6), NCI = LI == 0 ? 1 : (LI == 1 ? 3 : 6);
  const double sqrtpi3 = 5.5683279968317078453;  // pi^(3/2)
  for (int ii = 0; ii < NSI; ++ii) {
    double acc[NCJ][10];
#pragma unroll
    for (int c = 0; c < NCJ; ++c)
#pragma unroll
      for (int k = 0; k < 10; ++k) acc[c][k] = 0.0;
    if (r2 <= 2000.0) {
      for (int ip = 0; ip < ei.nprim[is]; ++ip) {
        const double ai = ei.alp[is][ip];
        for (int jp = 0; jp < ej.nprim[js]; ++jp) {
          const double aj = ej.alp[js][jp];
          double est = ai * aj * r2;
          if (est > intcut * (ai + aj)) continue;
          const double g = ai + aj, ab = 1.0 / g;
          est = est * ab;
          const double kab = exp(-est) * (sqrtpi3 * ab * sqrt(ab));
          const double cc = kab * ei.cc[is][ip] * ej.cc[js][jp];
          double v[3][LI + 1][LJ + 1][3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double rp = (ai * ri[k] + aj * rj[k]) * ab;
            oned<LI, LJ>(rp - ri[k], rp - rj[k], rp, g, v[k]);
          }
#pragma unroll
          for (int ci = 0; ci < NCI; ++ci) {
            const double wi = sph_w(LI, ci, ii);
            if (wi == 0.0) continue;
            const int ax = cart_pow(LI, ci, 0), ay = cart_pow(LI, ci, 1), az = cart_pow(LI, ci, 2);
            const double f0 = cc * wi * cnorm(LI, ci);
#pragma unroll
            for (int cj = 0; cj < NCJ; ++cj) {
              const int bx = cart_pow(LJ, cj, 0), by = cart_pow(LJ, cj, 1), bz = cart_pow(LJ, cj, 2);
              const double f = f0 * cnorm(LJ, cj);
              const double* vx = v[0][ax][bx];
              const double* vy = v[1][ay][by];
              const double* vz = v[2][az][bz];
              acc[cj][0] += f * vx[0] * vy[0] * vz[0];
              acc[cj][1] += f * vx[1] * vy[0] * vz[0];
              acc[cj][2] += f * vx[0] * vy[1] * vz[0];
              acc[cj][3] += f * vx[0] * vy[0] * vz[1];
              acc[cj][4] += f * vx[2] * vy[0] * vz[0];
              acc[cj][5] += f * vx[0] * vy[2] * vz[0];
              acc[cj][6] += f * vx[0] * vy[0] * vz[2];
              acc[cj][7] += f * vx[1] * vy[1] * vz[0];
              acc[cj][8] += f * vx[1] * vy[0] * vz[1];
              acc[cj][9] += f * vx[0] * vy[1] * vz[1];
            }
          }
        }
      }
    }
    for (int jj = 0; jj < NSJ; ++jj) {
      double o[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        double s = 0.0;
#pragma unroll
        for (int cj = 0; cj < NCJ; ++cj) s += sph_w(LJ, cj, jj) * acc[cj][k];
        o[k] = s;
      }
      // row aoi + ii (shell i, the atom later in the order), column aoj + jj:
      // the lower triangle only; k_sym_lower mirrors it afterwards with
      // coalesced tile transposes (a thread's mirrored stores would hit a
      // different cache line each, stride N)
      const long long p1 = (long long)(aoi + ii) * N + (aoj + jj);
      if (H0) {
        H0[p1] = hav_sp * o[0];
        H0n[p1] = hav_sp;
        S[p1] = o[0];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) D[k * NN + p1] = o[1 + k];
#pragma unroll
      for (int k = 0; k < 6; ++k) Q[k * NN + p1] = o[4 + k];
    }
  }
}

// LMAX 1: s and p shells only (no d-shell code, far fewer registers)
template <int LMAX = 2>
__device__ void pair_dispatch(int li, int lj, const Element& ei, int is, const Element& ej, int js, const double* ri,
                              const double* rj, double intcut, double hav_sp, int aoi, int aoj, int N, long long NN,
                              double* S, double* H0, double* H0n, double* D, double* Q) {
#define PB(a, b) pair_block<a, b>(ei, is, ej, js, ri, rj, intcut, hav_sp, aoi, aoj, N, NN, S, H0, H0n, D, Q)
  if (LMAX < 2) {
    switch (li * 2 + lj) {
      case 0: PB(0, 0); break;
      case 1: PB(0, 1); break;
      case 2: PB(1, 0); break;
      default: PB(1, 1); break;
    }
    return;
  }
  switch (li * 3 + lj) {
    case 0: PB(0, 0); break;
    case 1: PB(0, 1); break;
    case 2: PB(0, 2); break;
    case 3: PB(1, 0); break;
    case 4: PB(1, 1); break;
    case 5: PB(1, 2); break;
    case 6: PB(2, 0); break;
    case 7: PB(2, 1); break;
    default: PB(2, 2); break;
  }
#undef PB
}

// one thread per atom pair (i > j), pairs taken from the molecule's list
// grouped by element pair (one code path per wave); LMAX 1 takes the list's
// leading s/p-only pairs, LMAX 2 the rest
template <int LMAX>
__global__ void __launch_bounds__(64) k_sdqh0(const MolDesc* __restrict__ desc, const Element* __restrict__ el, const int* __restrict__ at,
                        const double* __restrict__ xyz, const int* __restrict__ at_sh0, const int* __restrict__ at_ao0,
                        const double* __restrict__ selfe, double intcut, const int2* __restrict__ pairs, double* S,
                        double* H0, double* H0n, double* D, double* Q) {
  const MolDesc md = desc[blockIdx.y];
  const long long npair = (long long)md.n * (md.n - 1) / 2;
  const long long t = (LMAX < 2 ? 0 : md.npsp) + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (LMAX < 2 ? md.npsp : npair)) return;
  const int2 pr = pairs[md.p0 + t];
  const int i = pr.x, j = pr.y;
  const int* a = at + md.a0;
  const double* x = xyz + 3 * md.a0;
  const Element& ei = el[a[i]];
  const Element& ej = el[a[j]];
  const int N = md.nao;
  const long long NN = (long long)N * N;
  double* Sm = S + md.m0;
  double* H0m = H0 + md.m0;
  double* Hnm = H0n + md.m0;
  double* Dm = D + 3 * md.m0;
  double* Qm = Q + 6 * md.m0;
  const double* ri = x + 3 * i;
  const double* rj = x + 3 * j;
  const double rab = sqrt((ri[0] - rj[0]) * (ri[0] - rj[0]) + (ri[1] - rj[1]) * (ri[1] - rj[1]) + (ri[2] - rj[2]) * (ri[2] - rj[2]));
  const double den = (ei.en - ej.en) * (ei.en - ej.en);
  const double enpoly = 1.0 + 0.02 * den * (1.0 + 0.0 * den);
  const double rr = rab / (ej.rad + ei.rad);
  int aoi = at_ao0[md.a0 + i];
  for (int is = 0; is < ei.nsh; ++is) {
    const int li = ei.l[is];
    int aoj = at_ao0[md.a0 + j];
    for (int js = 0; js < ej.nsh; ++js) {
      const int lj = ej.l[js];
      const double hii = selfe[md.s0 + at_sh0[md.a0 + i] + is];
      const double hjj = selfe[md.s0 + at_sh0[md.a0 + j] + js];
      const double zi = ei.zeta[is], zj = ej.zeta[js];
      const double zetaij = sqrt(2 * sqrt(zi * zj) / (zi + zj));
      // kScale: 0.5*(kshell_i + kshell_j), s-d and p-d pairs 2.0 (src/xtb/gfn2.f90:823-833)
      const double ks[3] = {1.85, 2.23, 2.23};
      double kscale = 0.5 * (ks[li] + ks[lj]);
      if ((li == 0 && lj == 2) || (li == 2 && lj == 0) || (li == 1 && lj == 2) || (li == 2 && lj == 1)) kscale = 2.0;
      const double km = kscale * enpoly * 1.0;
      const double hav = 0.5 * km * (hii + hjj) * zetaij;
      const double rf1 = 1.0 + 0.01 * ei.shpoly[li] * sqrt(rr);
      const double rf2 = 1.0 + 0.01 * ej.shpoly[lj] * sqrt(rr);
      const double shpoly = rf1 * rf2;
      pair_dispatch<LMAX>(li, lj, ei, is, ej, js, ri, rj, intcut, hav * shpoly, aoi, aoj, N, NN, Sm, H0m, Hnm, Dm,
                          Qm);
      aoj += 2 * lj + 1;
    }
    aoi += 2 * li + 1;
  }
}

// on-site blocks: S = 1, H0 = self energy, dipole/quadrupole moments of the
// atom's own shell pairs about the origin (build_SDQH0 :298-369); atoms
// without d shells take the s/p-only code (no d-shell temporaries)
__global__ void __launch_bounds__(64) k_sdq_onsite(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                             const int* __restrict__ at, const double* __restrict__ xyz, const int* __restrict__ at_sh0,
                             const int* __restrict__ at_ao0, const double* __restrict__ selfe, double intcut, double* S,
                             double* H0, double* H0n, double* D, double* Q) {
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.n) return;
  const Element& e = el[at[md.a0 + i]];
  const double* ri = xyz + 3 * (md.a0 + i);
  const int N = md.nao;
  const long long NN = (long long)N * N;
  double* Sm = S + md.m0;
  double* H0m = H0 + md.m0;
  double* Hnm = H0n + md.m0;
  double* Dm = D + 3 * md.m0;
  double* Qm = Q + 6 * md.m0;
  const int ao0 = at_ao0[md.a0 + i];
  // zero the whole on-site square first (S and H0 stay zero off the diagonal)
  int nao_at = 0;
  for (int s = 0; s < e.nsh; ++s) nao_at += 2 * e.l[s] + 1;
  for (int u = 0; u < nao_at; ++u)
    for (int v = 0; v < nao_at; ++v) {
      const long long p = (long long)(ao0 + u) * N + ao0 + v;
      Sm[p] = (u == v) ? 1.0 : 0.0;
      H0m[p] = 0.0;
      Hnm[p] = 0.0;
    }
  bool has_d = false;
  for (int s = 0; s < e.nsh; ++s) has_d = has_d || e.l[s] > 1;
  int aoi = ao0;
  for (int is = 0; is < e.nsh; ++is) {
    const double se = selfe[md.s0 + at_sh0[md.a0 + i] + is];
    for (int k = 0; k < 2 * e.l[is] + 1; ++k) {
      const long long p = (long long)(aoi + k) * N + aoi + k;
      H0m[p] = se;
      Hnm[p] = se;
    }
    int aoj = ao0;
    for (int js = 0; js < e.nsh; ++js) {
      if (js <= is) {
        if (has_d)
          pair_dispatch<2>(e.l[is], e.l[js], e, is, e, js, ri, ri, intcut, 0.0, aoi, aoj, N, NN, nullptr, nullptr,
                           nullptr, Dm, Qm);
        else
          pair_dispatch<1>(e.l[is], e.l[js], e, is, e, js, ri, ri, intcut, 0.0, aoi, aoj, N, NN, nullptr, nullptr,
                           nullptr, Dm, Qm);
      }
      aoj += 2 * e.l[js] + 1;
    }
    aoi += 2 * e.l[is] + 1;
  }
}

// The 12 integral matrices (S, H0, H0n, D x3, Q x6) of each molecule from
// their lower triangle: workgroup (tile pair, molecule, matrix) reads the
// 32 x 32 tile (I, J), I >= J, row by row and writes its transpose to tile
// (J, I) row by row through LDS (in a diagonal tile only the entries above
// the diagonal); matrix k0 + blockIdx.z of the 12.  Launched after k_sdqh0 and
// k_sdq_onsite.
__global__ void __launch_bounds__(256) k_sym_lower(const MolDesc* __restrict__ desc, double* S, double* H0,
                                                   double* H0n, double* D, double* Q, int k0) {
  __shared__ double t[32][33];
  const MolDesc md = desc[blockIdx.y];
  const int N = md.nao, nt = (N + 31) / 32;
  // tile pair index -> (I, J), I >= J, row-major over the lower triangle
  const int pidx = blockIdx.x;
  if (pidx >= nt * (nt + 1) / 2) return;
  int I = (int)((sqrt(8.0 * pidx + 1.0) - 1.0) * 0.5);
  while (I * (I + 1) / 2 > pidx) --I;
  while ((I + 1) * (I + 2) / 2 <= pidx) ++I;
  const int J = pidx - I * (I + 1) / 2;
  const long long NN = (long long)N * N;
  const int k = k0 + (int)blockIdx.z;
  double* M = k == 0 ? S + md.m0 : k == 1 ? H0 + md.m0 : k == 2 ? H0n + md.m0
            : k < 6 ? D + 3 * md.m0 + (k - 3) * NN : Q + 6 * md.m0 + (k - 6) * NN;
  const int c = threadIdx.x & 31, r8 = threadIdx.x >> 5;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = r8 + 8 * q, gr = 32 * I + r, gc = 32 * J + c;
    t[r][c] = (gr < N && gc < N) ? M[(long long)gr * N + gc] : 0.0;
  }
  __syncthreads();
  // element (row 32 J + rr, column 32 I + cc) of the upper tile takes the
  // lower element (32 I + cc, 32 J + rr)
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int rr = r8 + 8 * q, cc = c;
    const int gr = 32 * J + rr, gc = 32 * I + cc;
    if (gr < N && gc < N && (I > J || cc > rr)) M[(long long)gr * N + gc] = t[cc][rr];
  }
}

// Klopman-Ohno Coulomb matrix with arithmetic-mean hardness and exponent 2
// (getCoulombMatrixCluster: src/coulomb/klopmanohno.F90:257-312)
__global__ void k_coulomb(const MolDesc* __restrict__ desc, const Element* __restrict__ el, const int* __restrict__ at,
                          const double* __restrict__ xyz, const int* __restrict__ sh_at, const int* __restrict__ sh_m,
                          double* J) {
  const MolDesc md = desc[blockIdx.y];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ns = md.nsh;
  if (t >= ns * ns) return;
  const int s1 = (int)(t / ns), s2 = (int)(t % ns);
  const int ia = sh_at[md.s0 + s1], ib = sh_at[md.s0 + s2];
  const Element& ea = el[at[md.a0 + ia]];
  const Element& eb = el[at[md.a0 + ib]];
  const double ha = ea.hard[sh_m[md.s0 + s1]], hb = eb.hard[sh_m[md.s0 + s2]];
  double v;
  if (ia == ib) {
    v = (s1 == s2) ? ha : 0.5 * (ha + hb);
  } else {
    const double* xa = xyz + 3 * (md.a0 + ia);
    const double* xb = xyz + 3 * (md.a0 + ib);
    const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
    const double r1 = sqrt(dx * dx + dy * dy + dz * dz);
    const double gij = 0.5 * (ha + hb);
    v = 1.0 / sqrt(r1 * r1 + 1.0 / (gij * gij));
  }
  J[md.j0 + t] = v;
}

// weighted dispersion matrix (build_wdispmat: src/disp/dftd4.F90:660-732)
__global__ void k_dispmat(const MolDesc* __restrict__ desc, const Element* __restrict__ el, const int* __restrict__ at,
                          const double* __restrict__ xyz, const int* __restrict__ d4_at, const int* __restrict__ d4_ref,
                          const double* __restrict__ gw, const double* __restrict__ c6ref, double* dispmat) {
  const MolDesc md = desc[blockIdx.y];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nd = md.nd4;
  if (t >= nd * nd) return;
  const int k = (int)(t / nd), l = (int)(t % nd);
  const int ia = d4_at[md.d0 + k], ja = d4_at[md.d0 + l];
  double v = 0.0;
  if (ia != ja) {
    const int za = at[md.a0 + ia], zb = at[md.a0 + ja];
    const Element& ea = el[za];
    const Element& eb = el[zb];
    const double* xa = xyz + 3 * (md.a0 + ia);
    const double* xb = xyz + 3 * (md.a0 + ja);
    const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double gwgw = gw[md.d0 + k] * gw[md.d0 + l];
    if (r2 <= 72.0 * 72.0 && gwgw >= 1.0e-7) {
      const double r4r2ij = 3.0 * ea.sqrtzr4r2 * eb.sqrtzr4r2;
      const double cut = kD4a1 * sqrt(r4r2ij) + kD4a2;
      const double oor6 = 1.0 / (r2 * r2 * r2 + ipowd(cut, 6));
      const double oor8 = 1.0 / (r2 * r2 * r2 * r2 + ipowd(cut, 8));
      const double oor10 = 1.0 / (ipowd(r2, 5) + ipowd(cut, 10));
      const double c6 = c6ref[((long long)za * (XTB_MAX_Z + 1) + zb) * 49 + d4_ref[md.d0 + k] * 7 + d4_ref[md.d0 + l]];
      const double c8 = r4r2ij * c6;
      const double c10 = 49.0 / 40.0 * r4r2ij * r4r2ij * c6;
      v = gw[md.d0 + k] * gw[md.d0 + l] * (-kD4s6 * (c6 * oor6) - kD4s8 * (c8 * oor8) - kD4s10 * (c10 * oor10));
    }
  }
  dispmat[md.x0 + t] = v;
}

// EEQ system with the charge constraint (eeq_chrgeq_core / get_coulomb_matrix_0d:
// src/eeq_model.f90:60-80,678-860)
__global__ void k_eeq_matrix(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                             const int* __restrict__ at, const double* __restrict__ xyz, const double* __restrict__ cn_eeq,
                             double* E, double* xE) {
  const MolDesc md = desc[blockIdx.y];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ne = md.ne;
  if (t >= ne * ne) return;
  const int i = (int)(t / ne), j = (int)(t % ne);
  const int n = md.n;
  double v;
  if (i == n || j == n) {
    v = (i == n && j == n) ? 0.0 : 1.0;
  } else {
    const Element& ei = el[at[md.a0 + i]];
    if (i == j) {
      v = ei.eeq_gam + 0.79788456080286535588 / ei.eeq_alpha;  // sqrt(2/pi)
    } else {
      const Element& ej = el[at[md.a0 + j]];
      const double* xa = xyz + 3 * (md.a0 + i);
      const double* xb = xyz + 3 * (md.a0 + j);
      const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
      const double r1 = sqrt(dx * dx + dy * dy + dz * dz);
      const double gamij = 1.0 / sqrt(ei.eeq_alpha * ei.eeq_alpha + ej.eeq_alpha * ej.eeq_alpha);
      v = erf(gamij * r1) / r1;
    }
  }
  E[md.e0 + t] = v;
  if (j == 0) {
    if (i < n) {
      const Element& ei = el[at[md.a0 + i]];
      const double c = cn_eeq[md.a0 + i];
      const double tmp = ei.eeq_kappa / (sqrt(c) + 1e-14);
      xE[md.y0 + i] = -ei.eeq_en + tmp * c;
    } else {
      xE[md.y0 + n] = md.charge;
    }
  }
}

// CN-dependent AES radii and damped kernels (get_radcn, mmomgabzero:
// src/aespot.F90:1219-1281)
__global__ void k_aes_setup(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                            const int* __restrict__ at, const double* __restrict__ xyz, const double* __restrict__ cn,
                            double* radcn, double* gab3, double* gab5) {
  const MolDesc md = desc[blockIdx.y];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = md.n;
  if (t >= n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  const Element& ei = el[at[md.a0 + i]];
  const Element& ej = el[at[md.a0 + j]];
  const double ri = ei.mrad + (5.0 - ei.mrad) / (1.0 + exp(-4.0 * (cn[md.a0 + i] - ei.valcn - 1.2)));
  const double rj = ej.mrad + (5.0 - ej.mrad) / (1.0 + exp(-4.0 * (cn[md.a0 + j] - ej.valcn - 1.2)));
  if (j == 0) radcn[md.a0 + i] = ri;
  double g3 = 0.0, g5 = 0.0;
  if (i != j) {
    const double* xa = xyz + 3 * (md.a0 + i);
    const double* xb = xyz + 3 * (md.a0 + j);
    const double dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
    const double rinv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    const double rco = 0.5 * (ri + rj);
    const double d3 = 1.0 / (1.0 + 6.0 * pow(rco * rinv, 3.0));
    const double d5 = 1.0 / (1.0 + 6.0 * pow(rco * rinv, 4.0));
    g3 = d3 * pow(rinv, 3.0);
    g5 = d5 * pow(rinv, 5.0);
  }
  gab3[md.p0 + t] = g3;
  gab5[md.p0 + t] = g5;
}

}  // namespace xtbamd
