// Analytic gradient of the GFN2 single point (src/scf_module.F90:665-800).
// Every kernel runs one workgroup per atom and reduces its own partner terms
// in a fixed order, so the gradient is deterministic; pair terms are evaluated
// from both sides instead of being scattered with atomics.
#include "device_util.h"
#include "kernels.h"

namespace xtbamd {

__device__ __forceinline__ double exp_count_g(double k, double r, double r0) { return 1.0 / (1.0 + exp(-k * (r0 / r - 1.0))); }
__device__ __forceinline__ double dexp_count_g(double k, double r, double r0) {
  const double ex = exp(-k * (r0 / r - 1.0));
  return (-k * r0 * ex) / (r * r * ((ex + 1.0) * (ex + 1.0)));
}
__device__ __forceinline__ double erf_count_g(double k, double r, double r0) { return 0.5 * (1.0 + erf(-k * (r - r0) / r0)); }
__device__ __forceinline__ double derf_count_g(double k, double r, double r0) {
  return -k / 1.7724538509055160273 / r0 * exp(-k * k * (r - r0) * (r - r0) / (r0 * r0));
}

struct GradPairIn {
  const double *P, *Pew, *Hn, *S;  // molecule matrices
  int N;
  double ves_i, ves_j, vs_i, vs_j;
  const double *vd_i, *vd_j, *vq_i, *vq_j;
};

// Derivative of S/D/Q of a shell pair w.r.t. the centre of shell i contracted
// with the energy weights (build_dSDQH0_noreset: src/xtb/hamiltonian.F90:575-792,
// get_grad_multiint src/intgrad.f90:958-1035, shiftintg :703-728).  Weights are
// transformed to the cartesian basis once, then contracted per primitive pair.
template <int LI, int LJ>
__device__ void grad_pair(const Element& ei, int is, const Element& ej, int js, const double* ri, const double* rj,
                          double intcut, int aoi, int aoj, const GradPairIn& in, double* g) {
  constexpr int NCI = LI == 0 ? 1 : (LI == 1 ? 3 : 6), NCJ = LJ == 0 ? 1 : (LJ == 1 ? 3 : 6);
  constexpr int NSI = 2 * LI + 1, NSJ = 2 * LJ + 1;
  const double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
  const double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
  if (r2 > 2000.0) return;
  // spherical weights -> cartesian
  double wf[NCI][NCJ], wp[NCI][NCJ];
#pragma unroll
  for (int ci = 0; ci < NCI; ++ci)
#pragma unroll
    for (int cj = 0; cj < NCJ; ++cj) { wf[ci][cj] = 0.0; wp[ci][cj] = 0.0; }
  for (int ii = 0; ii < NSI; ++ii)
    for (int jj = 0; jj < NSJ; ++jj) {
      // lower triangle (row of atom i, the later one): P, Pew, S and H0n are
      // exactly symmetric, and the lanes of a wave (same i, consecutive j)
      // then read neighbouring columns of one row instead of rows N apart
      const long long pos = (long long)(aoi + ii) * in.N + (aoj + jj);
      const double pij = in.P[pos];
      const double hp = in.Hn[pos] * kEVtoAU * pij;  // H0 without overlap, in Eh
      const double fac = 2 * hp - 2 * in.Pew[pos] - pij * (in.ves_i + in.ves_j) + pij * (in.vs_i + in.vs_j);
#pragma unroll
      for (int ci = 0; ci < NCI; ++ci) {
        const double a = sph_w(LI, ci, ii);
        if (a == 0.0) continue;
#pragma unroll
        for (int cj = 0; cj < NCJ; ++cj) {
          const double b = sph_w(LJ, cj, jj);
          if (b == 0.0) continue;
          wf[ci][cj] += a * b * fac;
          wp[ci][cj] += a * b * pij;
        }
      }
    }
  const double sqrtpi3 = 5.5683279968317078453;
  for (int ip = 0; ip < ei.nprim[is]; ++ip) {
    const double ai = ei.alp[is][ip];
    for (int jp = 0; jp < ej.nprim[js]; ++jp) {
      const double aj = ej.alp[js][jp];
      double est = ai * aj * r2;
      if (est > intcut * (ai + aj)) continue;
      const double gg = ai + aj, ab = 1.0 / gg;
      est = est * ab;
      const double kab = exp(-est) * (sqrtpi3 * ab * sqrt(ab));
      const double cc = kab * ei.cc[is][ip] * ej.cc[js][jp];
      double v[3][LI + 2][LJ + 1][3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double rp = (ai * ri[k] + aj * rj[k]) * ab;
        oned<LI + 1, LJ>(rp - ri[k], rp - rj[k], rp - rj[k], gg, v[k]);
      }
#pragma unroll
      for (int ci = 0; ci < NCI; ++ci) {
        const int a0 = cart_pow(LI, ci, 0), a1 = cart_pow(LI, ci, 1), a2 = cart_pow(LI, ci, 2);
        const int ak[3] = {a0, a1, a2};
#pragma unroll
        for (int cj = 0; cj < NCJ; ++cj) {
          const double wfc = wf[ci][cj], wpc = wp[ci][cj];
          if (wfc == 0.0 && wpc == 0.0) continue;
          const int bk[3] = {cart_pow(LJ, cj, 0), cart_pow(LJ, cj, 1), cart_pow(LJ, cj, 2)};
          const double f = cc * cnorm(LI, ci) * cnorm(LJ, cj);
          double val[3][3], gra[3][3];
#pragma unroll
          for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
              val[k][m] = v[k][ak[k]][bk[k]][m];
              double t = 2.0 * ai * v[k][ak[k] + 1][bk[k]][m];
              if (ak[k] > 0) t -= ak[k] * v[k][ak[k] - 1][bk[k]][m];
              gra[k][m] = t;
            }
          const double s0 = f * val[0][0] * val[1][0] * val[2][0];
          const double dj[3] = {f * val[0][1] * val[1][0] * val[2][0], f * val[0][0] * val[1][1] * val[2][0],
                                f * val[0][0] * val[1][0] * val[2][1]};
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            double w[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
              for (int m = 0; m < 3; ++m) w[k][m] = (k == x) ? gra[k][m] : val[k][m];
            const double gs = f * w[0][0] * w[1][0] * w[2][0];
            double gdj[3], gqj[6];
            gdj[0] = f * w[0][1] * w[1][0] * w[2][0];
            gdj[1] = f * w[0][0] * w[1][1] * w[2][0];
            gdj[2] = f * w[0][0] * w[1][0] * w[2][1];
            gqj[0] = f * w[0][2] * w[1][0] * w[2][0];
            gqj[1] = f * w[0][0] * w[1][2] * w[2][0];
            gqj[2] = f * w[0][0] * w[1][0] * w[2][2];
            gqj[3] = f * w[0][1] * w[1][1] * w[2][0];
            gqj[4] = f * w[0][1] * w[1][0] * w[2][1];
            gqj[5] = f * w[0][0] * w[1][1] * w[2][1];
            const double* r = rij;
            double gdi[3], gqi[6];
#pragma unroll
            for (int k = 0; k < 3; ++k) gdi[k] = gdj[k] - r[k] * gs;
            gqi[0] = gqj[0] - 2 * r[0] * gdj[0] + r[0] * r[0] * gs;
            gqi[1] = gqj[1] - 2 * r[1] * gdj[1] + r[1] * r[1] * gs;
            gqi[2] = gqj[2] - 2 * r[2] * gdj[2] + r[2] * r[2] * gs;
            gqi[3] = gqj[3] - r[0] * gdj[1] - r[1] * gdj[0] + r[0] * r[1] * gs;
            gqi[4] = gqj[4] - r[0] * gdj[2] - r[2] * gdj[0] + r[0] * r[2] * gs;
            gqi[5] = gqj[5] - r[1] * gdj[2] - r[2] * gdj[1] + r[1] * r[2] * gs;
            if (x == 0) { gdi[0] -= s0; gqi[0] += -2 * dj[0] + 2 * r[0] * s0; gqi[3] += -dj[1] + r[1] * s0; gqi[4] += -dj[2] + r[2] * s0; }
            if (x == 1) { gdi[1] -= s0; gqi[1] += -2 * dj[1] + 2 * r[1] * s0; gqi[3] += -dj[0] + r[0] * s0; gqi[5] += -dj[2] + r[2] * s0; }
            if (x == 2) { gdi[2] -= s0; gqi[2] += -2 * dj[2] + 2 * r[2] * s0; gqi[4] += -dj[0] + r[0] * s0; gqi[5] += -dj[1] + r[1] * s0; }
            double dt = 0.0, qt = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) dt += gdi[k] * in.vd_i[k] + gdj[k] * in.vd_j[k];
#pragma unroll
            for (int k = 0; k < 6; ++k) qt += gqi[k] * in.vq_i[k] + gqj[k] * in.vq_j[k];
            g[x] += gs * wfc + wpc * (dt + qt);
          }
        }
      }
    }
  }
}

// LMAX 1: s and p shells only (no d-shell code, far fewer registers)
template <int LMAX>
__device__ void grad_dispatch(int li, int lj, const Element& ei, int is, const Element& ej, int js, const double* ri,
                              const double* rj, double intcut, int aoi, int aoj, const GradPairIn& in, double* g) {
#define GP(a, b) grad_pair<a, b>(ei, is, ej, js, ri, rj, intcut, aoi, aoj, in, g)
  if (LMAX < 2) {
    switch (li * 2 + lj) {
      case 0: GP(0, 0); break;
      case 1: GP(0, 1); break;
      case 2: GP(1, 0); break;
      default: GP(1, 1); break;
    }
    return;
  }
  switch (li * 3 + lj) {
    case 0: GP(0, 0); break;
    case 1: GP(0, 1); break;
    case 2: GP(0, 2); break;
    case 3: GP(1, 0); break;
    case 4: GP(1, 1); break;
    case 5: GP(1, 2); break;
    case 6: GP(2, 0); break;
    case 7: GP(2, 1); break;
    default: GP(2, 2); break;
  }
#undef GP
}

// Hamiltonian/overlap/multipole gradient and dE/dCN of the self energies,
// in two passes.  k_grad_h0_pairs evaluates every atom pair i > j within
// range once, with the reference's orientation: its gradient g (the term
// added to i and subtracted from j) and the dE/dCN sums of either side,
// stored at gpair[5 * (p0 + i n + j)].  k_grad_h0 then gives atom a (one
// workgroup, thread = partner b) the sum over its pairs in a fixed order.
template <int LMAX>
__global__ void __launch_bounds__(64) k_grad_h0_pairs(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                      const int* __restrict__ at, const double* __restrict__ xyz,
                                                      const int* __restrict__ at_sh0, const int* __restrict__ at_ao0,
                                                      const double* __restrict__ selfe,
                                                      const double* __restrict__ dsedcn, const double* __restrict__ P,
                                                      const double* __restrict__ Pew, const double* __restrict__ H0n,
                                                      const double* __restrict__ S, const double* __restrict__ shift,
                                                      const double* __restrict__ vs, const double* __restrict__ vd,
                                                      const double* __restrict__ vq, double intcut,
                                                      const int2* __restrict__ pairs, double* gpair) {
  const MolDesc md = desc[blockIdx.y];
  // LMAX 1 takes the pair list's leading s/p-only pairs, LMAX 2 the rest
  const long long t = (LMAX < 2 ? 0 : md.npsp) + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (LMAX < 2 ? (long long)md.npsp : (long long)md.n * (md.n - 1) / 2)) return;
  const int2 pr = pairs[md.p0 + t];   // grouped by element pair: one code path per wave
  const int iat = pr.x, jat = pr.y;
  const int N = md.nao;
  const double* x = xyz + 3 * md.a0;
  const double* ri = x + 3 * iat;
  const double* rj = x + 3 * jat;
  const double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
  const double rij2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
  if (rij2 > 1600.0) return;
  const Element& ei = el[at[md.a0 + iat]];
  const Element& ej = el[at[md.a0 + jat]];
  GradPairIn in;
  in.P = P + md.m0; in.Pew = Pew + md.m0; in.S = S + md.m0; in.Hn = H0n + md.m0; in.N = N;
  in.vs_i = vs[md.a0 + iat]; in.vs_j = vs[md.a0 + jat];
  in.vd_i = vd + 3 * (md.a0 + iat); in.vd_j = vd + 3 * (md.a0 + jat);
  in.vq_i = vq + 6 * (md.a0 + iat); in.vq_j = vq + 6 * (md.a0 + jat);
  const double rab = sqrt(rij2), rr = rab / (ei.rad + ej.rad);
  const double t14 = sqrt(rr), t17 = 1.0 / rij2;
  double gp[3] = {0, 0, 0}, dcn_i = 0.0, dcn_j = 0.0;
  int aoi = at_ao0[md.a0 + iat];
  for (int is = 0; is < ei.nsh; ++is) {
    const int li = ei.l[is];
    const int shi = md.s0 + at_sh0[md.a0 + iat] + is;
    int aoj = at_ao0[md.a0 + jat];
    for (int js = 0; js < ej.nsh; ++js) {
      const int lj = ej.l[js];
      const int shj = md.s0 + at_sh0[md.a0 + jat] + js;
      const double k1 = ei.shpoly[li] * 0.01, k2 = ej.shpoly[lj] * 0.01;
      const double t15 = k1 * t14, t23 = k2 * t14;
      const double shpoly = (1.0 + t15) * (1.0 + k2 * t14);
      const double dfac = (t15 * (1.0 + t23) + (1.0 + t15) * k2 * t14) * 0.5 * t17;
      const double hav = 0.5 * (selfe[shi] + selfe[shj]);
      in.ves_i = shift[shi];
      in.ves_j = shift[shj];
      double gx[3] = {0, 0, 0}, hs = 0.0, dcn = 0.0;
      for (int ii = 0; ii < 2 * li + 1; ++ii)
        for (int jj = 0; jj < 2 * lj + 1; ++jj) {
          const long long pos = md.m0 + (long long)(aoi + ii) * N + (aoj + jj);   // lower triangle
          const double hp = H0n[pos] * kEVtoAU * P[pos];
          hs += 2 * hp * S[pos];
          dcn += hp / hav * S[pos];
        }
      grad_dispatch<LMAX>(li, lj, ei, is, ej, js, ri, rj, intcut, aoi, aoj, in, gx);
      for (int k = 0; k < 3; ++k) gp[k] += gx[k] + hs * dfac * rij[k] / shpoly;
      dcn_i += dcn * dsedcn[shi];
      dcn_j += dcn * dsedcn[shj];
      aoj += 2 * lj + 1;
    }
    aoi += 2 * li + 1;
  }
  double* o = gpair + 5 * (md.p0 + (long long)iat * md.n + jat);
  o[0] = gp[0];
  o[1] = gp[1];
  o[2] = gp[2];
  o[3] = dcn_i;
  o[4] = dcn_j;
}

__global__ void __launch_bounds__(64) k_grad_h0(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                const int* __restrict__ at, const double* __restrict__ xyz,
                                                const int* __restrict__ at_sh0, const int* __restrict__ at_ao0,
                                                const double* __restrict__ dsedcn, const double* __restrict__ P,
                                                const double* __restrict__ gpair, double* grad, double* dhdcn) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  const int a = blockIdx.x;
  if (a >= md.n) return;
  const int N = md.nao;
  const double* x = xyz + 3 * md.a0;
  double ga[3] = {0, 0, 0}, dcn_a = 0.0;
  for (int b = threadIdx.x; b < md.n; b += blockDim.x) {
    if (b == a) continue;
    const int iat = a > b ? a : b, jat = a > b ? b : a;
    const double* ri = x + 3 * iat;
    const double* rj = x + 3 * jat;
    const double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
    const double rij2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    if (rij2 > 1600.0) continue;
    const double* gp = gpair + 5 * (md.p0 + (long long)iat * md.n + jat);
    const double sgn = (a == iat) ? 1.0 : -1.0;
    for (int k = 0; k < 3; ++k) ga[k] += sgn * gp[k];
    dcn_a += (a == iat) ? gp[3] : gp[4];
  }
  for (int k = 0; k < 3; ++k) ga[k] = block_sum<64>(ga[k], sh);
  dcn_a = block_sum<64>(dcn_a, sh);
  if (threadIdx.x == 0) {
    // diagonal self-energy contribution
    const Element& e = el[at[md.a0 + a]];
    int mu = at_ao0[md.a0 + a];
    double dd = 0.0;
    for (int s = 0; s < e.nsh; ++s) {
      const int sh_ = md.s0 + at_sh0[md.a0 + a] + s;
      for (int k = 0; k < 2 * e.l[s] + 1; ++k, ++mu) dd += P[md.m0 + (long long)mu * N + mu] * dsedcn[sh_] * kEVtoAU;
    }
    for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + a) + k] += ga[k];
    dhdcn[md.a0 + a] = dcn_a + dd;
  }
}

// atomic C6 and dC6/dCN for atom pairs (get_atomic_c6: src/disp/dftd4.F90:1069-1147),
// for the charge-scaled (c6, dc6) and the zero-charge (c60, dc60) weights
// PART 1: the zero-charge C6 and its CN derivatives (c60, dc60), PART 2:
// those at the SCC charges (c6, dc6)
template <int PART>
__global__ void __launch_bounds__(256) k_d4_c6(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                               const int* __restrict__ at,
                        const double* __restrict__ c6ref, const double* __restrict__ zetavec,
                        const double* __restrict__ zerovec, const double* __restrict__ zetadcn,
                        const double* __restrict__ zerodcn, double* c6, double* dc6, double* c60, double* dc60) {
  // one atom pair i >= j per thread (grid-stride over the lower triangle):
  // C6 is symmetric, its CN derivatives are formed for both atoms of the pair
  const MolDesc md = desc[blockIdx.y];
  const long long n = md.n, npair = n * (n + 1) / 2;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < npair;
       t += (long long)gridDim.x * blockDim.x) {
    int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long long)i * (i + 1) / 2 > t) --i;
    while ((long long)(i + 1) * (i + 2) / 2 <= t) ++i;
    const int j = (int)(t - (long long)i * (i + 1) / 2);
    const int zi = at[md.a0 + i], zj = at[md.a0 + j];
    const int nri = el[zi].nref, nrj = el[zj].nref;
    const int bi = (md.a0 + i) * XTB_MAX_REF, bj = (md.a0 + j) * XTB_MAX_REF;
    const double* tab = c6ref + ((long long)zi * (XTB_MAX_Z + 1) + zj) * 49;
    // every operand loaded before the sums (fixed trip counts; the table is
    // zero beyond each element's references, and so are the loads here)
    const double* wv = PART == 1 ? zerovec : zetavec;   // weights of this part
    const double* wd = PART == 1 ? zerodcn : zetadcn;
    double zI[XTB_MAX_REF], zdI[XTB_MAX_REF], zJ[XTB_MAX_REF], zdJ[XTB_MAX_REF];
#pragma unroll
    for (int r = 0; r < XTB_MAX_REF; ++r) {
      const bool oni = r < nri, onj = r < nrj;
      zI[r] = oni ? wv[bi + r] : 0.0;
      zdI[r] = oni ? wd[bi + r] : 0.0;
      zJ[r] = onj ? wv[bj + r] : 0.0;
      zdJ[r] = onj ? wd[bj + r] : 0.0;
    }
    double s1 = 0, di = 0, dj = 0;
#pragma unroll
    for (int r = 0; r < XTB_MAX_REF; ++r)
#pragma unroll
      for (int q = 0; q < XTB_MAX_REF; ++q) {
        const double rc = tab[r * 7 + q];
        const double zq = zJ[q] * rc;
        s1 = fma(zI[r], zq, s1);
        di = fma(zdI[r], zq, di);
        dj = fma(zI[r], zdJ[q] * rc, dj);
      }
    const long long ij = md.p0 + (long long)i * n + j, ji = md.p0 + (long long)j * n + i;
    double* cc = PART == 1 ? c60 : c6;
    double* dc = PART == 1 ? dc60 : dc6;
    cc[ij] = s1;
    cc[ji] = s1;
    dc[ij] = di;   // dC6(i,j)/dCN_i
    dc[ji] = dj;   // dC6(j,i)/dCN_j
  }
}

// D4 two-body gradient at fixed charges (disp_gradient_latp: src/disp/dftd4.F90:1832-1933)
__global__ void __launch_bounds__(64) k_d4_grad2(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                 const int* __restrict__ at, const double* __restrict__ xyz,
                                                 const double* __restrict__ c6, const double* __restrict__ dc6,
                                                 double* grad, double* dedcn) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= md.n) return;
  const int n = md.n;
  const double* x = xyz + 3 * md.a0;
  const Element& ei = el[at[md.a0 + i]];
  double g[3] = {0, 0, 0}, de = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    if (j == i) continue;
    const Element& ej = el[at[md.a0 + j]];
    const double r4r2ij = 3 * ei.sqrtzr4r2 * ej.sqrtzr4r2;
    const double r0 = kD4a1 * sqrt(r4r2ij) + kD4a2;
    const double rij[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
    const double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    if (r2 > 3600.0 || r2 < 1.0e-10) continue;
    const double t6 = 1. / (r2 * r2 * r2 + ipowd(r0, 6)), t8 = 1. / (ipowd(r2, 4) + ipowd(r0, 8));
    const double t10 = 1. / (ipowd(r2, 5) + ipowd(r0, 10));
    const double d6 = -6 * r2 * r2 * t6 * t6, d8 = -8 * r2 * r2 * r2 * t8 * t8, d10 = -10 * ipowd(r2, 4) * t10 * t10;
    const double disp = kD4s6 * t6 + kD4s8 * r4r2ij * t8 + kD4s10 * 49.0 / 40.0 * r4r2ij * r4r2ij * t10;
    const double ddisp = kD4s6 * d6 + kD4s8 * r4r2ij * d8 + kD4s10 * 49.0 / 40.0 * r4r2ij * r4r2ij * d10;
    const double cij = c6[md.p0 + (long long)i * n + j];
    for (int k = 0; k < 3; ++k) g[k] += -cij * ddisp * rij[k];
    de -= dc6[md.p0 + (long long)i * n + j] * disp;
  }
  for (int k = 0; k < 3; ++k) g[k] = block_sum<64>(g[k], sh);
  de = block_sum<64>(de, sh);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + i) + k] += g[k];
    dedcn[md.a0 + i] = de;
  }
}

// ATM three-body term with zero-charge C6 (atm_gradient_latp, molecular branch:
// src/disp/dftd4.F90:2121-2174, deriv_atm_triple :2354-2433): the triples of atom a;
// writes atom a's gradient, dE/dCN and energy terms (no accumulation: it runs
// beside the SCC, and the gradient adds them later)
constexpr int kAtmLds = 4096;   // molecules up to this size keep per-atom factors in LDS

__global__ void __launch_bounds__(256) k_d4_atm(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                const int* __restrict__ at, const double* __restrict__ xyz,
                                                const double* __restrict__ c6, const double* __restrict__ dc6,
                                                double* grad, double* dedcn, double* eatm) {
  __shared__ double sh[4];
  __shared__ double sq[kAtmLds];   // sqrt of each atom's sqrt(Z) <r4>/<r2> factor
  const MolDesc md = desc[blockIdx.y];
  const int a = blockIdx.x;
  if (a >= md.n) return;
  const int n = md.n;
  const double* x = xyz + 3 * md.a0;
  const double* C = c6 + md.p0;
  const double* dC = dc6 + md.p0;
  const double cut2 = 1600.0;
  const bool cached = n <= kAtmLds;
  if (cached)
    for (int b = threadIdx.x; b < n; b += blockDim.x) sq[b] = sqrt(el[at[md.a0 + b]].sqrtzr4r2);
  __syncthreads();
  // damping radius of a pair: a1 sqrt(3 s_i s_j) + a2 = a1 sqrt(3) sqrt(s_i) sqrt(s_j) + a2
  auto rad = [&](int i, int j, double si, double sj) {
    return cached ? kD4a1 * 1.7320508075688772 * sq[i] * sq[j] + kD4a2 : kD4a1 * sqrt(3.0 * si * sj) + kD4a2;
  };
  const double ra[3] = {x[3 * a], x[3 * a + 1], x[3 * a + 2]};
  const double sa = el[at[md.a0 + a]].sqrtzr4r2;
  double g[3] = {0, 0, 0}, e = 0.0, dcn = 0.0;
  const long long npair = (long long)n * (n - 1) / 2;
  for (long long t = threadIdx.x; t < npair; t += blockDim.x) {
    int b, c;
    tri_index(t, b, c);  // b > c
    if (b == a || c == a) continue;
    const double rij[3] = {x[3 * b] - ra[0], x[3 * b + 1] - ra[1], x[3 * b + 2] - ra[2]};
    const double r2ij = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    if (r2ij > cut2) continue;
    const double rik[3] = {x[3 * c] - ra[0], x[3 * c + 1] - ra[1], x[3 * c + 2] - ra[2]};
    const double r2ik = rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2];
    if (r2ik > cut2) continue;
    const double rjk[3] = {x[3 * c] - x[3 * b], x[3 * c + 1] - x[3 * b + 1], x[3 * c + 2] - x[3 * b + 2]};
    const double r2jk = rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2];
    if (r2jk > cut2) continue;
    const double sb = el[at[md.a0 + b]].sqrtzr4r2, sc = el[at[md.a0 + c]].sqrtzr4r2;
    const double cij = rad(a, b, sa, sb);
    const double cik = rad(a, c, sa, sc);
    const double cjk = rad(b, c, sb, sc);
    const double c6ij = C[(long long)a * n + b], c6ik = C[(long long)a * n + c], c6jk = C[(long long)b * n + c];
    const double c9 = -sqrt(c6ij * c6ik * c6jk);
    const double ccc1 = cij * cjk * cik;
    const double rrr2 = r2ij * r2jk * r2ik;
    const double rrr1 = sqrt(rrr2);
    const double rrr3 = rrr1 * rrr2;
    const double ang_fact = 0.375 / (rrr2 * rrr3);
    const double ang = ang_fact * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) * (-r2ij + r2jk + r2ik) + 1.0 / rrr3;
    // (ccc1 / rrr1)^(alp / 3) with alp = 16: x^5 cbrt(x)
    static_assert(kD4alp == 16, "ATM damping exponent");
    const double qd = ccc1 / rrr1, qd2 = qd * qd;
    const double cralp = qd2 * qd2 * qd * cbrt(qd);
    const double fdmp = 1.0 / (1.0 + 6.0 * cralp);
    const double dfdmp = -(2.0 * kD4alp * cralp) * fdmp * fdmp;
    const double dE = -fdmp * ang * c9;
    double dang = -ang_fact * (r2ij * r2ij * r2ij + r2ij * r2ij * (r2jk + r2ik) +
                               r2ij * (3.0 * r2jk * r2jk + 2.0 * r2jk * r2ik + 3.0 * r2ik * r2ik) -
                               5.0 * (r2jk - r2ik) * (r2jk - r2ik) * (r2jk + r2ik));
    const double dGr1 = (-dang * c9 * fdmp + dfdmp * c9 * ang) / r2ij;
    dang = -ang_fact * (r2ik * r2ik * r2ik + r2ik * r2ik * (r2jk + r2ij) +
                        r2ik * (3.0 * r2jk * r2jk + 2.0 * r2jk * r2ij + 3.0 * r2ij * r2ij) -
                        5.0 * (r2jk - r2ij) * (r2jk - r2ij) * (r2jk + r2ij));
    const double dGr2 = (-dang * c9 * fdmp + dfdmp * c9 * ang) / r2ik;
    for (int k = 0; k < 3; ++k) g[k] += (-dGr1 * rij[k] - dGr2 * rik[k]) * kD4s9;
    const double dc9 = 0.5 * c9 * (dC[(long long)a * n + b] / c6ij + dC[(long long)a * n + c] / c6ik);
    dcn += -ang * fdmp * dc9 * kD4s9;
    e += dE * kD4s9 / 3.0;
  }
  for (int k = 0; k < 3; ++k) g[k] = block_sum<256>(g[k], sh);
  dcn = block_sum<256>(dcn, sh);
  e = block_sum<256>(e, sh);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + a) + k] = g[k];
    dedcn[md.a0 + a] = dcn;
    eatm[md.a0 + a] = e;
  }
}

// isotropic Coulomb gradient (getCoulombDerivsCluster: src/coulomb/klopmanohno.F90:471-573)
// and repulsion energy/gradient (repulsionEnGrad: src/xtb/repulsion.F90:41-160)
__global__ void __launch_bounds__(64) k_grad_pairs(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                   const int* __restrict__ at, const double* __restrict__ xyz,
                                                   const int* __restrict__ at_sh0, const double* __restrict__ qsh,
                                                   double* grad, double* erep) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= md.n) return;
  const int n = md.n;
  const double* x = xyz + 3 * md.a0;
  const Element& ei = el[at[md.a0 + i]];
  const int zi = at[md.a0 + i];
  double g[3] = {0, 0, 0}, e = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    if (j == i) continue;
    const Element& ej = el[at[md.a0 + j]];
    const int zj = at[md.a0 + j];
    const double rij[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
    const double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    const double r1 = sqrt(r2);
    // Coulomb
    double cq = 0.0;
    for (int is = 0; is < ei.nsh; ++is) {
      const double qi = qsh[md.s0 + at_sh0[md.a0 + i] + is];
      for (int js = 0; js < ej.nsh; ++js) {
        const double gij = 0.5 * (ei.hard[is] + ej.hard[js]);
        const double g1 = 1.0 / (r1 * r1 + 1.0 / (gij * gij));
        cq += g1 * sqrt(g1) * qi * qsh[md.s0 + at_sh0[md.a0 + j] + js];
      }
    }
    for (int k = 0; k < 3; ++k) g[k] += -rij[k] * cq;
    // repulsion
    if (r2 <= 1600.0 && r2 >= 1.0e-8) {
      const double alpha = sqrt(ei.rep_alpha * ej.rep_alpha);
      const double zeff = ei.rep_zeff * ej.rep_zeff;
      const double kexp = (zi > 2 || zj > 2) ? 1.5 : 1.0;
      const double t16 = (zi > 2 || zj > 2) ? r1 * sqrt(r1) : r1, t26 = exp(-alpha * t16);
      const double dE = zeff * t26 / r1;
      e += 0.5 * dE;
      for (int k = 0; k < 3; ++k) g[k] += -(alpha * t16 * kexp + 1.0) * dE * rij[k] / r2;
    }
  }
  for (int k = 0; k < 3; ++k) g[k] = block_sum<64>(g[k], sh);
  e = block_sum<64>(e, sh);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + i) + k] += g[k];
    erep[md.a0 + i] = e;
  }
}

// CN chain rules: GFN CN with the self-energy and AES-radius weights, covalent
// CN with the D4 weights (src/scf_module.F90:726-756; dradcn src/aespot.F90:1292-1307)
__global__ void __launch_bounds__(64) k_cn_chain(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                 const int* __restrict__ at, const double* __restrict__ xyz,
                                                 const double* __restrict__ cn, const double* __restrict__ dhdcn,
                                                 const double* __restrict__ tmp2, const double* __restrict__ dedcn,
                                                 double* grad) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= md.n) return;
  const int n = md.n;
  const double* x = xyz + 3 * md.a0;
  auto wgt = [&](int a) {
    const Element& e = el[at[md.a0 + a]];
    const double t1 = exp(-4.0 * (cn[md.a0 + a] - e.valcn - 1.2));
    double t2 = (5.0 - e.mrad) / (1.0 + 2.0 * t1 + t1 * t1);
    t2 = t2 * 4.0 * t1;
    return dhdcn[md.a0 + a] - tmp2[md.a0 + a] * t2;
  };
  const Element& ei = el[at[md.a0 + i]];
  const double wi = wgt(i), di = dedcn[md.a0 + i];
  const double k4 = 4.10451, k5 = 19.08857, k6 = 2 * 11.28174 * 11.28174;
  double g[3] = {0, 0, 0};
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    if (j == i) continue;
    const Element& ej = el[at[md.a0 + j]];
    const double rij[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
    const double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    if (r2 > 1600.0 || r2 < 1.0e-12) continue;
    const double r1 = sqrt(r2);
    const double rc = ei.rcov_d3 + ej.rcov_d3;
    const double dg = dexp_count_g(10.0, r1, rc) * exp_count_g(20.0, r1, rc + 2) +
                      exp_count_g(10.0, r1, rc) * dexp_count_g(20.0, r1, rc + 2);
    const double d = fabs(ei.en - ej.en) + k5;
    const double dc = k4 * exp(-(d * d) / k6) * derf_count_g(7.5, r1, rc);
    const double f = dg * (wi + wgt(j)) + dc * (di + dedcn[md.a0 + j]);
    for (int k = 0; k < 3; ++k) g[k] += f * rij[k] / r1;
  }
  for (int k = 0; k < 3; ++k) g[k] = block_sum<64>(g[k], sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + i) + k] += g[k];
}

// per-AO weights focc * emo * evtoau for the energy-weighted density
__global__ void k_pew_scale(const MolDesc* __restrict__ desc, const double* __restrict__ focc,
                            const double* __restrict__ emo, double* out) {
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.nao) return;
  out[md.o0 + i] = focc[md.o0 + i] * emo[md.o0 + i] * kEVtoAU;
}

// Wiberg/Mayer bond orders from PS (get_wiberg: src/scc_core.f90:1208-1245)
// mode 0: wbo = sum (closed shell, src/scc_core.f90:1208-1245); mode 1:
// wbo = 2 sum over (Pa S) and mode 2: wbo += 2 sum over (Pb S) (open shell,
// src/scc_core.f90:1248-1295), only for molecules with mask[slot] set.
__global__ void k_wbo(const MolDesc* __restrict__ desc, const double* __restrict__ xyz, const int* __restrict__ at_ao0,
                      const int* __restrict__ at, const Element* __restrict__ el, const double* __restrict__ PS,
                      double* wbo, int mode, const int* __restrict__ mask) {
  const MolDesc md = desc[blockIdx.y];
  if (mask && !mask[md.mol]) return;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = md.n;
  if (t >= n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  double xs = 0.0;
  if (i != j) {
    const double* x = xyz + 3 * md.a0;
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) { const double d = x[3 * i + k] - x[3 * j + k]; r2 += d * d; }
    if (r2 < 100.0) {
      const int N = md.nao;
      const int ia0 = at_ao0[md.a0 + i], ja0 = at_ao0[md.a0 + j];
      const Element& ei = el[at[md.a0 + i]];
      const Element& ej = el[at[md.a0 + j]];
      int ni = 0, nj = 0;
      for (int s = 0; s < ei.nsh; ++s) ni += 2 * ei.l[s] + 1;
      for (int s = 0; s < ej.nsh; ++s) nj += 2 * ej.l[s] + 1;
      const int hi = i > j ? i : j;
      const int ka0 = hi == i ? ia0 : ja0, kn = hi == i ? ni : nj;
      const int ma0 = hi == i ? ja0 : ia0, mn = hi == i ? nj : ni;
      for (int k = ka0; k < ka0 + kn; ++k)
        for (int m = ma0; m < ma0 + mn; ++m)
          xs += PS[md.m0 + (long long)k * N + m] * PS[md.m0 + (long long)m * N + k];
    }
  }
  if (mode == 0) wbo[md.p0 + t] = xs;
  else if (mode == 1) wbo[md.p0 + t] = 2.0 * xs;
  else wbo[md.p0 + t] += 2.0 * xs;
}

}  // namespace xtbamd
