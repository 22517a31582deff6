// Anisotropic electrostatics (AES) of GFN2-xTB on the device.
//
// Model (C. Bannwarth, S. Ehlert, S. Grimme, J. Chem. Theory Comput. 15,
// 1652 (2019), section 2.2): every atom A carries its cumulative atomic
// multipoles, the charge q_A, the dipole mu_A and the traceless quadrupole
// Theta_A (stored packed xx, xy, yy, xz, yz, zz).  For an atom pair with bond
// vector R = r_B - r_A the interaction is
//
//   E_AB = f3(R) P1 + f5(R) P2,
//   P1 = q_B (mu_A.R) - q_A (mu_B.R),
//   P2 = q_B R.Theta_A.R + q_A R.Theta_B.R + (mu_A.mu_B) R^2 - 3 (mu_A.R)(mu_B.R),
//
// with the damped kernels f3 = d3 / R^3, f5 = d5 / R^5, d3 = 1 / (1 + 6 x^3),
// d5 = 1 / (1 + 6 x^4), x = rco / R and rco the mean of the two atoms'
// CN-dependent radii (gab3, gab5 of k_aes_setup).  E_AB is symmetric in A and
// B, so E_AES = 1/2 sum over ordered pairs.  The on-site exchange-correlation
// term is E_XC = sum_A dipk_A |mu_A|^2 + quadk_A |Theta_A|^2.
//
// Reference routines whose results these kernels reproduce: setvsdq
// (src/aespot.F90:727-848) and setdvsdq (:861-952) for the potentials,
// aniso_electro (:569-662) for the energy, aniso_grad (:1070-1165) for the
// gradient.  The formulation here is a different one: each kernel evaluates
// the pair function above (or its derivatives) in tensor form, one wave per
// atom gathering over its partners in lanes, and the moment-origin shift of
// the SCC potentials is applied afterwards as the re-expansion of the
// potential about the coordinate origin.
#include "device_util.h"
#include "kernels.h"

namespace xtbamd {

// packed symmetric 3x3 tensors (xx, xy, yy, xz, yz, zz): t . r and r . t . r
__device__ __forceinline__ void sym_apply(const double* t, const double r[3], double out[3]) {
  out[0] = t[0] * r[0] + t[1] * r[1] + t[3] * r[2];
  out[1] = t[1] * r[0] + t[2] * r[1] + t[4] * r[2];
  out[2] = t[3] * r[0] + t[4] * r[1] + t[5] * r[2];
}
__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// The potential the electrons of atom A see from the multipoles (AO
// contributions vs S + vd . D + vq : Q of the Fock matrix), one wave per atom.
//
// About A's own centre the potential is the negative derivative of E_AES +
// E_XC with respect to A's moments, written as an expansion
//   s + v . (r - r_A) + (r - r_A) . W . (r - r_A):
//   s = -dE/dq_A         = sum_B f3 (mu_B.R) - f5 R.Theta_B.R
//   v = -dE/dmu_A        = sum_B -f3 q_B R + f5 (3 (mu_B.R) R - R^2 mu_B)  - 2 dipk mu_A
//   W = -dE/d(r r^T)_A   = sum_B f5 q_B (R^2/2 I - 3/2 R R^T) - 3 quadk (Theta_A - tr(Theta_A)/3 I)
// (W is taken against the raw second moment r r^T, of which Theta is
// 3/2 times the traceless part).  With shift = 1 the dipole and quadrupole
// integrals are about the coordinate origin (the SCC's integrals), and the
// same quadratic function re-expanded about the origin gives
//   vs = s - v . r_A + r_A . W . r_A,   vd = v - 2 W r_A,   vq = W;
// with shift = 0 (the gradient's atom-centred integrals) vs = s, vd = v.
// vq is stored in the integral order xx, yy, zz, xy, xz, yz with the
// off-diagonal elements doubled (each appears twice in W : Q).
__global__ void __launch_bounds__(64) k_aes_pot(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                const int* __restrict__ at, const double* __restrict__ xyz,
                                                const double* __restrict__ q, const double* __restrict__ dipm,
                                                const double* __restrict__ qp, const double* __restrict__ gab3,
                                                const double* __restrict__ gab5, int shift, double* vs, double* vd,
                                                double* vq, const MolState* __restrict__ mst) {
  const MolDesc md = desc[blockIdx.y];
  if (mst && !mst[blockIdx.y].active) return;
  const int A = blockIdx.x;
  if (A >= md.n) return;
  const int n = md.n, lane = threadIdx.x;
  const double* x = xyz + 3 * md.a0;
  const double ra[3] = {x[3 * A], x[3 * A + 1], x[3 * A + 2]};
  const double* f3row = gab3 + md.p0 + (long long)A * n;
  const double* f5row = gab5 + md.p0 + (long long)A * n;
  // s, v[3], W packed (xx, xy, yy, xz, yz, zz)
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int B = lane; B < n; B += 64) {
    if (B == A) continue;
    const double R[3] = {x[3 * B] - ra[0], x[3 * B + 1] - ra[1], x[3 * B + 2] - ra[2]};
    const double f3 = f3row[B], f5 = f5row[B];
    const double qb = q[md.a0 + B];
    const double* mub = dipm + 3 * (md.a0 + B);
    const double mu[3] = {mub[0], mub[1], mub[2]};
    double tr[3];
    sym_apply(qp + 6 * (md.a0 + B), R, tr);
    const double b = dot3(mu, R), th = dot3(R, tr), r2 = dot3(R, R);
    acc[0] += f3 * b - f5 * th;
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[1 + k] += f5 * (3.0 * b * R[k] - r2 * mu[k]) - f3 * qb * R[k];
    const double h = 0.5 * f5 * qb, g = 1.5 * f5 * qb;
    acc[4] += h * r2 - g * R[0] * R[0];
    acc[5] -= g * R[0] * R[1];
    acc[6] += h * r2 - g * R[1] * R[1];
    acc[7] -= g * R[0] * R[2];
    acc[8] -= g * R[1] * R[2];
    acc[9] += h * r2 - g * R[2] * R[2];
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) acc[k] = wave_sum(acc[k]);
  if (lane != 0) return;
  const Element& e = el[at[md.a0 + A]];
  const double* mua = dipm + 3 * (md.a0 + A);
  const double* tha = qp + 6 * (md.a0 + A);
  // on-site exchange-correlation kernels
  double v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = acc[1 + k] - 2.0 * e.dipk * mua[k];
  const double third = (tha[0] + tha[2] + tha[5]) / 3.0;
  double W[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) W[k] = acc[4 + k] - 3.0 * e.quadk * tha[k];
  W[0] += 3.0 * e.quadk * third;
  W[2] += 3.0 * e.quadk * third;
  W[5] += 3.0 * e.quadk * third;
  double s = acc[0];
  if (shift) {
    double wr[3];
    sym_apply(W, ra, wr);
    s = s - dot3(v, ra) + dot3(ra, wr);
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] -= 2.0 * wr[k];
  }
  vs[md.a0 + A] = s;
#pragma unroll
  for (int k = 0; k < 3; ++k) vd[3 * (md.a0 + A) + k] = v[k];
  double* o = vq + 6 * (md.a0 + A);
  o[0] = W[0];
  o[1] = W[2];
  o[2] = W[5];
  o[3] = 2.0 * W[1];
  o[4] = 2.0 * W[3];
  o[5] = 2.0 * W[4];
}

// AES and XC energy parts of atom A, one wave per atom:
// ae[2A] = 1/2 sum_{B != A} E_AB, ae[2A + 1] = dipk |mu_A|^2 + quadk |Theta_A|^2
// (the full tensor norm: off-diagonal elements count twice).
__global__ void __launch_bounds__(64) k_aes_energy(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                                   const int* __restrict__ at, const double* __restrict__ xyz,
                                                   const double* __restrict__ q, const double* __restrict__ dipm,
                                                   const double* __restrict__ qp, const double* __restrict__ gab3,
                                                   const double* __restrict__ gab5, double* ae,
                                                   const MolState* __restrict__ state) {
  const MolDesc md = desc[blockIdx.y];
  if (!state[blockIdx.y].active) return;
  const int A = blockIdx.x;
  if (A >= md.n) return;
  const int n = md.n, lane = threadIdx.x;
  const double* x = xyz + 3 * md.a0;
  const double ra[3] = {x[3 * A], x[3 * A + 1], x[3 * A + 2]};
  const double qa = q[md.a0 + A];
  const double mua[3] = {dipm[3 * (md.a0 + A)], dipm[3 * (md.a0 + A) + 1], dipm[3 * (md.a0 + A) + 2]};
  const double* tha = qp + 6 * (md.a0 + A);
  const double* f3row = gab3 + md.p0 + (long long)A * n;
  const double* f5row = gab5 + md.p0 + (long long)A * n;
  double e = 0.0;
  for (int B = lane; B < n; B += 64) {
    if (B == A) continue;
    const double R[3] = {x[3 * B] - ra[0], x[3 * B + 1] - ra[1], x[3 * B + 2] - ra[2]};
    const double qb = q[md.a0 + B];
    const double mub[3] = {dipm[3 * (md.a0 + B)], dipm[3 * (md.a0 + B) + 1], dipm[3 * (md.a0 + B) + 2]};
    double ta[3], tb[3];
    sym_apply(tha, R, ta);
    sym_apply(qp + 6 * (md.a0 + B), R, tb);
    const double a = dot3(mua, R), b = dot3(mub, R);
    const double p1 = qb * a - qa * b;
    const double p2 = qb * dot3(R, ta) + qa * dot3(R, tb) + dot3(mua, mub) * dot3(R, R) - 3.0 * a * b;
    e += f3row[B] * p1 + f5row[B] * p2;
  }
  e = wave_sum(e);
  if (lane != 0) return;
  const Element& el_a = el[at[md.a0 + A]];
  const double t2 = tha[0] * tha[0] + tha[2] * tha[2] + tha[5] * tha[5] +
                    2.0 * (tha[1] * tha[1] + tha[3] * tha[3] + tha[4] * tha[4]);
  ae[2 * (md.a0 + A)] = 0.5 * e;
  ae[2 * (md.a0 + A) + 1] = el_a.dipk * dot3(mua, mua) + el_a.quadk * t2;
}

// AES gradient, one wave per atom: grad_A += dE_AES/dr_A at fixed radii, and
// tmp2_A = -dE_AES/drad_A (the dependence on A's CN-dependent radius, which
// k_cn_chain turns into a CN derivative).  With R = r_B - r_A,
//   dE_AB/dR = f3 (q_B mu_A - q_A mu_B)
//            + f5 (2 q_B Theta_A R + 2 q_A Theta_B R + 2 (mu_A.mu_B) R - 3 (mu_B.R) mu_A - 3 (mu_A.R) mu_B)
//            + (f3' P1 + f5' P2) R / |R|,
//   f3' = -3 f3 d3 / |R|,  f5' = -f5 (1 + 4 d5) / |R|,
//   dE_AB/drco = -[3 f3 (1 - d3) P1 + 4 f5 (1 - d5) P2] / rco,  drco/drad_A = 1/2,
// and dE/dr_A = -sum_B dE_AB/dR.
__global__ void __launch_bounds__(64) k_aes_grad(const MolDesc* __restrict__ desc, const double* __restrict__ xyz,
                                                 const double* __restrict__ q, const double* __restrict__ dipm,
                                                 const double* __restrict__ qp, const double* __restrict__ radcn,
                                                 double* grad, double* tmp2) {
  const MolDesc md = desc[blockIdx.y];
  const int A = blockIdx.x;
  if (A >= md.n) return;
  const int n = md.n, lane = threadIdx.x;
  const double* x = xyz + 3 * md.a0;
  const double ra[3] = {x[3 * A], x[3 * A + 1], x[3 * A + 2]};
  const double qa = q[md.a0 + A], rada = radcn[md.a0 + A];
  const double mua[3] = {dipm[3 * (md.a0 + A)], dipm[3 * (md.a0 + A) + 1], dipm[3 * (md.a0 + A) + 2]};
  const double* tha = qp + 6 * (md.a0 + A);
  double g[3] = {0, 0, 0}, dr = 0.0;
  for (int B = lane; B < n; B += 64) {
    if (B == A) continue;
    const double R[3] = {x[3 * B] - ra[0], x[3 * B + 1] - ra[1], x[3 * B + 2] - ra[2]};
    const double r2 = dot3(R, R), rinv = 1.0 / sqrt(r2);
    const double rco = 0.5 * (rada + radcn[md.a0 + B]);
    const double xr = rco * rinv, x3 = xr * xr * xr;
    const double d3 = 1.0 / (1.0 + 6.0 * x3), d5 = 1.0 / (1.0 + 6.0 * x3 * xr);
    const double ri2 = rinv * rinv, ri3 = ri2 * rinv;
    const double f3 = d3 * ri3, f5 = d5 * ri3 * ri2;
    const double qb = q[md.a0 + B];
    const double mub[3] = {dipm[3 * (md.a0 + B)], dipm[3 * (md.a0 + B) + 1], dipm[3 * (md.a0 + B) + 2]};
    double ta[3], tb[3];
    sym_apply(tha, R, ta);
    sym_apply(qp + 6 * (md.a0 + B), R, tb);
    const double a = dot3(mua, R), b = dot3(mub, R), mm = dot3(mua, mub);
    const double p1 = qb * a - qa * b;
    const double p2 = qb * dot3(R, ta) + qa * dot3(R, tb) + mm * r2 - 3.0 * a * b;
    const double radial = (-3.0 * f3 * d3 * p1 - f5 * (1.0 + 4.0 * d5) * p2) * ri2;   // (f3' P1 + f5' P2) / |R|
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double dk = f3 * (qb * mua[k] - qa * mub[k]) +
                        f5 * (2.0 * (qb * ta[k] + qa * tb[k] + mm * R[k]) - 3.0 * (b * mua[k] + a * mub[k])) +
                        radial * R[k];
      g[k] -= dk;
    }
    dr += (3.0 * f3 * (1.0 - d3) * p1 + 4.0 * f5 * (1.0 - d5) * p2) / (2.0 * rco);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = wave_sum(g[k]);
  dr = wave_sum(dr);
  if (lane != 0) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) grad[3 * (md.a0 + A) + k] += g[k];
  tmp2[md.a0 + A] = dr;
}

}  // namespace xtbamd
