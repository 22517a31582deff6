// SCC-iteration kernels (scc: src/scc_core.f90:255-640).  Per iteration:
// potentials -> Fock matrix -> generalized eigensolve (k_linalg.hip) ->
// Fermi occupations -> density -> Mulliken/CAMM populations -> energies ->
// convergence test and Broyden mixing.
#include "device_util.h"
#include "kernels.h"

namespace xtbamd {

// shift[s] = third * qsh[s]^2 * gam[s] + sum_t J[s][t] qsh[t]   (one wavefront per row)
// (ies%addShift: src/xtb/coulomb.f90:89-114, src/xtb/thirdorder.f90:100-131)
__global__ void __launch_bounds__(256) k_shell_pot(const MolDesc* __restrict__ desc, const double* __restrict__ J,
                                                   const double* __restrict__ qsh, const double* __restrict__ gam,
                                                   double* out, int third, const MolState* __restrict__ mst) {
  const MolDesc md = desc[blockIdx.y];
  if (mst && !mst[blockIdx.y].active) return;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= md.nsh) return;
  const double* row = J + md.j0 + (long long)s * md.nsh;
  const double* qv = qsh + md.s0;
  double acc = 0.0;
  for (int t = lane; t < md.nsh; t += 64) acc += row[t] * qv[t];
  acc = wave_sum(acc);
  if (lane == 0) {
    const double qs = qv[s];
    out[md.s0 + s] = third ? (qs * qs * gam[md.s0 + s] + acc) : acc;
  }
}

// zeta scaling of the D4 references (disppot / edisp_scc: src/disp/dftd4.F90:735-864)
__global__ void k_d4_zeta(const MolDesc* __restrict__ desc, const Element* __restrict__ el, const int* __restrict__ at,
                          const int* __restrict__ d4_at, const int* __restrict__ d4_ref, const double* __restrict__ gw,
                          const double* __restrict__ q, double* zv, double* dzv, const MolState* __restrict__ mst) {
  const MolDesc md = desc[blockIdx.y];
  if (mst && !mst[blockIdx.y].active) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= md.nd4) return;
  const int ia = d4_at[md.d0 + k], r = d4_ref[md.d0 + k];
  const Element& e = el[at[md.a0 + ia]];
  double z = 0.0, dz = 0.0;
  if (gw[md.d0 + k] >= 1.0e-7) {
    const double iz = e.d4zeff, c = e.d4gam * kD4gc, qa = q[md.a0 + ia];
    z = zeta_d(kD4ga, c, e.refq[r] + iz, qa + iz);
    dz = dzeta_d(kD4ga, c, e.refq[r] + iz, qa + iz);
  }
  zv[md.d0 + k] = z;
  if (dzv) dzv[md.d0 + k] = dz;
}

// y = dispmat * x  (one wavefront per row)
__global__ void __launch_bounds__(256) k_d4_symv(const MolDesc* __restrict__ desc, const double* __restrict__ M,
                                                 const double* __restrict__ x, double* y,
                                                 const MolState* __restrict__ mst) {
  const MolDesc md = desc[blockIdx.y];
  if (mst && !mst[blockIdx.y].active) return;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= md.nd4) return;
  const double* row = M + md.x0 + (long long)k * md.nd4;
  const double* xv = x + md.d0;
  double acc = 0.0;
  for (int t = lane; t < md.nd4; t += 64) acc += row[t] * xv[t];
  acc = wave_sum(acc);
  if (lane == 0) y[md.d0 + k] = acc;
}

// atomic D4 potential and the final shell shift (addToShellShift: src/scc_core.f90:779-796)
__global__ void k_shift_final(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                              const int* __restrict__ at, const int* __restrict__ sh_at, const double* __restrict__ d4dum,
                              const double* __restrict__ dzv, const double* __restrict__ solv_pot, double* shift,
                              const MolState* __restrict__ mst) {
  const MolDesc md = desc[blockIdx.y];
  if (mst && !mst[blockIdx.y].active) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= md.nsh) return;
  const int ia = sh_at[md.s0 + s];
  int k0 = 0;
  for (int a = 0; a < ia; ++a) k0 += el[at[md.a0 + a]].nref;
  const int nr = el[at[md.a0 + ia]].nref;
  double v = 0.0;
  for (int r = 0; r < nr; ++r) v += d4dum[md.d0 + k0 + r] * dzv[md.d0 + k0 + r];
  // atomic shift: solvation first, then dispersion (src/scc_core.f90:440-452)
  if (solv_pot) v = solv_pot[md.a0 + ia] + v;
  shift[md.s0 + s] += v;
}

// Fock matrix element (i, j) in eV (buildIsoAnisotropicH1: src/scc_core.f90:154-251),
// with the dipole/quadrupole neglect lists of setdqlist (src/aespot.F90:34-85)
// applied per element
__device__ __forceinline__ double fock_elem(const MolDesc& md, int i, int j, const double* __restrict__ S,
                                           const double* __restrict__ H0, const double* __restrict__ D,
                                           const double* __restrict__ Q, const double* __restrict__ shift,
                                           const double* __restrict__ vs, const double* __restrict__ vd,
                                           const double* __restrict__ vq, const int* __restrict__ ao_sh,
                                           const int* __restrict__ ao_at, double thr2) {
  const long long N = md.nao, NN = N * N;
  const long long t = (long long)i * N + j;
  const long long p = md.m0 + t;
  const double dum = S[p] * kAutoEV * 0.5;
  const int si = ao_sh[md.o0 + i], sj = ao_sh[md.o0 + j];
  double h = H0[p] - dum * (shift[md.s0 + si] + shift[md.s0 + sj]);
  const int ia = md.a0 + ao_at[md.o0 + i], ja = md.a0 + ao_at[md.o0 + j];
  h = h + dum * (vs[ia] + vs[ja]);
  const double* Dm = D + 3 * md.m0;
  const double* Qm = Q + 6 * md.m0;
  double d[3], qv[6];
  double t1 = 0.0, t2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) { d[k] = Dm[k * NN + t]; t1 += d[k] * d[k]; }
#pragma unroll
  for (int k = 0; k < 6; ++k) qv[k] = Qm[k * NN + t];
#pragma unroll
  for (int k = 0; k < 3; ++k) t2 -= qv[k] * qv[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) t2 += 2.0 * qv[k] * qv[k];
  if (t1 > thr2) {
    double e1 = 0.0;
#pragma unroll
    for (int l = 0; l < 3; ++l) e1 += d[l] * (vd[3 * ia + l] + vd[3 * ja + l]);
    h += 0.50 * e1 * kAutoEV;
  }
  if (t2 > thr2) {
    double e1 = 0.0;
#pragma unroll
    for (int l = 0; l < 6; ++l) e1 += qv[l] * (vq[6 * ia + l] + vq[6 * ja + l]);
    h += 0.50 * e1 * kAutoEV;
  }
  return h;
}

// The Fock matrix is symmetric bit for bit: S, H0, D and Q are mirrored
// copies of their lower triangles and every term adds the two AOs' potentials
// in a commutative sum.  Workgroup (tile pair, molecule) forms the 32 x 32
// tile (I, J), I >= J, of H from the lower triangles of the 11 inputs (rows
// of 32 consecutive elements, each read once) and writes it and, through LDS,
// its transpose (I, J) -> (J, I); a diagonal tile is formed whole.
__global__ void __launch_bounds__(256) k_fock(const MolDesc* __restrict__ desc, const double* __restrict__ S,
                                              const double* __restrict__ H0, const double* __restrict__ D,
                                              const double* __restrict__ Q, const double* __restrict__ shift,
                                              const double* __restrict__ vs, const double* __restrict__ vd,
                                              const double* __restrict__ vq, const int* __restrict__ ao_sh,
                                              const int* __restrict__ ao_at, double thr2, double* H,
                                              const MolState* __restrict__ state) {
  __shared__ double tile[32][33];
  const MolDesc md = desc[blockIdx.y];
  if (!state[blockIdx.y].active) return;
  const int N = md.nao;
  const int nt = (N + 31) / 32;
  const int npairs = nt * (nt + 1) / 2;
  const int c = threadIdx.x & 31, r0 = threadIdx.x >> 5;
  double* Hm = H + md.m0;
  for (int tp = blockIdx.x; tp < npairs; tp += gridDim.x) {
    // tp = I (I + 1) / 2 + J, 0 <= J <= I
    int I = (int)((sqrt(8.0 * tp + 1.0) - 1.0) * 0.5);
    while (I * (I + 1) / 2 > tp) --I;
    while ((I + 1) * (I + 2) / 2 <= tp) ++I;
    const int J = tp - I * (I + 1) / 2;
    for (int rr = r0; rr < 32; rr += 8) {
      const int i = 32 * I + rr, j = 32 * J + c;
      double h = 0.0;
      if (i < N && j < N) {
        h = fock_elem(md, i, j, S, H0, D, Q, shift, vs, vd, vq, ao_sh, ao_at, thr2);
        Hm[(long long)i * N + j] = h;
      }
      tile[rr][c] = h;
    }
    if (I != J) {
      __syncthreads();
      for (int rr = r0; rr < 32; rr += 8) {
        const int i = 32 * J + rr, j = 32 * I + c;
        if (i < N && j < N) Hm[(long long)i * N + j] = tile[c][rr];
      }
    }
    __syncthreads();
  }
}

// Aufbau occupation of orbital i (occu: src/scc_core.f90:1112-1170)
__device__ __forceinline__ void occu_i(int i, int nel, int nopen, double& fa, double& fb) {
  int f = 0;
  if ((nel & 1) == 0) {
    const int ihomo = nel / 2;
    if (i < ihomo) f = 2;
    if (nopen > 1)
      for (int k = 1; k <= nopen / 2; ++k) {
        if (i == ihomo - k) f -= 1;
        if (i == ihomo + k - 1) f += 1;
      }
  } else {
    const int na = nel / 2 + (nopen - 1) / 2 + 1, nb = nel / 2 - (nopen - 1) / 2;
    if (i < na) f += 1;
    if (i < nb) f += 1;
  }
  fa = (f >= 1) ? 1.0 : 0.0;
  fb = (f == 2) ? 1.0 : 0.0;
}

// Fermi smearing of one spin channel (fermismear: src/scc_core.f90:983-1064)
__device__ void fermi_channel(int norbs, int nel, double t, const double* eig, double* occ, double& ef_out,
                              double& s_out, double* sh) {
  const double boltz = kBoltzmann * kAutoEV, thr = 1e-9;
  const double bkt = boltz * t;
  double ef;
  if (nel + 1 > norbs) {
    ef = eig[nel - 1];
  } else if (nel == 0) {
    ef = eig[0];
  } else {
    ef = 0.5 * (eig[nel - 1] + eig[nel]);
    const double occt = nel;
    for (int cyc = 0; cyc < 200; ++cyc) {
      double tot = 0.0, totd = 0.0;
      for (int i = threadIdx.x; i < norbs; i += blockDim.x) {
        double ff = 0.0, df = 0.0;
        if ((eig[i] - ef) / bkt < 50) {
          const double ex = exp((eig[i] - ef) / bkt);
          ff = 1.0 / (ex + 1.0);
          df = ex / (bkt * (ex + 1.0) * (ex + 1.0));
        }
        occ[i] = ff;
        tot += ff;
        totd += df;
      }
      tot = block_sum<256>(tot, sh);
      totd = block_sum<256>(totd, sh);
      ef = ef + (occt - tot) / totd;
      if (fabs(occt - tot) <= thr) break;
    }
  }
  __syncthreads();
  double s = 0.0;
  for (int i = threadIdx.x; i < norbs; i += blockDim.x) {
    const double o = occ[i];
    if (o > thr && 1.0 - o > thr) s += o * log(o) + (1.0 - o) * log(1.0 - o);
  }
  s = block_sum<256>(s, sh);
  ef_out = ef;
  s_out = s * kBoltzmann * t;
}

// occupations, Fermi smearing, HOMO-LUMO gap and the Broyden input vector
__global__ void __launch_bounds__(256) k_occ_fermi(const MolDesc* __restrict__ desc, MolState* state,
                                                   const double* __restrict__ emo, double* focca, double* foccb,
                                                   double* focc, double etemp, const double* __restrict__ qsh,
                                                   const double* __restrict__ dipm, const double* __restrict__ qp,
                                                   double* q_in) {
  __shared__ double sh[4];
  const MolDesc md = desc[blockIdx.x];
  MolState& st = state[blockIdx.x];
  if (!st.active) return;
  const int N = md.nao;
  const double* e = emo + md.o0;
  double* fa = focca + md.o0;
  double* fb = foccb + md.o0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) occu_i(i, md.nel, md.nopen, fa[i], fb[i]);
  __syncthreads();
  double efa = 0, efb = 0, ga = 0, gb = 0;
  if (etemp > 0.1) {
    if (md.nel > 0) {
      fermi_channel(N, st.ihomoa, etemp, e, fa, efa, ga, sh);
      fermi_channel(N, st.ihomob, etemp, e, fb, efb, gb, sh);
    } else {
      for (int i = threadIdx.x; i < N; i += blockDim.x) fa[i] = fb[i] = 0.0;
      fermi_channel(N, 0, etemp, e, fa, efa, ga, sh);
      fermi_channel(N, 0, etemp, e, fb, efb, gb, sh);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) focc[md.o0 + i] = fa[i] + fb[i];
  // save the input charges / moments (gfn2broyden_save: src/aespot.F90:1384-1403)
  double* qi = q_in + md.v0;
  for (int k = threadIdx.x; k < md.nbr; k += blockDim.x) {
    double v;
    if (k < md.nsh) v = qsh[md.s0 + k];
    else {
      const int r = k - md.nsh, a = r / 9, c = r % 9;
      v = c < 3 ? dipm[3 * (md.a0 + a) + c] : qp[6 * (md.a0 + a) + c - 3];
    }
    qi[k] = v;
  }
  if (threadIdx.x == 0) {
    if (st.ihomo + 1 <= N && st.ihomo >= 1) st.egap = e[st.ihomo] - e[st.ihomo - 1];
    st.ga = ga;
    st.gb = gb;
    st.efa = efa;
    st.efb = efb;
  }
}

// row sums of the populations (mpopsh, mmompop, electro: src/scc_core.f90:686-722,
// 1408-1429; src/aespot.F90:289-427), one wavefront per AO row
__global__ void __launch_bounds__(256) k_population(const MolDesc* __restrict__ desc, const double* __restrict__ P,
                                                    const double* __restrict__ S, const double* __restrict__ H0,
                                                    const double* __restrict__ D, const double* __restrict__ Q,
                                                    double* row_q, double* row_e, double* row_d, double* row_qp,
                                                    const MolState* __restrict__ state, int force) {
  const MolDesc md = desc[blockIdx.y];
  if (!force && !state[blockIdx.y].active) return;
  const int mu = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (mu >= md.nao) return;
  const long long N = md.nao, NN = N * N;
  const long long r0 = md.m0 + (long long)mu * N;
  const double* Dm = D + 3 * md.m0;
  const double* Qm = Q + 6 * md.m0;
  double ps = 0, pe = 0, pd[3] = {0, 0, 0}, pq[6] = {0, 0, 0, 0, 0, 0};
  for (int nu = lane; nu < N; nu += 64) {
    const double p = P[r0 + nu];
    ps += p * S[r0 + nu];
    pe += p * H0[r0 + nu];
    const long long t = (long long)mu * N + nu;
#pragma unroll
    for (int k = 0; k < 3; ++k) pd[k] += p * Dm[k * NN + t];
#pragma unroll
    for (int k = 0; k < 6; ++k) pq[k] += p * Qm[k * NN + t];
  }
  ps = wave_sum(ps);
  pe = wave_sum(pe);
#pragma unroll
  for (int k = 0; k < 3; ++k) pd[k] = wave_sum(pd[k]);
#pragma unroll
  for (int k = 0; k < 6; ++k) pq[k] = wave_sum(pq[k]);
  if (lane == 0) {
    row_q[md.o0 + mu] = ps;
    row_e[md.o0 + mu] = pe;
    for (int k = 0; k < 3; ++k) row_d[3 * (md.o0 + mu) + k] = pd[k];
    for (int k = 0; k < 6; ++k) row_qp[6 * (md.o0 + mu) + k] = pq[k];
  }
}

// shell charges, atomic charges and cumulative atomic multipoles from the row
// sums; one thread per atom
__global__ void k_pop_reduce(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                             const int* __restrict__ at, const double* __restrict__ xyz, const int* __restrict__ at_sh0,
                             const int* __restrict__ at_ao0, const double* __restrict__ zsh,
                             const double* __restrict__ row_q, const double* __restrict__ row_d,
                             const double* __restrict__ row_qp, double* qsh, double* q, double* dipm, double* qp,
                             const MolState* __restrict__ state, int force) {
  const MolDesc md = desc[blockIdx.y];
  if (!force && !state[blockIdx.y].active) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.n) return;
  const Element& e = el[at[md.a0 + i]];
  const double* R = xyz + 3 * (md.a0 + i);
  int mu = md.o0 + at_ao0[md.a0 + i];
  double qa = 0.0;
  double d[3] = {0, 0, 0}, qq[6] = {0, 0, 0, 0, 0, 0};
  for (int s = 0; s < e.nsh; ++s) {
    double ps = 0.0;
    for (int k = 0; k < 2 * e.l[s] + 1; ++k, ++mu) {
      const double rq = row_q[mu];
      ps += rq;
      const double* pd = row_d + 3 * mu;
      const double* pq = row_qp + 6 * mu;  // xx, yy, zz, xy, xz, yz
      for (int c = 0; c < 3; ++c) d[c] += R[c] * rq - pd[c];
      // lin order xx, xy, yy, xz, yz, zz
      qq[0] += 2.0 * pd[0] * R[0] - R[0] * R[0] * rq - pq[0];
      qq[2] += 2.0 * pd[1] * R[1] - R[1] * R[1] * rq - pq[1];
      qq[5] += 2.0 * pd[2] * R[2] - R[2] * R[2] * rq - pq[2];
      qq[1] += pd[1] * R[0] + pd[0] * R[1] - R[0] * R[1] * rq - pq[3];
      qq[3] += pd[2] * R[0] + pd[0] * R[2] - R[0] * R[2] * rq - pq[4];
      qq[4] += pd[2] * R[1] + pd[1] * R[2] - R[1] * R[2] * rq - pq[5];
    }
    const int sh = md.s0 + at_sh0[md.a0 + i] + s;
    const double v = zsh[sh] - ps;
    qsh[sh] = v;
    qa += v;
  }
  q[md.a0 + i] = qa;
  for (int c = 0; c < 3; ++c) dipm[3 * (md.a0 + i) + c] = d[c];
  const double tii = 0.50 * (qq[0] + qq[2] + qq[5]);
  for (int c = 0; c < 6; ++c) qq[c] = 1.50 * qq[c];
  qq[0] -= tii;
  qq[2] -= tii;
  qq[5] -= tii;
  for (int c = 0; c < 6; ++c) qp[6 * (md.a0 + i) + c] = qq[c];
}

// Energy, convergence test and charge mixing of one molecule (one workgroup):
// src/scc_core.f90:523-632 with broyden (src/broyden.f90:37-153).
__global__ void __launch_bounds__(256) k_scc_control(
    const MolDesc* __restrict__ desc, MolState* state, const double* __restrict__ jq, const double* __restrict__ gam,
    const double* __restrict__ row_e, const double* __restrict__ ae, const double* __restrict__ zv,
    const double* __restrict__ d4dum, const double* __restrict__ solv_qbq, const double* __restrict__ solv_gsasa,
    double solv_gshift, const int* __restrict__ sh_at, double* qsh, double* q, double* dipm, double* qp,
    double* q_in, double* dq, double* qlast, double* dqlast, double* bdf, double* bu, double* ba, double* bomega,
    double scfconv, double qconv) {
  __shared__ double sh[4];
  __shared__ double red[8];
  const MolDesc md = desc[blockIdx.x];
  MolState& st = state[blockIdx.x];
  if (!st.active) return;
  const int tid = threadIdx.x;
  // --- energies
  double ej = 0.0, e3 = 0.0;
  for (int s = tid; s < md.nsh; s += blockDim.x) {
    const double qs = qsh[md.s0 + s];
    ej += jq[md.s0 + s] * qs;
    e3 += qs * qs * qs * gam[md.s0 + s] / 3.0;
  }
  ej = block_sum<256>(ej, sh);
  e3 = block_sum<256>(e3, sh);
  double eh = 0.0;
  for (int i = tid; i < md.nao; i += blockDim.x) eh += row_e[md.o0 + i];
  eh = block_sum<256>(eh, sh);
  double eaes = 0, apol = 0;
  for (int i = tid; i < md.n; i += blockDim.x) {
    eaes += ae[2 * (md.a0 + i)];
    apol += ae[2 * (md.a0 + i) + 1];
  }
  eaes = block_sum<256>(eaes, sh);
  apol = block_sum<256>(apol, sh);
  double edisp = 0.0;
  for (int k = tid; k < md.nd4; k += blockDim.x) edisp += zv[md.d0 + k] * (0.5 * d4dum[md.d0 + k]);
  edisp = block_sum<256>(edisp, sh);
  const double ees = 0.5 * ej + e3;
  double eel = ees + eh * kEVtoAU;
  eel = eel + eaes + apol;
  eel = eel + edisp;
  // solvation free energy (src/scc_core.f90:544-548, src/solv/gbsa.f90:525-548)
  if (solv_qbq) eel = eel + ((solv_qbq[blockIdx.x] + solv_gsasa[blockIdx.x]) + solv_gshift);
  eel = eel + st.ga + st.gb;
  const double eold = st.eel;
  const bool econv = fabs(eel - eold) < scfconv;
  // --- charge difference and rms
  double* qi = q_in + md.v0;
  double* dqv = dq + md.v0;
  double ss = 0.0;
  for (int k = tid; k < md.nbr; k += blockDim.x) {
    double v;
    if (k < md.nsh) v = qsh[md.s0 + k];
    else {
      const int r = k - md.nsh, a = r / 9, c = r % 9;
      v = c < 3 ? dipm[3 * (md.a0 + a) + c] : qp[6 * (md.a0 + a) + c - 3];
    }
    const double d = v - qi[k];
    dqv[k] = d;
    ss += d * d;
  }
  ss = block_sum<256>(ss, sh);
  const double rmsq = sqrt(ss / (double)md.n);
  const bool qok = rmsq < qconv;
  const int iter = st.iter + 1;
  const double alpha = 0.40;  // broydamp (src/setparam.f90:235)
  const int nbr = md.nbr;
  double* ql = qlast + md.v0;
  double* dql = dqlast + md.v0;
  if (md.n <= 1) {
    // simple damping (src/scc_core.f90:574-602)
    const double damp = st.damp;
    for (int k = tid; k < nbr; k += blockDim.x) {
      double v = qi[k] + dqv[k];  // output value
      qi[k] = damp * v + (1.0 - damp) * qi[k];
    }
    __syncthreads();
    if (tid == 0) {
      double dn = (eel - eold < 0.0) ? damp * 1.15 : 0.20;
      dn = fmin(dn, 1.0);
      if (st.egap < 1.0) dn = fmin(dn, 0.5);
      st.damp = dn;
    }
  } else if (iter == 1) {
    for (int k = tid; k < nbr; k += blockDim.x) {
      dql[k] = dqv[k];
      ql[k] = qi[k];
      qi[k] = qi[k] + alpha * dqv[k];
    }
  } else {
    const int it1 = iter - 1, i1 = it1 - 1;
    const int mx = md.maxiter;
    double* df = bdf + md.h0;
    double* u = bu + md.h0;
    double* a = ba + md.g0;
    double* om = bomega + (long long)blockIdx.x * mx;
    double* dfi = df + (long long)i1 * nbr;
    // omega weight
    double w = sqrt(ss);
    w = (w > 0.01 / 100000.0) ? 0.01 / w : 100000.0;
    if (w < 1.0) w = 1.0;
    double nn = 0.0;
    for (int k = tid; k < nbr; k += blockDim.x) {
      const double v = dqv[k] - dql[k];
      dfi[k] = v;
      nn += v * v;
    }
    nn = block_sum<256>(nn, sh);
    const double inv = 1.0 / sqrt(nn);
    for (int k = tid; k < nbr; k += blockDim.x) dfi[k] = inv * dfi[k];
    if (tid == 0) om[i1] = w;
    __syncthreads();
    // a(i, it1) and c(i)
    double* c = a + (long long)mx * mx;           // scratch: c[mx], beta[mx*mx], gamma[mx]
    double* beta = c + mx;
    double* gmm = beta + (long long)mx * mx;
    for (int i = 0; i < it1; ++i) {
      const double* dfk = df + (long long)i * nbr;
      double s1 = 0.0, s2 = 0.0;
      for (int k = tid; k < nbr; k += blockDim.x) {
        s1 += dfk[k] * dfi[k];
        s2 += dfk[k] * dqv[k];
      }
      s1 = block_sum<256>(s1, sh);
      s2 = block_sum<256>(s2, sh);
      if (tid == 0) {
        a[(long long)i * mx + i1] = s1;
        a[(long long)i1 * mx + i] = s1;
        c[i] = om[i] * s2;
      }
    }
    __syncthreads();
    for (int e = tid; e < it1 * it1; e += blockDim.x) {
      const int i = e / it1, j = e % it1;
      double v = om[j] * om[i] * a[(long long)j * mx + i];
      if (i == j) v += 0.01 * 0.01;
      beta[e] = v;
    }
    __syncthreads();
    // Gauss-Jordan inverse of beta (matinv: src/broyden.f90:158-184), in place with an identity copy in gmm area
    double* inv_m = gmm + mx;
    for (int e = tid; e < it1 * it1; e += blockDim.x) inv_m[e] = (e / it1 == e % it1) ? 1.0 : 0.0;
    __syncthreads();
    for (int k = 0; k < it1; ++k) {
      if (tid == 0) {
        int p = k;
        for (int i = k + 1; i < it1; ++i)
          if (fabs(beta[(long long)i * it1 + k]) > fabs(beta[(long long)p * it1 + k])) p = i;
        red[0] = p;
      }
      __syncthreads();
      const int p = (int)red[0];
      if (p != k)
        for (int j = tid; j < it1; j += blockDim.x) {
          double t = beta[(long long)k * it1 + j];
          beta[(long long)k * it1 + j] = beta[(long long)p * it1 + j];
          beta[(long long)p * it1 + j] = t;
          t = inv_m[(long long)k * it1 + j];
          inv_m[(long long)k * it1 + j] = inv_m[(long long)p * it1 + j];
          inv_m[(long long)p * it1 + j] = t;
        }
      __syncthreads();
      const double piv = beta[(long long)k * it1 + k];
      __syncthreads();
      for (int j = tid; j < it1; j += blockDim.x) {
        beta[(long long)k * it1 + j] /= piv;
        inv_m[(long long)k * it1 + j] /= piv;
      }
      __syncthreads();
      // eliminate column k: each thread owns full rows i
      for (int i = tid; i < it1; i += blockDim.x) {
        if (i == k) continue;
        const double f = beta[(long long)i * it1 + k];
        if (f == 0.0) continue;
        for (int j = 0; j < it1; ++j) {
          beta[(long long)i * it1 + j] -= f * beta[(long long)k * it1 + j];
          inv_m[(long long)i * it1 + j] -= f * inv_m[(long long)k * it1 + j];
        }
      }
      __syncthreads();
    }
    // gamma(j) = sum_i c(i) beta^-1(i, j)
    for (int j = tid; j < it1; j += blockDim.x) {
      double s = 0.0;
      for (int i = 0; i < it1; ++i) s += c[i] * inv_m[(long long)i * it1 + j];
      gmm[j] = s;
    }
    __syncthreads();
    double* ui = u + (long long)i1 * nbr;
    for (int k = tid; k < nbr; k += blockDim.x) {
      ui[k] = alpha * dfi[k] + inv * (qi[k] - ql[k]);
      dql[k] = dqv[k];
      ql[k] = qi[k];
    }
    __syncthreads();
    for (int k = tid; k < nbr; k += blockDim.x) {
      double v = qi[k] + alpha * dqv[k];
      for (int i = 0; i < it1; ++i) v = v - om[i] * gmm[i] * u[(long long)i * nbr + k];
      qi[k] = v;
    }
  }
  __syncthreads();
  // unpack mixed charges and moments (gfn2broyden_out: src/aespot.F90:1405-1425), q from qsh
  for (int k = tid; k < nbr; k += blockDim.x) {
    if (k < md.nsh) qsh[md.s0 + k] = qi[k];
    else {
      const int r = k - md.nsh, a = r / 9, c = r % 9;
      if (c < 3) dipm[3 * (md.a0 + a) + c] = qi[k];
      else qp[6 * (md.a0 + a) + c - 3] = qi[k];
    }
  }
  __syncthreads();
  for (int i = tid; i < md.n; i += blockDim.x) q[md.a0 + i] = 0.0;
  __syncthreads();
  if (tid == 0) {
    for (int s = 0; s < md.nsh; ++s) q[md.a0 + sh_at[md.s0 + s]] += qsh[md.s0 + s];
    st.eel = eel;
    st.eold = eold;
    st.ees = ees;
    st.eaes = eaes;
    st.epol = apol;
    st.ed = edisp;
    st.eh0 = eh;
    st.rmsq = rmsq;
    st.iter = iter;
    if (econv && qok) {
      st.converged = 1;
      if (st.lastdiag) st.active = 0;
      st.lastdiag = 1;
    }
    if (iter >= md.maxiter) st.active = 0;
  }
}

}  // namespace xtbamd
