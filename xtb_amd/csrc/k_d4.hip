// D4 reference weights of GFN2-xTB's dispersion on the device.
//
// Model (E. Caldeweyher, C. Bannwarth, S. Grimme, J. Chem. Phys. 147, 034112
// (2017); E. Caldeweyher et al., J. Chem. Phys. 150, 154122 (2019)): the
// dynamic polarizability of an atom interpolates between reference systems r
// of the element with Gaussian weights in the coordination number,
//   W_r(c) = sum_{k=1..n_r} exp(-k wf (c - c_r)^2),   w_r = W_r / sum_r' W_r',
// and each reference is scaled by the charge function zeta(q) of the atom.
// When every Gaussian has underflowed the whole weight goes to the
// reference(s) of largest CN.  The same weights appear in the reference's d4
// (src/disp/dftd4.F90:588-657) and weight_references (:948-1065).
#include "device_util.h"
#include "kernels.h"

namespace xtbamd {

// normalised weights w[r] and their CN derivatives dw[r] of one atom at CN c
__device__ __forceinline__ void d4_cn_weights(const Element& e, double c, double* w, double* dw) {
  double sum = 0.0, dsum = 0.0, top = -1e300;
  for (int r = 0; r < e.nref; ++r) {
    const double d = c - e.refcn[r];
    double g = 0.0, dg = 0.0;
    for (int k = 1; k <= e.ncount[r]; ++k) {
      const double a = k * kD4wf, t = a * d * d;
      if (t <= 200.0) {
        const double ex = exp(-t);
        g += ex;
        dg -= 2.0 * a * d * ex;
      }
    }
    w[r] = g;
    dw[r] = dg;
    sum += g;
    dsum += dg;
    top = fmax(top, e.refcn[r]);
  }
  if (sum > 1e-80) {
    const double inv = 1.0 / sum;
    for (int r = 0; r < e.nref; ++r) {
      dw[r] = (dw[r] - w[r] * dsum * inv) * inv;
      w[r] *= inv;
    }
  } else {
    for (int r = 0; r < e.nref; ++r) {
      w[r] = fabs(top - e.refcn[r]) < 1e-12 ? 1.0 : 0.0;
      dw[r] = 0.0;
    }
  }
}

// weights at the D4 coordination number, one thread per reference of the
// molecule (gw[d0 + k] for the reference k of atom d4_at[k])
__global__ void k_d4_weights(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                             const int* __restrict__ at, const int* __restrict__ d4_at, const int* __restrict__ d4_ref,
                             const double* __restrict__ cn_d4, double* gw) {
  const MolDesc md = desc[blockIdx.y];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= md.nd4) return;
  const int a = d4_at[md.d0 + k];
  double w[XTB_MAX_REF], dw[XTB_MAX_REF];
  d4_cn_weights(el[at[md.a0 + a]], cn_d4[md.a0 + a], w, dw);
  gw[md.d0 + k] = w[d4_ref[md.d0 + k]];
}

// charge-scaled and zero-charge weights at the covalent CN with their CN
// derivatives, XTB_MAX_REF slots per atom (unused slots zero), one thread
// per atom: zeta(q) w_r, zeta(0) w_r, zeta(q) dw_r/dc, zeta(0) dw_r/dc
__global__ void k_d4_refweights(const MolDesc* __restrict__ desc, const Element* __restrict__ el,
                                const int* __restrict__ at, const double* __restrict__ cn, const double* __restrict__ q,
                                double* zetavec, double* zerovec, double* zetadcn, double* zerodcn) {
  const MolDesc md = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= md.n) return;
  const Element& e = el[at[md.a0 + i]];
  double w[XTB_MAX_REF], dw[XTB_MAX_REF];
  d4_cn_weights(e, cn[md.a0 + i], w, dw);
  // q null: the zero-charge weights only (they do not depend on the charges)
  const double z = e.d4zeff, hard = kD4gc * e.d4gam, qi = q ? q[md.a0 + i] : 0.0;
  const long long base = (long long)(md.a0 + i) * XTB_MAX_REF;
  for (int r = 0; r < XTB_MAX_REF; ++r) {
    double sq = 0.0, s0 = 0.0;
    if (r < e.nref) {
      if (q) sq = zeta_d(kD4ga, hard, e.refq[r] + z, qi + z);
      s0 = zeta_d(kD4ga, hard, e.refq[r] + z, z);
    }
    const double wr = r < e.nref ? w[r] : 0.0, dr = r < e.nref ? dw[r] : 0.0;
    if (q) {
      zetavec[base + r] = sq * wr;
      zetadcn[base + r] = sq * dr;
    }
    zerovec[base + r] = s0 * wr;
    zerodcn[base + r] = s0 * dr;
  }
}

}  // namespace xtbamd
