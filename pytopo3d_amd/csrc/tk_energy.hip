// Element energies, compliance and the compliance / volume sensitivities.
//   ce[e] = u_e . KE . u_e                        pytopo3d/core/compliance.py:18-75
//   c     = sum_e (Emin + (1-Emin) rho^p) ce[e]   pytopo3d/core/optimizer.py:596-605
//   dc/drho = -p (1-Emin) rho^(p-1) ce            pytopo3d/core/optimizer.py:607-612
//   chain-rule weights (g * drho)/Hs for g = dc/drho and g = free mask
//                                                 pytopo3d/utils/filter.py:115-118
// The transposed filter that finishes the chain rule is k_filter_transpose2.
// When KE has the mirror symmetry of the element kernel the energy is formed in
// the Walsh-Hadamard sector form, u.KE.u = a.B.a with a the corner transform
// of u and B the 45 sector entries (21 modes, ~170 operations per element
// instead of the 1152 of the dense product).
#include "tk_plan.hpp"
#include "tk_sector.hpp"

namespace tk {

constexpr double kEminE = 1.0e-9;

__global__ __launch_bounds__(kBlock) void k_energy(
    Grid g, const double* __restrict__ KE, const double* __restrict__ U,
    const double* __restrict__ rho, const double* __restrict__ drho,
    const double* __restrict__ Hs, const uint8_t* __restrict__ fl, double penal, double c0,
    double* __restrict__ ce_out, double* __restrict__ dc_out, double* __restrict__ wc,
    double* __restrict__ wv, double* partials, Scalars* sc) {
  __shared__ double ke[576];
  for (int i = threadIdx.x; i < 576; i += blockDim.x) ke[i] = KE[i];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double part[1] = {0.0};
  if (e < g.nele) {
    const int iy = e % g.nely, t = e / g.nely, ix = t % g.nelx, iz = t / g.nelx;
    double u[24];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int m = (iy + corner_y(a)) + (ix + corner_x(a)) * g.nny +
                    (iz + corner_z(a)) * g.nnx * g.nny;
      u[3 * a] = U[3 * m];
      u[3 * a + 1] = U[3 * m + 1];
      u[3 * a + 2] = U[3 * m + 2];
    }
    double ce = 0.0;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
      double row = 0.0;
#pragma unroll
      for (int j = 0; j < 24; ++j) row += ke[i * 24 + j] * u[j];
      ce += u[i] * row;
    }
    const double r = rho[e];
    const double s = kEminE + (1.0 - kEminE) * pow(r, penal);
    part[0] = s * ce;
    const double dc = c0 * pow(r, penal - 1.0) * ce;
    if (ce_out) ce_out[e] = ce;
    if (dc_out) dc_out[e] = dc;
    if (wc) {
      const double d = drho[e], h = Hs[e];
      wc[e] = (dc * d) / h;
      wv[e] = ((fl[e] == 0 ? 1.0 : 0.0) * d) / h;
    }
  }
  block_partials<1>(part, partials);
}

__global__ __launch_bounds__(kBlock) void k_energy_sector(
    Grid g, SparseKE<double> kb, const double* __restrict__ U, const double* __restrict__ rho,
    const double* __restrict__ drho, const double* __restrict__ Hs,
    const uint8_t* __restrict__ fl, double penal, double c0, double* __restrict__ ce_out,
    double* __restrict__ dc_out, double* __restrict__ wc, double* __restrict__ wv,
    double* partials, Scalars* sc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double part[1] = {0.0};
  if (e < g.nele) {
    const int iy = e % g.nely, t = e / g.nely, ix = t % g.nelx, iz = t / g.nelx;
    const int nxy = g.nnx * g.nny;
    double Fb[3][4], Ft[3][4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int m0 = (iy + (f >> 1)) + (ix + (f & 1)) * g.nny + iz * nxy;
      const int m1 = m0 + nxy;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Fb[c][f] = U[3 * m0 + c];
        Ft[c][f] = U[3 * m1 + c];
      }
    }
    double a[3][8], o[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      face_fwd(Fb[c]);
      face_fwd(Ft[c]);
      a[c][0] = 0.0;
#pragma unroll
      for (int f = 1; f < 4; ++f) a[c][f] = Fb[c][f] + Ft[c][f];
#pragma unroll
      for (int f = 0; f < 4; ++f) a[c][f + 4] = Ft[c][f] - Fb[c][f];
    }
    sector_apply<double>(kb.v, 1.0, a, o);
    double ce = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int m = 1; m < 8; ++m) ce += a[c][m] * o[c][m];
    const double r = rho[e];
    const double s = kEminE + (1.0 - kEminE) * pow(r, penal);
    part[0] = s * ce;
    const double dc = c0 * pow(r, penal - 1.0) * ce;
    if (ce_out) ce_out[e] = ce;
    if (dc_out) dc_out[e] = dc;
    if (wc) {
      const double d = drho[e], h = Hs[e];
      wc[e] = (dc * d) / h;
      wv[e] = ((fl[e] == 0 ? 1.0 : 0.0) * d) / h;
    }
  }
  block_partials<1>(part, partials);
}

// k_energy_sector without a plan (c's partials and the finalize scalars in
// caller-provided device memory).
void launch_energy_sector_raw(const Grid& g, const double* sec, const double* U,
                              const double* rho, double penal, double* ce, double* dc_drho,
                              double* partials, Scalars* sc, hipStream_t s) {
  SparseKE<double> kb;
  for (int k = 0; k < kSecN; ++k) kb.v[k] = sec[k];
  hipLaunchKernelGGL(k_energy_sector, dim3(nblocks(g.nele)), dim3(kBlock), 0, s, g, kb, U, rho,
                     nullptr, nullptr, nullptr, penal, -penal * (1.0 - kEminE), ce, dc_drho,
                     nullptr, nullptr, partials, sc);
}

void launch_energy(Plan& P, const double* U, const double* rho, const double* drho, double penal,
                   double* ce, double* dc_drho, double* wc, double* wv, hipStream_t s) {
  const double c0 = -penal * (1.0 - kEminE);
  if (P.sector_ok) {
    SparseKE<double> kb;
    for (int k = 0; k < kSecN; ++k) kb.v[k] = P.sec64[k];
    hipLaunchKernelGGL(k_energy_sector, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g, kb, U,
                       rho, drho, P.Hs, P.eflags, penal, c0, ce, dc_drho, wc, wv, P.partials,
                       P.sc);
  } else {
    hipLaunchKernelGGL(k_energy, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g, P.KE64, U,
                       rho, drho, P.Hs, P.eflags, penal, c0, ce, dc_drho, wc, wv, P.partials,
                       P.sc);
  }
  finalize(P, kFinC, nblocks(P.g.nele), s);
}

}  // namespace tk

namespace tk {

// wc = (g_c * drho)/Hs, wv = (g_v * drho)/Hs with g_v = free indicator when null
// (pytopo3d/utils/filter.py:115-118).
__global__ __launch_bounds__(kBlock) void k_chain_weights(int n, const double* __restrict__ gc,
                                                           const double* __restrict__ gv,
                                                           const double* __restrict__ drho,
                                                           const double* __restrict__ Hs,
                                                           const uint8_t* __restrict__ fl,
                                                           double* __restrict__ wc,
                                                           double* __restrict__ wv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double d = drho[e], h = Hs[e];
  wc[e] = (gc[e] * d) / h;
  if (wv) {
    const double g = gv ? gv[e] : (fl[e] == 0 ? 1.0 : 0.0);
    wv[e] = (g * d) / h;
  }
}

void launch_chain_weights(Plan& P, const double* gc, const double* gv, const double* drho,
                          double* wc, double* wv, hipStream_t s) {
  hipLaunchKernelGGL(k_chain_weights, dim3(nblocks(P.g.nele)), dim3(kBlock), 0, s, P.g.nele, gc,
                     gv, drho, P.Hs, P.eflags, wc, wv);
}

}  // namespace tk
