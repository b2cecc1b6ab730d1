// Galerkin coarse operator on the third multigrid level.
//
// The hierarchy rediscretises: a coarse element's operator is KE times the
// mean of its 8 children's scales (times 2 for the doubled size).  Where the
// density varies inside a coarse element (struts next to voids), that average
// misrepresents which of the coarse element's modes are stiff, and the coarse
// correction weakens as the design gains contrast.  On the third level
// (level 2, coarse elements of 4x4x4 fine elements) the operator here is the
// Galerkin product instead: with the second level rediscretised, each
// third-level element is the union of 8 second-level elements c with scales
// s_c, and P^T A P restricted to it is exactly
//   K_E = sum_c s_c M_c,   M_c = P_c^T KE P_c,
// P_c the trilinear interpolation from the coarse corners to child c's
// corners.  In the Walsh-Hadamard basis of the element kernel this is
// sum_w shat_w G_w with shat = the Walsh transform of (s_c) over the child
// positions and G_w coupling parity sector chi with chi XOR w only
// (tk_galerkin_pattern.hpp, 279 entries against the sector form's 45).  So the
// level stores 8 Walsh coefficients per element instead of one scale, and its
// sweeps apply the 279-entry table.  For uniform children it is the
// rediscretised operator (sum_c M_c = 2 KE).
//
// Jacobi scaling: diag(K_E) = sum_w shat_w Dw_w (Dw = Walsh transform of the
// M_c diagonals), formed per node at set-up.  Chebyshev bound: with
// D_c = diag(M_c), x^T K_E x <= max_c lambda_max(D_c^-1 M_c) x^T diag(K_E) x
// for s_c >= 0, and summing over elements gives
// lambda_max(D^-1 A) <= max_c lambda_max(D_c^-1 M_c) (6.46 for the default KE
// against 5.24 for KE itself), a fixed bound computed on the host.
#include "tk_plan.hpp"
#include "tk_galerkin_pattern.hpp"

#include <cmath>
#include <vector>

namespace tk {

// Largest eigenvalue of D^-1/2 M D^-1/2 over the rows of M with a positive
// diagonal (24 x 24, symmetric), by cyclic Jacobi rotations.
double scaled_lambda_max(const double* M) {
  double dmax = 0.0;
  for (int i = 0; i < 24; ++i) dmax = std::fmax(dmax, M[i * 25]);
  int idx[24], n = 0;
  for (int i = 0; i < 24; ++i)
    if (M[i * 25] > 1e-14 * dmax) idx[n++] = i;
  double a[24][24];
  double dinv[24];
  for (int i = 0; i < n; ++i) dinv[i] = 1.0 / std::sqrt(M[idx[i] * 25]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[i][j] = M[idx[i] * 24 + idx[j]] * dinv[i] * dinv[j];
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a[i][j] * a[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (std::fabs(a[p][q]) < 1e-300) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t =
            (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  double m = 0.0;
  for (int i = 0; i < n; ++i) m = std::fmax(m, a[i][i]);
  return m;
}

namespace {

inline int walsh_sign(int w, int c) { return (__builtin_popcount(w & c) & 1) ? -1 : 1; }

// child c's corner DOFs from the coarse element's corner DOFs (24 x 24)
void child_interp(int c, double P[24][24]) {
  const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
  auto w1 = [](int f, int b) { return f == 2 * b ? 1.0 : (f == 1 ? 0.5 : 0.0); };
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) P[i][j] = 0.0;
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      const double w = w1(cx + corner_x(a), corner_x(b)) * w1(cy + corner_y(a), corner_y(b)) *
                       w1(cz + corner_z(a), corner_z(b));
      for (int k = 0; k < 3; ++k) P[3 * a + k][3 * b + k] = w;
    }
}

}  // namespace

// Builds the coupling table, the Walsh-domain diagonal and the bound for the
// plan's KE; false when KE does not have the expected pattern.
bool galerkin_setup(Plan& P) {
  // W[(a, k), (m, k)]: the kernel's Walsh basis (offset 0 -> -1, 1 -> +1 per mode bit)
  double W[24][24] = {};
  for (int a = 0; a < 8; ++a) {
    const int o[3] = {corner_x(a), corner_y(a), corner_z(a)};
    for (int m = 0; m < 8; ++m) {
      double v = 1.0;
      for (int d = 0; d < 3; ++d)
        if ((m >> d) & 1) v *= o[d] ? 1.0 : -1.0;
      for (int k = 0; k < 3; ++k) W[3 * a + k][3 * m + k] = v;
    }
  }
  std::vector<double> G(8 * 576, 0.0), Dw(8 * 8 * 3, 0.0);
  double lam = 0.0;
  for (int c = 0; c < 8; ++c) {
    double Pc[24][24], T1[24][24], M[576];
    child_interp(c, Pc);
    for (int i = 0; i < 24; ++i)          // T1 = KE Pc
      for (int j = 0; j < 24; ++j) {
        double acc = 0.0;
        for (int k = 0; k < 24; ++k) acc += P.KE[i * 24 + k] * Pc[k][j];
        T1[i][j] = acc;
      }
    for (int i = 0; i < 24; ++i)          // M = Pc^T KE Pc
      for (int j = 0; j < 24; ++j) {
        double acc = 0.0;
        for (int k = 0; k < 24; ++k) acc += Pc[k][i] * T1[k][j];
        M[i * 24 + j] = acc;
      }
    lam = std::fmax(lam, scaled_lambda_max(M));
    double T2[24][24];
    for (int i = 0; i < 24; ++i)          // T2 = M W
      for (int j = 0; j < 24; ++j) {
        double acc = 0.0;
        for (int k = 0; k < 24; ++k) acc += M[i * 24 + k] * W[k][j];
        T2[i][j] = acc;
      }
    for (int w = 0; w < 8; ++w) {
      const double sg = walsh_sign(w, c) / 8.0;
      for (int i = 0; i < 24; ++i)        // G_w += sign W^T M W / 64 / 8
        for (int j = 0; j < 24; ++j) {
          double acc = 0.0;
          for (int k = 0; k < 24; ++k) acc += W[k][i] * T2[k][j];
          G[w * 576 + i * 24 + j] += sg * acc / 64.0;
        }
      for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k) Dw[(w * 8 + a) * 3 + k] += sg * M[(3 * a + k) * 25];
    }
  }
  // the table in pattern order; everything outside the pattern must vanish
  double gmax = 0.0;
  for (double v : G) gmax = std::fmax(gmax, std::fabs(v));
  std::vector<char> used(8 * 576, 0);
  double tab64[kGalN];
  for (int g = 0; g < kGalGroups; ++g)
    for (int k = kGalStart[g]; k < kGalStart[g + 1]; ++k) {
      const int e = kGalW[g] * 576 + kGalRow[g] * 24 + kGalCol[k];
      tab64[k] = G[e];
      used[e] = 1;
    }
  for (int e = 0; e < 8 * 576; ++e)
    if (!used[e] && std::fabs(G[e]) > 1e-12 * gmax) return false;
  float tab32[kGalN];
  for (int k = 0; k < kGalN; ++k) tab32[k] = (float)tab64[k];
  if (hipMalloc((void**)&P.gal_tab64, sizeof(double) * kGalN) != hipSuccess ||
      hipMalloc((void**)&P.gal_tab32, sizeof(float) * kGalN) != hipSuccess ||
      hipMalloc((void**)&P.gal_dw, sizeof(double) * Dw.size()) != hipSuccess ||
      hipMemcpy(P.gal_tab64, tab64, sizeof(double) * kGalN, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P.gal_tab32, tab32, sizeof(float) * kGalN, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P.gal_dw, Dw.data(), sizeof(double) * Dw.size(), hipMemcpyHostToDevice) !=
          hipSuccess)
    return false;
  P.lam_gal = lam;
  return true;
}

// Walsh coefficients of the 8 children's scales of every element of level
// c (children on level f; children outside f's grid count as 0).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gal_coeffs(Grid f, Grid c, const T* __restrict__ sf,
                                                        T* __restrict__ sh) {
  const int E = blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= c.nele) return;
  const int Ey = E % c.nely, t = E / c.nely, Ex = t % c.nelx, Ez = t / c.nelx;
  double v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ex = 2 * Ex + (k & 1), ey = 2 * Ey + ((k >> 1) & 1), ez = 2 * Ez + (k >> 2);
    v[k] = (ex < f.nelx && ey < f.nely && ez < f.nelz)
               ? (double)sf[ey + ex * f.nely + ez * f.nelx * f.nely]
               : 0.0;
  }
  // Walsh butterflies over the child bits x, y, z: v[w] = sum_c (-1)^{w.c} s_c
#pragma unroll
  for (int h = 1; h < 8; h <<= 1)
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (!(k & h)) {
        const double a = v[k], b = v[k | h];
        v[k] = a + b;
        v[k | h] = a - b;
      }
  T* o = sh + 8 * (size_t)E;
#pragma unroll
  for (int k = 0; k < 8; ++k) o[k] = (T)v[k];
}

// Inverse diagonal of the Galerkin level (1 on fixed DOFs): per node, the
// corner diagonals of its (up to 8) elements, diag(K_E) = sum_w shat_w Dw_w.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gal_diag(Grid g, const T* __restrict__ sh,
                                                      const double* __restrict__ dw,
                                                      const uint8_t* __restrict__ fixed,
                                                      T* __restrict__ dinv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnode) return;
  const int iy = n % g.nny, t = n / g.nny, ix = t % g.nnx, iz = t / g.nnx;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < 8; ++a) {
    const int ey = iy - corner_y(a), ex = ix - corner_x(a), ez = iz - corner_z(a);
    if ((unsigned)ey >= (unsigned)g.nely || (unsigned)ex >= (unsigned)g.nelx ||
        (unsigned)ez >= (unsigned)g.nelz)
      continue;
    const T* s = sh + 8 * (size_t)(ey + ex * g.nely + ez * g.nelx * g.nely);
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      const double sw = (double)s[w];
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] += sw * dw[(w * 8 + a) * 3 + k];
    }
  }
  const uint8_t f = fixed[n];
#pragma unroll
  for (int k = 0; k < 3; ++k) dinv[3 * n + k] = (T)(((f >> k) & 1) ? 1.0 : 1.0 / acc[k]);
}

template <typename T> static T* gal_sh(Plan& P);
template <> double* gal_sh<double>(Plan& P) { return P.gal_sh64; }
template <> float* gal_sh<float>(Plan& P) { return P.gal_sh32; }
template <typename T> static T* scales_of(Level& L);
template <> double* scales_of<double>(Level& L) { return L.f64.s; }
template <> float* scales_of<float>(Level& L) { return L.f32.s; }
template <typename T> static T* dinv_of(Level& L);
template <> double* dinv_of<double>(Level& L) { return L.f64.dinv; }
template <> float* dinv_of<float>(Level& L) { return L.f32.dinv; }

template <typename T>
void launch_gal_coeffs(Plan& P, hipStream_t s) {
  const int l = P.gal_level;
  const Level& C = P.lv[l];
  hipLaunchKernelGGL(k_gal_coeffs<T>, dim3(nblocks(C.g.nele)), dim3(kBlock), 0, s, P.lv[l - 1].g,
                     C.g, scales_of<T>(P.lv[l - 1]), gal_sh<T>(P));
}

template <typename T>
void launch_gal_diag(Plan& P, hipStream_t s) {
  Level& L = P.lv[P.gal_level];
  hipLaunchKernelGGL(k_gal_diag<T>, dim3(nblocks(L.g.nnode)), dim3(kBlock), 0, s, L.g,
                     gal_sh<T>(P), P.gal_dw, L.fixed, dinv_of<T>(L));
}

template void launch_gal_coeffs<double>(Plan&, hipStream_t);
template void launch_gal_coeffs<float>(Plan&, hipStream_t);
template void launch_gal_diag<double>(Plan&, hipStream_t);
template void launch_gal_diag<float>(Plan&, hipStream_t);

}  // namespace tk
