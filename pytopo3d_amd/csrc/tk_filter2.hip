// LDS-tiled density filter (fast path of H x, used for the forward filter,
// its transpose and every OC bisection volume).
//
// A workgroup owns a 32 (y) x 16 (x) x 8 (z) tile of outputs and stages the
// tile plus an H-element halo (H = largest stencil offset, <= 3) in LDS, with
// zero padding outside the domain -- padding with zeros is exactly the
// boundary truncation of pytopo3d/utils/filter.py:122-228.  Each thread owns
// one (y, x) column of 8 outputs along z.  The stencil is grouped into (dy, dx)
// columns; a column with z half-width L is read once from LDS (8 + 2L values)
// and applied to all 8 outputs from registers, which cuts the LDS reads per
// output from 93 to about 34 at rmin = 3.
#include "tk_filter_dev.hpp"
#include "tk_plan.hpp"

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <map>
#include <utility>
#include <vector>

namespace tk {

constexpr int FTY = 32, FTX = 16, FTZ = 8, FHMAX = 3;
constexpr int FTHREADS = FTY * FTX;   // one thread per (y, x) output column
constexpr int FMAXCOL = 64;       // (2*3+1)^2 = 49 columns at most
constexpr int FMAXW = 343;        // weights of all columns, (2*3+1)^3 at most

struct FilterCols {
  int H;
  int ncol[FHMAX + 1];            // columns per half-width L = 0..3 (grouped in this order)
  int dydx[FMAXCOL];              // (dy + 16) | (dx + 16) << 8
  int woff[FMAXCOL];              // offset of the column's 2L+1 weights in w
  double w[FMAXW];
};

enum { kFProject = 0, kFSum = 1, kFOcVolume = 2, kFPlain = 3 };

template <int L>
__device__ __forceinline__ void col_accum(const double* __restrict__ tile, int base, int zs,
                                          const double* __restrict__ w, double acc[FTZ]) {
  double v[FTZ + 2 * L];
#pragma unroll
  for (int k = 0; k < FTZ + 2 * L; ++k) v[k] = tile[base + k * zs];
#pragma unroll
  for (int j = 0; j <= 2 * L; ++j) {
    const double wj = w[j];
#pragma unroll
    for (int z = 0; z < FTZ; ++z) acc[z] += wj * v[z + j];
  }
}

template <int MODE, int H>
__global__ __launch_bounds__(FTHREADS) void k_filter_tiled(
    Grid g, const FilterCols* __restrict__ colsg, const double* __restrict__ src,
    const double* __restrict__ Hs, const uint8_t* __restrict__ fl, double beta, double eta,
    double tb_eta, double den, double* __restrict__ o0, double* __restrict__ o1,
    double* __restrict__ o2, double* partials, Scalars* sc) {
  if (MODE == kFOcVolume && sc->oc_done) return;
  extern __shared__ double tile[];
  __shared__ FilterCols cols;
  {
    const int* a = reinterpret_cast<const int*>(colsg);
    int* b = reinterpret_cast<int*>(&cols);
    for (int i = threadIdx.x; i < (int)(sizeof(FilterCols) / sizeof(int)); i += blockDim.x)
      b[i] = a[i];
  }
  constexpr int TYH = FTY + 2 * H, TXH = FTX + 2 * H, TZH = FTZ + 2 * H;
  constexpr int zs = TYH * TXH;
  // XCD-aware order: z, then y, then x neighbours of a tile share an L2
  const int lg = xcd_logical(block_id(), block_count());
  const int bz = lg % gridDim.z, rr = lg / gridDim.z;
  const int y0 = (rr % gridDim.x) * FTY, x0 = (rr / gridDim.x) * FTX, z0 = bz * FTZ;
  const int exy = g.nelx * g.nely;
  constexpr int ntile = zs * TZH;
  // all global loads of the tile are issued before the first LDS store, so
  // their latencies overlap instead of adding up
  constexpr int NT = (ntile + FTHREADS - 1) / FTHREADS;
  double stg[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const int t = threadIdx.x + FTHREADS * k;
    const int yy = t % TYH, r = t / TYH, xx = r % TXH, zz = r / TXH;
    const int ey = y0 - H + yy, ex = x0 - H + xx, ez = z0 - H + zz;
    stg[k] = (t < ntile && (unsigned)ey < (unsigned)g.nely && (unsigned)ex < (unsigned)g.nelx &&
              (unsigned)ez < (unsigned)g.nelz)
                 ? src[ey + ex * g.nely + ez * exy]
                 : 0.0;
  }
#pragma unroll
  for (int k = 0; k < NT; ++k)
    if (threadIdx.x + FTHREADS * k < ntile) tile[threadIdx.x + FTHREADS * k] = stg[k];
  __syncthreads();
  const int ty = threadIdx.x % FTY, tx = threadIdx.x / FTY;
  double acc[FTZ];
#pragma unroll
  for (int z = 0; z < FTZ; ++z) acc[z] = 0.0;
  int c = 0;
#define TK_COLS(LL)                                                                      \
  if (H >= LL)                                                                            \
    for (int i = 0; i < cols.ncol[LL]; ++i, ++c) {                                        \
      const int dy = (cols.dydx[c] & 255) - 16, dx = ((cols.dydx[c] >> 8) & 255) - 16;    \
      const int base = (H - LL) * zs + (tx + H + dx) * TYH + (ty + H + dy);               \
      col_accum<LL>(tile, base, zs, cols.w + cols.woff[c], acc);                          \
    }
  TK_COLS(0)
  TK_COLS(1)
  TK_COLS(2)
  TK_COLS(3)
#undef TK_COLS
  const int ey = y0 + ty, ex = x0 + tx;
  const bool col_in = ey < g.nely && ex < g.nelx;
  double part[1] = {0.0};
#pragma unroll
  for (int z = 0; z < FTZ; ++z) {
    const int ez = z0 + z;
    if (!col_in || ez >= g.nelz) continue;
    const int e = ey + ex * g.nely + ez * exy;
    if (MODE == kFPlain) {
      o0[e] = acc[z];
    } else if (MODE == kFProject) {
      const double f = acc[z] / Hs[e];
      if (o0) o0[e] = f;
      double rr, dd;
      project(f, beta, eta, tb_eta, den, fl[e], rr, dd);
      if (o1) o1[e] = rr;
      if (o2) o2[e] = dd;
    } else {
      if (fl[e] == 0) {
        double rr, dd;
        project(acc[z] / Hs[e], beta, eta, tb_eta, den, 0, rr, dd);
        part[0] += rr;
      }
    }
  }
  if (MODE == kFSum || MODE == kFOcVolume) block_partials<1>(part, partials);
}

// ---------------------------------------------------------------------------
// Host side
bool filter_tiled_setup(Plan& P, const std::vector<int>& off, const std::vector<double>& w) {
  // group entries by (dy, dx); each column must be dz = -L..L
  std::map<std::pair<int, int>, std::vector<std::pair<int, double>>> colmap;
  int H = 0;
  for (size_t k = 0; k < off.size(); ++k) {
    const int dy = (off[k] & 255) - 16, dx = ((off[k] >> 8) & 255) - 16,
              dz = ((off[k] >> 16) & 255) - 16;
    H = std::max(H, std::max(std::abs(dy), std::max(std::abs(dx), std::abs(dz))));
    colmap[{dy, dx}].push_back({dz, w[k]});
  }
  if (H > FHMAX || (int)colmap.size() > FMAXCOL) return false;
  FilterCols fc{};
  fc.H = H;
  std::vector<std::pair<std::pair<int, int>, std::vector<std::pair<int, double>>>> byL[FHMAX + 1];
  for (auto& kv : colmap) {
    auto& v = kv.second;
    std::sort(v.begin(), v.end());
    const int L = -v.front().first;
    if (v.back().first != L || (int)v.size() != 2 * L + 1) return false;
    byL[L].push_back(kv);
  }
  int c = 0, wo = 0;
  for (int L = 0; L <= FHMAX; ++L) {
    fc.ncol[L] = (int)byL[L].size();
    for (auto& kv : byL[L]) {
      fc.dydx[c] = (kv.first.first + 16) | ((kv.first.second + 16) << 8);
      fc.woff[c] = wo;
      for (auto& e : kv.second) fc.w[wo++] = e.second;
      ++c;
    }
  }
  if (hipMalloc((void**)&P.fcols, sizeof(FilterCols)) != hipSuccess) return false;
  if (hipMemcpy(P.fcols, &fc, sizeof(FilterCols), hipMemcpyHostToDevice) != hipSuccess) return false;
  P.fhalo = H;
  const size_t lds = sizeof(double) * (FTY + 2 * H) * (FTX + 2 * H) * (FTZ + 2 * H);
  // tiles above 64 KiB need the kernels' dynamic LDS limit raised
  auto allow = [&](const void* f) {
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) ==
           hipSuccess;
  };
  bool ok = true;
#define TK_ALLOW(HH)                                                                   \
  ok = ok && allow((const void*)k_filter_tiled<kFProject, HH>) &&                      \
       allow((const void*)k_filter_tiled<kFSum, HH>) &&                                \
       allow((const void*)k_filter_tiled<kFOcVolume, HH>) &&                           \
       allow((const void*)k_filter_tiled<kFPlain, HH>)
  switch (H) {
    case 0: TK_ALLOW(0); break;
    case 1: TK_ALLOW(1); break;
    case 2: TK_ALLOW(2); break;
    default: TK_ALLOW(3); break;
  }
#undef TK_ALLOW
  if (!ok) return false;
  P.flds = lds;
  P.ftiled = true;
  return true;
}

size_t filter_cols_bytes() { return sizeof(FilterCols); }

static dim3 fgrid(const Grid& g) {
  return dim3((g.nely + FTY - 1) / FTY, (g.nelx + FTX - 1) / FTX, (g.nelz + FTZ - 1) / FTZ);
}

int filter_blocks(const Plan& P) {
  const dim3 gg = fgrid(P.g);
  return (int)(gg.x * gg.y * gg.z);
}

template <int MODE>
static void launch_tiled(Plan& P, const double* src, double beta, double eta, double tb,
                         double den, double* o0, double* o1, double* o2, hipStream_t s) {
  const FilterCols* fc = (const FilterCols*)P.fcols;
  const dim3 gg = fgrid(P.g);
#define TK_LAUNCH_H(HH)                                                                      \
  hipLaunchKernelGGL((k_filter_tiled<MODE, HH>), gg, dim3(FTHREADS), P.flds, s, P.g, fc, src,    \
                     P.Hs, \
                     P.eflags, beta, eta, tb, den, o0, o1, o2, P.partials, P.sc)
  switch (P.fhalo) {
    case 0: TK_LAUNCH_H(0); break;
    case 1: TK_LAUNCH_H(1); break;
    case 2: TK_LAUNCH_H(2); break;
    default: TK_LAUNCH_H(3); break;
  }
#undef TK_LAUNCH_H
}

void tiled_filter_project(Plan& P, const double* x, double beta, double eta, double tb, double den,
                          double* rf, double* rho, double* drho, hipStream_t s) {
  launch_tiled<kFProject>(P, x, beta, eta, tb, den, rf, rho, drho, s);
}

void tiled_filter_plain(Plan& P, const double* v, double* out, hipStream_t s) {
  launch_tiled<kFPlain>(P, v, 0.0, 0.0, 0.0, 1.0, out, nullptr, nullptr, s);
}

void tiled_free_sum(Plan& P, const double* x, double beta, double eta, double tb, double den,
                    int slot, hipStream_t s) {
  launch_tiled<kFSum>(P, x, beta, eta, tb, den, nullptr, nullptr, nullptr, s);
  finalize(P, slot == 0 ? kFinSum0 : kFinSum1, filter_blocks(P), s);
}

// Volume of one bisection candidate: filtered, projected and summed over free
// elements, then kFinOcVolume.
void tiled_oc_volume(Plan& P, const double* cand, double beta, double eta, double tb, double den,
                     hipStream_t s) {
  launch_tiled<kFOcVolume>(P, cand, beta, eta, tb, den, nullptr, nullptr, nullptr, s);
  finalize(P, kFinOcVolume, filter_blocks(P), s);
}

}  // namespace tk
