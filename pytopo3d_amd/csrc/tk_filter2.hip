// LDS-tiled density filter (fast path of H x, used for the forward filter,
// its transpose and every OC bisection volume).
//
// A workgroup owns a 32 (y) x 16 (x) x 8 (z) tile of outputs and stages the
// tile plus an H-element halo (H = largest stencil offset, <= 3) in LDS, with
// zero padding outside the domain -- padding with zeros is exactly the
// boundary truncation of pytopo3d/utils/filter.py:122-228.  Each thread owns
// one (y, x) column of 8 outputs along z.  The stencil is grouped into (dy, dx)
// columns, and columns with the same z weights (for the distance-based weights
// of the reference, those with the same dy^2 + dx^2) form one group: the
// group's source columns are summed first (8 + 2L values each, read once from
// LDS), then its 2L + 1 weights are applied to the sum from registers.  At
// rmin = 3 the 93 products per output become 24 products and 25 additions.
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
  int ngrp[FHMAX + 1];            // column groups per half-width L = 0..3 (stored in this order)
  int gfirst[FMAXCOL];            // a group's first member in dydx
  int gcount[FMAXCOL];            // its member count
  int woff[FMAXCOL];              // offset of its 2L+1 weights in w
  int dydx[FMAXCOL];              // members, group by group: (dy + 16) | (dx + 16) << 8
  double w[FMAXW];
};

enum { kFProject = 0, kFSum = 1, kFOcVolume = 2, kFPlain = 3 };

// One column group: the members' source columns summed in member order, then
// the group's z weights applied to the sum.  base: the thread's tile offset
// at the first source plane; the members' (dy, dx) are uniform.
template <int L, int TYH>
__device__ __forceinline__ void grp_accum(const double* __restrict__ tile, int base, int zs,
                                          const FilterCols* __restrict__ fc, int gi,
                                          double acc[FTZ]) {
  double v[FTZ + 2 * L];
  const int m0 = fc->gfirst[gi], mn = fc->gcount[gi];
  for (int m = 0; m < mn; ++m) {
    const int dd = fc->dydx[m0 + m];
    const int off = base + ((dd >> 8) & 255) * TYH + (dd & 255) - 16 * (TYH + 1);
    if (m == 0) {
#pragma unroll
      for (int k = 0; k < FTZ + 2 * L; ++k) v[k] = tile[off + k * zs];
    } else {
#pragma unroll
      for (int k = 0; k < FTZ + 2 * L; ++k) v[k] += tile[off + k * zs];
    }
  }
  const double* w = fc->w + fc->woff[gi];
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
  constexpr int TYH = FTY + 2 * H, TXH = FTX + 2 * H, TZH = FTZ + 2 * H;
  constexpr int zs = TYH * TXH;
  // XCD-aware order: z, then y, then x neighbours of a tile share an L2
  const int lg = xcd_logical(block_id(), block_count());
  const int bz = lg % gridDim.z, rr = lg / gridDim.z;
  const int y0 = (rr % gridDim.x) * FTY, x0 = (rr / gridDim.x) * FTX, z0 = bz * FTZ;
  const int exy = g.nelx * g.nely;
  // staging: each thread takes one or two (y, x) positions of the padded tile
  // and its column of TZH planes; all global loads are issued before the
  // first LDS store, so their latencies overlap instead of adding up
  constexpr int NP = (zs + FTHREADS - 1) / FTHREADS;
  double stg[NP][TZH];
  bool zin[TZH];
#pragma unroll
  for (int zz = 0; zz < TZH; ++zz) zin[zz] = (unsigned)(z0 - H + zz) < (unsigned)g.nelz;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = threadIdx.x + FTHREADS * j;
    const int ey = y0 - H + p % TYH, ex = x0 - H + p / TYH;
    const bool in = p < zs && (unsigned)ey < (unsigned)g.nely && (unsigned)ex < (unsigned)g.nelx;
    const double* col = src + (in ? ey + ex * g.nely + (long long)(z0 - H) * exy : 0);
#pragma unroll
    for (int zz = 0; zz < TZH; ++zz)
      stg[j][zz] = (in && zin[zz]) ? col[(long long)zz * exy] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = threadIdx.x + FTHREADS * j;
    if (p < zs)
#pragma unroll
      for (int zz = 0; zz < TZH; ++zz) tile[p + zz * zs] = stg[j][zz];
  }
  __syncthreads();
  const int ty = threadIdx.x % FTY, tx = threadIdx.x / FTY;
  double acc[FTZ];
#pragma unroll
  for (int z = 0; z < FTZ; ++z) acc[z] = 0.0;
  // the table is read with uniform (scalar) loads; groups are stored by L
  int gi = 0;
#define TK_GROUPS(LL)                                                                    \
  if (H >= LL)                                                                            \
    for (int i = 0; i < colsg->ngrp[LL]; ++i, ++gi)                                       \
      grp_accum<LL, TYH>(tile, (H - LL) * zs + (tx + H) * TYH + (ty + H), zs, colsg, gi, acc);
  TK_GROUPS(0)
  TK_GROUPS(1)
  TK_GROUPS(2)
  TK_GROUPS(3)
#undef TK_GROUPS
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
  // columns with bitwise equal z weights (hence equal L) share a group
  std::map<std::vector<double>, std::vector<std::pair<int, int>>> groups[FHMAX + 1];
  for (auto& kv : colmap) {
    auto& v = kv.second;
    std::sort(v.begin(), v.end());
    const int L = -v.front().first;
    if (v.back().first != L || (int)v.size() != 2 * L + 1) return false;
    std::vector<double> wz;
    for (auto& e : v) wz.push_back(e.second);
    groups[L][wz].push_back(kv.first);
  }
  int gi = 0, mi = 0, wo = 0;
  for (int L = 0; L <= FHMAX; ++L) {
    fc.ngrp[L] = (int)groups[L].size();
    for (auto& kv : groups[L]) {
      fc.gfirst[gi] = mi;
      fc.gcount[gi] = (int)kv.second.size();
      fc.woff[gi] = wo;
      for (double x : kv.first) fc.w[wo++] = x;
      for (auto& m : kv.second) fc.dydx[mi++] = (m.first + 16) | ((m.second + 16) << 8);
      ++gi;
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
