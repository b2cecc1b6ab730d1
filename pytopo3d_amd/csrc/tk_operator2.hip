// Element-centric matrix-free H8 operator (the fast path of K u).
//
// KE of an axis-aligned orthotropic cube commutes with the three mirror
// reflections of the element (the reference checks this invariance,
// tests/test_stiffness.py:69-78).  In the Walsh-Hadamard basis of the 8 corner
// values of each displacement component the 24x24 KE therefore splits into 8
// independent 3x3 blocks, one per parity sector chi: the members of sector chi
// are (component c, mode m = chi XOR e_c).  Orthotropy adds zeros: the three
// rigid translations (mode 0 of each component, in sectors 1, 2, 4) are null,
// and in sectors 3, 5, 6 one component decouples, leaving 45 non-zero entries
// (checked on the host, B = W^T KE W / 64).
//
// The Walsh-Hadamard transform factors into an in-plane (x, y) transform of
// each 2x2 face and a z butterfly between the two faces of an element.  A
// thread owns one element column and streams along z, so
//   - the forward face transform of node plane k+1 (top face of element layer
//     k) is reused as the bottom face of layer k+1, and
//   - the top-face spectrum of layer k and the bottom-face spectrum of layer
//     k+1 land on the same nodes, so they are summed before a single inverse
//     face transform per node plane.
// Per element layer: 24 (face forward) + 21 (z forward) + 45 FMA + 21 scale +
// 21 (z inverse) + 12 (face sum) + 24 (face inverse) = 168 operations, against
// 1152 for the dense 24x24 product.
//
// Scatter without atomics: a workgroup owns a 15 x 15 column of nodes in (y, x)
// and streams along z.  Its 16 x 16 threads evaluate the 16 x 16 elements
// touching those nodes (one halo row and column is re-evaluated by the
// neighbouring tile), exchange the three off-diagonal corner contributions of
// each node plane through LDS, and every node thread sums its 4 contributions
// in a fixed order, so results are deterministic.
#include "tk_galerkin_pattern.hpp"
#include "tk_plan.hpp"
#include "tk_sector.hpp"

#include <type_traits>
#include <vector>

namespace tk {

template <typename T>
struct Epilogue {
  T* y;            // MODE 0/1 output
  const T* b;      // MODE 1..4 right-hand side
  T* xo;           // MODE 2..4 updated iterate
  T* d;            // MODE 2..4 new Chebyshev direction (nullptr: not stored, last step)
  const T* dp;     // MODE 3 previous Chebyshev direction
  const T* dinv;   // MODE 4 inverse diagonal (MODE 2, 3, 6 form it from kdi and the element
                   // scales: KE's diagonal is the same at every corner)
  T c0, c1, c2;
  T kdi[3];        // reciprocal KE diagonal per component
  // MODE 6: coarse-grid correction, prolongated on the fly
  const T* xc;
  Grid cg;
  // MODE 5 (flexible CG direction): p = zs z32 + beta p32o, stored as p32n
  const float* z32;
  const float* p32o;
  float* p32n;
  const Scalars* sc;
  // MODE 5 with a direction ring (ring != nullptr): p32n is slot iter % ring_m
  // and p32o (when p32o != nullptr, used only as that flag) slot
  // (iter - 1) % ring_m, iter = sc->iter, slots ring_stride floats apart
  float* ring;
  long long ring_stride;
  int ring_m;
};

// kModeChebZero fuses the first two Chebyshev steps from x = 0: the node
// values are formed on the fly as x1 = c0 D^-1 b (= d1), then
// d2 = c1 d1 + c2 D^-1 (b - K x1), x2 = x1 + d2.
// kModeApplyDir (FP64 only) fuses the flexible-CG direction update into the
// operator: the node operand is p = zs z32 + beta p32o rounded to FP32 (the
// direction is stored in FP32; K p and p.Kp are FP64), written to p32n by the
// node owners; y = K p and the partials of p.Kp as in kModeApply.
enum {
  kModeApply = 0,
  kModeResidual = 1,
  kModeChebFirst = 2,
  kModeChebStep = 3,
  kModeChebZero = 4,
  kModeApplyDir = 5,
  kModeChebFirstP = 6
};
// kModeChebFirstP is kModeChebFirst on x + P xc (the coarse-grid correction
// prolongated on the fly while fetching the node values; it is added on free
// DOFs only), so the post-smoothing needs no separate prolongation sweep.

constexpr int kTile = 16;          // threads per tile edge
constexpr int kOut = kTile - 1;    // owned nodes per tile edge
constexpr int kPl = kTile + 1;     // node-plane edge held in LDS

// BF16 MFMA form of the FP32 sweeps (operator_kind 2).  Per element layer a
// wave evaluates its 64 elements as two 32-element blocks of the product
// Y = KE' U on v_mfma_f32_32x32x16_bf16 (two K steps each), FP32 accumulation:
//   rows/columns of KE' = 4 m + r, corner m = 4 oz + 2 oy + ox, component r
//   (r = 3 is a zero pad), so a K step of 16 holds one face (4 corners) and
//   lane half h = lane / 32 supplies the corners with oy = h of its element
//   column (two nodes, from the LDS node plane), and receives back the outputs
//   of the corners with ox = h (C rows 8 g + 4 h + r: corner 2 g + h).
// KE' is staged per lane in the order the A operand wants (kemf: 64 lanes x
// 2 steps x 8 bf16).  The outputs are scaled by the element scale in FP32 and
// summed per node through LDS in a fixed order.
typedef __bf16 mf_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mf_f32x16 __attribute__((ext_vector_type(16)));
typedef float mf_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mf_bf16x8 mf_pack(const float a[3], const float b[3]) {
  mf_bf16x8 v;
  v[0] = (__bf16)a[0]; v[1] = (__bf16)a[1]; v[2] = (__bf16)a[2]; v[3] = (__bf16)0.0f;
  v[4] = (__bf16)b[0]; v[5] = (__bf16)b[1]; v[6] = (__bf16)b[2]; v[7] = (__bf16)0.0f;
  return v;
}

// Registers of one node plane.  `fetch` only issues loads into them; the
// operand is formed and masked in `store`, a layer later, so no wave waits on
// loads that were just issued (the wait counter is in order: one early wait
// would also wait for every load issued before it).
template <typename T, int MODE>
struct PlaneRegs {
  // kModeApplyDir: the direction is an FP32 value (stored in FP32), so it is
  // held in FP32 and widened where it enters the FP64 operator
  using OwnT = std::conditional_t<MODE == kModeApplyDir, float, T>;
  OwnT own[3];   // own node operand, unmasked (formed in `store`; kModeChebZero
                 // keeps b and D^-1 instead and forms it again in the epilogue)
  uint8_t fix;   // own node fixed bits (7 outside the domain)
  T ext[3];      // extra halo node operand (kModeChebFirstP: formed and masked)
  uint8_t xfix;  // extra node fixed bits (7 outside the domain)
  T ob[3];       // kModeChebZero: own b and D^-1 (reused by the epilogue)
  T odi[3];
  T xb[3];       // kModeChebZero: extra node b and D^-1
  T xdi[3];
  float oz[3], op[3];   // kModeApplyDir: own z32 and previous direction
  float xz[3], xp[3];   // ... of the extra node
};

// Loads of element layer ez: its element scale and the epilogue operands of
// node plane ez (whose nodes the layer completes).  They are issued one layer
// ahead, like the node planes, so no layer waits on its own loads.
template <typename T>
struct LayerIn {
  T se;
  T eb[3], ed[3];
  T sh[8];   // Galerkin level: the element's Walsh coefficients
  T ev[3];   // Galerkin level: the node's inverse diagonal (Chebyshev sweeps)
};

// GAL: the Galerkin level (tk_galerkin.hip): s holds 8 Walsh coefficients per
// element, Kg the coupling table, and the Chebyshev sweeps read D^-1 from
// ep.dinv (the diagonal is no longer KE's times the scale sum).
template <typename T, int MODE, bool MF = false, bool GAL = false>
__device__ __forceinline__ void apply_el(const Grid& g, const SparseKE<float>& Kv,
                                         const T* __restrict__ Kg, const T* __restrict__ s,
                                         const uint8_t* __restrict__ fixed,
                                         const T* __restrict__ u, const Epilogue<T>& ep, int nty,
                                         int zc, int with_dot, const int* __restrict__ done,
                                         double* partials, const uint4* __restrict__ kemf) {
  static_assert(!MF || sizeof(T) == 4, "the MFMA form is an FP32 sweep");
  static_assert(!(MF && GAL), "the Galerkin level uses the sector-form sweep");
  __shared__ T un[2][3][kPl * kPl];       // two node planes of the tile (masked operand)
  // exchanged corner contributions: sector form (f = 1, 2, 3), or the MFMA
  // form's per-element corner outputs [oz][oy][ox][component]
  __shared__ T yo[MF ? 24 : 9][kTile * kTile];
  __shared__ T kb[kSecN];                 // FP64 sector entries
  __shared__ T ss[kTile * kTile];         // element scales of the layer (on-the-fly diagonal)
  __shared__ T gl[GAL ? kGalN : 1];       // Galerkin coupling table
  constexpr bool kCheb = MODE == kModeChebFirst || MODE == kModeChebStep || MODE == kModeChebFirstP;
  // Chebyshev sweeps form D^-1 from the element scales (the element kernel
  // is used only when KE's diagonal is the same at every corner)
  constexpr bool diag_fly = kCheb && !GAL;

  const int tid = threadIdx.x;
  // FP64 entries are staged in LDS (90 SGPRs would spill); FP32 entries stay
  // in scalar registers.
  if constexpr (GAL) {
    for (int i = tid; i < kGalN; i += blockDim.x) gl[i] = Kg[i];
  } else {
    if (sizeof(T) == 8 && tid < kSecN) kb[tid] = Kg[tid];
  }
  const int ly = tid & (kTile - 1), lx = tid >> 4;
  // XCD-aware order: z-chunks of a tile, then y, then x neighbours share an L2
  const int lg = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int chunk = lg % gridDim.y, tile = lg / gridDim.y;
  const int Y0 = (tile % nty) * kOut, X0 = (tile / nty) * kOut;
  const int Z0 = chunk * zc;
  const int Z1 = min(Z0 + zc, g.nnz);
  const int ey = Y0 - 1 + ly, ex = X0 - 1 + lx;   // element == owned node (tile coords)
  const bool col_el = (unsigned)ey < (unsigned)g.nely && (unsigned)ex < (unsigned)g.nelx;
  const bool col_node = (unsigned)ey < (unsigned)g.nny && (unsigned)ex < (unsigned)g.nnx;
  const bool owner = ly >= 1 && lx >= 1 && col_node;
  const int nxy = g.nnx * g.nny;
  const int exy = g.nelx * g.nely;

  // extra plane nodes loaded by threads 0..32: row i = 16 (j = 0..16), column j = 16 (i = 0..15)
  int xi = -1, xj = -1;
  if (tid < kPl) { xi = kTile; xj = tid; }
  else if (tid < kPl + kTile) { xi = tid - kPl; xj = kTile; }
  const int xny = Y0 - 1 + xi, xnx = X0 - 1 + xj;
  const bool ext_node = xi >= 0 && (unsigned)xny < (unsigned)g.nny && (unsigned)xnx < (unsigned)g.nnx;

  // kModeApplyDir: direction value at DOF 3n + c, formed in FP32 (the
  // direction is stored in FP32 anyway; K p is then taken in T).  Without a
  // previous direction the second operand is z32 itself with a zero
  // coefficient (finite whenever z32 is), so the loads stay unconditional.
  float dir_zs = 0.0f, dir_beta = 0.0f;
  const float* __restrict__ dir_p = ep.p32o ? ep.p32o : ep.z32;
  float* dir_out = ep.p32n;
  if constexpr (MODE == kModeApplyDir) {
    dir_zs = (float)ep.sc->zs;
    dir_beta = ep.p32o ? (float)ep.sc->beta : 0.0f;
    if (ep.ring) {
      const int it = ep.sc->iter, m = ep.ring_m;
      dir_out = ep.ring + (long long)(it % m) * ep.ring_stride;
      if (ep.p32o) dir_p = ep.ring + (long long)((it + m - 1) % m) * ep.ring_stride;
    }
  }
  // one explicit rounding order: the owner and the neighbouring tiles (halo
  // node) must form bitwise the same direction, or K p no longer matches the
  // stored p and the FP64 residual drifts from the true one; left to
  // contraction, the two call sites were fused differently
  auto dir_value = [&](float z, float p) -> float { return __builtin_fmaf(dir_zs, z, dir_beta * p); };
  // kModeChebFirstP: (P xc) at fine node (iy, ix, iz), trilinear weights (an
  // even fine index has one coarse parent per axis: weight 1, duplicate 0)
  auto prol = [&](int iy, int ix, int iz, T out[3]) {
    const Grid& c = ep.cg;
    const int y0 = iy >> 1, y1 = (iy + 1) >> 1, x0 = ix >> 1, x1 = (ix + 1) >> 1;
    const int z0 = iz >> 1, z1 = (iz + 1) >> 1;
    const T wy[2] = {(iy & 1) ? T(0.5) : T(1), (iy & 1) ? T(0.5) : T(0)};
    const T wx[2] = {(ix & 1) ? T(0.5) : T(1), (ix & 1) ? T(0.5) : T(0)};
    const T wz[2] = {(iz & 1) ? T(0.5) : T(1), (iz & 1) ? T(0.5) : T(0)};
    const int cxy = c.nnx * c.nny;
    out[0] = out[1] = out[2] = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int ky = k & 1, kx = (k >> 1) & 1, kz = k >> 2;
      const int I = (ky ? y1 : y0) + (kx ? x1 : x0) * c.nny + (kz ? z1 : z0) * cxy;
      const T w = wy[ky] * wx[kx] * wz[kz];
      out[0] += w * ep.xc[3 * I];
      out[1] += w * ep.xc[3 * I + 1];
      out[2] += w * ep.xc[3 * I + 2];
    }
  };
  // global loads of node plane pz (registers only; see PlaneRegs)
  auto fetch = [&](int pz, PlaneRegs<T, MODE>& R) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R.own[c] = 0;
      R.ob[c] = R.odi[c] = T(0);
      if constexpr (MODE == kModeApplyDir) R.oz[c] = R.op[c] = 0.0f;
    }
    R.fix = 7;
    if (col_node && (unsigned)pz < (unsigned)g.nnz) {
      const int n = ey + ex * g.nny + pz * nxy;
      R.fix = fixed[n];
      T pc[3] = {T(0), T(0), T(0)};
      if constexpr (MODE == kModeChebFirstP) prol(ey, ex, pz, pc);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if constexpr (MODE == kModeChebZero) {
          R.ob[c] = ep.b[3 * n + c];
          R.odi[c] = ep.dinv[3 * n + c];
        } else if constexpr (MODE == kModeApplyDir) {
          R.oz[c] = ep.z32[3 * n + c];
          R.op[c] = dir_p[3 * n + c];
        } else if constexpr (MODE == kModeChebFirstP) {
          R.own[c] = u[3 * n + c] + (((R.fix >> c) & 1) ? T(0) : pc[c]);
        } else {
          R.own[c] = u[3 * n + c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R.ext[c] = R.xb[c] = R.xdi[c] = T(0);
      if constexpr (MODE == kModeApplyDir) R.xz[c] = R.xp[c] = 0.0f;
    }
    R.xfix = 7;
    if (ext_node && (unsigned)pz < (unsigned)g.nnz) {
      const int n = xny + xnx * g.nny + pz * nxy;
      R.xfix = fixed[n];
      if constexpr (MODE == kModeChebFirstP) {
        // the gathered coarse values are summed at once (24 more registers
        // otherwise); this sweep runs on the small levels only
        T pc[3];
        prol(xny, xnx, pz, pc);
#pragma unroll
        for (int c = 0; c < 3; ++c) R.ext[c] = u[3 * n + c] + pc[c];
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if constexpr (MODE == kModeChebZero) {
            R.xb[c] = ep.b[3 * n + c];
            R.xdi[c] = ep.dinv[3 * n + c];
          } else if constexpr (MODE == kModeApplyDir) {
            R.xz[c] = ep.z32[3 * n + c];
            R.xp[c] = dir_p[3 * n + c];
          } else {
            R.ext[c] = u[3 * n + c];
          }
        }
      }
    }
  };
  // forms the plane's operands (kept in R.own for the epilogue) and writes the
  // masked values to LDS
  auto store = [&](int buf, PlaneRegs<T, MODE>& R) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T v;
      if constexpr (MODE == kModeChebZero) {
        v = ep.c0 * R.odi[c] * R.ob[c];
      } else {
        if constexpr (MODE == kModeApplyDir) R.own[c] = dir_value(R.oz[c], R.op[c]);
        v = (T)R.own[c];
      }
      un[buf][c][ly + kPl * lx] = ((R.fix >> c) & 1) ? T(0) : v;
    }
    if (xi >= 0)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T v = R.ext[c];
        if constexpr (MODE == kModeChebZero) v = ep.c0 * R.xdi[c] * R.xb[c];
        if constexpr (MODE == kModeApplyDir) v = (T)dir_value(R.xz[c], R.xp[c]);
        un[buf][c][xi + kPl * xj] = ((R.xfix >> c) & 1) ? T(0) : v;
      }
  };
  // face values of node plane buf at this thread's element column, then the
  // in-plane transform
  auto face_spectrum = [&](int buf, T F[3][4]) {
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int ox = f & 1, oy = f >> 1;
#pragma unroll
      for (int c = 0; c < 3; ++c) F[c][f] = un[buf][c][(ly + oy) + kPl * (lx + ox)];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) face_fwd(F[c]);
  };

  // element scale and epilogue operands of layer ez (see LayerIn)
  constexpr bool kNeedB = MODE == kModeResidual || MODE == kModeChebFirst ||
                          MODE == kModeChebStep || MODE == kModeChebFirstP;
  auto fetch_layer = [&](int ez, LayerIn<T>& L) {
    const bool el = col_el && (unsigned)ez < (unsigned)g.nelz;
    if constexpr (GAL) {
      const T* q = s + 8 * (size_t)(ey + ex * g.nely + ez * exy);
#pragma unroll
      for (int w = 0; w < 8; ++w) L.sh[w] = el ? q[w] : T(0);
      L.se = T(0);
    } else {
      L.se = el ? s[ey + ex * g.nely + ez * exy] : T(0);   // 0 gives zero contributions
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) L.eb[c] = L.ed[c] = L.ev[c] = T(0);
    if constexpr (kNeedB) {
      if (owner && ez >= Z0) {
        const int n_out = ey + ex * g.nny + ez * nxy;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          L.eb[c] = ep.b[3 * n_out + c];
          if (MODE == kModeChebStep) L.ed[c] = ep.dp[3 * n_out + c];
          if constexpr (GAL && kCheb) L.ev[c] = ep.dinv[3 * n_out + c];
        }
      }
    }
  };

  double dot = 0.0;
  PlaneRegs<T, MODE> RA, RB;
  LayerIn<T> LA, LB;
  fetch(Z0 - 1, RB);
  fetch(Z0, RA);
  fetch_layer(Z0 - 1, LB);
  // the convergence flag is read with the first loads in flight
  if (done && *done) return;
  store(0, RB);
  store(1, RA);
  // own node of the plane being finished (the bottom plane of the layer)
  typename PlaneRegs<T, MODE>::OwnT raw_prev[3] = {RB.own[0], RB.own[1], RB.own[2]};
  T pb_prev[3] = {RB.ob[0], RB.ob[1], RB.ob[2]};
  T pdi_prev[3] = {RB.odi[0], RB.odi[1], RB.odi[2]};
  uint8_t fix_prev = RB.fix;
  __syncthreads();
  T F0[3][4], F1[3][4];   // face spectra of two consecutive node planes
  T Gp[3][4];             // top-face output spectrum of the previous element layer
  T s_prev = T(0);        // on-the-fly diagonal: scale sum of the previous layer
  if constexpr (!MF) face_spectrum(0, F0);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int f = 0; f < 4; ++f) Gp[c][f] = T(0);
  // MFMA form: this lane supplies K half mh of element columns eA (block A)
  // and eB (block B); p* are the LDS plane offsets of node (ly_e + mh, lx_e)
  const int mlane = tid & 63, mh = mlane >> 5;
  const int eA = (tid & ~63) + (mlane & 31), eB = eA + 32;
  const int pA = (eA & (kTile - 1)) + mh + kPl * (eA >> 4);
  const int pB = (eB & (kTile - 1)) + mh + kPl * (eB >> 4);
  mf_bf16x8 ka0{}, ka1{}, bA{}, bB{};
  T top_prev[3] = {T(0), T(0), T(0)};   // top-corner outputs of the previous layer at the own node
  auto mf_face = [&](int b, int p) -> mf_bf16x8 {
    const float a0[3] = {(float)un[b][0][p], (float)un[b][1][p], (float)un[b][2][p]};
    const float a1[3] = {(float)un[b][0][p + kPl], (float)un[b][1][p + kPl],
                         (float)un[b][2][p + kPl]};
    return mf_pack(a0, a1);
  };
  if constexpr (MF) {
    const uint4 q0 = kemf[2 * mlane], q1 = kemf[2 * mlane + 1];
    __builtin_memcpy(&ka0, &q0, sizeof(q0));
    __builtin_memcpy(&ka1, &q1, sizeof(q1));
    bA = mf_face(0, pA);
    bB = mf_face(0, pB);
  }

  // Element layer ez.  Fb: spectrum of plane ez (bottom), Ft: filled with that
  // of plane ez + 1 (LDS buffer `buf`); nxt holds plane ez + 1 (registers),
  // pre receives plane ez + 2.  Consecutive layers swap the roles of the
  // buffer pairs, so the z-loop is unrolled by two and nothing is copied.
  auto layer = [&](int ez, int buf, T (&Fb)[3][4], T (&Ft)[3][4], const PlaneRegs<T, MODE>& nxt,
                   PlaneRegs<T, MODE>& pre, const LayerIn<T>& lin, LayerIn<T>& nx) {
    // the next plane and the next layer's scale and epilogue operands; this
    // layer's were loaded one layer ago (a load issued and used in the same
    // layer would, through the in-order load counter, wait for these too)
    const bool more = ez + 1 < Z1;
    if (more) {
      fetch(ez + 2, pre);
      fetch_layer(ez + 1, nx);
    }
    const bool out_now = owner && ez >= Z0;
    const int n_out = ey + ex * g.nny + ez * nxy;
    T eb[3], ed[3], edi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      eb[c] = (MODE == kModeChebZero) ? pb_prev[c] : lin.eb[c];
      edi[c] = (MODE == kModeChebZero) ? pdi_prev[c] : (GAL ? lin.ev[c] : T(0));   // else formed below
      ed[c] = lin.ed[c];
    }
    const T se = lin.se;

    // forward transform, sector blocks, inverse z butterfly.  FP64 operator
    // sweeps (3 waves per SIMD): the bottom face spectrum is formed again from
    // its plane, still in LDS, instead of being carried from the previous
    // layer (12 doubles fewer live across the sector blocks, which spilled);
    // the FP64 smoothing sweeps run at 2 waves without spilling and carry it
    T own0[3] = {T(0), T(0), T(0)};   // sector form: the own corner's output (bottom + previous top)
    if constexpr (MF) {
      // the two blocks' products on the matrix cores, the top face read now
      // and the bottom face carried from the previous layer
      const mf_bf16x8 tA = mf_face(buf, pA), tB = mf_face(buf, pB);
      mf_f32x16 yA{}, yB{};
      yA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ka0, bA, yA, 0, 0, 0);
      yB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ka0, bB, yB, 0, 0, 0);
      yA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ka1, tA, yA, 0, 0, 0);
      yB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ka1, tB, yB, 0, 0, 0);
      bA = tA;
      bB = tB;
      // element scales of the two columns (the partner lane holds the other)
      const T sp = __shfl_xor(se, 32, 64);
      const T sA = mh ? sp : se, sB = mh ? se : sp;
      // C row 8 g + 4 mh + r is component r of corner m = 2 g + mh
#pragma unroll
      for (int gq = 0; gq < 4; ++gq)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          yo[(2 * gq + mh) * 3 + r][eA] = sA * yA[4 * gq + r];
          yo[(2 * gq + mh) * 3 + r][eB] = sB * yB[4 * gq + r];
        }
    } else {
      if constexpr (sizeof(T) == 8 && (MODE == kModeApply || MODE == kModeApplyDir))
        face_spectrum(buf ^ 1, Fb);
      face_spectrum(buf, Ft);
      T a[3][8], o[3][8];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int f = 1; f < 4; ++f) a[c][f] = Fb[c][f] + Ft[c][f];
#pragma unroll
        for (int f = 0; f < 4; ++f) a[c][f + 4] = Ft[c][f] - Fb[c][f];
      }
      if constexpr (GAL) galerkin_apply<T>(gl, lin.sh, a, o);
      else if constexpr (sizeof(T) == 8) sector_apply<T>(kb, se, a, o);
      else sector_apply<T>(Kv.v, se, a, o);
      // bottom plane: sum with the previous layer's top, inverse face transform
      T nb[3][4];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        nb[c][0] = Gp[c][0] - o[c][4];
        Gp[c][0] = o[c][4];
#pragma unroll
        for (int f = 1; f < 4; ++f) {
          nb[c][f] = Gp[c][f] + (o[c][f] - o[c][f + 4]);
          Gp[c][f] = o[c][f] + o[c][f + 4];
        }
        face_inv(nb[c]);
        own0[c] = nb[c][0];
      }
#pragma unroll
      for (int f = 1; f < 4; ++f)
#pragma unroll
        for (int c = 0; c < 3; ++c) yo[(f - 1) * 3 + c][tid] = nb[c][f];
    }
    if (diag_fly) ss[tid] = se;
    __syncthreads();
    // MFMA form: node (ly, lx) of plane ez sums the bottom corners m = 2 oy + ox
    // of elements (ly - oy, lx - ox) of this layer and the top corners of the
    // previous layer's (carried), and carries this layer's top corners
    T kmf[3] = {T(0), T(0), T(0)};
    if constexpr (MF) {
      if (ly >= 1 && lx >= 1) {
        const int t1 = tid - kTile, t2 = tid - 1, t3 = tid - 1 - kTile;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          kmf[c] = top_prev[c] + yo[c][tid] + yo[3 + c][t1] + yo[6 + c][t2] + yo[9 + c][t3];
          top_prev[c] = yo[12 + c][tid] + yo[15 + c][t1] + yo[18 + c][t2] + yo[21 + c][t3];
        }
      }
    }
    if (diag_fly && owner) {
      // diag(K) at node (ly, lx) of plane ez: kd[c] times the scales of the 8
      // elements around it (4 in layer ez - 1, carried in s_prev, 4 in layer ez)
      const T scur = se + ss[tid - 1] + ss[tid - kTile] + ss[tid - kTile - 1];
      const T dsum = s_prev + scur;
      s_prev = scur;
      const T inv = recip(dsum);
#pragma unroll
      for (int c = 0; c < 3; ++c) edi[c] = ((fix_prev >> c) & 1) ? T(1) : inv * ep.kdi[c];
    }
    // node (ly, lx) of plane ez: own corner f = 0, element (ly-1, lx) its f = 2
    // (oy = 1), element (ly, lx-1) its f = 1, element (ly-1, lx-1) its f = 3
    if (out_now) {
      const int tym = tid - 1, txm = tid - kTile, tbm = tid - kTile - 1;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T ku = MF ? kmf[c] : own0[c] + yo[3 + c][tym] + yo[c][txm] + yo[6 + c][tbm];
        const int i = 3 * n_out + c;
        const bool fx = (fix_prev >> c) & 1;
        // the node operand (kModeChebZero: x1 = c0 D^-1 b, formed again as in `store`)
        const T xn = (MODE == kModeChebZero) ? ep.c0 * edi[c] * eb[c] : (T)raw_prev[c];
        const T val = fx ? xn : ku;
        if (MODE == kModeApply || MODE == kModeApplyDir) {
          ep.y[i] = val;
          dot += (double)xn * (double)val;
          if (MODE == kModeApplyDir) dir_out[i] = (float)raw_prev[c];
        } else if (MODE == kModeResidual) {
          // zero on fixed DOFs: the restriction reads no fine mask
          ep.y[i] = fx ? T(0) : eb[c] - val;
        } else if (MODE == kModeChebFirst || MODE == kModeChebFirstP) {
          const T dd = ep.c0 * edi[c] * (eb[c] - val);
          if (ep.d) ep.d[i] = dd;
          ep.xo[i] = xn + dd;
        } else if (MODE == kModeChebStep) {
          const T dd = ep.c1 * ed[c] + ep.c2 * edi[c] * (eb[c] - val);
          if (ep.d) ep.d[i] = dd;
          const T xnew = xn + dd;
          ep.xo[i] = xnew;
          if (with_dot) dot += (double)eb[c] * (double)xnew;   // r.z of the Krylov loop
        } else {   // kModeChebZero: d1 = x1
          const T dd = ep.c1 * xn + ep.c2 * edi[c] * (eb[c] - val);
          if (ep.d) ep.d[i] = dd;
          ep.xo[i] = xn + dd;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if constexpr (MODE != kModeChebZero) raw_prev[c] = nxt.own[c];
      pb_prev[c] = nxt.ob[c];
      pdi_prev[c] = nxt.odi[c];
    }
    fix_prev = nxt.fix;
    // unconditional: after the last layer `pre` still holds an earlier plane
    // and the LDS buffer is not read again; a conditional store would leave a
    // path without the load wait, and the compiler then waits for all loads
    // in flight before the next layer's first use of a prefetched value
    store(buf ^ 1, pre);
    __syncthreads();
  };

  for (int ez = Z0 - 1; ez < Z1; ez += 2) {
    layer(ez, 1, F0, F1, RA, RB, LB, LA);
    if (ez + 1 >= Z1) break;
    layer(ez + 1, 0, F1, F0, RB, RA, LA, LB);
  }
  if ((MODE == kModeApply || MODE == kModeApplyDir || MODE == kModeChebStep) && with_dot) {
    double v[1] = {dot};
    block_partials<1>(v, partials);
  }
}

// Two entry points with one body: the operator-only sweeps (FP32 and FP64 K u,
// FP64 K p with the direction update) are compiled for 3 waves per SIMD, the
// epilogue-heavy FP64 smoothing sweeps keep the compiler's register budget
// (3 waves would spill them).
#define TK_EL_ARGS                                                                              \
  Grid g, SparseKE<float> Kv, const T* __restrict__ Kg, const T* __restrict__ s,                \
      const uint8_t* __restrict__ fixed, const T* __restrict__ u, Epilogue<T> ep, int nty, int zc, \
      int with_dot, const int* __restrict__ done, double* partials, const uint4* __restrict__ kemf
template <typename T, int MODE, bool MF = false, bool GAL = false>
__global__ __launch_bounds__(256) void k_apply_el(TK_EL_ARGS) {
  apply_el<T, MODE, MF, GAL>(g, Kv, Kg, s, fixed, u, ep, nty, zc, with_dot, done, partials, kemf);
}
template <typename T, int MODE, bool MF = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_apply_el_w3(
    TK_EL_ARGS) {
  apply_el<T, MODE, MF>(g, Kv, Kg, s, fixed, u, ep, nty, zc, with_dot, done, partials, kemf);
}
#undef TK_EL_ARGS

// ---------------------------------------------------------------------------
// Host: sector decomposition of KE.  Returns false when KE is not
// reflection-symmetric with the orthotropic zero pattern (then the dense node
// kernel is used).
static bool build_sectors(const double* KE, double out[kSecN]) {
  // W[(a,c),(m,c)] = prod_d sigma_{a,d}^{m_d}, sigma = -1 (offset 0) / +1 (offset 1)
  auto psi = [](int a, int m) {
    const int o[3] = {corner_x(a), corner_y(a), corner_z(a)};
    double v = 1.0;
    for (int d = 0; d < 3; ++d)
      if ((m >> d) & 1) v *= o[d] ? 1.0 : -1.0;
    return v;
  };
  // B[(m,c),(m',c')] = sum_{a,b} psi(a,m) KE[(a,c),(b,c')] psi(b,m') / 64
  double B[24][24];
  double bmax = 0.0;
  for (int m = 0; m < 8; ++m)
    for (int c = 0; c < 3; ++c)
      for (int m2 = 0; m2 < 8; ++m2)
        for (int c2 = 0; c2 < 3; ++c2) {
          double acc = 0.0;
          for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b)
              acc += psi(a, m) * KE[(3 * a + c) * 24 + 3 * b + c2] * psi(b, m2);
          B[3 * m + c][3 * m2 + c2] = acc / 64.0;
          bmax = fmax(bmax, fabs(acc / 64.0));
        }
  // the 45 kept entries, as (chi, c, c') in the kernel's order
  static const int keep[kSecN][3] = {
      {0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 0}, {0, 1, 1}, {0, 1, 2}, {0, 2, 0}, {0, 2, 1},
      {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 1}, {1, 2, 2}, {2, 0, 0}, {2, 0, 2}, {2, 2, 0},
      {2, 2, 2}, {3, 0, 0}, {3, 0, 1}, {3, 1, 0}, {3, 1, 1}, {3, 2, 2}, {4, 0, 0}, {4, 0, 1},
      {4, 1, 0}, {4, 1, 1}, {5, 0, 0}, {5, 0, 2}, {5, 2, 0}, {5, 2, 2}, {5, 1, 1}, {6, 1, 1},
      {6, 1, 2}, {6, 2, 1}, {6, 2, 2}, {6, 0, 0}, {7, 0, 0}, {7, 0, 1}, {7, 0, 2}, {7, 1, 0},
      {7, 1, 1}, {7, 1, 2}, {7, 2, 0}, {7, 2, 1}, {7, 2, 2}};
  bool kept[24][24] = {};
  for (int k = 0; k < kSecN; ++k) {
    const int chi = keep[k][0], c = keep[k][1], c2 = keep[k][2];
    const int m = chi ^ (1 << c), m2 = chi ^ (1 << c2);
    out[k] = B[3 * m + c][3 * m2 + c2];
    kept[3 * m + c][3 * m2 + c2] = true;
  }
  double off = 0.0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j)
      if (!kept[i][j]) off = fmax(off, fabs(B[i][j]));
  return off <= 1e-12 * bmax;
}

// The 45 sector entries of KE (host), false when KE lacks the element's
// mirror symmetry.
bool sector_form(const double* KE, double out[kSecN]) { return build_sectors(KE, out); }

bool operator_setup(Plan& P) {
  double sec[kSecN];
  P.sector_ok = build_sectors(P.KE, sec);
  // KE diagonal: the same at every corner for each component?
  double dmax = 0.0, dev = 0.0;
  for (int c = 0; c < 3; ++c) P.kd[c] = P.KE[c * 24 + c];
  for (int a = 0; a < 8; ++a)
    for (int c = 0; c < 3; ++c) {
      const double v = P.KE[(3 * a + c) * 24 + 3 * a + c];
      dmax = fmax(dmax, fabs(v));
      dev = fmax(dev, fabs(v - P.kd[c]));
    }
  P.kd_ok = dev <= 1e-13 * dmax && dmax > 0.0;
  for (int k = 0; k < kSecN; ++k) {
    P.sec64[k] = sec[k];
    P.sec32[k] = (float)sec[k];
  }
  // MFMA form: KE' (rows / columns 4 m + r, corner m = 4 oz + 2 oy + ox, r = 3 a
  // zero pad) in BF16, laid out per lane as the A operand of the 32x32x16
  // product takes it: lane l, K step t holds row l % 32, columns 16 t + 8 (l / 32) + 0..7
  {
    int ref_corner[8];
    for (int a = 0; a < 8; ++a)
      ref_corner[4 * corner_z(a) + 2 * corner_y(a) + corner_x(a)] = a;
    auto kep = [&](int i, int k) -> double {
      const int mi = i >> 2, ri = i & 3, mk = k >> 2, rk = k & 3;
      if (ri == 3 || rk == 3) return 0.0;
      return P.KE[(3 * ref_corner[mi] + ri) * 24 + 3 * ref_corner[mk] + rk];
    };
    std::vector<__bf16> lanes(64 * 2 * 8);
    for (int l = 0; l < 64; ++l)
      for (int t = 0; t < 2; ++t)
        for (int q = 0; q < 8; ++q)
          lanes[(l * 2 + t) * 8 + q] = (__bf16)(float)kep(l % 32, 16 * t + 8 * (l / 32) + q);
    if (hipMalloc(&P.kemf, lanes.size() * sizeof(__bf16)) != hipSuccess ||
        hipMemcpy(P.kemf, lanes.data(), lanes.size() * sizeof(__bf16), hipMemcpyHostToDevice) !=
            hipSuccess)
      P.kemf = nullptr;
  }
  if (hipMalloc((void**)&P.sec64_dev, sizeof(double) * kSecN) != hipSuccess ||
      hipMalloc((void**)&P.sec32_dev, sizeof(float) * kSecN) != hipSuccess ||
      hipMemcpy(P.sec64_dev, P.sec64, sizeof(double) * kSecN, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P.sec32_dev, P.sec32, sizeof(float) * kSecN, hipMemcpyHostToDevice) != hipSuccess)
    P.sector_ok = false;
  return P.sector_ok;
}

template <typename T> static const T* sectors_of(const Plan& P);
template <> const double* sectors_of<double>(const Plan& P) { return P.sec64_dev; }
template <> const float* sectors_of<float>(const Plan& P) { return P.sec32_dev; }

// Launch shape of an element sweep: 15x15-node column tiles in x-y, cut into
// z chunks.  A workgroup streams zc node planes and re-evaluates one element
// layer, so its time grows as zc + 1, and the workgroups run in rounds of
// `slots` (compute units x resident workgroups per CU of the kernel).  The
// chunk count minimises rounds x (zc + 1): a last round that is only partly
// filled costs as much as a full one.  On the fine grid this gives one or two
// full rounds of 11-22 planes; on the coarse grids, whose launches are
// latency-bound, chunks of 1-3 planes.
static void el_grid(const Grid& g, int slots, long long max_blocks, int& nty, int& zc,
                    dim3& grid) {
  nty = (g.nny + kOut - 1) / kOut;
  const int ntx = (g.nnx + kOut - 1) / kOut;
  const long long tiles = (long long)nty * ntx;
  if (slots < 1) slots = 1;
  long long best = -1;
  int best_zc = g.nnz, best_n = 1;
  for (int z = 1; z <= g.nnz; ++z) {
    const int n = (g.nnz + z - 1) / z;
    if (n > 1 && tiles * n > max_blocks) continue;
    const long long rounds = (tiles * n + slots - 1) / slots;
    const long long cost = rounds * (z + 1);
    if (best < 0 || cost < best) {
      best = cost;
      best_zc = z;
      best_n = n;
    }
  }
  zc = best_zc;
  grid = dim3((unsigned)tiles, (unsigned)best_n);
}

// resident workgroups per compute unit of one element-kernel instantiation
// (the Galerkin level's operator-only sweep is the plain kernel, not the
// 3-wave one)
template <typename T, int MODE, bool MF, bool GAL = false>
static int el_resident() {
  static int v = 0;
  if (v == 0) {
    constexpr bool w3 = (MODE == kModeApply || MODE == kModeApplyDir) && !GAL;
    int n = 0;
    const void* fn = w3 ? (const void*)k_apply_el_w3<T, w3 ? MODE : kModeApply, MF>
                        : (const void*)k_apply_el<T, w3 ? kModeResidual : MODE, MF, GAL>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 256, 0) != hipSuccess || n < 1) n = 1;
    v = n;
  }
  return v;
}

template <typename T> static const T* gal_table(const Plan& P);
template <> const double* gal_table<double>(const Plan& P) { return P.gal_tab64; }
template <> const float* gal_table<float>(const Plan& P) { return P.gal_tab32; }
template <typename T> static const T* gal_coeffs(const Plan& P);
template <> const double* gal_coeffs<double>(const Plan& P) { return P.gal_sh64; }
template <> const float* gal_coeffs<float>(const Plan& P) { return P.gal_sh32; }

// GAL: sl is ignored; the level's Walsh coefficients and the coupling table
// take the place of its scales and sector entries.
template <typename T, int MODE, bool MF, bool GAL = false>
static int launch_el_as(Plan& P, const Level& L, const T* sl, const T* u, const Epilogue<T>& ep,
                        int with_dot, const int* done, hipStream_t s) {
  int nty, zc;
  dim3 grid;
  // the reduction scratch holds 4 values per block of the largest 1-D launch
  el_grid(L.g, P.ncu * el_resident<T, MODE, MF, GAL>(), P.partials_len / 4, nty, zc, grid);
  SparseKE<float> kv;
  for (int i = 0; i < kSecN; ++i) kv.v[i] = P.sec32[i];
  const uint4* km = (const uint4*)P.kemf;
  constexpr bool w3 = (MODE == kModeApply || MODE == kModeApplyDir) && !GAL;
  const T* kg = GAL ? gal_table<T>(P) : sectors_of<T>(P);
  const T* sc = GAL ? gal_coeffs<T>(P) : sl;
  if (w3)
    hipLaunchKernelGGL((k_apply_el_w3<T, w3 ? MODE : kModeApply, MF>), grid, dim3(256), 0, s, L.g,
                       kv, kg, sc, L.fixed, u, ep, nty, zc, with_dot, done, P.partials, km);
  else
    hipLaunchKernelGGL((k_apply_el<T, w3 ? kModeResidual : MODE, MF, GAL>), grid, dim3(256), 0, s,
                       L.g, kv, kg, sc, L.fixed, u, ep, nty, zc, with_dot, done, P.partials, km);
  return (int)(grid.x * grid.y);
}

// Returns the number of workgroups (= partials per reduced quantity).  FP32
// sweeps take the BF16 MFMA form when the plan asks for it (operator_kind 2).
template <typename T, int MODE>
static int launch_el(Plan& P, const Level& L, const T* sl, const T* u, const Epilogue<T>& ep,
                     int with_dot, const int* done, hipStream_t s, bool finalize_alpha = true) {
  int G;
  const bool gal = P.gal_level > 0 && &L == &P.lv[P.gal_level];
  if constexpr (MODE != kModeApplyDir) {
    if (gal) {
      // the Galerkin level's smoothing sweeps read D^-1 (and the zero-start
      // sweep always does): never launch them without it
      constexpr bool needs_dinv = MODE == kModeChebFirst || MODE == kModeChebStep ||
                                  MODE == kModeChebFirstP || MODE == kModeChebZero;
      if (needs_dinv && !ep.dinv) {
        set_error("internal: Galerkin-level smoothing sweep without an inverse diagonal");
        P.launch_refused = true;
        return 0;
      }
      G = launch_el_as<T, MODE, false, true>(P, L, sl, u, ep, with_dot, done, s);
      if (MODE == kModeApply && with_dot && finalize_alpha) finalize(P, kFinAlpha, G, s);
      return G;
    }
  }
  if constexpr (sizeof(T) == 4 && MODE != kModeApplyDir) {
    G = P.mfma ? launch_el_as<T, MODE, true>(P, L, sl, u, ep, with_dot, done, s)
               : launch_el_as<T, MODE, false>(P, L, sl, u, ep, with_dot, done, s);
  } else {
    G = launch_el_as<T, MODE, false>(P, L, sl, u, ep, with_dot, done, s);
  }
  if ((MODE == kModeApply || MODE == kModeApplyDir) && with_dot && finalize_alpha)
    finalize(P, kFinAlpha, G, s);
  return G;
}

// Flexible-CG direction and operator in one sweep (see kModeApplyDir):
// p32n = fp32(zs z32 + beta p32o) (p32o == nullptr: beta = 0), y = K p, and
// the p.Kp partials finalised into alpha.
void el_apply_dir(Plan& P, const float* z32, const float* p32o, float* p32n, double* y,
                  hipStream_t s) {
  const Level& L = P.lv[0];
  P.cnt.applies64++;
  Epilogue<double> ep{};
  ep.y = y;
  ep.z32 = z32;
  ep.p32o = p32o;
  ep.p32n = p32n;
  ep.sc = P.sc;
  launch_el<double, kModeApplyDir>(P, L, L.f64.s, nullptr, ep, 1, &P.sc->done, s);
}

// The same with the directions in the plan's ring (slots chosen on the device
// from the iteration counter); have_old = 0: beta = 0 (first direction).  The
// p.Kp partials are left for the update pass, which forms alpha itself;
// returns their count.
int el_apply_dir_ring(Plan& P, const float* z32, int have_old, double* y, hipStream_t s) {
  const Level& L = P.lv[0];
  P.cnt.applies64++;
  Epilogue<double> ep{};
  ep.y = y;
  ep.z32 = z32;
  ep.p32o = have_old ? P.pring : nullptr;
  ep.p32n = nullptr;
  ep.sc = P.sc;
  ep.ring = P.pring;
  ep.ring_stride = P.g.ndof;
  ep.ring_m = P.ring_m;
  return launch_el<double, kModeApplyDir>(P, L, L.f64.s, nullptr, ep, 1, &P.sc->done, s, false);
}

template <typename T>
void el_apply(Plan& P, const Level& L, const T* sl, const T* u, T* y, int with_dot,
              const int* done, hipStream_t s) {
  Epilogue<T> ep{};
  ep.y = y;
  launch_el<T, kModeApply>(P, L, sl, u, ep, with_dot, done, s);
}

template <typename T>
void el_residual(Plan& P, const Level& L, const T* sl, const T* b, const T* x, T* r,
                 const int* done, hipStream_t s) {
  Epilogue<T> ep{};
  ep.y = r;
  ep.b = b;
  launch_el<T, kModeResidual>(P, L, sl, x, ep, 0, done, s);
}

// dot_blocks (Chebyshev step only): the sweep also leaves the partials of
// b.xo, the Krylov r.z when this is the last step of the V-cycle.
template <typename T>
void el_cheb(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b, const T* x, T* xo,
             T* d, double c0, double c1, double c2, int first, int write_d, const int* done,
             hipStream_t s, int* dot_blocks) {
  Epilogue<T> ep{};
  ep.b = b;
  ep.xo = xo;
  ep.d = write_d ? d : nullptr;
  ep.dp = d;
  ep.dinv = dinv;   // read on the Galerkin level (elsewhere D^-1 is formed from the scales)
  for (int c = 0; c < 3; ++c) ep.kdi[c] = (T)(1.0 / P.kd[c]);
  ep.c0 = (T)c0;
  ep.c1 = (T)c1;
  ep.c2 = (T)c2;
  if (first) {
    launch_el<T, kModeChebFirst>(P, L, sl, x, ep, 0, done, s);
    if (dot_blocks) *dot_blocks = 0;
  } else {
    const int G = launch_el<T, kModeChebStep>(P, L, sl, x, ep, dot_blocks ? 1 : 0, done, s);
    if (dot_blocks) *dot_blocks = G;
  }
}

// First post-smoothing step on x + P xc (see kModeChebFirstP).
template <typename T>
void el_cheb_prolong(Plan& P, const Level& L, const Level& C, const T* xc, const T* sl,
                     const T* dinv, const T* b, const T* x, T* xo, T* d, double c0, int write_d,
                     const int* done, hipStream_t s) {
  Epilogue<T> ep{};
  ep.b = b;
  ep.xo = xo;
  ep.d = write_d ? d : nullptr;
  ep.dp = d;
  ep.dinv = dinv;   // read on the Galerkin level
  for (int c = 0; c < 3; ++c) ep.kdi[c] = (T)(1.0 / P.kd[c]);
  ep.c0 = (T)c0;
  ep.xc = xc;
  ep.cg = C.g;
  launch_el<T, kModeChebFirstP>(P, L, sl, x, ep, 0, done, s);
}

// First two Chebyshev steps from x = 0 in one sweep (see kModeChebZero).
template <typename T>
void el_cheb_zero2(Plan& P, const Level& L, const T* sl, const T* dinv, const T* b, T* xo, T* d,
                   double c0, double c1, double c2, int write_d, const int* done, hipStream_t s) {
  Epilogue<T> ep{};
  ep.b = b;
  ep.xo = xo;
  ep.d = write_d ? d : nullptr;
  ep.dinv = dinv;
  ep.c0 = (T)c0;
  ep.c1 = (T)c1;
  ep.c2 = (T)c2;
  launch_el<T, kModeChebZero>(P, L, sl, nullptr, ep, 0, done, s);
}

#define TK_INST2(T)                                                                              \
  template void el_apply<T>(Plan&, const Level&, const T*, const T*, T*, int, const int*,       \
                            hipStream_t);                                                       \
  template void el_residual<T>(Plan&, const Level&, const T*, const T*, const T*, T*,           \
                               const int*, hipStream_t);                                        \
  template void el_cheb<T>(Plan&, const Level&, const T*, const T*, const T*, const T*, T*, T*, \
                           double, double, double, int, int, const int*, hipStream_t, int*); \
  template void el_cheb_zero2<T>(Plan&, const Level&, const T*, const T*, const T*, T*, T*,     \
                                 double, double, double, int, const int*, hipStream_t);         \
  template void el_cheb_prolong<T>(Plan&, const Level&, const Level&, const T*, const T*,       \
                                   const T*, const T*, const T*, T*, T*, double, int,           \
                                   const int*, hipStream_t);
TK_INST2(double)
TK_INST2(float)
#undef TK_INST2

}  // namespace tk
