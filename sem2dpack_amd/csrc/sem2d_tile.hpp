// Fused tile kernels for structured (MESH_CART) meshes.
//
// Fields live in node-grid order: node (X, Z) of the (4nx+1) x (4nz+1[+1]) GLL grid
// at index Z*NX1 + X (a split fault duplicates one grid row).  The element grid is
// cut into tiles of up to TX x TZ elements; one workgroup owns one tile and, in a
// single pass over HBM, does for the tile's patch of (4*nxe+1) x (4*nze+1) nodes:
//
//   1. load d, v, a; Newmark predictor (SRC/solver.f90:59-60); predicted d in LDS;
//   2. element forces f = -K d (ELAST_KD2_*, SRC/mat_elastic.f90:589-641,739-775)
//      assembled into an LDS node array, colour by colour (2x2 checkerboard of the
//      tile's elements: elements of one colour share no node, so the LDS updates
//      need no atomics and the summation order is fixed);
//   3. interior patch nodes: a = f*rmass, corrector (SRC/solver.f90:78-82), store;
//      patch-perimeter nodes and "special" nodes (point sources): store the
//      partial force of this tile to fpart and leave the node untouched.
//
// k_deferred then finishes every perimeter / special node: it sums the partial
// forces of the tiles that share it (fixed tile order), adds sources and absorbing
// boundaries (in BC order) and applies the corrector; split-fault nodes are handed
// to the fault kernel instead (predicted d, v and the assembled force written to
// d, v, a).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sem2d_kernels.hpp"

namespace sem2d {

constexpr int TX = 8, TZ = 8;
constexpr int PXM = 4 * TX + 1, PZM = 4 * TZ + 1;
constexpr int PMAX = PXM * PZM;
constexpr int TILE_THREADS = 256;
constexpr int NODE_ITERS = (PMAX + TILE_THREADS - 1) / TILE_THREADS;
constexpr int PERIM_MAX = 2 * PXM + 2 * (PZM - 2);
constexpr int MAXSPECIAL = 16;
constexpr int SLOTS = PERIM_MAX + MAXSPECIAL;

struct TileDev {
  int ntiles;
  int per_xcd;        // tiles per XCD slice of the launch
  const int4* tiles;  // (tx0, tz0, nxe | nze << 8 | nx << 16, Z0), nx = elements per grid row
  const int4* grids;  // per tile: (node_base, ROW, NZ1, cell_base) of the grid the tile lies in
  const int* spec_cnt;
  const int* spec_idx;  // [ntiles][MAXSPECIAL] local patch indices
  size_t NGp;           // component stride of the field arrays
  size_t FPS;           // component stride of fpart (ntiles*SLOTS)
};

// Storage index of patch node (X, Z) of the tile at (X0, Z0) whose left and right
// columns lie on vertical tile-edge lines kL and kR.  A grid of NZ1 node rows starting
// at node_base stores, row after row, the nodes strictly between its vertical
// tile-edge lines: each tile column's 4*TX-1 nodes in a segment of 4*TX slots (ROW
// slots per row), so that the segments start on cache-line boundaries; the nodes of
// the tile-edge lines follow column by column (NZ1 each, unit stride along Z).
// Several grids (one per scenario of a batch) can share the field arrays, each at its
// own node_base.
__device__ __forceinline__ size_t node_index(const int4 gr, int X, int Z, int PX, int X0, int Z0, int kL,
                                             int kR) {
  const size_t base = (size_t)gr.x, grid = (size_t)gr.y * gr.z;
  if (X == 0) return base + grid + (size_t)kL * gr.z + (Z0 + Z);
  if (X == PX - 1) return base + grid + (size_t)kR * gr.z + (Z0 + Z);
  return base + (size_t)(Z0 + Z) * gr.y + X0 + X - 1;
}

// element (cell) index of local element (lx, lz) of a tile at (tx0, tz0).  The
// fourth descriptor field is the grid's first cell for per-cell coefficients; with
// shared coefficients (one element array per grid) it is the offset of the grid's
// array instead and the cell index is not used.
__device__ __forceinline__ size_t cell_index(const int4 gr, int nx, int tx0, int tz0, int lx, int lz) {
  return (size_t)gr.w + (size_t)(tz0 + lz) * nx + tx0 + lx;
}

template <typename R, int MODE>
__device__ __forceinline__ const R* grid_coefs(const R* coefs, const int4 gr) {
  return MODE == COEF_SHARED ? coefs + gr.w : coefs;
}

__device__ __forceinline__ int perim_slot(int X, int Z, int PX, int PZ) {
  // branch-free form of: Z == 0 -> X; Z == PZ-1 -> PX + X; X == 0 -> 2PX + Z-1;
  // X == PX-1 -> 2PX + PZ-2 + Z-1; otherwise -1
  int s = X == PX - 1 ? 2 * PX + (PZ - 2) + (Z - 1) : -1;
  s = X == 0 ? 2 * PX + (Z - 1) : s;
  s = Z == PZ - 1 ? PX + X : s;
  s = Z == 0 ? X : s;
  return s;
}

template <typename R, int NDOF, int NELAST, int MODE, bool OOP = false>
__global__ __launch_bounds__(TILE_THREADS) void k_tile(TileDev T, R* __restrict__ d, R* __restrict__ v,
                                                       R* __restrict__ a, const R* __restrict__ rmass0,
                                                       const R* __restrict__ coefs, const R* __restrict__ geom,
                                                       R* __restrict__ fpart, R dt, R c1, R c2, R c3, R c4,
                                                       const R* __restrict__ din = nullptr,
                                                       const R* __restrict__ vin = nullptr,
                                                       const R* __restrict__ ain = nullptr) {
  // out of place: a null vin / ain reads v / a (their step is then in place)
  const R* __restrict__ dsrc = OOP ? din : d;
  const R* vsrc = OOP && vin ? vin : v;
  const R* asrc = OOP && ain ? ain : a;
  __shared__ R Ud[NDOF][PMAX];
  __shared__ R Fn[NDOF][PMAX];
  __shared__ R T1[FE_EPB][NDOF][NP];
  __shared__ R T2[FE_EPB][NDOF][NP];
  __shared__ short sflag[PMAX];

  // XCD-aware order: the 8 XCDs receive workgroups round-robin, so give each XCD a
  // contiguous range of tiles (neighbouring tiles share patch rows in its L2).
  const int bid = blockIdx.x;
  const int t = (bid & 7) * T.per_xcd + (bid >> 3);
  if (t >= T.ntiles) return;
  const int tid = threadIdx.x;
  const int4 tg = T.tiles[t];
  const int4 gr = T.grids[t];
  const int tx0 = tg.x, tz0 = tg.y, nxe = tg.z & 0xff, nze = (tg.z >> 8) & 0xff, Z0 = tg.w;
  const int PX = 4 * nxe + 1, PZ = 4 * nze + 1, NPT = PX * PZ;
  const int X0 = 4 * tx0;
  const size_t S = T.NGp;
  const int kL = tx0 / TX, kR = (tx0 + nxe + TX - 1) / TX;

  R vp[NODE_ITERS][NDOF];
#pragma unroll
  for (int r = 0; r < NODE_ITERS; ++r) {
    const int p = tid + TILE_THREADS * r;
    if (p < NPT) {
      const int Z = p / PX, X = p - Z * PX;
      const size_t g = node_index(gr, X, Z, PX, X0, Z0, kL, kR);
#pragma unroll
      for (int c = 0; c < NDOF; ++c) {
        const R dk = dsrc[c * S + g], vk = vsrc[c * S + g], ak = asrc[c * S + g];
        Ud[c][p] = dk + dt * vk + c1 * ak;
        vp[r][c] = vk + c2 * ak;
        Fn[c][p] = 0;
      }
      sflag[p] = -1;
    }
  }
  __syncthreads();
  const int nspec = T.spec_cnt[t];
  if (tid < nspec) sflag[T.spec_idx[t * MAXSPECIAL + tid]] = (short)tid;

  // element forces, colour by colour
  const int le = tid / NP;
  const int pe = tid - le * NP;
  const int ii = pe % NG, jj = pe / NG;
#pragma unroll 1
  for (int cc = 0; cc < 4; ++cc) {
    const int cx = cc & 1, cz = cc >> 1;
    const int ncx = (nxe - cx + 1) >> 1, ncz = (nze - cz + 1) >> 1;
    const int nce = ncx * ncz;
#pragma unroll 1
    for (int b0 = 0; b0 < nce; b0 += FE_EPB) {
      const int ce = b0 + le;
      const bool active = (tid < FE_THREADS) && (ce < nce);
      int base = 0;
      if (active) {
        const int lx = cx + 2 * (ce % ncx), lz = cz + 2 * (ce / ncx);
        base = 4 * lz * PX + 4 * lx;
        const size_t cell = cell_index(gr, tg.z >> 16, tx0, tz0, lx, lz);
        R A[NELAST];
        load_coefs<R, NDOF, NELAST, MODE>(grid_coefs<R, MODE>(coefs, gr), geom, cell, pe, A);
        R dxi[NDOF], deta[NDOF];
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          R s1 = 0, s2 = 0;
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            s1 += Coef<R>::h(k, ii) * Ud[c][base + jj * PX + k];
            s2 += Ud[c][base + k * PX + ii] * Coef<R>::h(k, jj);
          }
          dxi[c] = s1;
          deta[c] = s2;
        }
        if constexpr (NDOF == 2) {
          if constexpr (NELAST == 6) {
            T1[le][0][pe] = A[0] * dxi[0] + A[1] * deta[1];
            T2[le][0][pe] = A[3] * (deta[0] + dxi[1]);
            T1[le][1][pe] = A[4] * deta[0] + A[5] * dxi[1];
            T2[le][1][pe] = A[1] * dxi[0] + A[2] * deta[1];
          } else {
            T1[le][0][pe] = A[0] * dxi[0] + A[6] * deta[0] + A[7] * dxi[1] + A[1] * deta[1];
            T2[le][0][pe] = A[6] * dxi[0] + A[3] * deta[0] + A[4] * dxi[1] + A[8] * deta[1];
            T1[le][1][pe] = A[7] * dxi[0] + A[4] * deta[0] + A[5] * dxi[1] + A[9] * deta[1];
            T2[le][1][pe] = A[1] * dxi[0] + A[8] * deta[0] + A[9] * dxi[1] + A[2] * deta[1];
          }
        } else {
          if constexpr (NELAST == 2) {
            T1[le][0][pe] = A[0] * dxi[0];
            T2[le][0][pe] = A[1] * deta[0];
          } else {
            T1[le][0][pe] = A[0] * dxi[0] + A[2] * deta[0];
            T2[le][0][pe] = A[2] * dxi[0] + A[1] * deta[0];
          }
        }
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          R s1 = 0, s2 = 0;
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            s1 += Coef<R>::h(ii, k) * T1[le][c][jj * NG + k];
            s2 += T2[le][c][k * NG + ii] * Coef<R>::h(jj, k);
          }
          R* fp = &Fn[c][base + jj * PX + ii];
          *fp = *fp + (s1 + s2);
        }
      }
      __syncthreads();
    }
  }

  // finish interior nodes, hand perimeter / special nodes over
  const size_t fbase = (size_t)t * SLOTS;
#pragma unroll
  for (int r = 0; r < NODE_ITERS; ++r) {
    const int p = tid + TILE_THREADS * r;
    if (p < NPT) {
      const int Z = p / PX, X = p - Z * PX;
      int slot = perim_slot(X, Z, PX, PZ);
      if (slot < 0 && sflag[p] >= 0) slot = PERIM_MAX + sflag[p];
      if (slot >= 0) {
#pragma unroll
        for (int c = 0; c < NDOF; ++c) fpart[c * T.FPS + fbase + slot] = Fn[c][p];
      } else {
        const size_t g = node_index(gr, X, Z, PX, X0, Z0, kL, kR);
        const R rm = rmass0[g];
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          const R ak = Fn[c][p] * rm;
          a[c * S + g] = ak;
          v[c * S + g] = vp[r][c] + c3 * ak;
          d[c * S + g] = Ud[c][p] + c4 * ak;
        }
      }
    }
  }
}

// LDS node value of NDOF components stored together (one vector access for P-SV)
template <typename R, int N>
struct LVec {
  using T = R;
};
template <typename R>
struct LVec<R, 2> {
  typedef R T __attribute__((ext_vector_type(2)));
};
template <typename R, int N>
__device__ __forceinline__ typename LVec<R, N>::T lv_make(const R* x) {
  if constexpr (N == 2) return typename LVec<R, N>::T{x[0], x[1]};
  else return x[0];
}
template <typename R, int N>
__device__ __forceinline__ R lv_get(const typename LVec<R, N>::T& v, int c) {
  if constexpr (N == 2) return c == 0 ? v.x : v.y;
  else return v;
}

// Element i of an array below 4 GiB, addressed as a uniform base plus a 32-bit byte
// offset (the scalar-base / vector-offset form of the memory instructions, with no
// 64-bit address arithmetic per access).
template <typename T>
__device__ __forceinline__ T& at32(T* base, uint32_t i) {
  using C = std::conditional_t<std::is_const<T>::value, const char, char>;
  return *reinterpret_cast<T*>(reinterpret_cast<C*>(base) + i * (uint32_t)sizeof(T));
}

// p, as a value the compiler cannot look through: a uniform base kept in scalar
// registers, so that base + 32-bit offset is not re-associated into 64-bit vector
// address arithmetic
template <typename T>
__device__ __forceinline__ T* sgpr_base(T* p) {
  using G = __attribute__((address_space(1))) T*;  // global memory
  G g = (G)p;
  asm("" : "+s"(g));
  return (T*)g;
}

// node_index for a full tile (PX = 4*TX+1, right edge on line kL+1), in 32 bits and
// branch-free (both forms computed, one selected)
__device__ __forceinline__ uint32_t node_index_full(const int4 gr, int X, int Z, int X0, int Z0, int kL) {
  const uint32_t edge = (uint32_t)gr.x + (uint32_t)gr.y * (uint32_t)gr.z + (uint32_t)(Z0 + Z) +
                        (uint32_t)(X == 0 ? kL : kL + 1) * (uint32_t)gr.z;
  const uint32_t inner = (uint32_t)gr.x + (uint32_t)(Z0 + Z) * (uint32_t)gr.y + (uint32_t)(X0 + X - 1);
  return (X == 0 || X == 4 * TX) ? edge : inner;
}

// Persistent, software-pipelined variant of k_tile.  Each workgroup walks a strided
// sequence of tiles inside its XCD's share (see below).  No tile waits on a
// global load issued during that tile: everything a tile reads comes from registers
// filled while the previous tile was being processed.
// Vector memory operations retire in issue order (loads and stores share one
// counter), so a wait for a load also waits for every store issued before it, and the
// compiler counts the stores issued after a load as the minimum over the branches.
// Per tile, therefore:
//   1. the predictor takes every wait, on loads issued a whole element phase earlier;
//   2. the previous tile's results are stored from LDS (each thread stores the patch
//      slots it overwrites next, so no barrier separates 2 and 3);
//   3. the predicted values go to LDS;
//   4. all loads of the next tile are issued in one group (d, v, a, inverse mass,
//      moduli or point coefficients, geometric factors, special-node index);
//   5. the element phase runs while they and the stores of 2 complete.
// Tile descriptors are scalar loads, counted apart from vector memory.
// Results are identical to k_tile: same predictor, same colour order and same
// summation order.  NT threads per workgroup, EPB elements per element batch
// (EPB*25 <= NT).  The launch covers the tiles tlist[0 .. nlist).  ttiles, tgrids and
// tspc are T.tiles, T.grids and T.spec_cnt, passed as restricted arguments so that the
// compiler may read them through the scalar cache.
//
// FULL: every listed tile has TX x TZ elements, so patch and colour shapes are
// compile-time constants (no run-time divisions in the index arithmetic), and the
// field, inverse-mass, partial-force and per-cell arrays are addressed with 32-bit
// offsets from per-component bases (the caller guarantees that every component of
// these arrays is below 4 GiB).
//
// EP: element-point terms, per element point in cell order [cell][25] (either may be
// null): Kelvin-Voigt viscosity eta, the element displacement being d + eta v at each
// point (MAT_KV_add_etav, SRC/mat_kelvin_voigt.f90:139-150), and the 2.5D coefficient
// b25, the point force then losing b25 (d + eta v) (MAT_ELAST_add_25D_f,
// SRC/mat_elastic.f90:447-459).  The predicted velocity then passes through LDS.
template <typename R, int NDOF, int NELAST, int MODE, int NT, int EPB, bool OOP = false, bool FULL = false,
          bool EP = false>
// Register budget: FP32 fields at 512 threads run two workgroups per CU (4 waves per
// SIMD, 128 VGPRs); every other shape (and the element-point variant, whose LDS allows
// one workgroup per CU) runs at most 2 waves per SIMD of 512-thread workgroups or 3 of
// 256-thread ones, so the kernel may use the registers of 2.
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 512 && sizeof(R) == 4 && !EP ? 4 : 2))) void k_tile_pipe(TileDev T, const int* __restrict__ tlist, int nlist,
                                                            const int4* __restrict__ ttiles,
                                                            const int4* __restrict__ tgrids,
                                                            const int* __restrict__ tspc, R* __restrict__ d,
                                                            R* __restrict__ v, R* __restrict__ a,
                                                            const R* __restrict__ rmass0,
                                                            const R* __restrict__ coefs, const R* __restrict__ geom,
                                                            R* __restrict__ fpart, R dt, R c1, R c2, R c3, R c4,
                                                            const R* __restrict__ din = nullptr,
                                                            const R* __restrict__ vin = nullptr,
                                                            const R* __restrict__ ain = nullptr,
                                                            const R* __restrict__ eta = nullptr,
                                                            const R* __restrict__ b25 = nullptr) {
  // OOP: the step reads the state (din, vin, ain) and writes (d, v, a); otherwise in place
  // out of place: a null vin / ain reads v / a (their step is then in place)
  const R* __restrict__ dsrc = OOP ? din : d;
  const R* vsrc = OOP && vin ? vin : v;
  const R* asrc = OOP && ain ? ain : a;
  // node values with both components side by side: one LDS access moves both
  using LV = typename LVec<R, NDOF>::T;
  __shared__ LV Ud[PMAX];
  __shared__ LV Fn[PMAX];
  constexpr int NITER = (PMAX + NT - 1) / NT;                       // patch nodes per thread
  constexpr int NBATCH = (TX * TZ / 4 + EPB - 1) / EPB;              // batches per colour
  constexpr int PB_MAXB = 4 * NBATCH;                               // point tasks per thread
  constexpr int ETHREADS = EPB * NP;
  static_assert(ETHREADS <= NT, "element batch larger than the workgroup");
  static_assert(3 * TX * TZ <= NT && MAXSPECIAL <= NT, "per-tile tables larger than the workgroup");
  __shared__ LV Vp[PMAX];  // predicted velocity
  __shared__ R Rm[PMAX];        // inverse mass of the patch nodes
  __shared__ LV T1b[2][EPB][NP];  // double-buffered element stresses
  __shared__ LV T2b[2][EPB][NP];
  __shared__ short sflag[PMAX];
  // H, GLL weight products and per-element geometric factors live in LDS (no global
  // load inside the element phase).
  __shared__ double sH[NP];
  __shared__ double sW2[NP];
  __shared__ R Gm[TX * TZ][3];
  // moduli of each thread's point tasks (compact mode), staged from the prefetch
  // registers by the predictor
  __shared__ LV ML[MODE == COEF_COMPACT ? PB_MAXB : 1][MODE == COEF_COMPACT ? ETHREADS : 1];
  // COEF_NODE: the moduli of the patch nodes (coefs is then a per-node array in the
  // field layout, component stride T.NGp), loaded with the node values
  constexpr bool NODEM = MODE == COEF_NODE;
  __shared__ LV Mn[NODEM ? PMAX : 1];
  // EP: eta and b25 of the tile's element points, [local element lz*nxe+lx][point]
  constexpr int NEPT = EP ? TX * TZ * NP : 1;
  __shared__ R Eta[NEPT];
  __shared__ R Bet[NEPT];

  const int nwg = gridDim.x;
  const int bid = blockIdx.x;
  // Workgroups reach the 8 XCDs round-robin (XCD = bid & 7).  XCD x owns the x-th
  // eighth of the listed tiles [s0, s1); its W workgroups take them interleaved
  // (workgroup k: s0 + k, s0 + k + W, ...), so at any time one XCD works on consecutive
  // tiles and the cache lines and halo nodes shared by x-neighbours are fetched once
  // into its L2.
  const int W = nwg >> 3;  // nwg is a multiple of 8
  const int slice = (nlist + 7) >> 3;
  const int s0 = (bid & 7) * slice;
  const int s1 = min(s0 + slice, nlist);
  const int q0 = s0 + (bid >> 3);
  if (q0 >= s1) return;
  const int tid = threadIdx.x;
  // lane index of the node loops, re-materialised per full tile (below) so that their
  // per-lane index arithmetic is recomputed rather than held in registers across tiles
  int tnode = tid;
  const size_t S = T.NGp;
  const int le = tid / NP;
  const int pe = tid - le * NP;
  const int ii = pe % NG, jj = pe / NG;
  if (tid < NP) {
    sH[tid] = c_H[tid];
    sW2[tid] = c_W2[tid];
  }
  // this lane's element point (ii, jj) is the same for every tile: keep the H entries
  // its contractions use in registers
  const R w2 = (R)c_W2[pe];
  R hA[NG], hB[NG], hC[NG], hD[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    hA[k] = (R)c_H[k * NG + ii];
    hB[k] = (R)c_H[k * NG + jj];
    hC[k] = (R)c_H[ii * NG + k];
    hD[k] = (R)c_H[jj * NG + k];
  }
  // tile geometry: compile-time constants for full tiles
  struct Geo {
    int tx0, tz0, nxe, nze, Z0, PX, PZ, NPT, X0, kL, kR;
  };
  auto geo = [&](const int4 tg) {
    Geo g;
    g.tx0 = tg.x;
    g.tz0 = tg.y;
    g.nxe = FULL ? TX : (tg.z & 0xff);
    g.nze = FULL ? TZ : ((tg.z >> 8) & 0xff);
    g.Z0 = tg.w;
    g.PX = 4 * g.nxe + 1;
    g.PZ = 4 * g.nze + 1;
    g.NPT = g.PX * g.PZ;
    g.X0 = 4 * g.tx0;
    g.kL = g.tx0 / TX;
    g.kR = FULL ? g.kL + 1 : (g.tx0 + g.nxe + TX - 1) / TX;
    return g;
  };
  // storage index of patch node (X, Z) and field access (32-bit offsets for full tiles)
  auto nidx = [&](const int4 gr, const Geo& G, int X, int Z) {
    if constexpr (FULL) return node_index_full(gr, X, Z, G.X0, G.Z0, G.kL);
    else return node_index(gr, X, Z, G.PX, G.X0, G.Z0, G.kL, G.kR);
  };
  auto fld = [&](auto* cbase, auto g) -> auto& {
    if constexpr (FULL) return at32(cbase, (uint32_t)g);
    else return cbase[g];
  };
  auto ubase = [&](auto* p) {
    if constexpr (FULL) return sgpr_base(p);
    else return p;
  };
  // per-component array bases, formed once
  const R *bD[NDOF], *bV[NDOF], *bA[NDOF];
  R *oD[NDOF], *oV[NDOF], *oA[NDOF], *bF[NDOF];
#pragma unroll
  for (int c = 0; c < NDOF; ++c) {
    bD[c] = ubase(dsrc + c * S);
    bV[c] = ubase(vsrc + c * S);
    bA[c] = ubase(asrc + c * S);
    if constexpr (OOP) {
      oD[c] = ubase(d + c * S);
      oV[c] = ubase(v + c * S);
      oA[c] = ubase(a + c * S);
    } else {
      oD[c] = const_cast<R*>(bD[c]);
      oV[c] = const_cast<R*>(bV[c]);
      oA[c] = const_cast<R*>(bA[c]);
    }
    bF[c] = ubase(fpart + c * T.FPS);
  }
  const R* bRm = ubase(rmass0);
  const R* bCo = ubase(coefs);
  const R* bMo[NDOF];
#pragma unroll
  for (int c = 0; c < NDOF; ++c) bMo[c] = NODEM ? ubase(coefs + c * S) : bCo;
  const R* bGe = ubase(geom);
  const R* bEt = EP && eta ? ubase(eta) : nullptr;
  const R* bBe = EP && b25 ? ubase(b25) : nullptr;
  // cell index in the element arrays (32 bits for full tiles)
  auto cidx = [&](const int4 gr, int nx, const Geo& G, int lx, int lz) {
    if constexpr (FULL) return (uint32_t)gr.w + (uint32_t)(G.tz0 + lz) * (uint32_t)nx + (uint32_t)(G.tx0 + lx);
    else return cell_index(gr, nx, G.tx0, G.tz0, lx, lz);
  };

  // Registers of the next tile, all issued in one group at the start of the current
  // tile's element phase: node values, inverse mass of the patch nodes, the moduli of
  // this thread's point tasks (compact mode) or its point coefficients (shared mode: the
  // same for every element of the tile's grid), one geometric factor and one
  // special-node index.
  R pd[NITER][NDOF], pv[NITER][NDOF], pa[NITER][NDOF];
  R pm[NITER];
  R pmod[NODEM ? NITER : 1][NDOF] = {};
  R lm[PB_MAXB][NDOF] = {};
  R sa[NELAST] = {}, sac[NELAST] = {};
  R pg = 0;
  int psp = -1;
  constexpr int NEP = EP ? (TX * TZ * NP + NT - 1) / NT : 1;  // element points per thread
  R peta[NEP] = {}, pbet[NEP] = {};
  auto issue = [&](int t) {
    const int4 tg = ttiles[t];
    const int4 gr = tgrids[t];
    const Geo G = geo(tg);
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + NT * r;
      if (p < G.NPT) {
        const int Z = p / G.PX, X = p - Z * G.PX;
        const auto g = nidx(gr, G, X, Z);
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          pd[r][c] = fld(bD[c], g);
          pv[r][c] = fld(bV[c], g);
          pa[r][c] = fld(bA[c], g);
        }
        pm[r] = fld(bRm, g);
        if constexpr (NODEM) {
#pragma unroll
          for (int c = 0; c < NDOF; ++c) pmod[r][c] = fld(bMo[c], g);
        }
      }
    }
    if constexpr (MODE == COEF_COMPACT || NODEM) {
      if (tid < G.nxe * G.nze * 3) {
        const int e = tid / 3, k = tid - 3 * (tid / 3);
        const int lx = e % G.nxe, lz = e / G.nxe;
        pg = fld(bGe, cidx(gr, tg.z >> 16, G, lx, lz) * 3 + k);
      }
    }
    if constexpr (MODE == COEF_COMPACT) {
#pragma unroll
      for (int q = 0; q < PB_MAXB; ++q) {
        const int cc = q / NBATCH, b0 = (q % NBATCH) * EPB;
        const int cx = cc & 1, cz = cc >> 1;
        const int ncx = (G.nxe - cx + 1) >> 1, ncz = (G.nze - cz + 1) >> 1;
        const int ce = b0 + le;
        if (tid < ETHREADS && ce < ncx * ncz) {
          const int lx = cx + 2 * (ce % ncx), lz = cz + 2 * (ce / ncx);
          const auto cell = cidx(gr, tg.z >> 16, G, lx, lz);
#pragma unroll
          for (int c = 0; c < NDOF; ++c) lm[q][c] = fld(bCo, cell * NDOF * NP + c * NP + pe);
        }
      }
    } else if constexpr (MODE == COEF_SHARED) {
      if (tid < ETHREADS) load_coefs<R, NDOF, NELAST, MODE>(grid_coefs<R, MODE>(coefs, gr), geom, 0, pe, sa);
    }
    if constexpr (EP) {
#pragma unroll
      for (int j = 0; j < NEP; ++j) {
        const int idx = tid + NT * j;
        if (idx < G.nxe * G.nze * NP) {
          const int e = idx / NP, p = idx - NP * (idx / NP);
          const auto cell = cidx(gr, tg.z >> 16, G, e % G.nxe, e / G.nxe);
          if (bEt) peta[j] = fld(bEt, cell * NP + p);
          if (bBe) pbet[j] = fld(bBe, cell * NP + p);
        }
      }
    }
    if (tid < tspc[t]) psp = T.spec_idx[t * MAXSPECIAL + tid];
  };

  // FP64 fields at 512 threads: a thread keeps the predicted velocity and inverse mass
  // of its patch nodes in registers from its predictor to its stores (the registers are
  // there at two waves per SIMD); other shapes pass them through LDS.
  constexpr bool REGV = sizeof(R) == 8 && NT == 512 && !EP;
  R vkeep[REGV ? NITER : 1][NDOF] = {}, rkeep[REGV ? NITER : 1] = {};
  // likewise the moduli of the current tile's point tasks (compact mode): copied from
  // the prefetch registers by the predictor instead of staged through LDS
  constexpr bool REGM = REGV && MODE == COEF_COMPACT;
  R lmc[REGM ? PB_MAXB : 1][NDOF] = {};

  // Stores of a finished tile (tile tp, descriptors tgp, grp) from LDS: a thread
  // stores the patch nodes it owns in the node loops (p = tnode + NT r), so the next
  // tile's predictor, run by the same thread on the same slots after this, needs no
  // barrier in between.  Partial forces for perimeter and special nodes, new a, v, d
  // for the others.
  auto store_tile = [&](const int tp, const int4 tgp, const int4 grp) {
    const Geo G = geo(tgp);
    const uint32_t fbase = (uint32_t)tp * SLOTS;
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + NT * r;
      if (p < G.NPT) {
        const int Z = p / G.PX, X = p - Z * G.PX;
        int slot = perim_slot(X, Z, G.PX, G.PZ);
        if (slot < 0 && sflag[p] >= 0) slot = PERIM_MAX + sflag[p];
        if (slot >= 0) {
#pragma unroll
          for (int c = 0; c < NDOF; ++c) fld(bF[c], fbase + slot) = lv_get<R, NDOF>(Fn[p], c);
        } else {
          const auto g = nidx(grp, G, X, Z);
          const R rm = REGV ? rkeep[r] : Rm[p];
#pragma unroll
          for (int c = 0; c < NDOF; ++c) {
            const R ak = lv_get<R, NDOF>(Fn[p], c) * rm;
            fld(oA[c], g) = ak;
            fld(oV[c], g) = (REGV ? vkeep[r][c] : lv_get<R, NDOF>(Vp[p], c)) + c3 * ak;
            fld(oD[c], g) = lv_get<R, NDOF>(Ud[p], c) + c4 * ak;
          }
        }
      }
    }
  };

  issue(tlist[q0]);
  // the previous tile of this workgroup, whose results are still in LDS
  int tp = -1;
  int4 tgp = {}, grp = {};

#pragma unroll 1
  for (int q0i = q0; q0i < s1; q0i += W) {
    int lelem = le;  // likewise for the element-batch lane index (full tiles)
    if constexpr (FULL) {
      tnode = tid;
      asm volatile("" : "+v"(tnode));
      asm volatile("" : "+v"(lelem));
      // keep the ranges known, so that loop bounds that hold for every lane fold away
      __builtin_assume(tnode >= 0 && tnode < NT);
      __builtin_assume(lelem >= 0 && lelem <= (NT - 1) / NP);
    }
    const int t = tlist[q0i];
    const int tn = q0i + W < s1 ? tlist[q0i + W] : -1;  // next tile of this workgroup
    const int4 tg = ttiles[t];
    const int4 gr = tgrids[t];
    const Geo G = geo(tg);
    const int PX = G.PX, PZ = G.PZ, NPT = G.NPT;

    // predictor from the prefetched values (every wait on a global load of this tile
    // is here, ahead of any store) ...
    R u[NITER][NDOF], w[NITER][NDOF];
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + NT * r;
      if (p < NPT) {
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          u[r][c] = pd[r][c] + dt * pv[r][c] + c1 * pa[r][c];
          w[r][c] = pv[r][c] + c2 * pa[r][c];
        }
      }
    }
    // ... then the previous tile's stores (this tile's element phase covers their
    // completion, and the next wait on a load is for loads issued after them) ...
    if (tp >= 0) store_tile(tp, tgp, grp);
    // ... then this tile's node values into LDS (FP64: predicted velocity and inverse
    // mass stay in this thread's registers until its stores)
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + NT * r;
      if (p < NPT) {
        Ud[p] = lv_make<R, NDOF>(u[r]);
        if constexpr (REGV) {
#pragma unroll
          for (int c = 0; c < NDOF; ++c) vkeep[r][c] = w[r][c];
          rkeep[r] = pm[r];
        } else {
          Vp[p] = lv_make<R, NDOF>(w[r]);
          Rm[p] = pm[r];
        }
        Fn[p] = LV{};
        sflag[p] = -1;
        if constexpr (NODEM) Mn[p] = lv_make<R, NDOF>(pmod[r]);
      }
    }
    if constexpr (NODEM) {
      if (tid < G.nxe * G.nze * 3) Gm[tid / 3][tid - 3 * (tid / 3)] = pg;
    }
    if constexpr (EP) {
#pragma unroll
      for (int j = 0; j < NEP; ++j) {
        const int idx = tid + NT * j;
        if (idx < G.nxe * G.nze * NP) {
          Eta[idx] = peta[j];
          Bet[idx] = pbet[j];
        }
      }
    }
    if constexpr (MODE == COEF_COMPACT) {
      if (tid < G.nxe * G.nze * 3) Gm[tid / 3][tid - 3 * (tid / 3)] = pg;
      if constexpr (REGM) {
#pragma unroll
        for (int q = 0; q < PB_MAXB; ++q)
#pragma unroll
          for (int c = 0; c < NDOF; ++c) lmc[q][c] = lm[q][c];
      } else if (tid < ETHREADS) {
#pragma unroll
        for (int q = 0; q < PB_MAXB; ++q) ML[q][tid] = lv_make<R, NDOF>(lm[q]);
      }
    } else if constexpr (MODE == COEF_SHARED) {
#pragma unroll
      for (int k = 0; k < NELAST; ++k) sac[k] = sa[k];
    }
    __syncthreads();
    if (psp >= 0) sflag[psp] = (short)tid;
    psp = -1;
    // overlap: all loads of the next tile fly during this tile's element phase
    if (tn >= 0) issue(tn);

    // Element batches, software-pipelined over one barrier per batch: in phase q the
    // lanes compute the stresses of batch q (into T[q & 1]) and assemble batch q-1 (from
    // T[(q-1) & 1]) into Fn.  Batch q-1's nodes belong to elements of one colour, and
    // the elements of one colour share no node, so the per-node summation order is the
    // colour order, as in the two-barrier form.
    bool pv_valid = false, pv_active = false;
    int pv_base = 0, pv_el = 0, nvb = 0;
#pragma unroll
    for (int q = 0; q <= PB_MAXB; ++q) {
      bool has = false, active = false;
      int base = 0, elc = 0;
      if (q < PB_MAXB) {
        const int cc = q / NBATCH, b0 = (q % NBATCH) * EPB;
        const int cx = cc & 1, cz = cc >> 1;
        const int ncx = (G.nxe - cx + 1) >> 1, ncz = (G.nze - cz + 1) >> 1;
        const int nce = ncx * ncz;
        has = b0 < nce;  // uniform across the workgroup
        const int ce = b0 + lelem;
        active = has && (tid < ETHREADS) && (ce < nce);
        if (active) {
          const int lx = cx + 2 * (ce % ncx), lz = cz + 2 * (ce / ncx);
          base = 4 * lz * PX + 4 * lx;
          elc = lz * G.nxe + lx;
          R A[NELAST];
          if constexpr (MODE == COEF_COMPACT || NODEM) {
            const int el = lz * G.nxe + lx;
            const R g0 = Gm[el][0], g1 = Gm[el][1], g2 = Gm[el][2];
            R m0, m1;
            if constexpr (NODEM) {
              const LV mm = Mn[base + jj * PX + ii];
              m0 = lv_get<R, NDOF>(mm, 0);
              m1 = lv_get<R, NDOF>(mm, NDOF - 1);
            } else if constexpr (REGM) {
              m0 = lmc[q][0];
              m1 = lmc[q][NDOF - 1];
            } else {
              const LV mm = ML[q][tid];
              m0 = lv_get<R, NDOF>(mm, 0);
              m1 = lv_get<R, NDOF>(mm, NDOF - 1);
            }
            if constexpr (NDOF == 2) {
              const R la = m0, mu = m1, kx = la + 2 * mu;
              A[0] = w2 * kx * g0;
              A[1] = w2 * la * g2;
              A[2] = w2 * kx * g1;
              A[3] = w2 * mu * g1;
              A[4] = w2 * mu * g2;
              A[5] = w2 * mu * g0;
            } else {
              A[0] = w2 * m0 * g0;
              A[1] = w2 * m0 * g1;
            }
          } else if constexpr (MODE == COEF_SHARED) {
#pragma unroll
            for (int k = 0; k < NELAST; ++k) A[k] = sac[k];
          } else {
            const size_t cell = cell_index(gr, tg.z >> 16, G.tx0, G.tz0, lx, lz);
            load_coefs<R, NDOF, NELAST, MODE>(grid_coefs<R, MODE>(coefs, gr), geom, cell, pe, A);
          }
          R dxi[NDOF], deta[NDOF];
#pragma unroll
          for (int c = 0; c < NDOF; ++c) dxi[c] = deta[c] = 0;
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            const int n1 = base + jj * PX + k, n2 = base + k * PX + ii;
            const LV u1 = Ud[n1], u2 = Ud[n2];
            if constexpr (EP) {  // d + eta v at the element's points (k, jj) and (ii, k)
              const LV w1 = Vp[n1], w2 = Vp[n2];
              const R e1 = Eta[elc * NP + jj * NG + k], e2 = Eta[elc * NP + k * NG + ii];
#pragma unroll
              for (int c = 0; c < NDOF; ++c) {
                dxi[c] += hA[k] * (lv_get<R, NDOF>(u1, c) + e1 * lv_get<R, NDOF>(w1, c));
                deta[c] += (lv_get<R, NDOF>(u2, c) + e2 * lv_get<R, NDOF>(w2, c)) * hB[k];
              }
            } else {
#pragma unroll
              for (int c = 0; c < NDOF; ++c) {
                dxi[c] += hA[k] * lv_get<R, NDOF>(u1, c);
                deta[c] += lv_get<R, NDOF>(u2, c) * hB[k];
              }
            }
          }
          LV(*T1) = T1b[nvb & 1][le];
          LV(*T2) = T2b[nvb & 1][le];
          R t1[NDOF], t2[NDOF];
          if constexpr (NDOF == 2) {
            if constexpr (NELAST == 6) {
              t1[0] = A[0] * dxi[0] + A[1] * deta[1];
              t2[0] = A[3] * (deta[0] + dxi[1]);
              t1[1] = A[4] * deta[0] + A[5] * dxi[1];
              t2[1] = A[1] * dxi[0] + A[2] * deta[1];
            } else {
              t1[0] = A[0] * dxi[0] + A[6] * deta[0] + A[7] * dxi[1] + A[1] * deta[1];
              t2[0] = A[6] * dxi[0] + A[3] * deta[0] + A[4] * dxi[1] + A[8] * deta[1];
              t1[1] = A[7] * dxi[0] + A[4] * deta[0] + A[5] * dxi[1] + A[9] * deta[1];
              t2[1] = A[1] * dxi[0] + A[8] * deta[0] + A[9] * dxi[1] + A[2] * deta[1];
            }
          } else {
            if constexpr (NELAST == 2) {
              t1[0] = A[0] * dxi[0];
              t2[0] = A[1] * deta[0];
            } else {
              t1[0] = A[0] * dxi[0] + A[2] * deta[0];
              t2[0] = A[2] * dxi[0] + A[1] * deta[0];
            }
          }
          T1[pe] = lv_make<R, NDOF>(t1);
          T2[pe] = lv_make<R, NDOF>(t2);
        }
      }
      if (pv_active) {  // assemble the previous batch
        const LV* T1 = T1b[(nvb - 1) & 1][le];
        const LV* T2 = T2b[(nvb - 1) & 1][le];
        R s1[NDOF], s2[NDOF], f[NDOF];
#pragma unroll
        for (int c = 0; c < NDOF; ++c) s1[c] = s2[c] = 0;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
          const LV a1 = T1[jj * NG + k], a2 = T2[k * NG + ii];
#pragma unroll
          for (int c = 0; c < NDOF; ++c) {
            s1[c] += hC[k] * lv_get<R, NDOF>(a1, c);
            s2[c] += lv_get<R, NDOF>(a2, c) * hD[k];
          }
        }
        LV* fp = &Fn[pv_base + jj * PX + ii];
        const LV f0 = *fp;
        if (EP && bBe) {  // 2.5D: the point force loses b25 (d + eta v) of its own point
          const int own = pv_base + jj * PX + ii;
          const LV uo = Ud[own], wo = Vp[own];
          const R eo = Eta[pv_el * NP + pe], bo = Bet[pv_el * NP + pe];
#pragma unroll
          for (int c = 0; c < NDOF; ++c)
            f[c] = lv_get<R, NDOF>(f0, c) +
                   ((s1[c] + s2[c]) - bo * (lv_get<R, NDOF>(uo, c) + eo * lv_get<R, NDOF>(wo, c)));
        } else {
#pragma unroll
          for (int c = 0; c < NDOF; ++c) f[c] = lv_get<R, NDOF>(f0, c) + (s1[c] + s2[c]);
        }
        *fp = lv_make<R, NDOF>(f);
      }
      if (has || pv_valid) __syncthreads();
      pv_valid = has;
      pv_active = active;
      pv_base = base;
      pv_el = elc;
      if (has) ++nvb;
    }
    // the last element batch ended with a barrier: Fn is complete
    tp = t;
    tgp = tg;
    grp = gr;
  }
  if (tp >= 0) store_tile(tp, tgp, grp);
}

// One deferred node: tile-perimeter, special or split-fault node.
struct DefNode {
  int g;        // grid index
  int pstart;   // partial offsets part_off[pstart .. pstart+pcnt)
  int pcnt;
  int sstart;   // sources src[sstart .. sstart+scnt)
  int scnt;
  int astart;   // absorbing coefficients abso[astart .. astart+acnt), [ndof] each
  int acnt;
  int fault;    // 1: split-fault node (finished by the fault kernels)
  int dmask;    // Dirichlet components zeroed in the force (bit c: component c), after the
                // absorbing terms; at a fault node, those of the Dirichlet boundaries that
                // act before the fault
  int pinl;     // device form only: 1 when the two partial offsets are poff (pcnt == 2)
  int poff[2];
};

// Device form of a deferred-node list: 16 bytes per node (one vector load) plus a side
// table for the few nodes with sources or absorbing terms.  cnt packs pcnt (bits 0-7),
// scnt (bits 8-15), acnt (bits 16-23), fault (bit 24), "offsets inline" (bit 25) and
// the Dirichlet mask (bits 26-27); x
// is the node's (sstart, astart) entry in ext, or -1 when it has neither.  A node with
// two partial forces and no ext entry (a tile-edge node) holds its two partial offsets
// in pstart and x instead (bit 25 set), so part_off is not read for it.
struct alignas(16) DefDev {
  int g, pstart, cnt, x;
};
struct DefTable {
  const DefDev* node = nullptr;
  const int2* ext = nullptr;
};
__device__ __forceinline__ DefNode def_unpack(const DefTable& D, int q) {
  const DefDev e = D.node[q];
  DefNode n;
  n.g = e.g;
  n.pstart = e.pstart;
  n.pcnt = e.cnt & 0xff;
  n.scnt = (e.cnt >> 8) & 0xff;
  n.acnt = (e.cnt >> 16) & 0xff;
  n.fault = (e.cnt >> 24) & 1;
  n.pinl = (e.cnt >> 25) & 1;
  n.dmask = (e.cnt >> 26) & 3;
  n.sstart = 0;
  n.astart = 0;
  n.poff[0] = n.poff[1] = 0;
  if (n.pinl) {
    n.poff[0] = e.pstart;
    n.poff[1] = e.x;
  } else if (e.x >= 0) {
    const int2 x = D.ext[e.x];
    n.sstart = x.x;
    n.astart = x.y;
  }
  return n;
}

// Stacey terms of the absorbing entries of the deferred nodes: entry q (the q-th
// absorbing coefficient set of the step, in boundary order) subtracts kxd[idx[t]],
// t = off[q] .. off[q+1]-1 (its boundary elements in order), after its -C v term.
struct StaceyRefs {
  const int* off = nullptr;
  const int* idx = nullptr;
  const double* kxd = nullptr;  // [boundary element][k][2], from k_stacey_pre
};

// Stacey term of flat P-SV absorbing boundaries (BC_ABSO_apply, SRC/bc_abso.f90:321-333)
// per boundary element e and edge point k: kxd[e][k][c] = K[e][k][c] * sum_j h'_j(x_k)
// D_c(node[e][j]), D the predicted displacement d + dt v + c1 a.  The boundary nodes are
// deferred nodes, which k_deferred predicts later in the step, so the prediction is formed
// here from the state of the previous step (the same expression as in k_deferred).
template <typename R>
__global__ void k_stacey_pre(int nbe, const int* __restrict__ enode, const double* __restrict__ K,
                             const R* __restrict__ d, const R* __restrict__ v, const R* __restrict__ a, size_t GS,
                             R dt, R c1, double* __restrict__ kxd) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;  // e * NG + k
  if (q >= nbe * NG) return;
  const int e = q / NG, k = q - NG * e;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      const size_t g = (size_t)c * GS + enode[e * NG + j];
      const R dp = d[g] + dt * v[g] + c1 * a[g];
      s += c_H[j * NG + k] * (double)dp;
    }
    kxd[(size_t)q * 2 + c] = K[(size_t)q * 2 + c] * s;
  }
}

template <typename R, int NDOF, bool OOP = false>
__global__ __launch_bounds__(256) void k_deferred(int ndef, DefTable defs,
                                                  const int* __restrict__ part_off, const R* __restrict__ fpart,
                                                  size_t FPS, const SrcEntry* __restrict__ src,
                                                  const double* __restrict__ ampl, int nt, int it_arg,
                                                  const int* __restrict__ itp,
                                                  const double* __restrict__ abso, const R* __restrict__ rmdef,
                                                  R* __restrict__ d, R* __restrict__ v, R* __restrict__ a, size_t GS,
                                                  R dt, R c1, R c2, R c3, R c4,
                                                  const R* __restrict__ din = nullptr,
                                                  const R* __restrict__ vin = nullptr,
                                                  const R* __restrict__ ain = nullptr, StaceyRefs SR = {}) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ndef) return;
  const DefNode n = def_unpack(defs, q);
  const size_t g = n.g;
  const int it = itp ? *itp : it_arg;  // step index: device counter under graph replay
  // out of place: a null vin / ain reads v / a (their step is then in place)
  const R* __restrict__ dsrc = OOP ? din : d;
  const R* vsrc = OOP && vin ? vin : v;
  const R* asrc = OOP && ain ? ain : a;
  R dp[NDOF], vpp[NDOF], f[NDOF];
#pragma unroll
  for (int c = 0; c < NDOF; ++c) {
    const R dk = dsrc[c * GS + g], vk = vsrc[c * GS + g], ak = asrc[c * GS + g];
    dp[c] = dk + dt * vk + c1 * ak;
    vpp[c] = vk + c2 * ak;
    R s = 0;
    if (n.pinl) {  // the loop below with pcnt = 2
      s = s + fpart[c * FPS + n.poff[0]];
      s = s + fpart[c * FPS + n.poff[1]];
    } else {
      for (int k = 0; k < n.pcnt; ++k) s = s + fpart[c * FPS + part_off[n.pstart + k]];
    }
    f[c] = s;
  }
  for (int k = 0; k < n.scnt; ++k) {
    const SrcEntry e = src[n.sstart + k];
    const double am = (it >= 1 && it <= nt) ? ampl[(size_t)e.src * nt + (it - 1)] : 0.0;
#pragma unroll
    for (int c = 0; c < NDOF; ++c)
      if (c == e.comp) f[c] = (R)((double)f[c] + e.coef * am);
  }
  for (int k = 0; k < n.acnt; ++k) {
#pragma unroll
    for (int c = 0; c < NDOF; ++c) f[c] = (R)((double)f[c] - abso[(size_t)(n.astart + k) * NDOF + c] * (double)vpp[c]);
    if (SR.off) {  // Stacey (P-SV)
      const int t0 = SR.off[n.astart + k], t1 = SR.off[n.astart + k + 1];
      if (t1 > t0) {
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          double s = 0.0;
          for (int t = t0; t < t1; ++t) s += SR.kxd[(size_t)SR.idx[t] * 2 + c];
          f[c] = (R)((double)f[c] - s);
        }
      }
    }
  }
  // Dirichlet components (BC_DIRNEU_apply, SRC/bc_dirneu.f90:148-169): absorbing
  // boundaries act first, then the others in their order
#pragma unroll
  for (int c = 0; c < NDOF; ++c)
    if (n.dmask >> c & 1) f[c] = 0;
  if (n.fault) {
#pragma unroll
    for (int c = 0; c < NDOF; ++c) {
      d[c * GS + g] = dp[c];
      v[c * GS + g] = vpp[c];
      a[c * GS + g] = f[c];
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < NDOF; ++c) {
    const R ak = f[c] * rmdef[(size_t)c * ndef + q];
    a[c * GS + g] = ak;
    v[c * GS + g] = vpp[c] + c3 * ak;
    d[c * GS + g] = dp[c] + c4 * ak;
  }
}

// ===========================================================================
// Adjoint (reverse-mode) step on structured grids, P-SV with compact coefficients.
//
// Forward step n (SRC/solver.f90:42-84 with beta = 0):
//   dp = d + dt v + c1 a,  vp = v + c2 a,  a' = (-K dp - C vp + s) / Mt,
//   v' = vp + c3 a',  d' = dp,                       Mt = M + c3 C (ABC-modified mass)
// Its transpose, from the adjoint (ad, av, aa) of step n to that of step n-1:
//   fb = (aa + c3 av) / Mt
//   ad <- ad - K^T fb,  vbp = av - C fb,  av <- vbp + dt ad,  aa <- c2 vbp + c1 ad
// and the sensitivities of J to the step's parameters:
//   dJ/dMt = -fb * a_n,  dJ/dC = -fb * vp_n,
//   dJ/dA_k = (B fb)^T dPhi/dA_k (B d_n)   per GLL point, A_k the KD2 coefficients
// with Phi the point flux matrix of ELAST_KD2_PSV (SRC/mat_elastic.f90:589-641):
//   flux(xi,x) = A0 gxx + A1 gez,  flux(eta,x) = A3 (gex + gxz),
//   flux(xi,z) = A4 gex + A5 gxz,  flux(eta,z) = A1 gxx + A2 gez,
// (g.. the xi/eta derivatives of the x/z components).  -K^T uses Phi^T:
//   flux(eta,x) = A3 gex + A4 gxz,  flux(xi,z) = A3 gex + A5 gxz  (others unchanged).
// A = W2 * modulus * geometric factor (A0 = W2 (l+2m) g0, A1 = W2 l g2, A2 = W2 (l+2m) g1,
// A3 = W2 m g1, A4 = W2 m g2, A5 = W2 m g0), so per point
//   dJ/dl = W2 [g0 pxx gxx + g2 (pxx gez + pez gxx) + g1 pez gez]
//   dJ/dm = W2 [2 g0 pxx gxx + 2 g1 pez gez + g1 pex (gex + gxz) + g2 pxz gex + g0 pxz gxz]
// (p.. the derivatives of fb, g.. those of d_n), summed over the element's points.
// One workgroup per tile; interior nodes finish here, perimeter nodes in k_adj_deferred.
// ===========================================================================
// Forward acceleration a_n at field index i: stored (a != nullptr), or, from a history
// of displacements, the second difference (d_{n+1} - 2 d_n + d_{n-1}) / dt^2, which equals
// a_n for the explicit Newmark scheme (beta = 0, gamma = 1/2): d_{n+1} - d_n = dt v_n +
// dt^2/2 a_n and v_n - v_{n-1} = dt/2 (a_{n-1} + a_n).
template <typename R>
__device__ __forceinline__ double fwd_accel(const R* __restrict__ a, const R* __restrict__ dprev,
                                            const R* __restrict__ dnext, size_t i, R dcur, double idt2) {
  if (a) return (double)a[i];
  const double dc = (double)dcur;
  return (((double)dnext[i] - dc) - (dc - (double)dprev[i])) * idt2;
}

template <typename R, bool FULL = false>
__global__ __launch_bounds__(TILE_THREADS) void k_adj_tile(
    TileDev T, const int* __restrict__ tlist, int nlist, R* __restrict__ ad, R* __restrict__ av,
    R* __restrict__ aa, const R* __restrict__ rmass,
    const R* __restrict__ dfw, const R* __restrict__ afw, const R* __restrict__ coefs, const R* __restrict__ geom,
    R* __restrict__ fpart, double* __restrict__ glam, double* __restrict__ gmu, double* __restrict__ mbar,
    R dt, R c1, R c2, R c3, const R* __restrict__ dprev = nullptr, const R* __restrict__ dnext = nullptr,
    double idt2 = 0.0) {
  constexpr int NDOF = 2;
  using LV = typename LVec<R, NDOF>::T;  // both components of a node side by side
  __shared__ LV Uf[PMAX];  // fb
  __shared__ LV Ud[PMAX];  // forward d_n
  __shared__ LV Fn[PMAX];
  __shared__ LV T1[FE_EPB][NP];
  __shared__ LV T2[FE_EPB][NP];
  __shared__ double Gr[FE_EPB][NP][2];
  __shared__ double Gs[FE_EPB][NG][2];  // row sums of Gr
  __shared__ double Ge[TX * TZ][2];     // dJ/dlambda, dJ/dmu of the tile's elements, accumulated
  __shared__ R Gm[TX * TZ][3];          // geometric factors of the tile's elements
  __shared__ short sflag[PMAX];

  // workgroup -> listed tile: each XCD (bid & 7) takes a contiguous eighth of the list
  const int bid = blockIdx.x;
  const int per = (nlist + 7) >> 3;
  const int qt = (bid & 7) * per + (bid >> 3);
  if (qt >= nlist) return;
  const int t = tlist[qt];
  const int tid = threadIdx.x;
  const int4 tg = T.tiles[t];
  const int4 gr = T.grids[t];
  const int tx0 = tg.x, tz0 = tg.y, Z0 = tg.w;
  const int nxe = FULL ? TX : (tg.z & 0xff), nze = FULL ? TZ : ((tg.z >> 8) & 0xff);
  const int PX = 4 * nxe + 1, PZ = 4 * nze + 1, NPT = PX * PZ;
  const int X0 = 4 * tx0;
  const size_t S = T.NGp;
  const int kL = tx0 / TX, kR = FULL ? kL + 1 : (tx0 + nxe + TX - 1) / TX;
  // full tiles: compile-time shapes, 32-bit offsets from per-component bases held in
  // scalar registers (every array below 4 GiB, checked by the caller)
  auto nidx = [&](int X, int Z) {
    if constexpr (FULL) return node_index_full(gr, X, Z, X0, Z0, kL);
    else return node_index(gr, X, Z, PX, X0, Z0, kL, kR);
  };
  auto cidx = [&](int lx, int lz) {
    if constexpr (FULL)
      return (uint32_t)gr.w + (uint32_t)(tz0 + lz) * (uint32_t)(tg.z >> 16) + (uint32_t)(tx0 + lx);
    else return cell_index(gr, tg.z >> 16, tx0, tz0, lx, lz);
  };
  auto ub = [&](auto* p) {
    if constexpr (FULL) return sgpr_base(p);
    else return p;
  };
  auto at = [&](auto* base, auto g) -> auto& {
    if constexpr (FULL) return at32(base, (uint32_t)g);
    else return base[g];
  };
  R *bad[NDOF], *bav[NDOF], *baa[NDOF], *bfp[NDOF];
  const R *brm[NDOF], *bdf[NDOF], *baf[NDOF], *bdp[NDOF], *bdn[NDOF];
#pragma unroll
  for (int c = 0; c < NDOF; ++c) {
    bad[c] = ub(ad + c * S);
    bav[c] = ub(av + c * S);
    baa[c] = ub(aa + c * S);
    brm[c] = ub(rmass + c * S);
    bdf[c] = ub(dfw + c * S);
    baf[c] = afw ? ub(afw + c * S) : nullptr;
    bdp[c] = dprev ? ub(dprev + c * S) : nullptr;
    bdn[c] = dnext ? ub(dnext + c * S) : nullptr;
    bfp[c] = ub(fpart + c * T.FPS);
  }
  double* bmb = ub(mbar);
  double* bgl = ub(glam);
  double* bgm = ub(gmu);
  const R* bco = ub(coefs);
  const R* bge = ub(geom);
  // forward acceleration at node g (as fwd_accel)
  auto accel = [&](int c, auto g, R dcur) -> double {
    if (afw) return (double)at(baf[c], g);
    const double dc = (double)dcur;
    return (((double)at(bdn[c], g) - dc) - (dc - (double)at(bdp[c], g))) * idt2;
  };
  const int le = tid / NP;
  const int pe = tid - le * NP;
  const int ii = pe % NG, jj = pe / NG;
  // the H entries and GLL weight of this lane's element point
  const R w = (R)c_W2[pe];
  R hA[NG], hB[NG], hC[NG], hD[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    hA[k] = (R)c_H[k * NG + ii];
    hB[k] = (R)c_H[k * NG + jj];
    hC[k] = (R)c_H[ii * NG + k];
    hD[k] = (R)c_H[jj * NG + k];
  }

  R adr[NODE_ITERS][NDOF], avr[NODE_ITERS][NDOF];
  double mbr[NODE_ITERS];  // mass sensitivity of the node, -fb . a_n (used at interior nodes)
#pragma unroll
  for (int r = 0; r < NODE_ITERS; ++r) {
    const int p = tid + TILE_THREADS * r;
    if (p < NPT) {
      const int Z = p / PX, X = p - Z * PX;
      const auto g = nidx(X, Z);
      R f[NDOF], u[NDOF];
      double mb = 0;
#pragma unroll
      for (int c = 0; c < NDOF; ++c) {
        const R dk = at(bad[c], g), vk = at(bav[c], g), ak = at(baa[c], g);
        f[c] = (ak + c3 * vk) * at(brm[c], g);
        u[c] = at(bdf[c], g);
        adr[r][c] = dk;
        avr[r][c] = vk;
        mb -= (double)f[c] * accel(c, g, u[c]);
      }
      // interior nodes: the running sum is read here, with the other loads, and written
      // back in the store phase (perimeter and special nodes are the deferred kernel's)
      mbr[r] = perim_slot(X, Z, PX, PZ) < 0 ? at(bmb, g) + mb : 0.0;
      Uf[p] = lv_make<R, NDOF>(f);
      Ud[p] = lv_make<R, NDOF>(u);
      Fn[p] = LV{};
      sflag[p] = -1;
    }
  }
  // FP64 fields: moduli of this lane's point in each batch loaded with the node values
  // (no global load inside the element batches; with FP32 fields the registers are
  // better spent on occupancy, and the batches load their own); geometric factors of the
  // elements in LDS
  constexpr bool PREF = sizeof(R) == 8;
  R lm[4 * ((TX * TZ / 4 + FE_EPB - 1) / FE_EPB)][2];
#pragma unroll
  for (int q = 0; q < (PREF ? 4 * ((TX * TZ / 4 + FE_EPB - 1) / FE_EPB) : 0); ++q) {
    constexpr int NB = (TX * TZ / 4 + FE_EPB - 1) / FE_EPB;
    const int cc = q / NB, b0 = (q % NB) * FE_EPB;
    const int cx = cc & 1, cz = cc >> 1;
    const int ncx = (nxe - cx + 1) >> 1, ncz = (nze - cz + 1) >> 1;
    const int ce = b0 + le;
    if (tid < FE_THREADS && ce < ncx * ncz) {
      const auto cell = cidx(cx + 2 * (ce % ncx), cz + 2 * (ce / ncx));
      lm[q][0] = at(bco, cell * 2 * NP + pe);
      lm[q][1] = at(bco, cell * 2 * NP + NP + pe);
    }
  }
  if (tid < nxe * nze * 3) {
    const int e = tid / 3, k = tid - 3 * (tid / 3);
    Gm[e][k] = at(bge, cidx(e % nxe, e / nxe) * 3 + k);
  }
  if (tid < nxe * nze) {  // the elements' running sensitivities (local index lz*nxe + lx)
    const auto cell = cidx(tid % nxe, tid / nxe);
    Ge[tid][0] = at(bgl, cell);
    Ge[tid][1] = at(bgm, cell);
  }
  __syncthreads();
  const int nspec = T.spec_cnt[t];
  if (tid < nspec) sflag[T.spec_idx[t * MAXSPECIAL + tid]] = (short)tid;

  // element batch q: colour cc = q / NBATCH, elements b0 .. b0+FE_EPB-1 of that colour
  constexpr int NBATCH = (TX * TZ / 4 + FE_EPB - 1) / FE_EPB;
  auto batch = [&](const int q) {
    const int cc = q / NBATCH, b0 = (q % NBATCH) * FE_EPB;
    const int cx = cc & 1, cz = cc >> 1;
    const int ncx = (nxe - cx + 1) >> 1, ncz = (nze - cz + 1) >> 1;
    const int nce = ncx * ncz;
    if (b0 < nce) {  // uniform
      const int ce = b0 + le;
      const bool active = (tid < FE_THREADS) && (ce < nce);
      int base = 0, el = 0;
      if (active) {
        const int lx = cx + 2 * (ce % ncx), lz = cz + 2 * (ce / ncx);
        base = 4 * lz * PX + 4 * lx;
        el = lz * nxe + lx;
        const R g0 = Gm[el][0], g1 = Gm[el][1], g2 = Gm[el][2];
        R la, mu;
        if constexpr (PREF) {
          la = lm[q][0];
          mu = lm[q][1];
        } else {
          const auto cell = cidx(lx, lz);
          la = at(bco, cell * 2 * NP + pe);
          mu = at(bco, cell * 2 * NP + NP + pe);
        }
        const R kx = la + 2 * mu;
        const R A0 = w * kx * g0, A1 = w * la * g2, A2 = w * kx * g1;
        const R A3 = w * mu * g1, A4 = w * mu * g2, A5 = w * mu * g0;
        R pxi[NDOF] = {0, 0}, peta[NDOF] = {0, 0}, gxi[NDOF] = {0, 0}, geta[NDOF] = {0, 0};
#pragma unroll
        for (int k = 0; k < NG; ++k) {
          const LV f1 = Uf[base + jj * PX + k], f2 = Uf[base + k * PX + ii];
          const LV u1 = Ud[base + jj * PX + k], u2 = Ud[base + k * PX + ii];
#pragma unroll
          for (int c = 0; c < NDOF; ++c) {
            pxi[c] += hA[k] * lv_get<R, NDOF>(f1, c);
            peta[c] += lv_get<R, NDOF>(f2, c) * hB[k];
            gxi[c] += hA[k] * lv_get<R, NDOF>(u1, c);
            geta[c] += lv_get<R, NDOF>(u2, c) * hB[k];
          }
        }
        const R pxx = pxi[0], pex = peta[0], pxz = pxi[1], pez = peta[1];
        const R gxx = gxi[0], gex = geta[0], gxz = gxi[1], gez = geta[1];
        T1[le][pe] = LV{A0 * pxx + A1 * pez, A3 * pex + A5 * pxz};
        T2[le][pe] = LV{A3 * pex + A4 * pxz, A1 * pxx + A2 * pez};
        const double wd = (double)w;
        Gr[le][pe][0] = wd * ((double)g0 * pxx * gxx + (double)g2 * (pxx * gez + pez * gxx) + (double)g1 * pez * gez);
        Gr[le][pe][1] = wd * (2.0 * g0 * pxx * gxx + 2.0 * g1 * pez * gez + (double)g1 * pex * (gex + gxz) +
                              (double)g2 * pxz * gex + (double)g0 * pxz * gxz);
      }
      __syncthreads();
      if (active) {
        R s1[NDOF] = {0, 0}, s2[NDOF] = {0, 0}, f[NDOF];
#pragma unroll
        for (int k = 0; k < NG; ++k) {
          const LV a1 = T1[le][jj * NG + k], a2 = T2[le][k * NG + ii];
#pragma unroll
          for (int c = 0; c < NDOF; ++c) {
            s1[c] += hC[k] * lv_get<R, NDOF>(a1, c);
            s2[c] += lv_get<R, NDOF>(a2, c) * hD[k];
          }
        }
        LV* fp = &Fn[base + jj * PX + ii];
        const LV f0 = *fp;
#pragma unroll
        for (int c = 0; c < NDOF; ++c) f[c] = lv_get<R, NDOF>(f0, c) + (s1[c] + s2[c]);
        *fp = lv_make<R, NDOF>(f);
      }
      // fixed-order sum over the element's points in two levels: lanes pe < 5 sum the
      // rows jj = pe into Gs, then (after the barrier) lane pe = 0 sums the five row
      // sums.  Gs is written again only after the next batch's first barrier, which
      // lane pe = 0 reaches after its reads; the barrier also orders this batch's Fn
      // updates and Gr, T1, T2 reads before the next batch's writes.
      if (active && pe < NG) {
        double rl = 0, rm = 0;
#pragma unroll
        for (int q = 0; q < NG; ++q) {
          rl += Gr[le][pe * NG + q][0];
          rm += Gr[le][pe * NG + q][1];
        }
        Gs[le][pe][0] = rl;
        Gs[le][pe][1] = rm;
      }
      __syncthreads();
      if (active && pe == 0) {
        double sl = 0, sm = 0;
#pragma unroll
        for (int q = 0; q < NG; ++q) {
          sl += Gs[le][q][0];
          sm += Gs[le][q][1];
        }
        Ge[el][0] += sl;
        Ge[el][1] += sm;
      }
    }
  };
  if constexpr (PREF) {  // moduli in registers: batches unrolled
#pragma unroll
    for (int q = 0; q < 4 * NBATCH; ++q) batch(q);
  } else {
#pragma unroll 1
    for (int q = 0; q < 4 * NBATCH; ++q) batch(q);
  }
  __syncthreads();  // every batch's element sums are in Ge
  if (tid < nxe * nze) {
    const auto cell = cidx(tid % nxe, tid / nxe);
    at(bgl, cell) = Ge[tid][0];
    at(bgm, cell) = Ge[tid][1];
  }
  const uint32_t fbase = (uint32_t)t * SLOTS;
  // lane index re-materialised for the store loop, so that the node-index arithmetic of
  // the load loop is recomputed here instead of being held in registers through the
  // element batches
  int tst = tid;
  asm volatile("" : "+v"(tst));
  __builtin_assume(tst >= 0 && tst < TILE_THREADS);
#pragma unroll
  for (int r = 0; r < NODE_ITERS; ++r) {
    const int p = tst + TILE_THREADS * r;
    if (p < NPT) {
      const int Z = p / PX, X = p - Z * PX;
      int slot = perim_slot(X, Z, PX, PZ);
      if (slot < 0 && sflag[p] >= 0) slot = PERIM_MAX + sflag[p];
      const LV fn = Fn[p];
      if (slot >= 0) {
#pragma unroll
        for (int c = 0; c < NDOF; ++c) at(bfp[c], fbase + slot) = lv_get<R, NDOF>(fn, c);
      } else {
        const auto g = nidx(X, Z);
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          const R dn = adr[r][c] + lv_get<R, NDOF>(fn, c);
          const R vbp = avr[r][c];
          at(bad[c], g) = dn;
          at(bav[c], g) = vbp + dt * dn;
          at(baa[c], g) = c2 * vbp + c1 * dn;
        }
        at(bmb, g) = mbr[r];
      }
    }
  }
}

// Adjoint step at the deferred nodes (tile perimeters, source nodes): the partial
// forces -K^T fb of the tiles, the absorbing term, the mass and ABC sensitivities.
// gabs[c][q] accumulates dJ/dC_c of the node (for nodes with absorbing entries):
// -fb_c (vp_c + c3 a_c), vp = v_{n-1} + c2 a_{n-1} the forward predicted velocity.
// dJ/d(recorded field) of one sample at the receiver nodes (see k_adj_inject); the
// injection of step n-1 can ride on the deferred kernel of step n: udef[q] is the
// injection node handled by deferred node q after its own update (or -1), uext the
// injection nodes that are not deferred nodes, done by extra lanes.
struct InjectPlan {
  int nu = 0, nrec = 0;
  const int *unode = nullptr, *off = nullptr, *rec = nullptr;
  const double* w = nullptr;
  const double* g = nullptr;  // the sample, [ndof][nrec]; null: no injection
  void* fld = nullptr;        // the adjoint field it is added to (bulk type)
  const int* udef = nullptr;
  const int* uext = nullptr;
  int next = 0;
};

template <typename R>
__device__ __forceinline__ void inject_node(const InjectPlan& P, int u, R* __restrict__ fld, size_t GS) {
  for (int c = 0; c < 2; ++c) {
    const size_t i = c * GS + P.unode[u];
    R x = fld[i];
    for (int k = P.off[u]; k < P.off[u + 1]; ++k) x = (R)((double)x + P.w[k] * P.g[(size_t)c * P.nrec + P.rec[k]]);
    fld[i] = x;
  }
}

// Persistent, pipelined form of k_adj_tile, with the same arithmetic and the same
// summation orders (results equal bit for bit): AP_NT threads, one colour of a tile's
// elements per batch (two barriers per batch), each workgroup on a run of tiles (the
// XCD slices of k_tile_pipe).  While a tile's element batches run, the adjoint state,
// inverse mass, forward displacements (or acceleration), mass sensitivity, moduli,
// geometric factors and element sensitivities of the workgroup's next tile load into
// registers; a finished tile is stored by the threads that own its patch slots, right
// before they overwrite those slots with the next tile's values.  The adjoint state of
// the patch and the running mass sensitivity wait in LDS for the store (the registers
// hold the prefetch).  The launch covers tlist[0 .. nlist); the grid is a multiple of 8.
constexpr int AP_NT = 512;
constexpr int AP_EPB = 20;  // elements per batch: a whole colour of a full tile (16)
template <typename R, bool FULL = false>
__global__ __launch_bounds__(AP_NT) __attribute__((amdgpu_waves_per_eu(2))) void k_adj_pipe(
    TileDev T, const int* __restrict__ tlist, int nlist, const int4* __restrict__ ttiles,
    const int4* __restrict__ tgrids, const int* __restrict__ tspc, R* __restrict__ ad, R* __restrict__ av,
    R* __restrict__ aa, const R* __restrict__ rmass, const R* __restrict__ dfw, const R* __restrict__ afw,
    const R* __restrict__ coefs, const R* __restrict__ geom, R* __restrict__ fpart, double* __restrict__ glam,
    double* __restrict__ gmu, double* __restrict__ mbar, R dt, R c1, R c2, R c3,
    const R* __restrict__ dprev = nullptr, const R* __restrict__ dnext = nullptr, double idt2 = 0.0) {
  constexpr int NDOF = 2;
  using LV = typename LVec<R, NDOF>::T;
  constexpr int NITER = (PMAX + AP_NT - 1) / AP_NT;        // patch nodes per thread
  constexpr int NBC = (TX * TZ / 4 + AP_EPB - 1) / AP_EPB;  // batches per colour
  constexpr int NB = 4 * NBC;                               // batches per tile
  constexpr int ETH = AP_EPB * NP;
  static_assert(ETH <= AP_NT && 3 * TX * TZ <= AP_NT && MAXSPECIAL <= AP_NT, "tables larger than the workgroup");
  __shared__ LV Uf[PMAX];  // fb
  __shared__ LV Ud[PMAX];  // forward d_n
  __shared__ LV Fn[PMAX];  // -K^T fb, assembled
  __shared__ LV Ad[PMAX];  // adjoint displacement and velocity, kept for the store
  __shared__ LV Av[PMAX];
  __shared__ double Mb[PMAX];  // running mass sensitivity of the interior nodes
  __shared__ LV T1[AP_EPB][NP];
  __shared__ LV T2[AP_EPB][NP];
  __shared__ double Gr[AP_EPB][NP][2];
  __shared__ double Gs[AP_EPB][NG][2];
  __shared__ double Ge[TX * TZ][2];
  __shared__ R Gm[TX * TZ][3];
  __shared__ short sflag[PMAX];
  __shared__ LV ML[NB][ETH];  // this tile's moduli (lambda, mu) of each lane's point, per batch
  __shared__ double sH[NP];

  const int nwg = gridDim.x;
  const int bid = blockIdx.x;
  const int W = nwg >> 3;
  const int slice = (nlist + 7) >> 3;
  const int s0 = (bid & 7) * slice;
  const int s1 = min(s0 + slice, nlist);
  const int q0 = s0 + (bid >> 3);
  if (q0 >= s1) return;
  const int tid = threadIdx.x;
  const size_t S = T.NGp;
  // lane indices of the node loops and of the element batches, re-materialised per
  // tile (below) so that the index arithmetic derived from them is recomputed in each
  // tile rather than held in registers through the loop
  int tnode = tid, lel = tid / NP;
  // H entries from LDS (the registers hold the prefetch)
  if (tid < NP) sH[tid] = c_H[tid];
  struct Geo {
    int tx0, tz0, nxe, nze, Z0, PX, PZ, NPT, X0, kL, kR;
  };
  auto geo = [&](const int4 tg) {
    Geo g;
    g.tx0 = tg.x;
    g.tz0 = tg.y;
    g.nxe = FULL ? TX : (tg.z & 0xff);
    g.nze = FULL ? TZ : ((tg.z >> 8) & 0xff);
    g.Z0 = tg.w;
    g.PX = 4 * g.nxe + 1;
    g.PZ = 4 * g.nze + 1;
    g.NPT = g.PX * g.PZ;
    g.X0 = 4 * g.tx0;
    g.kL = g.tx0 / TX;
    g.kR = FULL ? g.kL + 1 : (g.tx0 + g.nxe + TX - 1) / TX;
    return g;
  };
  auto nidx = [&](const int4 gr, const Geo& G, int X, int Z) {
    if constexpr (FULL) return node_index_full(gr, X, Z, G.X0, G.Z0, G.kL);
    else return node_index(gr, X, Z, G.PX, G.X0, G.Z0, G.kL, G.kR);
  };
  auto cidx = [&](const int4 gr, const int4 tg, const Geo& G, int lx, int lz) {
    if constexpr (FULL)
      return (uint32_t)gr.w + (uint32_t)(G.tz0 + lz) * (uint32_t)(tg.z >> 16) + (uint32_t)(G.tx0 + lx);
    else return cell_index(gr, tg.z >> 16, G.tx0, G.tz0, lx, lz);
  };
  auto ub = [&](auto* p) {
    if constexpr (FULL) return sgpr_base(p);
    else return p;
  };
  auto at = [&](auto* base, auto g) -> auto& {
    if constexpr (FULL) return at32(base, (uint32_t)g);
    else return base[g];
  };
  R *bad[NDOF], *bav[NDOF], *baa[NDOF], *bfp[NDOF];
  const R *brm[NDOF], *bdf[NDOF], *baf[NDOF], *bdp[NDOF], *bdn[NDOF];
#pragma unroll
  for (int c = 0; c < NDOF; ++c) {
    bad[c] = ub(ad + c * S);
    bav[c] = ub(av + c * S);
    baa[c] = ub(aa + c * S);
    brm[c] = ub(rmass + c * S);
    bdf[c] = ub(dfw + c * S);
    baf[c] = afw ? ub(afw + c * S) : nullptr;
    bdp[c] = dprev ? ub(dprev + c * S) : nullptr;
    bdn[c] = dnext ? ub(dnext + c * S) : nullptr;
    bfp[c] = ub(fpart + c * T.FPS);
  }
  // x1: the forward acceleration (stored) or d_{n-1}; x2: d_{n+1} (history of d only)
  const R *bx1[NDOF], *bx2[NDOF];
#pragma unroll
  for (int c = 0; c < NDOF; ++c) {
    bx1[c] = afw ? baf[c] : bdp[c];
    bx2[c] = bdn[c];
  }
  double* bmb = ub(mbar);
  double* bgl = ub(glam);
  double* bgm = ub(gmu);
  const R* bco = ub(coefs);
  const R* bge = ub(geom);

  // next tile's registers: node values (x1, x2: the forward acceleration, or d_{n-1}
  // and d_{n+1}), the mass sensitivity of interior nodes, this lane's moduli in each
  // batch, one geometric factor, one element's two sensitivities, one special node
  R pad[NITER][NDOF] = {}, pav[NITER][NDOF] = {}, paa[NITER][NDOF] = {}, prm[NITER][NDOF] = {};
  R pdf[NITER][NDOF] = {}, px1[NITER][NDOF] = {}, px2[NITER][NDOF] = {};
  double pmb[NITER] = {};
  R plm[NB][2] = {};
  R pg = 0;
  double pge0 = 0, pge1 = 0;
  int psp = -1;
  auto issue = [&](int t) {
    const int4 tg = ttiles[t];
    const int4 gr = tgrids[t];
    const Geo G = geo(tg);
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + AP_NT * r;
      if (p < G.NPT) {
        const int Z = p / G.PX, X = p - Z * G.PX;
        const auto g = nidx(gr, G, X, Z);
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          pad[r][c] = at(bad[c], g);
          pav[r][c] = at(bav[c], g);
          paa[r][c] = at(baa[c], g);
          prm[r][c] = at(brm[c], g);
          pdf[r][c] = at(bdf[c], g);
          px1[r][c] = at(bx1[c], g);
          if (!afw) px2[r][c] = at(bx2[c], g);
        }
        pmb[r] = perim_slot(X, Z, G.PX, G.PZ) < 0 ? at(bmb, g) : 0.0;
      }
    }
    const int pl = tnode - lel * NP;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int cc = q / NBC, b0 = (q % NBC) * AP_EPB;
      const int cx = cc & 1, cz = cc >> 1;
      const int ncx = (G.nxe - cx + 1) >> 1, ncz = (G.nze - cz + 1) >> 1;
      const int ce = b0 + lel;
      if (tnode < ETH && ce < ncx * ncz) {
        const auto cell = cidx(gr, tg, G, cx + 2 * (ce % ncx), cz + 2 * (ce / ncx));
        plm[q][0] = at(bco, cell * 2 * NP + pl);
        plm[q][1] = at(bco, cell * 2 * NP + NP + pl);
      }
    }
    if (tid < G.nxe * G.nze * 3) {
      const int e = tid / 3, k = tid - 3 * (tid / 3);
      pg = at(bge, cidx(gr, tg, G, e % G.nxe, e / G.nxe) * 3 + k);
    }
    if (tid < G.nxe * G.nze) {
      const auto cell = cidx(gr, tg, G, tid % G.nxe, tid / G.nxe);
      pge0 = at(bgl, cell);
      pge1 = at(bgm, cell);
    }
    if (tid < tspc[t]) psp = T.spec_idx[t * MAXSPECIAL + tid];
  };

  // Stores of a finished tile from LDS (the element sensitivities, the partial forces
  // of perimeter and special nodes, the new adjoint state and mass sensitivity of the
  // others), by the threads that own the slots
  auto store_tile = [&](const int tp, const int4 tgp, const int4 grp) {
    const Geo G = geo(tgp);
    if (tid < G.nxe * G.nze) {
      const auto cell = cidx(grp, tgp, G, tid % G.nxe, tid / G.nxe);
      at(bgl, cell) = Ge[tid][0];
      at(bgm, cell) = Ge[tid][1];
    }
    const uint32_t fbase = (uint32_t)tp * SLOTS;
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + AP_NT * r;
      if (p < G.NPT) {
        const int Z = p / G.PX, X = p - Z * G.PX;
        int slot = perim_slot(X, Z, G.PX, G.PZ);
        if (slot < 0 && sflag[p] >= 0) slot = PERIM_MAX + sflag[p];
        const LV fn = Fn[p];
        if (slot >= 0) {
#pragma unroll
          for (int c = 0; c < NDOF; ++c) at(bfp[c], fbase + slot) = lv_get<R, NDOF>(fn, c);
        } else {
          const auto g = nidx(grp, G, X, Z);
          const LV a0 = Ad[p], v0 = Av[p];
#pragma unroll
          for (int c = 0; c < NDOF; ++c) {
            const R dn = lv_get<R, NDOF>(a0, c) + lv_get<R, NDOF>(fn, c);
            const R vbp = lv_get<R, NDOF>(v0, c);
            at(bad[c], g) = dn;
            at(bav[c], g) = vbp + dt * dn;
            at(baa[c], g) = c2 * vbp + c1 * dn;
          }
          at(bmb, g) = Mb[p];
        }
      }
    }
  };

  issue(tlist[q0]);
  int tp = -1;
  int4 tgp = {}, grp = {};
#pragma unroll 1
  for (int qi = q0; qi < s1; qi += W) {
    tnode = tid;
    lel = tid / NP;
    asm volatile("" : "+v"(tnode));
    asm volatile("" : "+v"(lel));
    __builtin_assume(tnode >= 0 && tnode < AP_NT);
    __builtin_assume(lel >= 0 && lel <= (AP_NT - 1) / NP);
    const int pe = tnode - lel * NP;
    const int ii = pe % NG, jj = pe / NG;
    const R w = (R)c_W2[pe];
    const int t = tlist[qi];
    const int tn = qi + W < s1 ? tlist[qi + W] : -1;
    const int4 tg = ttiles[t];
    const int4 gr = tgrids[t];
    const Geo G = geo(tg);
    const int PX = G.PX, NPT = G.NPT;
    // this tile's fb, forward displacement and mass-sensitivity increment from the
    // prefetched values ...
    R f[NITER][NDOF], u[NITER][NDOF];
    double mb[NITER];
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + AP_NT * r;
      if (p < NPT) {
        double m = 0;
#pragma unroll
        for (int c = 0; c < NDOF; ++c) {
          f[r][c] = (paa[r][c] + c3 * pav[r][c]) * prm[r][c];
          u[r][c] = pdf[r][c];
          double acc;
          if (afw) {
            acc = (double)px1[r][c];
          } else {
            const double dc = (double)u[r][c];
            acc = (((double)px2[r][c] - dc) - (dc - (double)px1[r][c])) * idt2;
          }
          m -= (double)f[r][c] * acc;
        }
        mb[r] = m;
      }
    }
    // ... then the previous tile's stores ...
    if (tp >= 0) store_tile(tp, tgp, grp);
    // ... then this tile's values into the slots just stored
#pragma unroll
    for (int r = 0; r < NITER; ++r) {
      const int p = tnode + AP_NT * r;
      if (p < NPT) {
        const int Z = p / PX, X = p - Z * PX;
        Uf[p] = lv_make<R, NDOF>(f[r]);
        Ud[p] = lv_make<R, NDOF>(u[r]);
        Fn[p] = LV{};
        Ad[p] = lv_make<R, NDOF>(pad[r]);
        Av[p] = lv_make<R, NDOF>(pav[r]);
        Mb[p] = perim_slot(X, Z, PX, G.PZ) < 0 ? pmb[r] + mb[r] : 0.0;
        sflag[p] = -1;
      }
    }
    if (tid < G.nxe * G.nze * 3) Gm[tid / 3][tid - 3 * (tid / 3)] = pg;
    if (tid < G.nxe * G.nze) {
      Ge[tid][0] = pge0;
      Ge[tid][1] = pge1;
    }
    if (tnode < ETH) {
#pragma unroll
      for (int q = 0; q < NB; ++q) ML[q][tnode] = lv_make<R, NDOF>(plm[q]);
    }
    __syncthreads();
    if (psp >= 0) sflag[psp] = (short)tid;
    psp = -1;
    if (tn >= 0) issue(tn);

#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int cc = q / NBC, b0 = (q % NBC) * AP_EPB;
      const int cx = cc & 1, cz = cc >> 1;
      const int ncx = (G.nxe - cx + 1) >> 1, ncz = (G.nze - cz + 1) >> 1;
      const int nce = ncx * ncz;
      if (b0 < nce) {  // uniform
        const int ce = b0 + lel;
        const bool active = (tnode < ETH) && (ce < nce);
        int base = 0, el = 0;
        if (active) {
          const int lx = cx + 2 * (ce % ncx), lz = cz + 2 * (ce / ncx);
          base = 4 * lz * PX + 4 * lx;
          el = lz * G.nxe + lx;
          const R g0 = Gm[el][0], g1 = Gm[el][1], g2 = Gm[el][2];
          const LV mm = ML[q][tnode];
          const R la = lv_get<R, NDOF>(mm, 0), mu = lv_get<R, NDOF>(mm, 1);
          const R kx = la + 2 * mu;
          const R A0 = w * kx * g0, A1 = w * la * g2, A2 = w * kx * g1;
          const R A3 = w * mu * g1, A4 = w * mu * g2, A5 = w * mu * g0;
          R pxi[NDOF] = {0, 0}, peta[NDOF] = {0, 0}, gxi[NDOF] = {0, 0}, geta[NDOF] = {0, 0};
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            const LV f1 = Uf[base + jj * PX + k], f2 = Uf[base + k * PX + ii];
            const LV u1 = Ud[base + jj * PX + k], u2 = Ud[base + k * PX + ii];
            const R ha = (R)sH[k * NG + ii], hb = (R)sH[k * NG + jj];
#pragma unroll
            for (int c = 0; c < NDOF; ++c) {
              pxi[c] += ha * lv_get<R, NDOF>(f1, c);
              peta[c] += lv_get<R, NDOF>(f2, c) * hb;
              gxi[c] += ha * lv_get<R, NDOF>(u1, c);
              geta[c] += lv_get<R, NDOF>(u2, c) * hb;
            }
          }
          const R pxx = pxi[0], pex = peta[0], pxz = pxi[1], pez = peta[1];
          const R gxx = gxi[0], gex = geta[0], gxz = gxi[1], gez = geta[1];
          T1[lel][pe] = LV{A0 * pxx + A1 * pez, A3 * pex + A5 * pxz};
          T2[lel][pe] = LV{A3 * pex + A4 * pxz, A1 * pxx + A2 * pez};
          const double wd = (double)w;
          Gr[lel][pe][0] =
              wd * ((double)g0 * pxx * gxx + (double)g2 * (pxx * gez + pez * gxx) + (double)g1 * pez * gez);
          Gr[lel][pe][1] = wd * (2.0 * g0 * pxx * gxx + 2.0 * g1 * pez * gez + (double)g1 * pex * (gex + gxz) +
                                (double)g2 * pxz * gex + (double)g0 * pxz * gxz);
        }
        __syncthreads();
        if (active) {
          R s1[NDOF] = {0, 0}, s2[NDOF] = {0, 0}, fo[NDOF];
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            const LV a1 = T1[lel][jj * NG + k], a2 = T2[lel][k * NG + ii];
            const R hc = (R)sH[ii * NG + k], hd = (R)sH[jj * NG + k];
#pragma unroll
            for (int c = 0; c < NDOF; ++c) {
              s1[c] += hc * lv_get<R, NDOF>(a1, c);
              s2[c] += lv_get<R, NDOF>(a2, c) * hd;
            }
          }
          LV* fp = &Fn[base + jj * PX + ii];
          const LV f0 = *fp;
#pragma unroll
          for (int c = 0; c < NDOF; ++c) fo[c] = lv_get<R, NDOF>(f0, c) + (s1[c] + s2[c]);
          *fp = lv_make<R, NDOF>(fo);
        }
        if (active && pe < NG) {
          double rl = 0, rm = 0;
#pragma unroll
          for (int q2 = 0; q2 < NG; ++q2) {
            rl += Gr[lel][pe * NG + q2][0];
            rm += Gr[lel][pe * NG + q2][1];
          }
          Gs[lel][pe][0] = rl;
          Gs[lel][pe][1] = rm;
        }
        __syncthreads();
        if (active && pe == 0) {
          double sl = 0, sm = 0;
#pragma unroll
          for (int q2 = 0; q2 < NG; ++q2) {
            sl += Gs[lel][q2][0];
            sm += Gs[lel][q2][1];
          }
          Ge[el][0] += sl;
          Ge[el][1] += sm;
        }
      }
    }
    __syncthreads();  // the last batch's element sums are in Ge
    tp = t;
    tgp = tg;
    grp = gr;
  }
  if (tp >= 0) store_tile(tp, tgp, grp);
}

template <typename R>
__global__ __launch_bounds__(256) void k_adj_deferred(int ndef, DefTable defs,
                                                      const int* __restrict__ part_off, const R* __restrict__ fpart,
                                                      size_t FPS, const double* __restrict__ abso,
                                                      const R* __restrict__ rmdef, R* __restrict__ ad,
                                                      R* __restrict__ av, R* __restrict__ aa, size_t GS,
                                                      const R* __restrict__ afw, const R* __restrict__ vprev,
                                                      const R* __restrict__ aprev, double* __restrict__ mbar,
                                                      double* __restrict__ gabs, R dt, R c1, R c2, R c3,
                                                      const R* __restrict__ dcur = nullptr,
                                                      const R* __restrict__ dprev = nullptr,
                                                      const R* __restrict__ dnext = nullptr, double idt = 0.0,
                                                      double idt2 = 0.0, InjectPlan IP = {}) {
  constexpr int NDOF = 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ndef) {  // injection lanes (nodes that are not deferred)
    if (IP.g && q < ndef + IP.next) inject_node<R>(IP, IP.uext[q - ndef], (R*)IP.fld, GS);
    return;
  }
  const DefNode n = def_unpack(defs, q);
  const size_t g = n.g;
  double mb = 0;
#pragma unroll
  for (int c = 0; c < NDOF; ++c) {
    const R dk = ad[c * GS + g], vk = av[c * GS + g], ak = aa[c * GS + g];
    const R fb = (ak + c3 * vk) * rmdef[(size_t)c * ndef + q];
    R F = 0;
    if (n.pinl) {  // the loop below with pcnt = 2
      F = F + fpart[c * FPS + n.poff[0]];
      F = F + fpart[c * FPS + n.poff[1]];
    } else {
      for (int k = 0; k < n.pcnt; ++k) F = F + fpart[c * FPS + part_off[n.pstart + k]];
    }
    double Cc = 0;
    for (int k = 0; k < n.acnt; ++k) Cc += abso[(size_t)(n.astart + k) * NDOF + c];
    const R vbp = (R)((double)vk - Cc * (double)fb);
    const R dn = dk + F;
    ad[c * GS + g] = dn;
    av[c * GS + g] = vbp + dt * dn;
    aa[c * GS + g] = c2 * vbp + c1 * dn;
    // with a displacement history (dcur set): a_n as in fwd_accel, and the predicted
    // velocity v_{n-1} + dt/2 a_{n-1} = (d_n - d_{n-1}) / dt
    const double an = dcur ? fwd_accel(afw, dprev, dnext, c * GS + g, dcur[c * GS + g], idt2)
                           : (double)afw[c * GS + g];
    mb -= (double)fb * an;
    if (n.acnt > 0) {
      const double vp = dcur ? ((double)dcur[c * GS + g] - (double)dprev[c * GS + g]) * idt
                             : (double)vprev[c * GS + g] + (double)c2 * (double)aprev[c * GS + g];
      gabs[(size_t)c * ndef + q] -= (double)fb * (vp + (double)c3 * an);
    }
  }
  mbar[g] += mb;
  if (IP.g && IP.udef[q] >= 0) inject_node<R>(IP, IP.udef[q], (R*)IP.fld, GS);  // after the update
}

// dJ/d(recorded field) of one sample added at the receiver nodes, or spread over the
// element nodes with the interpolation weights (the transpose of the interpolation).
// One lane per distinct node u: its contributions (receiver rec[k], weight w[k], k in
// [off[u], off[u+1]), in receiver order) are added one after the other, which is the
// order a receiver-by-receiver loop would use.
template <typename R>
__global__ void k_adj_inject(int nu, const int* __restrict__ unode, const int* __restrict__ off,
                             const int* __restrict__ rec, const double* __restrict__ w, int nrec,
                             const double* __restrict__ g, R* __restrict__ fld, size_t GS) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  InjectPlan P;
  P.nrec = nrec;
  P.unode = unode;
  P.off = off;
  P.rec = rec;
  P.w = w;
  P.g = g;
  inject_node<R>(P, u, fld, GS);
}

}  // namespace sem2d
