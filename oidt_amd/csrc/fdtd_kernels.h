// HIP kernels for gfx950: fused Yee half-steps (curl + source + CPML psi +
// port-plane capture) and the chunked FP64 running DFT.
//
// One launch covers every tile of every run resident in a batch: a block walks
// a device tile list (slot, x0, y0, z0) with a grid-stride loop, so a plain run
// and an uneven ensemble use the same code.  Each thread owns one (j,k) column
// (3D) or one j row (2D) and marches XC cells along x, keeping the x-neighbour
// of the previous cell in registers.
//
// Reference semantics reproduced per cell, in the reference launch order:
//   TMz  fdtd_funcs.py:44-124 (FDTD_2D_TM), order fdtd_main.py:277-298
//   TEz  fdtd_funcs.py:44-124 (FDTD_2D_TE), order fdtd_main.py:280-304
//   3D   fdtd_funcs.py:44-266 (FDTD_3D),    order fdtd_main.py:359-392
// Every CPML psi cell lies inside its component's bulk update range, so the
// per-cell order "bulk curl, source, psi(axis a), psi(axis b)" equals the
// reference's kernel sequence.  The one reordering is the 3D TM source on Hy,
// applied at the start of the H half-step instead of between the E curl and
// the E-CPML; the host rejects source boxes that the E-CPML stencil reads, so
// the results are the same.
#pragma once
#include <type_traits>

#include "fdtd_slot.h"

namespace oidt {

// Blocks are (64, 4): threadIdx.y is the wavefront index, the same for every
// lane of a wave.  Reading it through readfirstlane lets the compiler keep
// everything derived from it in scalar registers: the fused steps take their
// tile index from it, and the 3D half-step kernels their y row (so the y
// conditions and the scalar part of every address are computed once per
// wave).  Splitting those kernels further into an interior body without the
// per-lane z conditions measured slower (its remaining conditions become
// scalar branches between the loads), so they keep one body.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.y); }

// Cross-lane moves that stay in registers (no LDS, so no wait on the
// counter that scalar memory loads also use): the value of lane l-1 (DPP
// wave shift right by one; lane 0 gets `lo`) and the value of a uniform lane.
__device__ __forceinline__ float lane_up1(float v, float lo) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lo), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_up1(double v, double lo) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(lo);
  const int blo = __builtin_amdgcn_update_dpp((int)o, (int)b, 0x138, 0xf, 0xf, false);
  const int bhi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), 0x138, 0xf, 0xf, false);
  return __longlong_as_double(((long long)bhi << 32) | (unsigned)blo);
}
__device__ __forceinline__ float lane_get(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_get(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Global-memory access by a 32-bit byte offset.  Every pointer of a run is a
// device-global allocation; saying so (address space 1) replaces flat
// instructions (64-bit per-lane addresses, counted on both the vector-memory
// and the LDS counter) with global ones addressed as an SGPR base plus a
// 32-bit VGPR offset.  Arrays addressed this way are below 4 GiB (checked
// when a run is described).
template <typename T>
__device__ __forceinline__ T ldg(const T* base, unsigned boff) {
  return *(const OIDT_GLOBAL T*)(reinterpret_cast<const char*>(base) + boff);
}
// Load that the caches may drop at once (non-temporal): for values a kernel
// reads exactly once when the batch's working set is far larger than the
// last-level cache, so they do not evict the neighbour values other waves
// re-read.
template <bool NT, typename T>
__device__ __forceinline__ T ldx(const T* base, unsigned boff) {
  if constexpr (NT)
    return __builtin_nontemporal_load((const OIDT_GLOBAL T*)(reinterpret_cast<const char*>(base) + boff));
  else
    return ldg<T>(base, boff);
}
template <typename T>
__device__ __forceinline__ void stg(T* base, unsigned boff, T v) {
  *(OIDT_GLOBAL T*)(reinterpret_cast<char*>(base) + boff) = v;
}
// Non-temporal store of a field value that no later access in the same
// launch reads.
template <typename T>
__device__ __forceinline__ void stg_nt(T* base, unsigned boff, T v) {
  __builtin_nontemporal_store(v, (OIDT_GLOBAL T*)(reinterpret_cast<char*>(base) + boff));
}

// Blocks are dispatched round-robin over the 8 XCDs; remap so that the blocks
// of one XCD take a contiguous range of tiles (shared L2 for x-neighbours).
__device__ __forceinline__ int xcd_block() {
  const int g = gridDim.x, b = blockIdx.x;
  if ((g & 7) != 0) return b;
  return (b & 7) * (g >> 3) + (b >> 3);
}

// E-type CPML slab along one axis: low cells [b, b+N) use coefficient N-1-d,
// high cells [e-N, e) use coefficient d (fdtd_funcs.py:85,87).
__device__ __forceinline__ int slab_e(int x, int b, int e, int N, int& d, int& ci) {
  if (x >= b && x < b + N) { d = x - b; ci = N - 1 - d; return 0; }
  if (x >= e - N && x < e) { d = x - (e - N); ci = d; return 1; }
  return -1;
}
// H-type CPML slab: low cells (b, b+N) use N-d, high cells (e-N, e) use d
// (fdtd_funcs.py:109,111).
__device__ __forceinline__ int slab_h(int x, int b, int e, int N, int& d, int& ci) {
  if (x > b && x < b + N) { d = x - b; ci = N - d; return 0; }
  if (x > e - N && x < e) { d = x - (e - N); ci = d; return 1; }
  return -1;
}

// cb of E component c for a cell: the material byte holds a 2-bit index per
// component into the run's (at most 4) distinct cb values of that component.
template <typename T>
__device__ __forceinline__ T cbsel(const Slot<T>& s, uint8_t m, int c) {
  const int q = (m >> (2 * c)) & 3;
  return q == 0 ? s.cbv[c][0] : (q == 1 ? s.cbv[c][1] : (q == 2 ? s.cbv[c][2] : s.cbv[c][3]));
}

template <typename P, typename T>
__device__ __forceinline__ T psi_upd(P p, T b, T c, T diff) {
  const T v = b * (*p) + c * diff;
  *p = v;
  return v;
}

template <typename T>
__device__ __forceinline__ void capture(const Slot<T>& s, int tstep, int plane, int cell, T v) {
  s.ring[((size_t)tstep * s.nplanes + plane) * s.pcells + cell] = v;
}

// ---------------------------------------------------------------- 2D TMz ---

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tmz_e(const Slot<T>* __restrict__ slots,
                                                    const int4* __restrict__ tiles,
                                                    const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    const int i1 = min(i0 + XC, s.nx);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    T* __restrict__ Ez = s.f[F_EZ];
    const T* __restrict__ Hx = s.f[F_HX];
    const T* __restrict__ Hy = s.f[F_HY];
    const bool jin = j >= jb && j < je;
    int dy, cy;
    const int sy = jin ? slab_e(j, jb, je, N, dy, cy) : -1;
    const T iky = jin ? s.ike[1][j] : T(0);
    const bool jsrc = s.src_field == F_EZ && j >= s.sy0 && j < s.sy1;
    const T jv = jsrc ? s.jsrc[n] : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    T hy_next = (i0 < i1) ? Hy[(size_t)i0 * ny + j] : T(0);
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * ny + j;
      const T hy0 = hy_next;
      hy_next = (i + 1 < s.nx) ? Hy[idx + ny] : T(0);
      const bool bulk = jin && i >= ib && i < ie;
      const bool src = jsrc && i >= s.sx0 && i < s.sx1;
      const bool cap0 = jcap && i == s.port[0];
      const bool cap1 = jcap && i == s.port[1];
      if (!(bulk || src || cap0 || cap1)) continue;
      T v = Ez[idx];
      if (bulk) {
        const T cb = cbsel(s, s.mat[idx], 2);
        const T dHy = hy_next - hy0;
        const T dHx = Hx[idx + 1] - Hx[idx];
        v = v + cb * (dHy * s.ike[0][i] - dHx * iky);
        if (src) v -= jv;
        int dx, cx;
        const int sx = slab_e(i, ib, ie, N, dx, cx);
        if (sx >= 0) v += cb * psi_upd(s.psi[psi_id(F_EZ, 0, sx)] + (size_t)dx * ny + j, s.be[cx], s.ceo[cx], dHy);
        if (sy >= 0) v -= cb * psi_upd(s.psi[psi_id(F_EZ, 1, sy)] + (size_t)i * N + dy, s.be[cy], s.ceo[cy], dHx);
        Ez[idx] = v;
      } else if (src) {
        v -= jv;
        Ez[idx] = v;
      }
      if (cap0) capture(s, tstep, 0, j - s.wy0, v);
      if (cap1) capture(s, tstep, 3, j - s.wy0, v);
    }
  }
}

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tmz_h(const Slot<T>* __restrict__ slots,
                                                    const int4* __restrict__ tiles,
                                                    const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    const int i1 = min(i0 + XC, s.nx);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    const T* __restrict__ Ez = s.f[F_EZ];
    T* __restrict__ Hx = s.f[F_HX];
    T* __restrict__ Hy = s.f[F_HY];
    const T db = s.db;
    const bool jin = j >= jb && j < je;
    const bool jin1 = j > jb && j < je;
    int dy, cy;
    const int sy = jin1 ? slab_h(j, jb, je, N, dy, cy) : -1;
    const T iky = jin1 ? s.ikh[1][j] : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    T ez_prev = (i0 > 0 && i0 < i1) ? Ez[(size_t)(i0 - 1) * ny + j] : T(0);
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * ny + j;
      const T ez = Ez[idx];
      const T ezm = ez_prev;
      ez_prev = ez;
      const bool iin = i >= ib && i < ie;
      const bool iin1 = i > ib && i < ie;
      // update_hx: i in [ib,ie), j in (jb,je)
      if (iin && jin1) {
        const T dEz = ez - Ez[idx - 1];
        T v = Hx[idx] + db * (-dEz) * iky;
        if (sy >= 0) v -= db * psi_upd(s.psi[psi_id(F_HX, 1, sy)] + (size_t)i * N + dy, s.bh[cy], s.cho[cy], dEz);
        Hx[idx] = v;
      }
      // update_hy: i in (ib,ie), j in [jb,je)
      const bool bulk = iin1 && jin;
      const bool c0 = jcap && i == s.port[0], c0b = jcap && i == s.port[0] + 1;
      const bool c1 = jcap && i == s.port[1], c1b = jcap && i == s.port[1] + 1;
      if (!(bulk || c0 || c0b || c1 || c1b)) continue;
      T v = Hy[idx];
      if (bulk) {
        const T dEz = ez - ezm;
        v = v + db * dEz * s.ikh[0][i];
        int dx, cx;
        const int sx = slab_h(i, ib, ie, N, dx, cx);
        if (sx >= 0) v += db * psi_upd(s.psi[psi_id(F_HY, 0, sx)] + (size_t)dx * ny + j, s.bh[cx], s.cho[cx], dEz);
        Hy[idx] = v;
      }
      const int c = j - s.wy0;
      if (c0) capture(s, tstep, 1, c, v);
      if (c0b) capture(s, tstep, 2, c, v);
      if (c1) capture(s, tstep, 4, c, v);
      if (c1b) capture(s, tstep, 5, c, v);
    }
  }
}

// ---------------------------------------------------------------- 2D TEz ---

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tez_e(const Slot<T>* __restrict__ slots,
                                                    const int4* __restrict__ tiles,
                                                    const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    const int i1 = min(i0 + XC, s.nx);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    T* __restrict__ Ex = s.f[F_EX];
    T* __restrict__ Ey = s.f[F_EY];
    const T* __restrict__ Hz = s.f[F_HZ];
    const bool jin = j >= jb && j < je;
    const bool jin1 = j > jb && j < je;
    int dy, cy;
    const int sy = jin ? slab_e(j, jb, je, N, dy, cy) : -1;
    const T iky = jin ? s.ike[1][j] : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    T hz_next = (i0 < i1) ? Hz[(size_t)i0 * ny + j] : T(0);
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * ny + j;
      const T hz0 = hz_next;
      hz_next = (i + 1 < s.nx) ? Hz[idx + ny] : T(0);
      const bool iin = i >= ib && i < ie;
      const bool iin1 = i > ib && i < ie;
      const bool bx = iin1 && jin, by = iin && jin1;
      const bool c0 = jcap && i == s.port[0], c1 = jcap && i == s.port[1];
      if (!(bx || by || c0 || c1)) continue;
      const uint8_t m = (bx || by) ? s.mat[idx] : 0;
      // update_ex: i in (ib,ie), j in [jb,je); psi ex_y (+)
      if (bx) {
        const T cb = cbsel(s, m, 0);
        const T dHz = Hz[idx + 1] - hz0;
        T v = Ex[idx] + cb * dHz * iky;
        if (sy >= 0) v += cb * psi_upd(s.psi[psi_id(F_EX, 1, sy)] + (size_t)i * N + dy, s.be[cy], s.ceo[cy], dHz);
        Ex[idx] = v;
      }
      // update_ey: i in [ib,ie), j in (jb,je); psi ey_x (-)
      T v = Ey[idx];
      if (by) {
        const T cb = cbsel(s, m, 1);
        const T dHz = hz_next - hz0;
        v = v + cb * (-dHz) * s.ike[0][i];
        int dx, cx;
        const int sx = slab_e(i, ib, ie, N, dx, cx);
        if (sx >= 0) v -= cb * psi_upd(s.psi[psi_id(F_EY, 0, sx)] + (size_t)dx * ny + j, s.be[cx], s.ceo[cx], dHz);
        Ey[idx] = v;
      }
      if (c0) capture(s, tstep, 0, j - s.wy0, v);
      if (c1) capture(s, tstep, 3, j - s.wy0, v);
    }
  }
}

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tez_h(const Slot<T>* __restrict__ slots,
                                                    const int4* __restrict__ tiles,
                                                    const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    const int i1 = min(i0 + XC, s.nx);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    const T* __restrict__ Ex = s.f[F_EX];
    const T* __restrict__ Ey = s.f[F_EY];
    T* __restrict__ Hz = s.f[F_HZ];
    const T db = s.db;
    const bool jin1 = j > jb && j < je;
    int dy, cy;
    const int sy = jin1 ? slab_h(j, jb, je, N, dy, cy) : -1;
    const T iky = jin1 ? s.ikh[1][j] : T(0);
    const bool jsrc = s.src_field == F_HZ && j >= s.sy0 && j < s.sy1;
    const T jv = jsrc ? s.jsrc[n] : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    T ey_prev = (i0 > 0 && i0 < i1) ? Ey[(size_t)(i0 - 1) * ny + j] : T(0);
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * ny + j;
      const T ey = Ey[idx];
      const T eym = ey_prev;
      ey_prev = ey;
      const bool bulk = jin1 && i > ib && i < ie;
      const bool src = jsrc && i >= s.sx0 && i < s.sx1;
      const bool c0 = jcap && i == s.port[0], c0b = jcap && i == s.port[0] + 1;
      const bool c1 = jcap && i == s.port[1], c1b = jcap && i == s.port[1] + 1;
      if (!(bulk || src || c0 || c0b || c1 || c1b)) continue;
      T v = Hz[idx];
      if (bulk) {
        const T dEx = Ex[idx] - Ex[idx - 1];
        const T dEy = ey - eym;
        v = v + db * (dEx * iky - dEy * s.ikh[0][i]);
        if (src) v -= jv;
        int dx, cx;
        const int sx = slab_h(i, ib, ie, N, dx, cx);
        if (sx >= 0) v -= db * psi_upd(s.psi[psi_id(F_HZ, 0, sx)] + (size_t)dx * ny + j, s.bh[cx], s.cho[cx], dEy);
        if (sy >= 0) v += db * psi_upd(s.psi[psi_id(F_HZ, 1, sy)] + (size_t)i * N + dy, s.bh[cy], s.cho[cy], dEx);
        Hz[idx] = v;
      } else if (src) {
        v -= jv;
        Hz[idx] = v;
      }
      const int c = j - s.wy0;
      if (c0) capture(s, tstep, 1, c, v);
      if (c0b) capture(s, tstep, 2, c, v);
      if (c1) capture(s, tstep, 4, c, v);
      if (c1b) capture(s, tstep, 5, c, v);
    }
  }
}

// -------------------------------------------------------------------- 3D ---
// 3D half-steps as a prefetching march.  A wave owns one (j row, 64 z-cells)
// strip of XC planes and walks it along x; the loads of plane i+1 are issued
// before plane i is updated, so every wave keeps a plane of loads in flight
// while it computes and stores.  The walk is fully unrolled over the tile's
// planes, so each plane's values land in registers of their own.  Each E
// plane record also holds Hy and Hz of plane i+1, which the next plane reuses
// as its own; the H half-step reuses Ey and Ez of plane i-1 from the previous
// record.  The row (y) values are wave-uniform; the plane's x-only
// coefficients are loaded once per tile, one plane per lane, and broadcast
// with readlane; the cb values are copied to registers first and the source
// and capture planes become bit masks over the tile (stores through the field
// pointers could otherwise alias the descriptor and force reloads).  With NT
// the values read only once (the updated component's own old value and its
// psi) are loaded non-temporally.

// Bits u of the tile's planes i0 + u (u < 64) that lie in [lo, hi).
__device__ __forceinline__ unsigned long long plane_bits(int i0, int lo, int hi) {
  const int a = min(max(lo - i0, 0), 64), b = min(max(hi - i0, 0), 64);
  if (b <= a) return 0ull;
  const unsigned long long upto = b >= 64 ? ~0ull : (1ull << b) - 1ull;
  return upto & ~((1ull << a) - 1ull);
}

template <typename T>
__device__ __forceinline__ T cb_of(T v0, T v1, T v2, T v3, unsigned m, int c) {
  const int q = (m >> (2 * c)) & 3;
  return q == 0 ? v0 : (q == 1 ? v1 : (q == 2 ? v2 : v3));
}

// What the E march loads for one cell of plane i.
template <typename T>
struct E3Plane {
  T ex, ey, ez, hx, hxk, hxj, hyk, hzj, hy1, hz1, qxy, qzy, qxz, qyz, qyx, qzx;
  unsigned m;   // material byte
};

template <typename T, int XC, bool NT>
__global__ __launch_bounds__(kThreads) void k_3d_e(const Slot<T>* __restrict__ slots,
                                                   const int4* __restrict__ tiles,
                                                   const int* __restrict__ counts, int tstep) {
  static_assert(XC <= kBX, "one plane's x-only coefficients per lane");
  constexpr unsigned es = sizeof(T);
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int nx = s.nx, ny = s.ny, nz = s.nz, N = s.N;
    const int j = tl.z + wave_id();
    if (j >= ny) continue;
    const int k = tl.w + lane, kc = min(k, nz - 1);   // lanes past nz load clamped and store nothing
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    T* __restrict__ Ex = s.f[F_EX];
    T* __restrict__ Ey = s.f[F_EY];
    T* __restrict__ Ez = s.f[F_EZ];
    const T* __restrict__ Hx = s.f[F_HX];
    const T* __restrict__ Hy = s.f[F_HY];
    const T* __restrict__ Hz = s.f[F_HZ];
    const uint8_t* __restrict__ mat = s.mat;
    const unsigned pl = (unsigned)ny * nz;
    const unsigned jk = (unsigned)j * nz + kc;
    const unsigned offk = (kc + 1 < nz) ? es : 0u;
    const unsigned offj = (j + 1 < ny) ? (unsigned)nz * es : 0u;
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dj = 0, cj = 0, dk = 0, ck = 0;
    const int sj = jin ? slab_e(j, jb, je, N, dj, cj) : -1;
    const int sk = kin ? slab_e(k, kb, ke, N, dk, ck) : -1;
    const T iky = jin ? ldg<T>(s.ike[1], j * es) : T(0);
    const T ikz = kin ? ldg<T>(s.ike[2], k * es) : T(0);
    const T bej = sj >= 0 ? ldg<T>(s.be, cj * es) : T(0), cej = sj >= 0 ? ldg<T>(s.ceo, cj * es) : T(0);
    const T bek = sk >= 0 ? ldg<T>(s.be, ck * es) : T(0), cek = sk >= 0 ? ldg<T>(s.ceo, ck * es) : T(0);
    T* __restrict__ pExy = sj >= 0 ? s.psi[psi_id(F_EX, 1, sj)] : nullptr;
    T* __restrict__ pEzy = sj >= 0 ? s.psi[psi_id(F_EZ, 1, sj)] : nullptr;
    // z- and x-slab psi: the low-side array, plus the side's byte distance
    T* __restrict__ pExz = s.psi[psi_id(F_EX, 2, 0)];
    T* __restrict__ pEyz = s.psi[psi_id(F_EY, 2, 0)];
    T* __restrict__ pEyx = s.psi[psi_id(F_EY, 0, 0)];
    T* __restrict__ pEzx = s.psi[psi_id(F_EZ, 0, 0)];
    const unsigned zside = sk == 1 ? s.psi_hi[2] : 0u, xhi = s.psi_hi[0];
    const bool jksrc = s.src_field == F_EY && j >= s.sy0 && j < s.sy1 && k >= s.sz0 && k < s.sz1;
    const T jv = jksrc ? ldg<T>(s.jsrc, n * es) : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    const int cap = !jcap ? -1 : (k == s.port[0] ? 0 : (k == s.port[1] ? 1 : -1));
    const int cap2 = (jcap && s.port[0] == s.port[1] && k == s.port[0]) ? 1 : -1;
    const T cx0 = s.cbv[0][0], cx1 = s.cbv[0][1], cx2 = s.cbv[0][2], cx3 = s.cbv[0][3];
    const T cy0 = s.cbv[1][0], cy1 = s.cbv[1][1], cy2 = s.cbv[1][2], cy3 = s.cbv[1][3];
    const T cz0 = s.cbv[2][0], cz1 = s.cbv[2][1], cz2 = s.cbv[2][2], cz3 = s.cbv[2][3];
    const int i0 = tl.y, iend = min(i0 + XC, nx);
    // planes of the tile inside the source box and the capture window
    const unsigned long long srcm = plane_bits(i0, s.sx0, s.sx1), capm = plane_bits(i0, s.wx0, s.wx1 + 1);
    // x-only coefficients: lane u holds plane i0 + u
    T lk_ikx = T(0), lk_bex = T(0), lk_cex = T(0);
    if (lane < XC && i0 + lane < nx) {
      const int i = i0 + lane;
      lk_ikx = ldg<T>(s.ike[0], i * es);
      int di = 0, ci = 0;
      if (slab_e(i, ib, ie, N, di, ci) >= 0) {
        lk_bex = ldg<T>(s.be, ci * es);
        lk_cex = ldg<T>(s.ceo, ci * es);
      }
    }
    auto load = [&](int i) {
      E3Plane<T> p;
      const unsigned idx = (unsigned)i * pl + jk;
      const unsigned bo = idx * es;
      const unsigned bo1 = (i + 1 < nx ? idx + pl : idx) * es;
      p.ex = ldx<NT, T>(Ex, bo);
      p.ey = ldx<NT, T>(Ey, bo);
      p.ez = ldx<NT, T>(Ez, bo);
      p.hx = ldg<T>(Hx, bo);
      p.hxk = ldg<T>(Hx, bo + offk);
      p.hxj = ldg<T>(Hx, bo + offj);
      p.hyk = ldg<T>(Hy, bo + offk);
      p.hzj = ldg<T>(Hz, bo + offj);
      p.hy1 = ldg<T>(Hy, bo1);
      p.hz1 = ldg<T>(Hz, bo1);
      p.m = ldg<uint8_t>(mat, idx);
      const unsigned iy = (((unsigned)i * N + dj) * nz + kc) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
      p.qxy = pExy ? ldx<NT, T>(pExy, iy) : T(0);
      p.qzy = pEzy ? ldx<NT, T>(pEzy, iy) : T(0);
      p.qxz = sk >= 0 ? ldx<NT, T>(pExz, iz + zside) : T(0);
      p.qyz = sk >= 0 ? ldx<NT, T>(pEyz, iz + zside) : T(0);
      int di = 0, ci = 0;
      const int si = slab_e(i, ib, ie, N, di, ci);
      const unsigned ix = (((unsigned)di * ny + j) * nz + kc) * es + (si == 1 ? xhi : 0u);
      p.qyx = si >= 0 ? ldx<NT, T>(pEyx, ix) : T(0);
      p.qzx = si >= 0 ? ldx<NT, T>(pEzx, ix) : T(0);
      return p;
    };
    const unsigned b0 = ((unsigned)i0 * pl + jk) * es;
    T hy0 = ldg<T>(Hy, b0), hz0 = ldg<T>(Hz, b0);   // Hy, Hz of the plane being updated
    auto update = [&](const int i, const E3Plane<T>& p0) {
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool bx = iin1 && jin && kin;
      const bool by = iin && jin1 && kin;
      const bool bz = iin && jin && kin1;
      const bool src = jksrc && ((srcm >> (i - i0)) & 1ull);
      const bool icap = cap >= 0 && ((capm >> (i - i0)) & 1ull);
      if (bx || by || bz || src || icap) {
        const int u = i - i0;
        int di = 0, ci = 0;
        const int si = slab_e(i, ib, ie, N, di, ci);
        const unsigned idx = (unsigned)i * pl + jk;
        const unsigned bo = idx * es;
        const unsigned iy = (((unsigned)i * N + dj) * nz + kc) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
        const unsigned ix = (((unsigned)di * ny + j) * nz + kc) * es + (si == 1 ? xhi : 0u);
        const T ikx = lane_get(lk_ikx, u);
        T vx = p0.ex, vy = p0.ey;
        // update_ex: i (ib,ie), j [jb,je), k [kb,ke); psi ex_y (+), ex_z (-)
        if (bx) {
          const T cb = cb_of(cx0, cx1, cx2, cx3, p0.m, 0);
          const T dHz = p0.hzj - hz0;
          const T dHy = p0.hyk - hy0;
          vx = vx + cb * (dHz * iky - dHy * ikz);
          if (sj >= 0) { const T q = bej * p0.qxy + cej * dHz; stg<T>(pExy, iy, q); vx += cb * q; }
          if (sk >= 0) { const T q = bek * p0.qxz + cek * dHy; stg<T>(pExz, iz + zside, q); vx -= cb * q; }
          stg_nt<T>(Ex, bo, vx);
        }
        // update_ey: i [ib,ie), j (jb,je), k [kb,ke); source (TE); psi ey_x (-), ey_z (+)
        if (by) {
          const T cb = cb_of(cy0, cy1, cy2, cy3, p0.m, 1);
          const T dHx = p0.hxk - p0.hx;
          const T dHz = p0.hz1 - hz0;
          vy = vy + cb * (dHx * ikz - dHz * ikx);
          if (src) vy -= jv;
          if (si >= 0) {
            const T q = lane_get(lk_bex, u) * p0.qyx + lane_get(lk_cex, u) * dHz;
            stg<T>(pEyx, ix, q);
            vy -= cb * q;
          }
          if (sk >= 0) { const T q = bek * p0.qyz + cek * dHx; stg<T>(pEyz, iz + zside, q); vy += cb * q; }
          stg_nt<T>(Ey, bo, vy);
        } else if (src) {
          vy -= jv;
          stg_nt<T>(Ey, bo, vy);
        }
        // update_ez: i [ib,ie), j [jb,je), k (kb,ke); psi ez_x (+), ez_y (-)
        if (bz) {
          const T cb = cb_of(cz0, cz1, cz2, cz3, p0.m, 2);
          const T dHy = p0.hy1 - hy0;
          const T dHx = p0.hxj - p0.hx;
          T v = p0.ez + cb * (dHy * ikx - dHx * iky);
          if (si >= 0) {
            const T q = lane_get(lk_bex, u) * p0.qzx + lane_get(lk_cex, u) * dHy;
            stg<T>(pEzx, ix, q);
            v += cb * q;
          }
          if (sj >= 0) { const T q = bej * p0.qzy + cej * dHx; stg<T>(pEzy, iy, q); v -= cb * q; }
          stg_nt<T>(Ez, bo, v);
        }
        if (icap) {
          const size_t c = (size_t)((i - s.wx0) * s.prow + (j - s.wy0));
          const size_t base = (size_t)tstep * s.nplanes;
          const size_t pcells = s.pcells;
          T* __restrict__ ring = s.ring;
          ring[(base + cap * 6 + 0) * pcells + c] = vx;
          ring[(base + cap * 6 + 1) * pcells + c] = vy;
          if (cap2 >= 0) {
            ring[(base + 6) * pcells + c] = vx;
            ring[(base + 7) * pcells + c] = vy;
          }
        }
      }
      hy0 = p0.hy1;
      hz0 = p0.hz1;
    };
    // Fully unrolled over the tile's planes, so the next plane's registers
    // are new ones (no copy waits for its loads).
    E3Plane<T> pc = load(i0);
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      const int i = i0 + u;
      if (i >= iend) break;
      E3Plane<T> pn;
      if (u + 1 < XC) pn = load(min(i + 1, iend - 1));
      update(i, pc);
      if (u + 1 < XC) pc = pn;
    }
  }
}

// What the H march loads for one cell of plane i (its Ey and Ez are the
// i-1 neighbours of plane i+1).
template <typename T>
struct H3Plane {
  T hx, hy, hz, ex, exk, exj, eyk, ezj, ey, ez, qxy, qzy, qxz, qyz, qyx, qzx;
};

template <typename T, int XC, bool NT>
__global__ __launch_bounds__(kThreads) void k_3d_h(const Slot<T>* __restrict__ slots,
                                                   const int4* __restrict__ tiles,
                                                   const int* __restrict__ counts, int tstep) {
  static_assert(XC <= kBX, "one plane's x-only coefficients per lane");
  constexpr unsigned es = sizeof(T);
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int nx = s.nx, ny = s.ny, nz = s.nz, N = s.N;
    const int j = tl.z + wave_id();
    if (j >= ny) continue;
    const int k = tl.w + lane, kc = min(k, nz - 1);   // lanes past nz load clamped and store nothing
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    const T* __restrict__ Ex = s.f[F_EX];
    const T* __restrict__ Ey = s.f[F_EY];
    const T* __restrict__ Ez = s.f[F_EZ];
    T* __restrict__ Hx = s.f[F_HX];
    T* __restrict__ Hy = s.f[F_HY];
    T* __restrict__ Hz = s.f[F_HZ];
    const T db = s.db;
    const unsigned pl = (unsigned)ny * nz;
    const unsigned jk = (unsigned)j * nz + kc;
    const unsigned offk = (kc > 0) ? es : 0u;
    const unsigned offj = (j > 0) ? (unsigned)nz * es : 0u;
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dj = 0, cj = 0, dk = 0, ck = 0;
    const int sj = jin1 ? slab_h(j, jb, je, N, dj, cj) : -1;
    const int sk = kin1 ? slab_h(k, kb, ke, N, dk, ck) : -1;
    const T iky = jin1 ? ldg<T>(s.ikh[1], j * es) : T(0);
    const T ikz = kin1 ? ldg<T>(s.ikh[2], k * es) : T(0);
    const T bhj = sj >= 0 ? ldg<T>(s.bh, cj * es) : T(0), chj = sj >= 0 ? ldg<T>(s.cho, cj * es) : T(0);
    const T bhk = sk >= 0 ? ldg<T>(s.bh, ck * es) : T(0), chk = sk >= 0 ? ldg<T>(s.cho, ck * es) : T(0);
    T* __restrict__ pHxy = sj >= 0 ? s.psi[psi_id(F_HX, 1, sj)] : nullptr;
    T* __restrict__ pHzy = sj >= 0 ? s.psi[psi_id(F_HZ, 1, sj)] : nullptr;
    // z- and x-slab psi: the low-side array, plus the side's byte distance
    T* __restrict__ pHxz = s.psi[psi_id(F_HX, 2, 0)];
    T* __restrict__ pHyz = s.psi[psi_id(F_HY, 2, 0)];
    T* __restrict__ pHyx = s.psi[psi_id(F_HY, 0, 0)];
    T* __restrict__ pHzx = s.psi[psi_id(F_HZ, 0, 0)];
    const unsigned zside = sk == 1 ? s.psi_hi[2] : 0u, xhi = s.psi_hi[0];
    // 3D TM source on Hy (fdtd_main.py:369-370), applied before the H curl
    const bool jksrc = s.src_field == F_HY && j >= s.sy0 && j < s.sy1 && k >= s.sz0 && k < s.sz1;
    const T jv = jksrc ? ldg<T>(s.jsrc, n * es) : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    // capture planes: Hx/Hy at kc (planes 2,4) and kc+1 (planes 3,5) of each port
    int capA = -1, capB = -1, capC = -1, capD = -1;
    if (jcap) {
      if (k == s.port[0]) capA = 0;
      if (k == s.port[0] + 1) capB = 0;
      if (k == s.port[1]) capC = 1;
      if (k == s.port[1] + 1) capD = 1;
    }
    const bool anycap = capA >= 0 || capB >= 0 || capC >= 0 || capD >= 0;
    const int i0 = tl.y, iend = min(i0 + XC, nx);
    // planes of the tile inside the source box and the capture window
    const unsigned long long srcm = plane_bits(i0, s.sx0, s.sx1), capm = plane_bits(i0, s.wx0, s.wx1 + 1);
    T lk_ikx = T(0), lk_bhx = T(0), lk_chx = T(0);
    if (lane < XC && i0 + lane < nx) {
      const int i = i0 + lane;
      lk_ikx = ldg<T>(s.ikh[0], i * es);
      int di = 0, ci = 0;
      if (slab_h(i, ib, ie, N, di, ci) >= 0) {
        lk_bhx = ldg<T>(s.bh, ci * es);
        lk_chx = ldg<T>(s.cho, ci * es);
      }
    }
    auto load = [&](int i) {
      H3Plane<T> p;
      const unsigned idx = (unsigned)i * pl + jk;
      const unsigned bo = idx * es;
      p.hx = ldx<NT, T>(Hx, bo);
      p.hy = ldx<NT, T>(Hy, bo);
      p.hz = ldx<NT, T>(Hz, bo);
      p.ex = ldg<T>(Ex, bo);
      p.exk = ldg<T>(Ex, bo - offk);
      p.exj = ldg<T>(Ex, bo - offj);
      p.eyk = ldg<T>(Ey, bo - offk);
      p.ezj = ldg<T>(Ez, bo - offj);
      p.ey = ldg<T>(Ey, bo);
      p.ez = ldg<T>(Ez, bo);
      const unsigned iy = (((unsigned)i * N + dj) * nz + kc) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
      p.qxy = pHxy ? ldx<NT, T>(pHxy, iy) : T(0);
      p.qzy = pHzy ? ldx<NT, T>(pHzy, iy) : T(0);
      p.qxz = sk >= 0 ? ldx<NT, T>(pHxz, iz + zside) : T(0);
      p.qyz = sk >= 0 ? ldx<NT, T>(pHyz, iz + zside) : T(0);
      int di = 0, ci = 0;
      const int si = slab_h(i, ib, ie, N, di, ci);
      const unsigned ix = (((unsigned)di * ny + j) * nz + kc) * es + (si == 1 ? xhi : 0u);
      p.qyx = si >= 0 ? ldx<NT, T>(pHyx, ix) : T(0);
      p.qzx = si >= 0 ? ldx<NT, T>(pHzx, ix) : T(0);
      return p;
    };
    const unsigned bm = ((unsigned)max(i0 - 1, 0) * pl + jk) * es;
    T eym = ldg<T>(Ey, bm), ezm = ldg<T>(Ez, bm);   // Ey, Ez of plane i-1
    auto update = [&](const int i, const H3Plane<T>& p0) {
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool bx = iin && jin1 && kin1;
      const bool by = iin1 && jin && kin1;
      const bool bz = iin1 && jin1 && kin;
      const bool src = jksrc && ((srcm >> (i - i0)) & 1ull);
      const bool icap = anycap && ((capm >> (i - i0)) & 1ull);
      if (bx || by || bz || src || icap) {
        const int u = i - i0;
        int di = 0, ci = 0;
        const int si = slab_h(i, ib, ie, N, di, ci);
        const unsigned idx = (unsigned)i * pl + jk;
        const unsigned bo = idx * es;
        const unsigned iy = (((unsigned)i * N + dj) * nz + kc) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
        const unsigned ix = (((unsigned)di * ny + j) * nz + kc) * es + (si == 1 ? xhi : 0u);
        const T ikx = lane_get(lk_ikx, u);
        // update_hx: i [ib,ie), j (jb,je), k (kb,ke); psi hx_y (-), hx_z (+)
        T vx = p0.hx;
        if (bx) {
          const T dEy = p0.ey - p0.eyk;
          const T dEz = p0.ez - p0.ezj;
          vx = vx + db * (dEy * ikz - dEz * iky);
          if (sj >= 0) { const T q = bhj * p0.qxy + chj * dEz; stg<T>(pHxy, iy, q); vx -= db * q; }
          if (sk >= 0) { const T q = bhk * p0.qxz + chk * dEy; stg<T>(pHxz, iz + zside, q); vx += db * q; }
          stg_nt<T>(Hx, bo, vx);
        }
        // update_hy: i (ib,ie), j [jb,je), k (kb,ke); psi hy_x (+), hy_z (-)
        T vy = p0.hy;
        if (src) vy -= jv;
        if (by) {
          const T dEz = p0.ez - ezm;
          const T dEx = p0.ex - p0.exk;
          vy = vy + db * (dEz * ikx - dEx * ikz);
          if (si >= 0) {
            const T q = lane_get(lk_bhx, u) * p0.qyx + lane_get(lk_chx, u) * dEz;
            stg<T>(pHyx, ix, q);
            vy += db * q;
          }
          if (sk >= 0) { const T q = bhk * p0.qyz + chk * dEx; stg<T>(pHyz, iz + zside, q); vy -= db * q; }
        }
        if (by || src) stg_nt<T>(Hy, bo, vy);
        // update_hz: i (ib,ie), j (jb,je), k [kb,ke); psi hz_x (-), hz_y (+)
        if (bz) {
          const T dEx = p0.ex - p0.exj;
          const T dEy = p0.ey - eym;
          T v = p0.hz + db * (dEx * iky - dEy * ikx);
          if (si >= 0) {
            const T q = lane_get(lk_bhx, u) * p0.qzx + lane_get(lk_chx, u) * dEy;
            stg<T>(pHzx, ix, q);
            v -= db * q;
          }
          if (sj >= 0) { const T q = bhj * p0.qzy + chj * dEx; stg<T>(pHzy, iy, q); v += db * q; }
          stg_nt<T>(Hz, bo, v);
        }
        if (icap) {
          const size_t c = (size_t)((i - s.wx0) * s.prow + (j - s.wy0));
          const size_t base = (size_t)tstep * s.nplanes;
          const size_t pcells = s.pcells;
          T* __restrict__ ring = s.ring;
          if (capA >= 0) { ring[(base + 2) * pcells + c] = vx; ring[(base + 4) * pcells + c] = vy; }
          if (capB >= 0) { ring[(base + 3) * pcells + c] = vx; ring[(base + 5) * pcells + c] = vy; }
          if (capC >= 0) { ring[(base + 8) * pcells + c] = vx; ring[(base + 10) * pcells + c] = vy; }
          if (capD >= 0) { ring[(base + 9) * pcells + c] = vx; ring[(base + 11) * pcells + c] = vy; }
        }
      }
      eym = p0.ey;
      ezm = p0.ez;
    };
    // Fully unrolled over the tile's planes, so the next plane's registers
    // are new ones (no copy waits for its loads).
    H3Plane<T> pc = load(i0);
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      const int i = i0 + u;
      if (i >= iend) break;
      H3Plane<T> pn;
      if (u + 1 < XC) pn = load(min(i + 1, iend - 1));
      update(i, pc);
      if (u + 1 < XC) pc = pn;
    }
  }
}


// ------------------------------------------------------ batched-load form ---
// Same per-cell arithmetic as the march kernels above, but every value a
// thread needs for its XC cells along x -- fields (clamped, unconditional),
// the material byte and the psi values of the CPML slabs it lies in -- is
// loaded before the first store.  A wave then keeps ~10*XC loads in flight
// instead of exposing one memory latency per cell (the compiler cannot move a
// later cell's loads above an earlier cell's stores through the run's field
// pointers), and no load depends on another load.

#define OIDT_TILE_PROLOGUE                                 \
  const int ntiles = counts[0];                            \
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) { \
    const int4 tl = tiles[tt];                             \
    const Slot<T>& s = slots[tl.x];                        \
    const int n = s.base + tstep;                          \
    if (!s.active || n >= s.nt) continue;

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tmz_e2(const Slot<T>* __restrict__ slots,
                                                     const int4* __restrict__ tiles,
                                                     const int* __restrict__ counts, int tstep) {
  OIDT_TILE_PROLOGUE
    const int nx = s.nx, ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    if (i0 >= nx) continue;
    const int cnt = min(XC, nx - i0);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    T* __restrict__ Ez = s.f[F_EZ];
    const T* __restrict__ Hx = s.f[F_HX];
    const T* __restrict__ Hy = s.f[F_HY];
    const bool jin = j >= jb && j < je;
    int dy = 0, cy = 0;
    const int sy = jin ? slab_e(j, jb, je, N, dy, cy) : -1;
    const T iky = jin ? s.ike[1][j] : T(0);
    const T bey = sy >= 0 ? s.be[cy] : T(0), cey = sy >= 0 ? s.ceo[cy] : T(0);
    T* __restrict__ py = sy >= 0 ? s.psi[psi_id(F_EZ, 1, sy)] : nullptr;
    const bool jsrc = s.src_field == F_EZ && j >= s.sy0 && j < s.sy1;
    const T jv = jsrc ? s.jsrc[n] : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    const unsigned offj = (j + 1 < ny) ? 1u : 0u;
    T ez[XC], hx[XC], hxj[XC], hy[XC + 1], pzy[XC], pzx[XC];
    uint8_t m[XC];
#pragma unroll
    for (int u = 0; u <= XC; ++u) {
      const int i = min(i0 + u, nx - 1);
      const unsigned idx = (unsigned)i * ny + j;
      hy[u] = Hy[idx];
      if (u < XC) {
        ez[u] = Ez[idx];
        hx[u] = Hx[idx];
        hxj[u] = Hx[idx + offj];
        m[u] = s.mat[idx];
        pzy[u] = py ? py[(size_t)i * N + dy] : T(0);
        int dx = 0, cx = 0;
        const int sx = slab_e(i, ib, ie, N, dx, cx);
        pzx[u] = sx >= 0 ? s.psi[psi_id(F_EZ, 0, sx)][(size_t)dx * ny + j] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      if (u >= cnt) break;
      const int i = i0 + u;
      const unsigned idx = (unsigned)i * ny + j;
      const bool bulk = jin && i >= ib && i < ie;
      const bool src = jsrc && i >= s.sx0 && i < s.sx1;
      const bool cap0 = jcap && i == s.port[0];
      const bool cap1 = jcap && i == s.port[1];
      if (!(bulk || src || cap0 || cap1)) continue;
      T v = ez[u];
      if (bulk) {
        const T cb = cbsel(s, m[u], 2);
        const T dHy = hy[u + 1] - hy[u];
        const T dHx = hxj[u] - hx[u];
        v = v + cb * (dHy * s.ike[0][i] - dHx * iky);
        if (src) v -= jv;
        int dx = 0, cx = 0;
        const int sx = slab_e(i, ib, ie, N, dx, cx);
        if (sx >= 0) {
          const T p = s.be[cx] * pzx[u] + s.ceo[cx] * dHy;
          s.psi[psi_id(F_EZ, 0, sx)][(size_t)dx * ny + j] = p;
          v += cb * p;
        }
        if (sy >= 0) {
          const T p = bey * pzy[u] + cey * dHx;
          py[(size_t)i * N + dy] = p;
          v -= cb * p;
        }
        Ez[idx] = v;
      } else if (src) {
        v -= jv;
        Ez[idx] = v;
      }
      if (cap0) capture(s, tstep, 0, j - s.wy0, v);
      if (cap1) capture(s, tstep, 3, j - s.wy0, v);
    }
  }
}

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tmz_h2(const Slot<T>* __restrict__ slots,
                                                     const int4* __restrict__ tiles,
                                                     const int* __restrict__ counts, int tstep) {
  OIDT_TILE_PROLOGUE
    const int nx = s.nx, ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    if (i0 >= nx) continue;
    const int cnt = min(XC, nx - i0);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    const T* __restrict__ Ez = s.f[F_EZ];
    T* __restrict__ Hx = s.f[F_HX];
    T* __restrict__ Hy = s.f[F_HY];
    const T db = s.db;
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    int dy = 0, cy = 0;
    const int sy = jin1 ? slab_h(j, jb, je, N, dy, cy) : -1;
    const T iky = jin1 ? s.ikh[1][j] : T(0);
    const T bhy = sy >= 0 ? s.bh[cy] : T(0), chy = sy >= 0 ? s.cho[cy] : T(0);
    T* __restrict__ py = sy >= 0 ? s.psi[psi_id(F_HX, 1, sy)] : nullptr;
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    const unsigned offj = (j > 0) ? 1u : 0u;
    T ez[XC + 1], ezj[XC], hx[XC], hy[XC], pxy[XC], pyx[XC];
#pragma unroll
    for (int u = 0; u <= XC; ++u) {
      ez[u] = Ez[(unsigned)min(max(i0 - 1 + u, 0), nx - 1) * ny + j];
      if (u < XC) {
        const int i = min(i0 + u, nx - 1);
        const unsigned idx = (unsigned)i * ny + j;
        ezj[u] = Ez[idx - offj];
        hx[u] = Hx[idx];
        hy[u] = Hy[idx];
        pxy[u] = py ? py[(size_t)i * N + dy] : T(0);
        int dx = 0, cx = 0;
        const int sx = slab_h(i, ib, ie, N, dx, cx);
        pyx[u] = sx >= 0 ? s.psi[psi_id(F_HY, 0, sx)][(size_t)dx * ny + j] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      if (u >= cnt) break;
      const int i = i0 + u;
      const unsigned idx = (unsigned)i * ny + j;
      const T ezc = ez[u + 1];
      if (i >= ib && i < ie && jin1) {
        const T dEz = ezc - ezj[u];
        T v = hx[u] + db * (-dEz) * iky;
        if (sy >= 0) {
          const T p = bhy * pxy[u] + chy * dEz;
          py[(size_t)i * N + dy] = p;
          v -= db * p;
        }
        Hx[idx] = v;
      }
      const bool bulk = i > ib && i < ie && jin;
      const bool jc = jcap;
      const bool c0 = jc && i == s.port[0], c0b = jc && i == s.port[0] + 1;
      const bool c1 = jc && i == s.port[1], c1b = jc && i == s.port[1] + 1;
      if (!(bulk || c0 || c0b || c1 || c1b)) continue;
      T v = hy[u];
      if (bulk) {
        const T dEz = ezc - ez[u];
        v = v + db * dEz * s.ikh[0][i];
        int dx = 0, cx = 0;
        const int sx = slab_h(i, ib, ie, N, dx, cx);
        if (sx >= 0) {
          const T p = s.bh[cx] * pyx[u] + s.cho[cx] * dEz;
          s.psi[psi_id(F_HY, 0, sx)][(size_t)dx * ny + j] = p;
          v += db * p;
        }
        Hy[idx] = v;
      }
      const int c = j - s.wy0;
      if (c0) capture(s, tstep, 1, c, v);
      if (c0b) capture(s, tstep, 2, c, v);
      if (c1) capture(s, tstep, 4, c, v);
      if (c1b) capture(s, tstep, 5, c, v);
    }
  }
}

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tez_e2(const Slot<T>* __restrict__ slots,
                                                     const int4* __restrict__ tiles,
                                                     const int* __restrict__ counts, int tstep) {
  OIDT_TILE_PROLOGUE
    const int nx = s.nx, ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    if (i0 >= nx) continue;
    const int cnt = min(XC, nx - i0);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    T* __restrict__ Ex = s.f[F_EX];
    T* __restrict__ Ey = s.f[F_EY];
    const T* __restrict__ Hz = s.f[F_HZ];
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    int dy = 0, cy = 0;
    const int sy = jin ? slab_e(j, jb, je, N, dy, cy) : -1;
    const T iky = jin ? s.ike[1][j] : T(0);
    const T bey = sy >= 0 ? s.be[cy] : T(0), cey = sy >= 0 ? s.ceo[cy] : T(0);
    T* __restrict__ py = sy >= 0 ? s.psi[psi_id(F_EX, 1, sy)] : nullptr;
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    const unsigned offj = (j + 1 < ny) ? 1u : 0u;
    T ex[XC], ey[XC], hz[XC + 1], hzj[XC], pxy[XC], pyx[XC];
    uint8_t m[XC];
#pragma unroll
    for (int u = 0; u <= XC; ++u) {
      const int i = min(i0 + u, nx - 1);
      const unsigned idx = (unsigned)i * ny + j;
      hz[u] = Hz[idx];
      if (u < XC) {
        ex[u] = Ex[idx];
        ey[u] = Ey[idx];
        hzj[u] = Hz[idx + offj];
        m[u] = s.mat[idx];
        pxy[u] = py ? py[(size_t)i * N + dy] : T(0);
        int dx = 0, cx = 0;
        const int sx = slab_e(i, ib, ie, N, dx, cx);
        pyx[u] = sx >= 0 ? s.psi[psi_id(F_EY, 0, sx)][(size_t)dx * ny + j] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      if (u >= cnt) break;
      const int i = i0 + u;
      const unsigned idx = (unsigned)i * ny + j;
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool bx = iin1 && jin, by = iin && jin1;
      const bool c0 = jcap && i == s.port[0], c1 = jcap && i == s.port[1];
      if (!(bx || by || c0 || c1)) continue;
      if (bx) {
        const T cb = cbsel(s, m[u], 0);
        const T dHz = hzj[u] - hz[u];
        T v = ex[u] + cb * dHz * iky;
        if (sy >= 0) {
          const T p = bey * pxy[u] + cey * dHz;
          py[(size_t)i * N + dy] = p;
          v += cb * p;
        }
        Ex[idx] = v;
      }
      T v = ey[u];
      if (by) {
        const T cb = cbsel(s, m[u], 1);
        const T dHz = hz[u + 1] - hz[u];
        v = v + cb * (-dHz) * s.ike[0][i];
        int dx = 0, cx = 0;
        const int sx = slab_e(i, ib, ie, N, dx, cx);
        if (sx >= 0) {
          const T p = s.be[cx] * pyx[u] + s.ceo[cx] * dHz;
          s.psi[psi_id(F_EY, 0, sx)][(size_t)dx * ny + j] = p;
          v -= cb * p;
        }
        Ey[idx] = v;
      }
      if (c0) capture(s, tstep, 0, j - s.wy0, v);
      if (c1) capture(s, tstep, 3, j - s.wy0, v);
    }
  }
}

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tez_h2(const Slot<T>* __restrict__ slots,
                                                     const int4* __restrict__ tiles,
                                                     const int* __restrict__ counts, int tstep) {
  OIDT_TILE_PROLOGUE
    const int nx = s.nx, ny = s.ny, N = s.N;
    const int j = tl.z + threadIdx.x;
    if (j >= ny) continue;
    const int i0 = tl.y + threadIdx.y * XC;
    if (i0 >= nx) continue;
    const int cnt = min(XC, nx - i0);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je;
    const T* __restrict__ Ex = s.f[F_EX];
    const T* __restrict__ Ey = s.f[F_EY];
    T* __restrict__ Hz = s.f[F_HZ];
    const T db = s.db;
    const bool jin1 = j > jb && j < je;
    int dy = 0, cy = 0;
    const int sy = jin1 ? slab_h(j, jb, je, N, dy, cy) : -1;
    const T iky = jin1 ? s.ikh[1][j] : T(0);
    const T bhy = sy >= 0 ? s.bh[cy] : T(0), chy = sy >= 0 ? s.cho[cy] : T(0);
    T* __restrict__ py = sy >= 0 ? s.psi[psi_id(F_HZ, 1, sy)] : nullptr;
    const bool jsrc = s.src_field == F_HZ && j >= s.sy0 && j < s.sy1;
    const T jv = jsrc ? s.jsrc[n] : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    const unsigned offj = (j > 0) ? 1u : 0u;
    T ey[XC + 1], ex[XC], exj[XC], hz[XC], pzx[XC], pzy[XC];
#pragma unroll
    for (int u = 0; u <= XC; ++u) {
      ey[u] = Ey[(unsigned)min(max(i0 - 1 + u, 0), nx - 1) * ny + j];
      if (u < XC) {
        const int i = min(i0 + u, nx - 1);
        const unsigned idx = (unsigned)i * ny + j;
        ex[u] = Ex[idx];
        exj[u] = Ex[idx - offj];
        hz[u] = Hz[idx];
        pzy[u] = py ? py[(size_t)i * N + dy] : T(0);
        int dx = 0, cx = 0;
        const int sx = slab_h(i, ib, ie, N, dx, cx);
        pzx[u] = sx >= 0 ? s.psi[psi_id(F_HZ, 0, sx)][(size_t)dx * ny + j] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      if (u >= cnt) break;
      const int i = i0 + u;
      const unsigned idx = (unsigned)i * ny + j;
      const bool bulk = jin1 && i > ib && i < ie;
      const bool src = jsrc && i >= s.sx0 && i < s.sx1;
      const bool c0 = jcap && i == s.port[0], c0b = jcap && i == s.port[0] + 1;
      const bool c1 = jcap && i == s.port[1], c1b = jcap && i == s.port[1] + 1;
      if (!(bulk || src || c0 || c0b || c1 || c1b)) continue;
      T v = hz[u];
      if (bulk) {
        const T dEx = ex[u] - exj[u];
        const T dEy = ey[u + 1] - ey[u];
        v = v + db * (dEx * iky - dEy * s.ikh[0][i]);
        if (src) v -= jv;
        int dx = 0, cx = 0;
        const int sx = slab_h(i, ib, ie, N, dx, cx);
        if (sx >= 0) {
          const T p = s.bh[cx] * pzx[u] + s.cho[cx] * dEy;
          s.psi[psi_id(F_HZ, 0, sx)][(size_t)dx * ny + j] = p;
          v -= db * p;
        }
        if (sy >= 0) {
          const T p = bhy * pzy[u] + chy * dEx;
          py[(size_t)i * N + dy] = p;
          v += db * p;
        }
        Hz[idx] = v;
      } else if (src) {
        v -= jv;
        Hz[idx] = v;
      }
      const int c = j - s.wy0;
      if (c0) capture(s, tstep, 1, c, v);
      if (c0b) capture(s, tstep, 2, c, v);
      if (c1) capture(s, tstep, 4, c, v);
      if (c1b) capture(s, tstep, 5, c, v);
    }
  }
}

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_3d_e2(const Slot<T>* __restrict__ slots,
                                                    const int4* __restrict__ tiles,
                                                    const int* __restrict__ counts, int tstep) {
  OIDT_TILE_PROLOGUE
    const int nx = s.nx, ny = s.ny, nz = s.nz, N = s.N;
    // a wave is one y row: j is wave-uniform (scalar registers)
    const int k = tl.w + threadIdx.x, j = tl.z + wave_id();
    if (j >= ny || k >= nz) continue;
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    T* __restrict__ Ex = s.f[F_EX];
    T* __restrict__ Ey = s.f[F_EY];
    T* __restrict__ Ez = s.f[F_EZ];
    const T* __restrict__ Hx = s.f[F_HX];
    const T* __restrict__ Hy = s.f[F_HY];
    const T* __restrict__ Hz = s.f[F_HZ];
    const uint8_t* __restrict__ mat = s.mat;
    const unsigned pl = (unsigned)ny * nz;
    const unsigned jk = (unsigned)j * nz + k;
    const unsigned offk = (k + 1 < nz) ? 1u : 0u;
    const unsigned offj = (j + 1 < ny) ? (unsigned)nz : 0u;
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dj = 0, cj = 0, dk = 0, ck = 0;
    const int sj = jin ? slab_e(j, jb, je, N, dj, cj) : -1;
    const int sk = kin ? slab_e(k, kb, ke, N, dk, ck) : -1;
    constexpr unsigned es = sizeof(T);
    const T iky = jin ? ldg<T>(s.ike[1], j * es) : T(0);
    const T ikz = kin ? ldg<T>(s.ike[2], k * es) : T(0);
    const T bej = sj >= 0 ? ldg<T>(s.be, cj * es) : T(0), cej = sj >= 0 ? ldg<T>(s.ceo, cj * es) : T(0);
    const T bek = sk >= 0 ? ldg<T>(s.be, ck * es) : T(0), cek = sk >= 0 ? ldg<T>(s.ceo, ck * es) : T(0);
    T* __restrict__ pExy = sj >= 0 ? s.psi[psi_id(F_EX, 1, sj)] : nullptr;
    T* __restrict__ pEzy = sj >= 0 ? s.psi[psi_id(F_EZ, 1, sj)] : nullptr;
    T* __restrict__ pExz = sk >= 0 ? s.psi[psi_id(F_EX, 2, sk)] : nullptr;
    T* __restrict__ pEyz = sk >= 0 ? s.psi[psi_id(F_EY, 2, sk)] : nullptr;
    const bool jksrc = s.src_field == F_EY && j >= s.sy0 && j < s.sy1 && k >= s.sz0 && k < s.sz1;
    const T jv = jksrc ? ldg<T>(s.jsrc, n * es) : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    const int cap = !jcap ? -1 : (k == s.port[0] ? 0 : (k == s.port[1] ? 1 : -1));
    const int cap2 = (jcap && s.port[0] == s.port[1] && k == s.port[0]) ? 1 : -1;
    const int i0 = tl.y;
    const int cnt = min(XC, nx - i0);
    T ex[XC], ey[XC], ez[XC], hx[XC], hxk[XC], hxj[XC], hyk[XC], hzj[XC], hy[XC + 1], hz[XC + 1];
    T qxy[XC], qzy[XC], qxz[XC], qyz[XC], qyx[XC], qzx[XC];
    uint8_t m[XC];
#pragma unroll
    for (int u = 0; u <= XC; ++u) {
      const int i = min(i0 + u, nx - 1);
      const unsigned idx = (unsigned)i * pl + jk;
      const unsigned bo = idx * es;
      hy[u] = ldg<T>(Hy, bo);
      hz[u] = ldg<T>(Hz, bo);
      if (u < XC) {
        ex[u] = ldg<T>(Ex, bo);
        ey[u] = ldg<T>(Ey, bo);
        ez[u] = ldg<T>(Ez, bo);
        hx[u] = ldg<T>(Hx, bo);
        hxk[u] = ldg<T>(Hx, bo + offk * es);
        hxj[u] = ldg<T>(Hx, bo + offj * es);
        hyk[u] = ldg<T>(Hy, bo + offk * es);
        hzj[u] = ldg<T>(Hz, bo + offj * es);
        m[u] = ldg<uint8_t>(mat, idx);
        const unsigned iy = (((unsigned)i * N + dj) * nz + k) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
        qxy[u] = pExy ? ldg<T>(pExy, iy) : T(0);
        qzy[u] = pEzy ? ldg<T>(pEzy, iy) : T(0);
        qxz[u] = pExz ? ldg<T>(pExz, iz) : T(0);
        qyz[u] = pEyz ? ldg<T>(pEyz, iz) : T(0);
        int di = 0, ci = 0;
        const int si = slab_e(i, ib, ie, N, di, ci);
        const unsigned ix = (((unsigned)di * ny + j) * nz + k) * es;
        qyx[u] = si >= 0 ? ldg<T>(s.psi[psi_id(F_EY, 0, si)], ix) : T(0);
        qzx[u] = si >= 0 ? ldg<T>(s.psi[psi_id(F_EZ, 0, si)], ix) : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      if (u >= cnt) break;
      const int i = i0 + u;
      const unsigned idx = (unsigned)i * pl + jk;
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool bx = iin1 && jin && kin;
      const bool by = iin && jin1 && kin;
      const bool bz = iin && jin && kin1;
      const bool src = jksrc && i >= s.sx0 && i < s.sx1;
      const bool icap = cap >= 0 && i >= s.wx0 && i <= s.wx1;
      if (!(bx || by || bz || src || icap)) continue;
      int di = 0, ci = 0;
      const int si = slab_e(i, ib, ie, N, di, ci);
      const unsigned iy = (((unsigned)i * N + dj) * nz + k) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
      const unsigned ix = (((unsigned)di * ny + j) * nz + k) * es;
      const unsigned bo = idx * es;
      const T ikx = ldg<T>(s.ike[0], i * es);
      T vx = ex[u], vy = ey[u];
      if (bx) {
        const T cb = cbsel(s, m[u], 0);
        const T dHz = hzj[u] - hz[u];
        const T dHy = hyk[u] - hy[u];
        vx = vx + cb * (dHz * iky - dHy * ikz);
        if (sj >= 0) { const T p = bej * qxy[u] + cej * dHz; stg<T>(pExy, iy, p); vx += cb * p; }
        if (sk >= 0) { const T p = bek * qxz[u] + cek * dHy; stg<T>(pExz, iz, p); vx -= cb * p; }
        stg_nt<T>(Ex, bo, vx);
      }
      if (by) {
        const T cb = cbsel(s, m[u], 1);
        const T dHx = hxk[u] - hx[u];
        const T dHz = hz[u + 1] - hz[u];
        vy = vy + cb * (dHx * ikz - dHz * ikx);
        if (src) vy -= jv;
        if (si >= 0) {
          const T p = ldg<T>(s.be, ci * es) * qyx[u] + ldg<T>(s.ceo, ci * es) * dHz;
          stg<T>(s.psi[psi_id(F_EY, 0, si)], ix, p);
          vy -= cb * p;
        }
        if (sk >= 0) { const T p = bek * qyz[u] + cek * dHx; stg<T>(pEyz, iz, p); vy += cb * p; }
        stg_nt<T>(Ey, bo, vy);
      } else if (src) {
        vy -= jv;
        stg_nt<T>(Ey, bo, vy);
      }
      if (bz) {
        const T cb = cbsel(s, m[u], 2);
        const T dHy = hy[u + 1] - hy[u];
        const T dHx = hxj[u] - hx[u];
        T v = ez[u] + cb * (dHy * ikx - dHx * iky);
        if (si >= 0) {
          const T p = ldg<T>(s.be, ci * es) * qzx[u] + ldg<T>(s.ceo, ci * es) * dHy;
          stg<T>(s.psi[psi_id(F_EZ, 0, si)], ix, p);
          v += cb * p;
        }
        if (sj >= 0) { const T p = bej * qzy[u] + cej * dHx; stg<T>(pEzy, iy, p); v -= cb * p; }
        stg_nt<T>(Ez, bo, v);
      }
      if (icap) {
        const int c = (i - s.wx0) * s.prow + (j - s.wy0);
        capture(s, tstep, cap * 6 + 0, c, vx);
        capture(s, tstep, cap * 6 + 1, c, vy);
        if (cap2 >= 0) {
          capture(s, tstep, 6, c, vx);
          capture(s, tstep, 7, c, vy);
        }
      }
    }
  }
}

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_3d_h2(const Slot<T>* __restrict__ slots,
                                                    const int4* __restrict__ tiles,
                                                    const int* __restrict__ counts, int tstep) {
  OIDT_TILE_PROLOGUE
    const int nx = s.nx, ny = s.ny, nz = s.nz, N = s.N;
    // a wave is one y row: j is wave-uniform (scalar registers)
    const int k = tl.w + threadIdx.x, j = tl.z + wave_id();
    if (j >= ny || k >= nz) continue;
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    const T* __restrict__ Ex = s.f[F_EX];
    const T* __restrict__ Ey = s.f[F_EY];
    const T* __restrict__ Ez = s.f[F_EZ];
    T* __restrict__ Hx = s.f[F_HX];
    T* __restrict__ Hy = s.f[F_HY];
    T* __restrict__ Hz = s.f[F_HZ];
    const T db = s.db;
    const unsigned pl = (unsigned)ny * nz;
    const unsigned jk = (unsigned)j * nz + k;
    const unsigned offk = (k > 0) ? 1u : 0u;
    const unsigned offj = (j > 0) ? (unsigned)nz : 0u;
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dj = 0, cj = 0, dk = 0, ck = 0;
    const int sj = jin1 ? slab_h(j, jb, je, N, dj, cj) : -1;
    const int sk = kin1 ? slab_h(k, kb, ke, N, dk, ck) : -1;
    constexpr unsigned es = sizeof(T);
    const T iky = jin1 ? ldg<T>(s.ikh[1], j * es) : T(0);
    const T ikz = kin1 ? ldg<T>(s.ikh[2], k * es) : T(0);
    const T bhj = sj >= 0 ? ldg<T>(s.bh, cj * es) : T(0), chj = sj >= 0 ? ldg<T>(s.cho, cj * es) : T(0);
    const T bhk = sk >= 0 ? ldg<T>(s.bh, ck * es) : T(0), chk = sk >= 0 ? ldg<T>(s.cho, ck * es) : T(0);
    T* __restrict__ pHxy = sj >= 0 ? s.psi[psi_id(F_HX, 1, sj)] : nullptr;
    T* __restrict__ pHzy = sj >= 0 ? s.psi[psi_id(F_HZ, 1, sj)] : nullptr;
    T* __restrict__ pHxz = sk >= 0 ? s.psi[psi_id(F_HX, 2, sk)] : nullptr;
    T* __restrict__ pHyz = sk >= 0 ? s.psi[psi_id(F_HY, 2, sk)] : nullptr;
    const bool jksrc = s.src_field == F_HY && j >= s.sy0 && j < s.sy1 && k >= s.sz0 && k < s.sz1;
    const T jv = jksrc ? ldg<T>(s.jsrc, n * es) : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    int capA = -1, capB = -1, capC = -1, capD = -1;
    if (jcap) {
      if (k == s.port[0]) capA = 0;
      if (k == s.port[0] + 1) capB = 0;
      if (k == s.port[1]) capC = 1;
      if (k == s.port[1] + 1) capD = 1;
    }
    const bool anycap = capA >= 0 || capB >= 0 || capC >= 0 || capD >= 0;
    const int i0 = tl.y;
    const int cnt = min(XC, nx - i0);
    // ey[u], ez[u] hold plane i0-1+u; the other arrays plane i0+u
    T hx[XC], hy[XC], hz[XC], ex[XC], exk[XC], exj[XC], eyk[XC], ezj[XC], ey[XC + 1], ez[XC + 1];
    T qxy[XC], qzy[XC], qxz[XC], qyz[XC], qyx[XC], qzx[XC];
#pragma unroll
    for (int u = 0; u <= XC; ++u) {
      const unsigned pidx = (unsigned)min(max(i0 - 1 + u, 0), nx - 1) * pl + jk;
      ey[u] = ldg<T>(Ey, pidx * es);
      ez[u] = ldg<T>(Ez, pidx * es);
      if (u < XC) {
        const int i = min(i0 + u, nx - 1);
        const unsigned idx = (unsigned)i * pl + jk;
        const unsigned bo = idx * es;
        hx[u] = ldg<T>(Hx, bo);
        hy[u] = ldg<T>(Hy, bo);
        hz[u] = ldg<T>(Hz, bo);
        ex[u] = ldg<T>(Ex, bo);
        exk[u] = ldg<T>(Ex, bo - offk * es);
        exj[u] = ldg<T>(Ex, bo - offj * es);
        eyk[u] = ldg<T>(Ey, bo - offk * es);
        ezj[u] = ldg<T>(Ez, bo - offj * es);
        const unsigned iy = (((unsigned)i * N + dj) * nz + k) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
        qxy[u] = pHxy ? ldg<T>(pHxy, iy) : T(0);
        qzy[u] = pHzy ? ldg<T>(pHzy, iy) : T(0);
        qxz[u] = pHxz ? ldg<T>(pHxz, iz) : T(0);
        qyz[u] = pHyz ? ldg<T>(pHyz, iz) : T(0);
        int di = 0, ci = 0;
        const int si = slab_h(i, ib, ie, N, di, ci);
        const unsigned ix = (((unsigned)di * ny + j) * nz + k) * es;
        qyx[u] = si >= 0 ? ldg<T>(s.psi[psi_id(F_HY, 0, si)], ix) : T(0);
        qzx[u] = si >= 0 ? ldg<T>(s.psi[psi_id(F_HZ, 0, si)], ix) : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < XC; ++u) {
      if (u >= cnt) break;
      const int i = i0 + u;
      const unsigned idx = (unsigned)i * pl + jk;
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool bx = iin && jin1 && kin1;
      const bool by = iin1 && jin && kin1;
      const bool bz = iin1 && jin1 && kin;
      const bool src = jksrc && i >= s.sx0 && i < s.sx1;
      const bool icap = anycap && i >= s.wx0 && i <= s.wx1;
      if (!(bx || by || bz || src || icap)) continue;
      int di = 0, ci = 0;
      const int si = slab_h(i, ib, ie, N, di, ci);
      const unsigned iy = (((unsigned)i * N + dj) * nz + k) * es, iz = (((unsigned)i * ny + j) * N + dk) * es;
      const unsigned ix = (((unsigned)di * ny + j) * nz + k) * es;
      const unsigned bo = idx * es;
      const T ikx = ldg<T>(s.ikh[0], i * es);
      const T eyc = ey[u + 1], ezc = ez[u + 1];
      T vx = hx[u];
      if (bx) {
        const T dEy = eyc - eyk[u];
        const T dEz = ezc - ezj[u];
        vx = vx + db * (dEy * ikz - dEz * iky);
        if (sj >= 0) { const T p = bhj * qxy[u] + chj * dEz; stg<T>(pHxy, iy, p); vx -= db * p; }
        if (sk >= 0) { const T p = bhk * qxz[u] + chk * dEy; stg<T>(pHxz, iz, p); vx += db * p; }
        stg_nt<T>(Hx, bo, vx);
      }
      T vy = hy[u];
      if (src) vy -= jv;
      if (by) {
        const T dEz = ezc - ez[u];
        const T dEx = ex[u] - exk[u];
        vy = vy + db * (dEz * ikx - dEx * ikz);
        if (si >= 0) {
          const T p = ldg<T>(s.bh, ci * es) * qyx[u] + ldg<T>(s.cho, ci * es) * dEz;
          stg<T>(s.psi[psi_id(F_HY, 0, si)], ix, p);
          vy += db * p;
        }
        if (sk >= 0) { const T p = bhk * qyz[u] + chk * dEx; stg<T>(pHyz, iz, p); vy -= db * p; }
      }
      if (by || src) stg_nt<T>(Hy, bo, vy);
      if (bz) {
        const T dEx = ex[u] - exj[u];
        const T dEy = eyc - ey[u];
        T v = hz[u] + db * (dEx * iky - dEy * ikx);
        if (si >= 0) {
          const T p = ldg<T>(s.bh, ci * es) * qzx[u] + ldg<T>(s.cho, ci * es) * dEy;
          stg<T>(s.psi[psi_id(F_HZ, 0, si)], ix, p);
          v -= db * p;
        }
        if (sj >= 0) { const T p = bhj * qzy[u] + chj * dEx; stg<T>(pHzy, iy, p); v += db * p; }
        stg_nt<T>(Hz, bo, v);
      }
      if (icap) {
        const int c = (i - s.wx0) * s.prow + (j - s.wy0);
        if (capA >= 0) { capture(s, tstep, 2, c, vx); capture(s, tstep, 4, c, vy); }
        if (capB >= 0) { capture(s, tstep, 3, c, vx); capture(s, tstep, 5, c, vy); }
        if (capC >= 0) { capture(s, tstep, 8, c, vx); capture(s, tstep, 10, c, vy); }
        if (capD >= 0) { capture(s, tstep, 9, c, vx); capture(s, tstep, 11, c, vy); }
      }
    }
  }
}

#undef OIDT_TILE_PROLOGUE


// ---------------------------------------------------------- fused Yee step ---
// One launch per time step for TMz: the E and H half-steps together, reading
// field/psi set (n & 1) and writing the other set, so no tile races another.
// Each wavefront owns a tile of 64 rows (y, one per lane) x XM planes (x) and
// marches along x: the new Ez of the previous plane stays in a register (the
// i-1 neighbour of Hy), the j-1 neighbour of Hx comes from the next-lower
// lane, and every load of plane u+1 (fields, psi, coefficients) is issued
// before the stores of plane u.  Before the march the wave computes, without
// storing, the plane below its tile (i0-1) and the XM cells of the row below
// it (j0-1, one per lane) with the same per-cell code as their owners, so
// they are bit-identical to them.  Waves never wait for each other (no LDS,
// no barriers); per cell and step each field is read once and written once.
// Field writes are non-temporal: nothing reads them again within the step
// (measured 4 % faster on the 64-run ensemble).
// The run's descriptor fields are copied to registers per tile (TmzView):
// stores through the field pointers could otherwise alias them and force
// reloads after every store.

template <typename T>
struct TmzView {
  int nx, ny, N, ib, ie, jb, je, sx0, sx1, sy0, sy1, port0, port1, wy0, wy1, nplanes, pcells;
  bool src_ez;
  T db, cb[kMatMax];
  const T *ike0, *ike1, *ikh0, *ikh1, *be, *ceo, *bh, *cho;
  const uint8_t* mat;
  T* ring;
};

template <typename T>
__device__ __forceinline__ TmzView<T> tmz_view(const Slot<T>& s) {
  TmzView<T> v;
  v.nx = s.nx; v.ny = s.ny; v.N = s.N;
  v.ib = s.ib; v.ie = s.ie; v.jb = s.jb; v.je = s.je;
  v.sx0 = s.sx0; v.sx1 = s.sx1; v.sy0 = s.sy0; v.sy1 = s.sy1;
  v.port0 = s.port[0]; v.port1 = s.port[1]; v.wy0 = s.wy0; v.wy1 = s.wy1;
  v.nplanes = s.nplanes; v.pcells = s.pcells;
  v.src_ez = s.src_field == F_EZ;
  v.db = s.db;
#pragma unroll
  for (int q = 0; q < kMatMax; ++q) v.cb[q] = s.cbv[2][q];
  v.ike0 = s.ike[0]; v.ike1 = s.ike[1]; v.ikh0 = s.ikh[0]; v.ikh1 = s.ikh[1];
  v.be = s.be; v.ceo = s.ceo; v.bh = s.bh; v.cho = s.cho;
  v.mat = s.mat;
  v.ring = s.ring;
  return v;
}

// Where a TMz cell's Ez is updated (index ranges fdtd_funcs.py:44-80 and the
// CPML slabs :82-100, FDTD_2D_TM).
struct TmzWhere {
  bool bulk, src;
  int sx, sy, dx, dy, cx, cy;
};

template <typename T>
__device__ __forceinline__ TmzWhere tmz_where(const TmzView<T>& v, int i, int j) {
  TmzWhere f;
  const bool in = i >= 0 && j >= 0 && i < v.nx && j < v.ny;
  f.bulk = in && i >= v.ib && i < v.ie && j >= v.jb && j < v.je;
  f.src = in && v.src_ez && i >= v.sx0 && i < v.sx1 && j >= v.sy0 && j < v.sy1;
  f.dx = f.dy = f.cx = f.cy = 0;
  f.sx = f.bulk ? slab_e(i, v.ib, v.ie, v.N, f.dx, f.cx) : -1;
  f.sy = f.bulk ? slab_e(j, v.jb, v.je, v.N, f.dy, f.cy) : -1;
  return f;
}

// Everything the new Ez of one cell reads from memory (indexes clamped into
// the grid; coefficients only where the cell uses them).  Marching callers
// skip hy: it is the previous plane's hyi.
template <typename T>
struct TmzIn {
  T ez, hx, hxj, hy, hyi, pzx, pzy, ikx, bex, cex;
  uint8_t m;
};

// Coefficients of the Ez update that depend on y only (fixed over a march).
template <typename T>
struct TmzJ {
  T iky, bey, cey;
};

template <typename T>
__device__ __forceinline__ TmzJ<T> tmz_jcoef(const TmzView<T>& v, const TmzWhere& f, int j) {
  TmzJ<T> c;
  const int jc = min(max(j, 0), v.ny - 1);
  c.iky = f.bulk ? v.ike1[jc] : T(0);
  c.bey = f.sy >= 0 ? v.be[f.cy] : T(0);
  c.cey = f.sy >= 0 ? v.ceo[f.cy] : T(0);
  return c;
}

template <typename T>
__device__ __forceinline__ TmzIn<T> tmz_load(const TmzView<T>& v, int i, int j, const TmzWhere& f,
                                             const T* __restrict__ EzA, const T* __restrict__ HxA,
                                             const T* __restrict__ HyA, OIDT_GLOBAL T* const* psiA, bool load_hy,
                                             const T* pzy_lane = nullptr) {
  TmzIn<T> a;
  const int nx = v.nx, ny = v.ny;
  const int ic = min(max(i, 0), nx - 1), jc = min(max(j, 0), ny - 1);
  const unsigned idx = (unsigned)ic * ny + jc;
  a.ez = EzA[idx];
  a.hx = HxA[idx];
  a.hxj = HxA[jc + 1 < ny ? idx + 1 : idx];
  a.hyi = HyA[ic + 1 < nx ? idx + ny : idx];
  a.hy = load_hy ? HyA[idx] : T(0);
  a.m = v.mat[idx];
  a.ikx = f.bulk ? v.ike0[ic] : T(0);
  a.pzx = f.sx >= 0 ? psiA[psi_id(F_EZ, 0, f.sx)][(size_t)f.dx * ny + jc] : T(0);
  a.bex = f.sx >= 0 ? v.be[f.cx] : T(0);
  a.cex = f.sx >= 0 ? v.ceo[f.cx] : T(0);
  a.pzy = f.sy >= 0 ? (pzy_lane ? pzy_lane : psiA[psi_id(F_EZ, 1, f.sy)])[(size_t)ic * v.N + f.dy] : T(0);
  return a;
}

template <typename T>
struct TmzE {
  T ez, pzx, pzy;
};

// New Ez in the reference order: bulk curl, source, psi x, psi y.
template <typename T>
__device__ __forceinline__ TmzE<T> tmz_e(const TmzView<T>& v, const TmzWhere& f, const TmzIn<T>& a,
                                         const TmzJ<T>& c, T jv) {
  TmzE<T> r;
  r.ez = a.ez;
  r.pzx = r.pzy = T(0);
  if (f.bulk) {
    const int q = (a.m >> 4) & 3;
    const T cb = q == 0 ? v.cb[0] : (q == 1 ? v.cb[1] : (q == 2 ? v.cb[2] : v.cb[3]));
    const T dHy = a.hyi - a.hy;
    const T dHx = a.hxj - a.hx;
    T e = a.ez + cb * (dHy * a.ikx - dHx * c.iky);
    if (f.src) e -= jv;
    if (f.sx >= 0) {
      r.pzx = a.bex * a.pzx + a.cex * dHy;
      e += cb * r.pzx;
    }
    if (f.sy >= 0) {
      r.pzy = c.bey * a.pzy + c.cey * dHx;
      e -= cb * r.pzy;
    }
    r.ez = e;
  } else if (f.src) {
    r.ez = a.ez - jv;
  }
  return r;
}

template <typename T>
__device__ __forceinline__ T tmz_ez_cell(const TmzView<T>& v, int i, int j, T jv, const T* __restrict__ EzA,
                                         const T* __restrict__ HxA, const T* __restrict__ HyA, OIDT_GLOBAL T* const* psiA) {
  const TmzWhere f = tmz_where(v, i, j);
  return tmz_e(v, f, tmz_load(v, i, j, f, EzA, HxA, HyA, psiA, true), tmz_jcoef(v, f, j), jv).ez;
}

// What the march loads for one plane x = i: the nine values each cell reads.
// The plane's x-only coefficients are loaded once per tile, one plane per
// lane, and broadcast with readlane when the plane is updated.
template <typename T>
struct TmzPlane {
  T ez, hx, hxj, hyi, pzx, pzy, qxy, qyx;
  uint8_t m;
};

// x-only flags of plane i (recomputed where used rather than carried).
struct TmzX {
  bool ibulk, isrc, hbulk;
  int sxe, dxe, cxe, sxh, dxh, cxh;
};

template <typename T>
__device__ __forceinline__ TmzX tmz_x(const TmzView<T>& v, int i) {
  TmzX x;
  x.ibulk = i >= v.ib && i < v.ie;
  x.isrc = i >= v.sx0 && i < v.sx1;
  x.hbulk = i > v.ib && i < v.ie;
  x.dxe = x.cxe = x.dxh = x.cxh = 0;
  x.sxe = x.ibulk ? slab_e(i, v.ib, v.ie, v.N, x.dxe, x.cxe) : -1;
  x.sxh = x.hbulk ? slab_h(i, v.ib, v.ie, v.N, x.dxh, x.cxh) : -1;
  return x;
}

template <typename T, int XM>
__global__ __launch_bounds__(kThreads) void k_tmz_step(const Slot<T>* __restrict__ slots,
                                                       const int4* __restrict__ tiles,
                                                       const int* __restrict__ counts, int tstep) {
  static_assert(XM <= kBX, "the j0-1 halo row takes one cell per lane");
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  const int stride = gridDim.x * kBY;
  for (int tt = xcd_block() * kBY + wave_id(); tt < ntiles; tt += stride) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const TmzView<T> v = tmz_view(s);
    const bool odd = (n & 1) != 0;
    const T* __restrict__ EzA = odd ? s.f2[F_EZ] : s.f[F_EZ];
    const T* __restrict__ HxA = odd ? s.f2[F_HX] : s.f[F_HX];
    const T* __restrict__ HyA = odd ? s.f2[F_HY] : s.f[F_HY];
    T* __restrict__ EzB = odd ? s.f[F_EZ] : s.f2[F_EZ];
    T* __restrict__ HxB = odd ? s.f[F_HX] : s.f2[F_HX];
    T* __restrict__ HyB = odd ? s.f[F_HY] : s.f2[F_HY];
    OIDT_GLOBAL T* const* psiAll = odd ? s.psi2 : s.psi;
    OIDT_GLOBAL T* const* psiBll = odd ? s.psi : s.psi2;
    const T jv = v.src_ez ? s.jsrc[n] : T(0);
    const int nx = v.nx, ny = v.ny, N = v.N;
    const int i0 = tl.y, j0 = tl.z;
    const int j = j0 + lane;
    const bool jok = j < ny;
    const int jc = min(j, ny - 1);
    // row (y-only) constants of this lane, fixed over the march
    const bool jbulk = jok && j >= v.jb && j < v.je;          // Ez and Hy rows
    const bool jin1 = jok && j > v.jb && j < v.je;            // Hx rows
    const bool jsrc = jok && v.src_ez && j >= v.sy0 && j < v.sy1;
    int dye = 0, cye = 0, dyh = 0, cyh = 0;
    const int sye = jbulk ? slab_e(j, v.jb, v.je, N, dye, cye) : -1;
    const int syh = jin1 ? slab_h(j, v.jb, v.je, N, dyh, cyh) : -1;
    TmzJ<T> jcf;
    jcf.iky = jbulk ? v.ike1[j] : T(0);
    jcf.bey = sye >= 0 ? v.be[cye] : T(0);
    jcf.cey = sye >= 0 ? v.ceo[cye] : T(0);
    const T ikyh = jin1 ? v.ikh1[j] : T(0);
    const T bhy = syh >= 0 ? v.bh[cyh] : T(0), chy = syh >= 0 ? v.cho[cyh] : T(0);
    const bool jcap = jok && j >= v.wy0 && j <= v.wy1;
    // y-slab psi arrays of this lane's row; the x-slab arrays are looked up
    // in the few planes that need them
    const T* pzyA = sye >= 0 ? psiAll[psi_id(F_EZ, 1, sye)] : nullptr;
    T* pzyB = sye >= 0 ? psiBll[psi_id(F_EZ, 1, sye)] : nullptr;
    const T* pxyA = syh >= 0 ? psiAll[psi_id(F_HX, 1, syh)] : nullptr;
    T* pxyB = syh >= 0 ? psiBll[psi_id(F_HX, 1, syh)] : nullptr;
    const int iend = min(i0 + XM, nx);
    auto load_plane = [&](int i) {
      TmzPlane<T> p;
      const TmzX x = tmz_x(v, i);
      const unsigned idx = (unsigned)i * ny + jc;
      constexpr unsigned es = sizeof(T);
      p.ez = ldg<T>(EzA, idx * es);
      p.hx = ldg<T>(HxA, idx * es);
      p.hxj = ldg<T>(HxA, (jc + 1 < ny ? idx + 1 : idx) * es);
      p.hyi = ldg<T>(HyA, (i + 1 < nx ? idx + ny : idx) * es);
      p.m = ldg<uint8_t>(v.mat, idx);
      const bool bulk = x.ibulk && jbulk;
      p.pzx = bulk && x.sxe >= 0 ? psiAll[psi_id(F_EZ, 0, x.sxe)][(size_t)x.dxe * ny + jc] : T(0);
      p.pzy = bulk && sye >= 0 ? pzyA[(size_t)i * N + dye] : T(0);
      p.qxy = syh >= 0 ? pxyA[(size_t)i * N + dyh] : T(0);
      p.qyx = jbulk && x.sxh >= 0 ? psiAll[psi_id(F_HY, 0, x.sxh)][(size_t)x.dxh * ny + jc] : T(0);
      return p;
    };
    // ---- x-only coefficients of the tile's planes: lane u holds plane i0+u
    T lk_ikx = T(0), lk_bex = T(0), lk_cex = T(0), lk_ikhx = T(0), lk_bhx = T(0), lk_chx = T(0);
    if (lane < XM && i0 + lane < nx) {
      const TmzX x = tmz_x(v, i0 + lane);
      lk_ikx = x.ibulk ? v.ike0[i0 + lane] : T(0);
      lk_bex = x.sxe >= 0 ? v.be[x.cxe] : T(0);
      lk_cex = x.sxe >= 0 ? v.ceo[x.cxe] : T(0);
      lk_ikhx = v.ikh0[i0 + lane];
      lk_bhx = x.sxh >= 0 ? v.bh[x.cxh] : T(0);
      lk_chx = x.sxh >= 0 ? v.cho[x.cxh] : T(0);
    }
    // ---- halos from set A, not stored: row j0-1 (lane u holds plane i0+u) and plane i0-1
    const T colE = lane < XM ? tmz_ez_cell(v, i0 + lane, j0 - 1, jv, EzA, HxA, HyA, psiAll) : T(0);
    T prevE = tmz_ez_cell(v, i0 - 1, j, jv, EzA, HxA, HyA, psiAll);
    // ---- march along x.  Loads run two planes ahead of the updates (planes
    // i+1 and i+2 are in flight while plane i is updated); Hy(i) is the
    // previous plane's Hy(i+1).
    T hyC = HyA[(unsigned)i0 * ny + jc];
    TmzPlane<T> p0 = load_plane(i0);
    TmzPlane<T> p1 = i0 + 1 < iend ? load_plane(i0 + 1) : p0;
#pragma unroll 1
    for (int u = 0; u < XM; ++u) {
      const int i = i0 + u;
      if (i >= iend) break;
      const TmzPlane<T> p2 = i + 2 < iend ? load_plane(i + 2) : p1;
      // E half-step of (i, j), through the same update as the halo cells
      const TmzX x = tmz_x(v, i);
      TmzWhere f;
      f.bulk = x.ibulk && jbulk;
      f.src = x.isrc && jsrc;
      f.sx = f.bulk ? x.sxe : -1;
      f.dx = x.dxe;
      f.cx = 0;
      f.sy = f.bulk ? sye : -1;
      f.dy = dye;
      f.cy = cye;
      TmzIn<T> a;
      a.ez = p0.ez;
      a.hx = p0.hx;
      a.hxj = p0.hxj;
      a.hy = hyC;
      a.hyi = p0.hyi;
      a.pzx = p0.pzx;
      a.pzy = p0.pzy;
      a.ikx = lane_get(lk_ikx, u);
      a.bex = lane_get(lk_bex, u);
      a.cex = lane_get(lk_cex, u);
      a.m = p0.m;
      const TmzE<T> e = tmz_e(v, f, a, jcf, jv);
      const unsigned idx = (unsigned)i * ny + jc;
      if (f.bulk) {
        __builtin_nontemporal_store(e.ez, (OIDT_GLOBAL T*)(reinterpret_cast<char*>(EzB) + idx * (unsigned)sizeof(T)));
        if (f.sx >= 0) psiBll[psi_id(F_EZ, 0, f.sx)][(size_t)f.dx * ny + j] = e.pzx;
        if (f.sy >= 0) pzyB[(size_t)i * N + f.dy] = e.pzy;
      } else if (f.src) {
        __builtin_nontemporal_store(e.ez, (OIDT_GLOBAL T*)(reinterpret_cast<char*>(EzB) + idx * (unsigned)sizeof(T)));
      }
      if (jcap) {
        const size_t c = (size_t)(j - v.wy0);
        if (i == v.port0) v.ring[((size_t)tstep * v.nplanes + 0) * v.pcells + c] = e.ez;
        if (i == v.port1) v.ring[((size_t)tstep * v.nplanes + 3) * v.pcells + c] = e.ez;
      }
      // H half-step of (i, j): Ez(j-1) from the lower lane, lane 0 from the halo row
      const T ezj = lane_up1(e.ez, lane_get(colE, u));
      if (jok) {
        const T db = v.db;
        if (x.ibulk && jin1) {
          const T dEz = e.ez - ezj;
          T hv = p0.hx + db * (-dEz) * ikyh;
          if (syh >= 0) {
            const T q = bhy * p0.qxy + chy * dEz;
            pxyB[(size_t)i * N + dyh] = q;
            hv -= db * q;
          }
          __builtin_nontemporal_store(hv, (OIDT_GLOBAL T*)(reinterpret_cast<char*>(HxB) + idx * (unsigned)sizeof(T)));
        }
        T hv = hyC;
        if (x.hbulk && jbulk) {
          const T dEz = e.ez - prevE;
          hv = hv + db * dEz * lane_get(lk_ikhx, u);
          if (x.sxh >= 0) {
            const T q = lane_get(lk_bhx, u) * p0.qyx + lane_get(lk_chx, u) * dEz;
            psiBll[psi_id(F_HY, 0, x.sxh)][(size_t)x.dxh * ny + j] = q;
            hv += db * q;
          }
          __builtin_nontemporal_store(hv, (OIDT_GLOBAL T*)(reinterpret_cast<char*>(HyB) + idx * (unsigned)sizeof(T)));
        }
        if (jcap) {
          const size_t c = (size_t)(j - v.wy0);
          const size_t base = (size_t)tstep * v.nplanes;
          if (i == v.port0) v.ring[(base + 1) * v.pcells + c] = hv;
          if (i == v.port0 + 1) v.ring[(base + 2) * v.pcells + c] = hv;
          if (i == v.port1) v.ring[(base + 4) * v.pcells + c] = hv;
          if (i == v.port1 + 1) v.ring[(base + 5) * v.pcells + c] = hv;
        }
      }
      prevE = e.ez;
      hyC = p0.hyi;
      p0 = p1;
      p1 = p2;
    }
  }
}

// ---------------------------------------------------------- fused TEz step ---
// The TEz counterpart of k_tmz_step: one launch per time step, set (n & 1)
// read and the other written, one wavefront per tile of 64 rows x XM planes
// marching along x.  The E half-step reads Hz at j+1 and i+1 (old set); the
// Hz update reads the new Ex at j-1 (next-lower lane, lane 0 from the row
// below the tile) and the new Ey at i-1 (the previous plane, in a register;
// the plane below the tile computed first).  Same register rules as TMz: the
// run's fields copied per tile, x-coefficients as per-tile lane vectors, DPP
// and readlane for cross-lane values, loads two planes ahead, non-temporal
// field stores.

template <typename T>
struct TezView {
  int nx, ny, N, ib, ie, jb, je, sx0, sx1, sy0, sy1, port0, port1, wy0, wy1, nplanes, pcells;
  bool src_hz;
  T db, cb0[kMatMax], cb1[kMatMax];
  const T *ike0, *ike1, *ikh0, *ikh1, *be, *ceo, *bh, *cho;
  const uint8_t* mat;
  T* ring;
};

template <typename T>
__device__ __forceinline__ TezView<T> tez_view(const Slot<T>& s) {
  TezView<T> v;
  v.nx = s.nx; v.ny = s.ny; v.N = s.N;
  v.ib = s.ib; v.ie = s.ie; v.jb = s.jb; v.je = s.je;
  v.sx0 = s.sx0; v.sx1 = s.sx1; v.sy0 = s.sy0; v.sy1 = s.sy1;
  v.port0 = s.port[0]; v.port1 = s.port[1]; v.wy0 = s.wy0; v.wy1 = s.wy1;
  v.nplanes = s.nplanes; v.pcells = s.pcells;
  v.src_hz = s.src_field == F_HZ;
  v.db = s.db;
#pragma unroll
  for (int q = 0; q < kMatMax; ++q) {
    v.cb0[q] = s.cbv[0][q];
    v.cb1[q] = s.cbv[1][q];
  }
  v.ike0 = s.ike[0]; v.ike1 = s.ike[1]; v.ikh0 = s.ikh[0]; v.ikh1 = s.ikh[1];
  v.be = s.be; v.ceo = s.ceo; v.bh = s.bh; v.cho = s.cho;
  v.mat = s.mat;
  v.ring = s.ring;
  return v;
}

// Where a TEz cell's E components are updated (index ranges of
// fdtd_funcs.py:44-80 and the CPML slabs :82-100, FDTD_2D_TE).
struct TezWhere {
  bool bx, by;
  int sx, sy, dx, dy, cx, cy;   // Ey x-slab, Ex y-slab
};

template <typename T>
__device__ __forceinline__ TezWhere tez_where(const TezView<T>& v, int i, int j) {
  TezWhere f;
  const bool in = i >= 0 && j >= 0 && i < v.nx && j < v.ny;
  const bool iin = in && i >= v.ib && i < v.ie, iin1 = in && i > v.ib && i < v.ie;
  const bool jin = in && j >= v.jb && j < v.je, jin1 = in && j > v.jb && j < v.je;
  f.bx = iin1 && jin;
  f.by = iin && jin1;
  f.dx = f.dy = f.cx = f.cy = 0;
  f.sy = f.bx ? slab_e(j, v.jb, v.je, v.N, f.dy, f.cy) : -1;
  f.sx = f.by ? slab_e(i, v.ib, v.ie, v.N, f.dx, f.cx) : -1;
  return f;
}

template <typename T>
struct TezIn {
  T ex, ey, hz, hzj, hzi, pxy, pyx, ikx, iky, bex, cex, bey, cey;
  uint8_t m;
};

template <typename T>
__device__ __forceinline__ TezIn<T> tez_load(const TezView<T>& v, int i, int j, const TezWhere& f,
                                             const T* __restrict__ ExA, const T* __restrict__ EyA,
                                             const T* __restrict__ HzA, OIDT_GLOBAL T* const* psiA) {
  TezIn<T> a;
  const int nx = v.nx, ny = v.ny;
  const int ic = min(max(i, 0), nx - 1), jc = min(max(j, 0), ny - 1);
  const unsigned idx = (unsigned)ic * ny + jc;
  a.ex = ExA[idx];
  a.ey = EyA[idx];
  a.hz = HzA[idx];
  a.hzj = HzA[jc + 1 < ny ? idx + 1 : idx];
  a.hzi = HzA[ic + 1 < nx ? idx + ny : idx];
  a.m = v.mat[idx];
  a.iky = f.bx ? v.ike1[jc] : T(0);
  a.ikx = f.by ? v.ike0[ic] : T(0);
  a.pxy = f.sy >= 0 ? psiA[psi_id(F_EX, 1, f.sy)][(size_t)ic * v.N + f.dy] : T(0);
  a.bey = f.sy >= 0 ? v.be[f.cy] : T(0);
  a.cey = f.sy >= 0 ? v.ceo[f.cy] : T(0);
  a.pyx = f.sx >= 0 ? psiA[psi_id(F_EY, 0, f.sx)][(size_t)f.dx * ny + jc] : T(0);
  a.bex = f.sx >= 0 ? v.be[f.cx] : T(0);
  a.cex = f.sx >= 0 ? v.ceo[f.cx] : T(0);
  return a;
}

template <typename T>
struct TezE {
  T ex, ey, pxy, pyx;
};

// New Ex and Ey in the reference order: bulk curl, then the CPML term.
template <typename T>
__device__ __forceinline__ TezE<T> tez_e(const TezView<T>& v, const TezWhere& f, const TezIn<T>& a) {
  TezE<T> r;
  r.ex = a.ex;
  r.ey = a.ey;
  r.pxy = r.pyx = T(0);
  if (f.bx) {
    const int q = a.m & 3;
    const T cb = q == 0 ? v.cb0[0] : (q == 1 ? v.cb0[1] : (q == 2 ? v.cb0[2] : v.cb0[3]));
    const T dHz = a.hzj - a.hz;
    T e = a.ex + cb * dHz * a.iky;
    if (f.sy >= 0) {
      r.pxy = a.bey * a.pxy + a.cey * dHz;
      e += cb * r.pxy;
    }
    r.ex = e;
  }
  if (f.by) {
    const int q = (a.m >> 2) & 3;
    const T cb = q == 0 ? v.cb1[0] : (q == 1 ? v.cb1[1] : (q == 2 ? v.cb1[2] : v.cb1[3]));
    const T dHz = a.hzi - a.hz;
    T e = a.ey + cb * (-dHz) * a.ikx;
    if (f.sx >= 0) {
      r.pyx = a.bex * a.pyx + a.cex * dHz;
      e -= cb * r.pyx;
    }
    r.ey = e;
  }
  return r;
}

template <typename T>
__device__ __forceinline__ TezE<T> tez_e_cell(const TezView<T>& v, int i, int j, const T* __restrict__ ExA,
                                              const T* __restrict__ EyA, const T* __restrict__ HzA,
                                              OIDT_GLOBAL T* const* psiA) {
  const TezWhere f = tez_where(v, i, j);
  return tez_e(v, f, tez_load(v, i, j, f, ExA, EyA, HzA, psiA));
}

template <typename T>
struct TezPlane {
  T ex, ey, hzj, hzi, pxy, pyx, qzx, qzy;
  uint8_t m;
};

template <typename T, int XM>
__global__ __launch_bounds__(kThreads) void k_tez_step(const Slot<T>* __restrict__ slots,
                                                       const int4* __restrict__ tiles,
                                                       const int* __restrict__ counts, int tstep) {
  static_assert(XM <= kBX, "the j0-1 halo row takes one cell per lane");
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  const int stride = gridDim.x * kBY;
  for (int tt = xcd_block() * kBY + wave_id(); tt < ntiles; tt += stride) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const TezView<T> v = tez_view(s);
    const bool odd = (n & 1) != 0;
    const T* __restrict__ ExA = odd ? s.f2[F_EX] : s.f[F_EX];
    const T* __restrict__ EyA = odd ? s.f2[F_EY] : s.f[F_EY];
    const T* __restrict__ HzA = odd ? s.f2[F_HZ] : s.f[F_HZ];
    T* __restrict__ ExB = odd ? s.f[F_EX] : s.f2[F_EX];
    T* __restrict__ EyB = odd ? s.f[F_EY] : s.f2[F_EY];
    T* __restrict__ HzB = odd ? s.f[F_HZ] : s.f2[F_HZ];
    OIDT_GLOBAL T* const* psiAll = odd ? s.psi2 : s.psi;
    OIDT_GLOBAL T* const* psiBll = odd ? s.psi : s.psi2;
    const T jv = v.src_hz ? s.jsrc[n] : T(0);
    const int nx = v.nx, ny = v.ny, N = v.N;
    const int i0 = tl.y, j0 = tl.z;
    const int j = j0 + lane;
    const bool jok = j < ny;
    const int jc = min(j, ny - 1);
    // row (y-only) constants of this lane, fixed over the march
    const bool jin = jok && j >= v.jb && j < v.je;     // Ex rows
    const bool jin1 = jok && j > v.jb && j < v.je;     // Ey and Hz rows
    const bool jsrc = jok && v.src_hz && j >= v.sy0 && j < v.sy1;
    int dye = 0, cye = 0, dyh = 0, cyh = 0;
    const int sye = jin ? slab_e(j, v.jb, v.je, N, dye, cye) : -1;
    const int syh = jin1 ? slab_h(j, v.jb, v.je, N, dyh, cyh) : -1;
    const T ikye = jin ? v.ike1[j] : T(0);
    const T bey = sye >= 0 ? v.be[cye] : T(0), cey = sye >= 0 ? v.ceo[cye] : T(0);
    const T ikyh = jin1 ? v.ikh1[j] : T(0);
    const T bhy = syh >= 0 ? v.bh[cyh] : T(0), chy = syh >= 0 ? v.cho[cyh] : T(0);
    const bool jcap = jok && j >= v.wy0 && j <= v.wy1;
    const T* pxyA = sye >= 0 ? psiAll[psi_id(F_EX, 1, sye)] : nullptr;
    T* pxyB = sye >= 0 ? psiBll[psi_id(F_EX, 1, sye)] : nullptr;
    const T* pzyA = syh >= 0 ? psiAll[psi_id(F_HZ, 1, syh)] : nullptr;
    T* pzyB = syh >= 0 ? psiBll[psi_id(F_HZ, 1, syh)] : nullptr;
    const int iend = min(i0 + XM, nx);
    auto load_plane = [&](int i) {
      TezPlane<T> p;
      const bool iin = i >= v.ib && i < v.ie, iin1 = i > v.ib && i < v.ie;
      int dx = 0, cx = 0, dxh = 0, cxh = 0;
      const int sxe = iin ? slab_e(i, v.ib, v.ie, N, dx, cx) : -1;
      const int sxh = iin1 ? slab_h(i, v.ib, v.ie, N, dxh, cxh) : -1;
      const unsigned idx = (unsigned)i * ny + jc;
      p.ex = ExA[idx];
      p.ey = EyA[idx];
      p.hzj = HzA[jc + 1 < ny ? idx + 1 : idx];
      p.hzi = HzA[i + 1 < nx ? idx + ny : idx];
      p.m = v.mat[idx];
      p.pxy = iin1 && sye >= 0 ? pxyA[(size_t)i * N + dye] : T(0);
      p.pyx = jin1 && sxe >= 0 ? psiAll[psi_id(F_EY, 0, sxe)][(size_t)dx * ny + jc] : T(0);
      p.qzx = jin1 && sxh >= 0 ? psiAll[psi_id(F_HZ, 0, sxh)][(size_t)dxh * ny + jc] : T(0);
      p.qzy = iin1 && syh >= 0 ? pzyA[(size_t)i * N + dyh] : T(0);
      return p;
    };
    // ---- x-only coefficients of the tile's planes: lane u holds plane i0+u
    T lk_ikx = T(0), lk_bex = T(0), lk_cex = T(0), lk_ikhx = T(0), lk_bhx = T(0), lk_chx = T(0);
    if (lane < XM && i0 + lane < nx) {
      const int i = i0 + lane;
      const bool iin = i >= v.ib && i < v.ie, iin1 = i > v.ib && i < v.ie;
      int dx = 0, cx = 0;
      lk_ikx = iin ? v.ike0[i] : T(0);
      if (iin && slab_e(i, v.ib, v.ie, N, dx, cx) >= 0) {
        lk_bex = v.be[cx];
        lk_cex = v.ceo[cx];
      }
      lk_ikhx = v.ikh0[i];
      if (iin1 && slab_h(i, v.ib, v.ie, N, dx, cx) >= 0) {
        lk_bhx = v.bh[cx];
        lk_chx = v.cho[cx];
      }
    }
    // ---- halos from set A, not stored: new Ex of row j0-1 (lane u holds
    // plane i0+u) and new Ey of plane i0-1
    const T colEx = lane < XM ? tez_e_cell(v, i0 + lane, j0 - 1, ExA, EyA, HzA, psiAll).ex : T(0);
    T prevEy = tez_e_cell(v, i0 - 1, j, ExA, EyA, HzA, psiAll).ey;
    // ---- march along x; Hz(i) is the previous plane's Hz(i+1)
    T hzC = HzA[(unsigned)i0 * ny + jc];
    TezPlane<T> p0 = load_plane(i0);
    TezPlane<T> p1 = i0 + 1 < iend ? load_plane(i0 + 1) : p0;
#pragma unroll 1
    for (int u = 0; u < XM; ++u) {
      const int i = i0 + u;
      if (i >= iend) break;
      const TezPlane<T> p2 = i + 2 < iend ? load_plane(i + 2) : p1;
      const bool iin = i >= v.ib && i < v.ie, iin1 = i > v.ib && i < v.ie;
      int dxe = 0, cxe = 0, dxh = 0, cxh = 0;
      const int sxe = iin ? slab_e(i, v.ib, v.ie, N, dxe, cxe) : -1;
      const int sxh = iin1 ? slab_h(i, v.ib, v.ie, N, dxh, cxh) : -1;
      // E half-step of (i, j), through the same update as the halo cells
      TezWhere f;
      f.bx = iin1 && jin;
      f.by = iin && jin1;
      f.sy = f.bx ? sye : -1;
      f.dy = dye;
      f.cy = cye;
      f.sx = f.by ? sxe : -1;
      f.dx = dxe;
      f.cx = cxe;
      TezIn<T> a;
      a.ex = p0.ex;
      a.ey = p0.ey;
      a.hz = hzC;
      a.hzj = p0.hzj;
      a.hzi = p0.hzi;
      a.m = p0.m;
      a.iky = ikye;
      a.ikx = lane_get(lk_ikx, u);
      a.pxy = p0.pxy;
      a.bey = bey;
      a.cey = cey;
      a.pyx = p0.pyx;
      a.bex = lane_get(lk_bex, u);
      a.cex = lane_get(lk_cex, u);
      const TezE<T> e = tez_e(v, f, a);
      const unsigned idx = (unsigned)i * ny + jc;
      if (f.bx) {
        __builtin_nontemporal_store(e.ex, ExB + idx);
        if (f.sy >= 0) pxyB[(size_t)i * N + f.dy] = e.pxy;
      }
      if (f.by) {
        __builtin_nontemporal_store(e.ey, EyB + idx);
        if (f.sx >= 0) psiBll[psi_id(F_EY, 0, f.sx)][(size_t)f.dx * ny + j] = e.pyx;
      }
      if (jcap) {
        const size_t c = (size_t)(j - v.wy0);
        if (i == v.port0) v.ring[((size_t)tstep * v.nplanes + 0) * v.pcells + c] = e.ey;
        if (i == v.port1) v.ring[((size_t)tstep * v.nplanes + 3) * v.pcells + c] = e.ey;
      }
      // Hz of (i, j): Ex(j-1) from the lower lane, lane 0 from the halo row
      const T exj = lane_up1(e.ex, lane_get(colEx, u));
      if (jok) {
        const T db = v.db;
        const bool bulk = jin1 && iin1;
        const bool src = jsrc && i >= v.sx0 && i < v.sx1;
        T hv = hzC;
        if (bulk) {
          const T dEx = e.ex - exj;
          const T dEy = e.ey - prevEy;
          hv = hv + db * (dEx * ikyh - dEy * lane_get(lk_ikhx, u));
          if (src) hv -= jv;
          if (sxh >= 0) {
            const T q = lane_get(lk_bhx, u) * p0.qzx + lane_get(lk_chx, u) * dEy;
            psiBll[psi_id(F_HZ, 0, sxh)][(size_t)dxh * ny + j] = q;
            hv -= db * q;
          }
          if (syh >= 0) {
            const T q = bhy * p0.qzy + chy * dEx;
            pzyB[(size_t)i * N + dyh] = q;
            hv += db * q;
          }
          __builtin_nontemporal_store(hv, HzB + idx);
        } else if (src) {
          hv -= jv;
          __builtin_nontemporal_store(hv, HzB + idx);
        }
        if (jcap) {
          const size_t c = (size_t)(j - v.wy0);
          const size_t base = (size_t)tstep * v.nplanes;
          if (i == v.port0) v.ring[(base + 1) * v.pcells + c] = hv;
          if (i == v.port0 + 1) v.ring[(base + 2) * v.pcells + c] = hv;
          if (i == v.port1) v.ring[(base + 4) * v.pcells + c] = hv;
          if (i == v.port1 + 1) v.ring[(base + 5) * v.pcells + c] = hv;
        }
      }
      prevEy = e.ey;
      hzC = p0.hzi;
      p0 = p1;
      p1 = p2;
    }
  }
}

// Workgroup barrier for LDS exchanges only: the fences order this wave's
// LDS accesses around the barrier (a wait on the LDS counter), not its global
// loads and stores.  In the kernels that use it no wave reads what another
// writes to global memory in the same launch (reads come from field set A,
// writes go to set B), so loads stay in flight across the barrier instead of
// being waited for as a full __syncthreads() would (its release fence waits
// for every outstanding global access of the wave).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------ temporally blocked TMz ---
// TB Yee steps per launch.  A block of W waves holds an extended tile of
// kTbPlanes planes (x) by 64 rows (y, one per lane) in registers: wave w
// owns K = kTbPlanes / W planes from ib0 + w*K, every lane its row of those
// planes (Ez, Hx, Hy, the four psi values and the material byte of each).
// The tile's own cells are the inner (kTbPlanes - 2 TB) x (64 - 2 TB); the
// TB-cell ghost frame on every side is loaded and stepped like the rest and
// discarded: each step (E then H) moves the region that depends on values
// outside the tile one cell inwards from each side, so after TB steps the
// inner cells are exact.  Per step the only exchanges are the y neighbours
// (DPP lane shifts) and, between waves, one plane of old Hy (for the E
// update of the next-lower wave's last plane) and one plane of new Ez (for
// the H update of the next-higher wave's first plane) through LDS, each
// behind a barrier.  The per-cell arithmetic is k_tmz_step's (tmz_e, then
// the Hx and Hy updates in its order), so results are bit-identical to TB
// launches of the one-step form.  The launch reads field/psi set
// ((n0 / TB) & 1) and writes the inner cells' state after its steps to the
// other set.
// HBM traffic per cell and launch: every field and psi value read once
// (ghost frame included) and written once, for TB steps.  Two block shapes
// share the tiles: 8 waves of 4 planes, or 16 waves of 2 planes.
constexpr int kTbPlanes = 32;   // planes of a block's extended tile

__device__ __forceinline__ float lane_down1(float v, float hi) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(hi), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_down1(double v, double hi) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(hi);
  const int blo = __builtin_amdgcn_update_dpp((int)o, (int)b, 0x130, 0xf, 0xf, false);
  const int bhi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), 0x130, 0xf, 0xf, false);
  return __longlong_as_double(((long long)bhi << 32) | (unsigned)blo);
}

template <typename T, int TB, int W>
__global__ __launch_bounds__(64 * W) void k_tmz_tb(const Slot<T>* __restrict__ slots,
                                                   const int4* __restrict__ tiles,
                                                   const int* __restrict__ counts, int tstep) {
  constexpr int K = kTbPlanes / W;
  constexpr int TX = kTbPlanes - 2 * TB, TY = 64 - 2 * TB;   // inner cells of a tile
  static_assert(TX > 0 && TY > 0 && kChunk % TB == 0, "blocking depth");
  constexpr unsigned es = sizeof(T);
  __shared__ T xhy[W][64];   // old Hy of each wave's first plane
  __shared__ T xez[W][64];   // new Ez of each wave's last plane
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  const int w = wave_id();
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n0 = s.base + tstep;
    if (!s.active || n0 >= s.nt) continue;      // uniform over the block
    const int nsteps = min(TB, s.nt - n0);
    const TmzView<T> v = tmz_view(s);
    const bool odd = ((n0 / TB) & 1) != 0;
    const T* __restrict__ EzA = odd ? s.f2[F_EZ] : s.f[F_EZ];
    const T* __restrict__ HxA = odd ? s.f2[F_HX] : s.f[F_HX];
    const T* __restrict__ HyA = odd ? s.f2[F_HY] : s.f[F_HY];
    T* __restrict__ EzB = odd ? s.f[F_EZ] : s.f2[F_EZ];
    T* __restrict__ HxB = odd ? s.f[F_HX] : s.f2[F_HX];
    T* __restrict__ HyB = odd ? s.f[F_HY] : s.f2[F_HY];
    OIDT_GLOBAL T* const* psiAll = odd ? s.psi2 : s.psi;
    OIDT_GLOBAL T* const* psiBll = odd ? s.psi : s.psi2;
    const int nx = v.nx, ny = v.ny, N = v.N;
    const int i0 = tl.y, j0 = tl.z;               // first inner cell
    const int ib0 = i0 - TB + w * K;              // this wave's first plane
    const int j = j0 - TB + lane;
    const bool jok = j >= 0 && j < ny;
    const int jc = min(max(j, 0), ny - 1);
    const bool jown = jok && j >= j0 && j < j0 + TY;
    // row (y-only) constants of this lane
    const bool jbulk = jok && j >= v.jb && j < v.je;          // Ez and Hy rows
    const bool jin1 = jok && j > v.jb && j < v.je;            // Hx rows
    const bool jsrc = jok && v.src_ez && j >= v.sy0 && j < v.sy1;
    int dye = 0, cye = 0, dyh = 0, cyh = 0;
    const int sye = jbulk ? slab_e(j, v.jb, v.je, N, dye, cye) : -1;
    const int syh = jin1 ? slab_h(j, v.jb, v.je, N, dyh, cyh) : -1;
    TmzJ<T> jcf;
    jcf.iky = jbulk ? v.ike1[j] : T(0);
    jcf.bey = sye >= 0 ? v.be[cye] : T(0);
    jcf.cey = sye >= 0 ? v.ceo[cye] : T(0);
    const T ikyh = jin1 ? v.ikh1[j] : T(0);
    const T bhy = syh >= 0 ? v.bh[cyh] : T(0), chy = syh >= 0 ? v.cho[cyh] : T(0);
    const bool jcap = jown && j >= v.wy0 && j <= v.wy1;
    const T* pzyA = sye >= 0 ? psiAll[psi_id(F_EZ, 1, sye)] : nullptr;
    T* pzyB = sye >= 0 ? psiBll[psi_id(F_EZ, 1, sye)] : nullptr;
    const T* pxyA = syh >= 0 ? psiAll[psi_id(F_HX, 1, syh)] : nullptr;
    T* pxyB = syh >= 0 ? psiBll[psi_id(F_HX, 1, syh)] : nullptr;
    // ---- the wave's cells, from set A (outside the grid: zero, never updated)
    T ez[K], hx[K], hy[K], pzx[K], pzy[K], qxy[K], qyx[K];
    unsigned mm[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = ib0 + q;
      const bool ok = jok && i >= 0 && i < nx;
      const TmzX x = tmz_x(v, i);
      const unsigned idx = (unsigned)min(max(i, 0), nx - 1) * ny + jc;
      ez[q] = ok ? ldg<T>(EzA, idx * es) : T(0);
      hx[q] = ok ? ldg<T>(HxA, idx * es) : T(0);
      hy[q] = ok ? ldg<T>(HyA, idx * es) : T(0);
      mm[q] = ok ? (unsigned)ldg<uint8_t>(v.mat, idx) : 0u;
      pzx[q] = ok && x.ibulk && jbulk && x.sxe >= 0 ? psiAll[psi_id(F_EZ, 0, x.sxe)][(size_t)x.dxe * ny + jc] : T(0);
      pzy[q] = ok && x.ibulk && sye >= 0 ? pzyA[(size_t)i * N + dye] : T(0);
      qxy[q] = ok && x.ibulk && syh >= 0 ? pxyA[(size_t)i * N + dyh] : T(0);
      qyx[q] = ok && x.hbulk && jbulk && x.sxh >= 0 ? psiAll[psi_id(F_HY, 0, x.sxh)][(size_t)x.dxh * ny + jc] : T(0);
    }
    // ---- per-plane constants, fixed over the launch: flags (wave-uniform)
    // and coefficients (lane q holds plane ib0 + q); the cb of each cell;
    // the source waveform (lane t holds step n0 + t)
    unsigned pf[K];
    T cbz[K];
    T lk_ikx = T(0), lk_bex = T(0), lk_cex = T(0), lk_ikh = T(0), lk_bhx = T(0), lk_chx = T(0);
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = ib0 + q;
      const TmzX x = tmz_x(v, i);
      const bool iok = i >= 0 && i < nx;
      const bool inner = iok && i >= i0 && i < i0 + TX;
      pf[q] = (x.ibulk ? 1u : 0u) | (x.hbulk ? 2u : 0u) | (iok && x.isrc ? 4u : 0u) | (x.sxe >= 0 ? 8u : 0u) |
              (x.sxh >= 0 ? 16u : 0u) | (inner && i == v.port0 ? 32u : 0u) | (inner && i == v.port1 ? 64u : 0u) |
              (inner && i == v.port0 + 1 ? 128u : 0u) | (inner && i == v.port1 + 1 ? 256u : 0u);
      if (lane == q) {
        lk_ikx = x.ibulk ? v.ike0[i] : T(0);
        lk_bex = x.sxe >= 0 ? v.be[x.cxe] : T(0);
        lk_cex = x.sxe >= 0 ? v.ceo[x.cxe] : T(0);
        lk_ikh = x.hbulk ? v.ikh0[i] : T(0);
        lk_bhx = x.sxh >= 0 ? v.bh[x.cxh] : T(0);
        lk_chx = x.sxh >= 0 ? v.cho[x.cxh] : T(0);
      }
      const int qm = (mm[q] >> 4) & 3;
      cbz[q] = qm == 0 ? v.cb[0] : (qm == 1 ? v.cb[1] : (qm == 2 ? v.cb[2] : v.cb[3]));
    }
    const T lk_jv = v.src_ez && lane < nsteps ? s.jsrc[n0 + lane] : T(0);
    const T db = v.db;
    T* const ring = v.ring;
    const int pcells = v.pcells;
    const size_t cw = (size_t)(j - v.wy0);
    // Interior tiles (tile flag set by the host): every cell of the extended
    // tile is in every component's update range, outside the CPML slabs, the
    // source box and the capture planes, so the update is the bulk curl
    // alone (the same expressions as below, whose other terms are absent
    // there).
    for (int t = 0; tl.w != 0 && t < nsteps; ++t) {
      xhy[w][lane] = hy[0];
      lds_barrier();
      const T hyup = w + 1 < W ? xhy[w + 1][lane] : hy[K - 1];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const T hxj = lane_down1(hx[q], hx[q]);
        const T hyi = q + 1 < K ? hy[q + 1] : hyup;
        const T dHy = hyi - hy[q];
        const T dHx = hxj - hx[q];
        ez[q] = ez[q] + cbz[q] * (dHy * lane_get(lk_ikx, q) - dHx * jcf.iky);
      }
      xez[w][lane] = ez[K - 1];
      lds_barrier();
      const T ezdn = w > 0 ? xez[w - 1][lane] : ez[0];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const T ezj = lane_up1(ez[q], ez[q]);
        const T prevE = q > 0 ? ez[q - 1] : ezdn;
        const T dEzy = ez[q] - ezj;
        hx[q] = hx[q] + db * (-dEzy) * ikyh;
        const T dEzx = ez[q] - prevE;
        hy[q] = hy[q] + db * dEzx * lane_get(lk_ikh, q);
      }
    }
    for (int t = 0; tl.w == 0 && t < nsteps; ++t) {
      const T jv = lane_get(lk_jv, t);
      const size_t rbase = (size_t)(tstep + t) * v.nplanes;
      // ---- E half-step (tmz_e per cell): Hy(i+1) of the last plane from the next wave
      xhy[w][lane] = hy[0];
      lds_barrier();
      const T hyup = w + 1 < W ? xhy[w + 1][lane] : hy[K - 1];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const unsigned f = pf[q];
        const bool bulk = (f & 1u) && jbulk;
        const bool src = (f & 4u) && jsrc;
        const T hxj = lane_down1(hx[q], hx[q]);
        const T hyi = q + 1 < K ? hy[q + 1] : hyup;
        T e = ez[q];
        if (bulk) {
          const T cb = cbz[q];
          const T dHy = hyi - hy[q];
          const T dHx = hxj - hx[q];
          e = e + cb * (dHy * lane_get(lk_ikx, q) - dHx * jcf.iky);
          if (src) e -= jv;
          if (f & 8u) {
            const T qv = lane_get(lk_bex, q) * pzx[q] + lane_get(lk_cex, q) * dHy;
            pzx[q] = qv;
            e += cb * qv;
          }
          if (sye >= 0) {
            const T qv = jcf.bey * pzy[q] + jcf.cey * dHx;
            pzy[q] = qv;
            e -= cb * qv;
          }
        } else if (src) {
          e = e - jv;
        }
        ez[q] = e;
        if (jcap && (f & 96u)) {
          if (f & 32u) ring[(rbase + 0) * pcells + cw] = e;
          if (f & 64u) ring[(rbase + 3) * pcells + cw] = e;
        }
      }
      // ---- H half-step: Ez(i-1) of the first plane from the previous wave
      xez[w][lane] = ez[K - 1];
      lds_barrier();
      const T ezdn = w > 0 ? xez[w - 1][lane] : ez[0];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const unsigned f = pf[q];
        const T ezj = lane_up1(ez[q], ez[q]);
        const T prevE = q > 0 ? ez[q - 1] : ezdn;
        if ((f & 1u) && jin1) {
          const T dEz = ez[q] - ezj;
          T hv = hx[q] + db * (-dEz) * ikyh;
          if (syh >= 0) {
            const T qv = bhy * qxy[q] + chy * dEz;
            qxy[q] = qv;
            hv -= db * qv;
          }
          hx[q] = hv;
        }
        T hv = hy[q];
        if ((f & 2u) && jbulk) {
          const T dEz = ez[q] - prevE;
          hv = hv + db * dEz * lane_get(lk_ikh, q);
          if (f & 16u) {
            const T qv = lane_get(lk_bhx, q) * qyx[q] + lane_get(lk_chx, q) * dEz;
            qyx[q] = qv;
            hv += db * qv;
          }
          hy[q] = hv;
        }
        if (jcap && (f & 480u)) {
          if (f & 32u) ring[(rbase + 1) * pcells + cw] = hv;
          if (f & 128u) ring[(rbase + 2) * pcells + cw] = hv;
          if (f & 64u) ring[(rbase + 4) * pcells + cw] = hv;
          if (f & 256u) ring[(rbase + 5) * pcells + cw] = hv;
        }
      }
    }
    // ---- the inner cells to set B (fields unconditionally: cells no step
    // updates hold the same value in both sets)
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = ib0 + q;
      if (!jown || i < i0 || i >= i0 + TX || i >= nx) continue;
      const TmzX x = tmz_x(v, i);
      const unsigned idx = (unsigned)i * ny + j;
      stg_nt<T>(EzB, idx * es, ez[q]);
      stg_nt<T>(HxB, idx * es, hx[q]);
      stg_nt<T>(HyB, idx * es, hy[q]);
      if (x.ibulk && jbulk && x.sxe >= 0) psiBll[psi_id(F_EZ, 0, x.sxe)][(size_t)x.dxe * ny + j] = pzx[q];
      if (x.ibulk && jbulk && sye >= 0) pzyB[(size_t)i * N + dye] = pzy[q];
      if (x.ibulk && jin1 && syh >= 0) pxyB[(size_t)i * N + dyh] = qxy[q];
      if (x.hbulk && jbulk && x.sxh >= 0) psiBll[psi_id(F_HY, 0, x.sxh)][(size_t)x.dxh * ny + j] = qyx[q];
    }
    lds_barrier();   // the next tile reuses the exchange planes
  }
}

// ------------------------------------------------ temporally blocked TEz ---
// The TEz counterpart of k_tmz_tb (same tiles, ghost frame, block shapes and
// field-set rule): per step the E update of a wave's last plane takes the old
// Hz of the next wave's first plane, and the Hz update of its first plane the
// new Ey of the previous wave's last plane, through LDS; the j+1 neighbour of
// Ex's update and the j-1 neighbour of Hz's come from DPP lane shifts.  The
// per-cell arithmetic is k_tez_step's (tez_e, then the Hz update in its order).

template <typename T, int TB, int W>
__global__ __launch_bounds__(64 * W) void k_tez_tb(const Slot<T>* __restrict__ slots,
                                                   const int4* __restrict__ tiles,
                                                   const int* __restrict__ counts, int tstep) {
  constexpr int K = kTbPlanes / W;
  constexpr int TX = kTbPlanes - 2 * TB, TY = 64 - 2 * TB;
  static_assert(TX > 0 && TY > 0 && kChunk % TB == 0, "blocking depth");
  constexpr unsigned es = sizeof(T);
  __shared__ T xhz[W][64];   // old Hz of each wave's first plane
  __shared__ T xey[W][64];   // new Ey of each wave's last plane
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  const int w = wave_id();
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n0 = s.base + tstep;
    if (!s.active || n0 >= s.nt) continue;      // uniform over the block
    const int nsteps = min(TB, s.nt - n0);
    const TezView<T> v = tez_view(s);
    const bool odd = ((n0 / TB) & 1) != 0;
    const T* __restrict__ ExA = odd ? s.f2[F_EX] : s.f[F_EX];
    const T* __restrict__ EyA = odd ? s.f2[F_EY] : s.f[F_EY];
    const T* __restrict__ HzA = odd ? s.f2[F_HZ] : s.f[F_HZ];
    T* __restrict__ ExB = odd ? s.f[F_EX] : s.f2[F_EX];
    T* __restrict__ EyB = odd ? s.f[F_EY] : s.f2[F_EY];
    T* __restrict__ HzB = odd ? s.f[F_HZ] : s.f2[F_HZ];
    OIDT_GLOBAL T* const* psiAll = odd ? s.psi2 : s.psi;
    OIDT_GLOBAL T* const* psiBll = odd ? s.psi : s.psi2;
    const int nx = v.nx, ny = v.ny, N = v.N;
    const int i0 = tl.y, j0 = tl.z;               // first inner cell
    const int ib0 = i0 - TB + w * K;              // this wave's first plane
    const int j = j0 - TB + lane;
    const bool jok = j >= 0 && j < ny;
    const int jc = min(max(j, 0), ny - 1);
    const bool jown = jok && j >= j0 && j < j0 + TY;
    // row (y-only) constants of this lane
    const bool jin = jok && j >= v.jb && j < v.je;     // Ex rows
    const bool jin1 = jok && j > v.jb && j < v.je;     // Ey and Hz rows
    const bool jsrc = jok && v.src_hz && j >= v.sy0 && j < v.sy1;
    int dye = 0, cye = 0, dyh = 0, cyh = 0;
    const int sye = jin ? slab_e(j, v.jb, v.je, N, dye, cye) : -1;
    const int syh = jin1 ? slab_h(j, v.jb, v.je, N, dyh, cyh) : -1;
    const T ikye = jin ? v.ike1[j] : T(0);
    const T bey = sye >= 0 ? v.be[cye] : T(0), cey = sye >= 0 ? v.ceo[cye] : T(0);
    const T ikyh = jin1 ? v.ikh1[j] : T(0);
    const T bhy = syh >= 0 ? v.bh[cyh] : T(0), chy = syh >= 0 ? v.cho[cyh] : T(0);
    const bool jcap = jown && j >= v.wy0 && j <= v.wy1;
    const T* pxyA = sye >= 0 ? psiAll[psi_id(F_EX, 1, sye)] : nullptr;
    T* pxyB = sye >= 0 ? psiBll[psi_id(F_EX, 1, sye)] : nullptr;
    const T* pzyA = syh >= 0 ? psiAll[psi_id(F_HZ, 1, syh)] : nullptr;
    T* pzyB = syh >= 0 ? psiBll[psi_id(F_HZ, 1, syh)] : nullptr;
    // ---- per-plane flags (bit 0 iin, 1 iin1, 2 source plane, 3 E x-slab,
    // 4 H x-slab, 5..8 capture at port0, port1, port0 + 1, port1 + 1); lane q
    // holds plane ib0 + q's coefficients.  The x-slab cell indexes are
    // recomputed where the psi values are loaded and stored.
    unsigned pf[K];
    T lk_ikx = T(0), lk_bex = T(0), lk_cex = T(0), lk_ikh = T(0), lk_bhx = T(0), lk_chx = T(0);
    auto xslabs = [&](int i, int& sxe, int& dxe, int& cxe, int& sxh, int& dxh, int& cxh) {
      const bool iok = i >= 0 && i < nx;
      const bool iin = iok && i >= v.ib && i < v.ie, iin1 = iok && i > v.ib && i < v.ie;
      dxe = cxe = dxh = cxh = 0;
      sxe = iin ? slab_e(i, v.ib, v.ie, N, dxe, cxe) : -1;
      sxh = iin1 ? slab_h(i, v.ib, v.ie, N, dxh, cxh) : -1;
    };
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = ib0 + q;
      const bool iok = i >= 0 && i < nx;
      const bool iin = iok && i >= v.ib && i < v.ie, iin1 = iok && i > v.ib && i < v.ie;
      const bool inner = iok && i >= i0 && i < i0 + TX;
      int sxe, dxe, cxe, sxh, dxh, cxh;
      xslabs(i, sxe, dxe, cxe, sxh, dxh, cxh);
      pf[q] = (iin ? 1u : 0u) | (iin1 ? 2u : 0u) | (iok && i >= v.sx0 && i < v.sx1 ? 4u : 0u) | (sxe >= 0 ? 8u : 0u) |
              (sxh >= 0 ? 16u : 0u) | (inner && i == v.port0 ? 32u : 0u) | (inner && i == v.port1 ? 64u : 0u) |
              (inner && i == v.port0 + 1 ? 128u : 0u) | (inner && i == v.port1 + 1 ? 256u : 0u);
      if (lane == q) {
        lk_ikx = iin ? v.ike0[i] : T(0);
        lk_bex = sxe >= 0 ? v.be[cxe] : T(0);
        lk_cex = sxe >= 0 ? v.ceo[cxe] : T(0);
        lk_ikh = iin1 ? v.ikh0[i] : T(0);
        lk_bhx = sxh >= 0 ? v.bh[cxh] : T(0);
        lk_chx = sxh >= 0 ? v.cho[cxh] : T(0);
      }
    }
    // ---- the wave's cells, from set A (outside the grid: zero, never updated)
    T ex[K], ey[K], hz[K], pxy[K], pyx[K], qzx[K], qzy[K], cb0[K], cb1[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = ib0 + q;
      const bool ok = jok && i >= 0 && i < nx;
      const unsigned idx = (unsigned)min(max(i, 0), nx - 1) * ny + jc;
      ex[q] = ok ? ldg<T>(ExA, idx * es) : T(0);
      ey[q] = ok ? ldg<T>(EyA, idx * es) : T(0);
      hz[q] = ok ? ldg<T>(HzA, idx * es) : T(0);
      const unsigned m = ok ? (unsigned)ldg<uint8_t>(v.mat, idx) : 0u;
      const int q0 = m & 3, q1 = (m >> 2) & 3;
      cb0[q] = q0 == 0 ? v.cb0[0] : (q0 == 1 ? v.cb0[1] : (q0 == 2 ? v.cb0[2] : v.cb0[3]));
      cb1[q] = q1 == 0 ? v.cb1[0] : (q1 == 1 ? v.cb1[1] : (q1 == 2 ? v.cb1[2] : v.cb1[3]));
      const unsigned f = pf[q];
      int sxe, dxe, cxe, sxh, dxh, cxh;
      xslabs(i, sxe, dxe, cxe, sxh, dxh, cxh);
      pxy[q] = ok && (f & 2u) && sye >= 0 ? pxyA[(size_t)i * N + dye] : T(0);
      pyx[q] = ok && jin1 && sxe >= 0 ? psiAll[psi_id(F_EY, 0, sxe)][(size_t)dxe * ny + jc] : T(0);
      qzx[q] = ok && jin1 && sxh >= 0 ? psiAll[psi_id(F_HZ, 0, sxh)][(size_t)dxh * ny + jc] : T(0);
      qzy[q] = ok && (f & 2u) && syh >= 0 ? pzyA[(size_t)i * N + dyh] : T(0);
    }
    const T lk_jv = v.src_hz && lane < nsteps ? s.jsrc[n0 + lane] : T(0);
    const T db = v.db;
    T* const ring = v.ring;
    const int pcells = v.pcells;
    const size_t cw = (size_t)(j - v.wy0);
    // Interior tiles (as in k_tmz_tb): the bulk curls alone.
    for (int t = 0; tl.w != 0 && t < nsteps; ++t) {
      xhz[w][lane] = hz[0];
      lds_barrier();
      const T hzup = w + 1 < W ? xhz[w + 1][lane] : hz[K - 1];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const T hzj = lane_down1(hz[q], hz[q]);
        const T hzi = q + 1 < K ? hz[q + 1] : hzup;
        const T dHzy = hzj - hz[q];
        ex[q] = ex[q] + cb0[q] * dHzy * ikye;
        const T dHzx = hzi - hz[q];
        ey[q] = ey[q] + cb1[q] * (-dHzx) * lane_get(lk_ikx, q);
      }
      xey[w][lane] = ey[K - 1];
      lds_barrier();
      const T eydn = w > 0 ? xey[w - 1][lane] : ey[0];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const T exj = lane_up1(ex[q], ex[q]);
        const T prevEy = q > 0 ? ey[q - 1] : eydn;
        const T dEx = ex[q] - exj;
        const T dEy = ey[q] - prevEy;
        hz[q] = hz[q] + db * (dEx * ikyh - dEy * lane_get(lk_ikh, q));
      }
    }
    for (int t = 0; tl.w == 0 && t < nsteps; ++t) {
      const T jv = lane_get(lk_jv, t);
      const size_t rbase = (size_t)(tstep + t) * v.nplanes;
      // ---- E half-step (tez_e per cell): Hz(i+1) of the last plane from the next wave
      xhz[w][lane] = hz[0];
      lds_barrier();
      const T hzup = w + 1 < W ? xhz[w + 1][lane] : hz[K - 1];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const unsigned f = pf[q];
        const T hzj = lane_down1(hz[q], hz[q]);
        const T hzi = q + 1 < K ? hz[q + 1] : hzup;
        if ((f & 2u) && jin) {       // Ex: i (ib, ie), j [jb, je)
          const T dHz = hzj - hz[q];
          T e = ex[q] + cb0[q] * dHz * ikye;
          if (sye >= 0) {
            const T qv = bey * pxy[q] + cey * dHz;
            pxy[q] = qv;
            e += cb0[q] * qv;
          }
          ex[q] = e;
        }
        if ((f & 1u) && jin1) {      // Ey: i [ib, ie), j (jb, je)
          const T dHz = hzi - hz[q];
          T e = ey[q] + cb1[q] * (-dHz) * lane_get(lk_ikx, q);
          if (f & 8u) {
            const T qv = lane_get(lk_bex, q) * pyx[q] + lane_get(lk_cex, q) * dHz;
            pyx[q] = qv;
            e -= cb1[q] * qv;
          }
          ey[q] = e;
        }
        if (jcap && (f & 96u)) {
          if (f & 32u) ring[(rbase + 0) * pcells + cw] = ey[q];
          if (f & 64u) ring[(rbase + 3) * pcells + cw] = ey[q];
        }
      }
      // ---- H half-step: Ey(i-1) of the first plane from the previous wave
      xey[w][lane] = ey[K - 1];
      lds_barrier();
      const T eydn = w > 0 ? xey[w - 1][lane] : ey[0];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const unsigned f = pf[q];
        const T exj = lane_up1(ex[q], ex[q]);
        const T prevEy = q > 0 ? ey[q - 1] : eydn;
        const bool src = jsrc && (f & 4u);
        T hv = hz[q];
        if ((f & 2u) && jin1) {
          const T dEx = ex[q] - exj;
          const T dEy = ey[q] - prevEy;
          hv = hv + db * (dEx * ikyh - dEy * lane_get(lk_ikh, q));
          if (src) hv -= jv;
          if (f & 16u) {
            const T qv = lane_get(lk_bhx, q) * qzx[q] + lane_get(lk_chx, q) * dEy;
            qzx[q] = qv;
            hv -= db * qv;
          }
          if (syh >= 0) {
            const T qv = bhy * qzy[q] + chy * dEx;
            qzy[q] = qv;
            hv += db * qv;
          }
        } else if (src) {
          hv -= jv;
        }
        hz[q] = hv;
        if (jcap && (f & 480u)) {
          if (f & 32u) ring[(rbase + 1) * pcells + cw] = hv;
          if (f & 128u) ring[(rbase + 2) * pcells + cw] = hv;
          if (f & 64u) ring[(rbase + 4) * pcells + cw] = hv;
          if (f & 256u) ring[(rbase + 5) * pcells + cw] = hv;
        }
      }
    }
    // ---- the inner cells to set B (fields unconditionally: cells no step
    // updates hold the same value in both sets)
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = ib0 + q;
      if (!jown || i < i0 || i >= i0 + TX || i >= nx) continue;
      const unsigned f = pf[q];
      int sxe, dxe, cxe, sxh, dxh, cxh;
      xslabs(i, sxe, dxe, cxe, sxh, dxh, cxh);
      const unsigned idx = (unsigned)i * ny + j;
      stg_nt<T>(ExB, idx * es, ex[q]);
      stg_nt<T>(EyB, idx * es, ey[q]);
      stg_nt<T>(HzB, idx * es, hz[q]);
      if ((f & 2u) && jin && sye >= 0) pxyB[(size_t)i * N + dye] = pxy[q];
      if ((f & 1u) && jin1 && sxe >= 0) psiBll[psi_id(F_EY, 0, sxe)][(size_t)dxe * ny + j] = pyx[q];
      if ((f & 2u) && jin1 && sxh >= 0) psiBll[psi_id(F_HZ, 0, sxh)][(size_t)dxh * ny + j] = qzx[q];
      if ((f & 2u) && jin1 && syh >= 0) pzyB[(size_t)i * N + dyh] = qzy[q];
    }
    lds_barrier();   // the next tile reuses the exchange planes
  }
}

// ------------------------------------------------------- fused 3D step ---
// One launch per 3D time step (form 3 in 3D), with no LDS and no barriers.  Each wavefront
// owns one row j and 63 z-cells of XM planes: lane l holds z-cell k0 - 1 + l,
// lane 0 being the cell below the tile, which only computes.  It reads
// field/psi set (n & 1), writes the other, and marches along x; per plane i:
//   - new E of row j, all three components (lane 0 included);
//   - new Ex and Ez of row j - 1, not stored: the j-1 neighbours of the H
//     update, recomputed from set A with the same per-cell code as their
//     owner computes them;
//   - new H of row j: the k-1 neighbours from the next-lower lane (DPP), the
//     x-1 neighbours (Ey, Ez of plane i-1) from the previous plane's
//     registers.
// Before the first plane it computes new E of plane i0 - 1 the same way.
// The k+1 neighbours of the E update come from the next-higher lane, the top
// lane loading its own.  Every load of plane i+1 is issued before plane i is
// updated.  The per-cell arithmetic and the update ranges are those of k_3d_e
// and k_3d_h.  HBM traffic per cell and step: each field read once and
// written once (48 B in FP32) plus psi and the material byte; the re-read
// row j-1 and z-cell k0 - 1 come from the caches.

// One plane's loads: row j (own), the top lane's k+1 values, row j-1.
template <typename T>
struct HaloPlane {
  T ex, ey, ez, hx, hy, hz, hxj, hzj, hy1, hz1;   // row j; hxj/hzj at j+1, hy1/hz1 at i+1
  T hxt, hyt;                                     // Hx, Hy at k+1 (top lane only)
  T qxy, qzy, qxz, qyz, qyx, qzx;                 // E psi ex_y, ez_y, ex_z, ey_z, ey_x, ez_x
  T rxy, rzy, rxz, ryz, ryx, rzx;                 // H psi hx_y, hz_y, hx_z, hy_z, hy_x, hz_x
  T ex_, ez_, hx_, hy_, hz_, hy1_, hyt_;          // row j-1
  T pxy_, pzy_, pxz_, pzx_;                       // its psi ex_y, ez_y, ex_z, ez_x
  unsigned m, m_;                                 // material bytes, rows j and j-1
};

template <typename T, int XM>
__global__ __launch_bounds__(kThreads, 4) void k_3d_halo(const Slot<T>* __restrict__ slots,
                                                          const int4* __restrict__ tiles,
                                                          const int* __restrict__ counts, int tstep) {
  static_assert(XM + 1 <= kBX, "one plane's x-only coefficients per lane");
  constexpr unsigned es = sizeof(T);
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  const int stride = gridDim.x * kBY;
  for (int tt = xcd_block() * kBY + wave_id(); tt < ntiles; tt += stride) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int nx = s.nx, ny = s.ny, nz = s.nz, N = s.N;
    const int i0 = tl.y, j = tl.z, k0 = tl.w;
    const int k = k0 - 1 + lane;
    const int kc = min(max(k, 0), nz - 1);
    const bool owner = lane >= 1 && k < nz;
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    const bool odd = (n & 1) != 0;
    OIDT_GLOBAL T* const* FA = odd ? s.f2 : s.f;
    OIDT_GLOBAL T* const* FB = odd ? s.f : s.f2;
    const T* __restrict__ ExA = FA[F_EX];
    const T* __restrict__ EyA = FA[F_EY];
    const T* __restrict__ EzA = FA[F_EZ];
    const T* __restrict__ HxA = FA[F_HX];
    const T* __restrict__ HyA = FA[F_HY];
    const T* __restrict__ HzA = FA[F_HZ];
    T* __restrict__ ExB = FB[F_EX];
    T* __restrict__ EyB = FB[F_EY];
    T* __restrict__ EzB = FB[F_EZ];
    T* __restrict__ HxB = FB[F_HX];
    T* __restrict__ HyB = FB[F_HY];
    T* __restrict__ HzB = FB[F_HZ];
    OIDT_GLOBAL T* const* psiA = odd ? s.psi2 : s.psi;
    OIDT_GLOBAL T* const* psiB = odd ? s.psi : s.psi2;
    const uint8_t* __restrict__ mat = s.mat;
    const T db = s.db;
    const unsigned pl = (unsigned)ny * nz;
    const int jm = max(j - 1, 0);
    const unsigned jk = (unsigned)j * nz + kc, jmk = (unsigned)jm * nz + kc;
    const unsigned offk = kc + 1 < nz ? es : 0u;                 // top lane's k+1 (clamped as k_3d_e)
    const unsigned offj = j + 1 < ny ? (unsigned)nz * es : 0u;
    // ---- row (wave-uniform) coefficients: E rows j and j-1, H row j
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    const bool jinm = j - 1 >= jb && j - 1 < je;
    int dj = 0, cj = 0, djm = 0, cjm = 0, djh = 0, cjh = 0;
    const int sj = jin ? slab_e(j, jb, je, N, dj, cj) : -1;
    const int sjm = jinm ? slab_e(j - 1, jb, je, N, djm, cjm) : -1;
    const int sjh = jin1 ? slab_h(j, jb, je, N, djh, cjh) : -1;
    const T iky = jin ? s.ike[1][j] : T(0), ikym = jinm ? s.ike[1][j - 1] : T(0);
    const T ikyh = jin1 ? s.ikh[1][j] : T(0);
    const T bej = sj >= 0 ? s.be[cj] : T(0), cej = sj >= 0 ? s.ceo[cj] : T(0);
    const T bejm = sjm >= 0 ? s.be[cjm] : T(0), cejm = sjm >= 0 ? s.ceo[cjm] : T(0);
    const T bhj = sjh >= 0 ? s.bh[cjh] : T(0), chj = sjh >= 0 ? s.cho[cjh] : T(0);
    // ---- lane (z) coefficients
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dk = 0, ck = 0, dkh = 0, ckh = 0;
    const int sk = kin ? slab_e(k, kb, ke, N, dk, ck) : -1;
    const int skh = kin1 ? slab_h(k, kb, ke, N, dkh, ckh) : -1;
    const T ikz = kin ? s.ike[2][kc] : T(0), ikzh = kin1 ? s.ikh[2][kc] : T(0);
    const T bek = sk >= 0 ? s.be[ck] : T(0), cek = sk >= 0 ? s.ceo[ck] : T(0);
    const T bhk = skh >= 0 ? s.bh[ckh] : T(0), chk = skh >= 0 ? s.cho[ckh] : T(0);
    const unsigned zside = sk == 1 ? s.psi_hi[2] : 0u, zsideh = skh == 1 ? s.psi_hi[2] : 0u, xhi = s.psi_hi[0];
    // sources: TE on Ey (E half-step), TM on Hy (start of the H half-step)
    const bool ksrc = k >= s.sz0 && k < s.sz1 && j >= s.sy0 && j < s.sy1;
    const bool srcE = s.src_field == F_EY && ksrc, srcH = s.src_field == F_HY && ksrc;
    const T jv = (srcE || srcH) ? s.jsrc[n] : T(0);
    const int sx0 = s.sx0, sx1 = s.sx1;
    // capture: E at k = port (planes cap*6 + 0/1; 6/7 when the ports coincide),
    // Hx/Hy at k = port (2/4, 8/10) and port + 1 (3/5, 9/11)
    const bool jcap = owner && j >= s.wy0 && j <= s.wy1;
    const int port0 = s.port[0], port1 = s.port[1];
    const int capE = !jcap ? -1 : (k == port0 ? 0 : (k == port1 ? 1 : -1));
    const bool capE2 = jcap && port0 == port1 && k == port0;
    const bool capA = jcap && k == port0, capB = jcap && k == port0 + 1;
    const bool capC = jcap && k == port1, capD = jcap && k == port1 + 1;
    const int wx0 = s.wx0, wx1 = s.wx1, prow = s.prow, pcells = s.pcells, wy0 = s.wy0;
    const size_t rbase = (size_t)tstep * s.nplanes;
    T* __restrict__ ring = s.ring;
    const T cx0 = s.cbv[0][0], cx1 = s.cbv[0][1], cx2 = s.cbv[0][2], cx3 = s.cbv[0][3];
    const T cy0 = s.cbv[1][0], cy1 = s.cbv[1][1], cy2 = s.cbv[1][2], cy3 = s.cbv[1][3];
    const T cz0 = s.cbv[2][0], cz1 = s.cbv[2][1], cz2 = s.cbv[2][2], cz3 = s.cbv[2][3];
    const int iend = min(i0 + XM, nx);
    // ---- x-only coefficients: lane u holds plane i0 - 1 + u
    T lx_ik = T(0), lx_be = T(0), lx_ce = T(0), lx_ikh = T(0), lx_bh = T(0), lx_ch = T(0);
    {
      const int ip = i0 - 1 + lane;
      if (lane <= XM && ip >= 0 && ip < nx) {
        int d = 0, ci = 0;
        lx_ik = s.ike[0][ip];
        if (slab_e(ip, ib, ie, N, d, ci) >= 0) { lx_be = s.be[ci]; lx_ce = s.ceo[ci]; }
        lx_ikh = s.ikh[0][ip];
        if (slab_h(ip, ib, ie, N, d, ci) >= 0) { lx_bh = s.bh[ci]; lx_ch = s.cho[ci]; }
      }
    }
    T* const pEyx = psiA[psi_id(F_EY, 0, 0)];
    T* const pEzx = psiA[psi_id(F_EZ, 0, 0)];
    T* const pExz = psiA[psi_id(F_EX, 2, 0)];
    T* const pEyz = psiA[psi_id(F_EY, 2, 0)];
    T* const pHyx = psiA[psi_id(F_HY, 0, 0)];
    T* const pHzx = psiA[psi_id(F_HZ, 0, 0)];
    T* const pHxz = psiA[psi_id(F_HX, 2, 0)];
    T* const pHyz = psiA[psi_id(F_HY, 2, 0)];
    T* const pExy = sj >= 0 ? psiA[psi_id(F_EX, 1, sj)] : nullptr;
    T* const pEzy = sj >= 0 ? psiA[psi_id(F_EZ, 1, sj)] : nullptr;
    T* const pExym = sjm >= 0 ? psiA[psi_id(F_EX, 1, sjm)] : nullptr;
    T* const pEzym = sjm >= 0 ? psiA[psi_id(F_EZ, 1, sjm)] : nullptr;
    T* const pHxy = sjh >= 0 ? psiA[psi_id(F_HX, 1, sjh)] : nullptr;
    T* const pHzy = sjh >= 0 ? psiA[psi_id(F_HZ, 1, sjh)] : nullptr;

    // plane i's loads; `full`: also row j-1 and the H psi (not needed for
    // the plane below the tile)
    auto load = [&](int i, bool full) {
      HaloPlane<T> p;
      const unsigned idx = (unsigned)i * pl + jk, bo = idx * es;
      const unsigned bo1 = (i + 1 < nx ? idx + pl : idx) * es;
      p.ex = ldg<T>(ExA, bo);
      p.ey = ldg<T>(EyA, bo);
      p.ez = ldg<T>(EzA, bo);
      p.hx = ldg<T>(HxA, bo);
      p.hy = ldg<T>(HyA, bo);
      p.hz = ldg<T>(HzA, bo);
      p.hxj = ldg<T>(HxA, bo + offj);
      p.hzj = ldg<T>(HzA, bo + offj);
      p.hy1 = ldg<T>(HyA, bo1);
      p.hz1 = ldg<T>(HzA, bo1);
      p.hxt = lane == 63 ? ldg<T>(HxA, bo + offk) : T(0);
      p.hyt = lane == 63 ? ldg<T>(HyA, bo + offk) : T(0);
      p.m = ldg<uint8_t>(mat, idx);
      int di = 0, ci = 0;
      const int si = slab_e(i, ib, ie, N, di, ci);
      const unsigned ix = (((unsigned)di * ny + j) * nz + kc) * es + (si == 1 ? xhi : 0u);
      const unsigned iy = (((unsigned)i * N + dj) * nz + kc) * es;
      const unsigned iz = (((unsigned)i * ny + j) * N + dk) * es + zside;
      p.qxy = sj >= 0 ? ldg<T>(pExy, iy) : T(0);
      p.qzy = sj >= 0 ? ldg<T>(pEzy, iy) : T(0);
      p.qxz = sk >= 0 ? ldg<T>(pExz, iz) : T(0);
      p.qyz = sk >= 0 ? ldg<T>(pEyz, iz) : T(0);
      p.qyx = si >= 0 ? ldg<T>(pEyx, ix) : T(0);
      p.qzx = si >= 0 ? ldg<T>(pEzx, ix) : T(0);
      p.rxy = p.rzy = p.rxz = p.ryz = p.ryx = p.rzx = T(0);
      p.ex_ = p.ez_ = p.hx_ = p.hy_ = p.hz_ = p.hy1_ = p.hyt_ = T(0);
      p.pxy_ = p.pzy_ = p.pxz_ = p.pzx_ = T(0);
      p.m_ = 0u;
      if (full) {
        int dih = 0, cih = 0;
        const int sih = slab_h(i, ib, ie, N, dih, cih);
        const unsigned ixh = (((unsigned)dih * ny + j) * nz + kc) * es + (sih == 1 ? xhi : 0u);
        const unsigned iyh = (((unsigned)i * N + djh) * nz + kc) * es;
        const unsigned izh = (((unsigned)i * ny + j) * N + dkh) * es + zsideh;
        p.rxy = sjh >= 0 ? ldg<T>(pHxy, iyh) : T(0);
        p.rzy = sjh >= 0 ? ldg<T>(pHzy, iyh) : T(0);
        p.rxz = skh >= 0 ? ldg<T>(pHxz, izh) : T(0);
        p.ryz = skh >= 0 ? ldg<T>(pHyz, izh) : T(0);
        p.ryx = sih >= 0 ? ldg<T>(pHyx, ixh) : T(0);
        p.rzx = sih >= 0 ? ldg<T>(pHzx, ixh) : T(0);
        if (j > 0) {
          const unsigned idm = (unsigned)i * pl + jmk, bm = idm * es;
          const unsigned bm1 = (i + 1 < nx ? idm + pl : idm) * es;
          p.ex_ = ldg<T>(ExA, bm);
          p.ez_ = ldg<T>(EzA, bm);
          p.hx_ = ldg<T>(HxA, bm);
          p.hy_ = ldg<T>(HyA, bm);
          p.hz_ = ldg<T>(HzA, bm);
          p.hy1_ = ldg<T>(HyA, bm1);
          p.hyt_ = lane == 63 ? ldg<T>(HyA, bm + offk) : T(0);
          p.m_ = ldg<uint8_t>(mat, idm);
          const unsigned ixm = (((unsigned)di * ny + jm) * nz + kc) * es + (si == 1 ? xhi : 0u);
          const unsigned iym = (((unsigned)i * N + djm) * nz + kc) * es;
          const unsigned izm = (((unsigned)i * ny + jm) * N + dk) * es + zside;
          p.pxy_ = sjm >= 0 ? ldg<T>(pExym, iym) : T(0);
          p.pzy_ = sjm >= 0 ? ldg<T>(pEzym, iym) : T(0);
          p.pxz_ = sk >= 0 ? ldg<T>(pExz, izm) : T(0);
          p.pzx_ = si >= 0 ? ldg<T>(pEzx, ixm) : T(0);
        }
      }
      return p;
    };

    T eyp = T(0), ezp = T(0);   // new Ey, Ez of plane i-1, row j
    const int ifirst = i0 > 0 ? i0 - 1 : i0;
    HaloPlane<T> pc = load(ifirst, ifirst >= i0);
#pragma unroll
    for (int u = 0; u <= XM; ++u) {
      const int i = i0 - 1 + u;
      if (i < ifirst) continue;
      if (i >= iend) break;
      HaloPlane<T> pn;
      if (u < XM && i + 1 < iend) pn = load(i + 1, true);
      const bool own = i >= i0;               // plane i0 - 1: new E of row j only
      const int ux = i - (i0 - 1);
      const T ikx = lane_get(lx_ik, ux), bex = lane_get(lx_be, ux), cex = lane_get(lx_ce, ux);
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      int di = 0, ci = 0;
      const int si = slab_e(i, ib, ie, N, di, ci);
      const unsigned idx = (unsigned)i * pl + jk, bo = idx * es;
      const unsigned ix = (((unsigned)di * ny + j) * nz + kc) * es + (si == 1 ? xhi : 0u);
      const unsigned iy = (((unsigned)i * N + dj) * nz + kc) * es;
      const unsigned iz = (((unsigned)i * ny + j) * N + dk) * es + zside;
      const bool st = own && owner;
      // ---- new E, row j (k_3d_e per cell).  The plane and row conditions
      // are wave-uniform branches; the z conditions select per lane (outside
      // the z slabs the z-psi coefficients and values are zero and add an
      // exact zero), so a wave does not split on them.
      const T hxk = lane_down1(pc.hx, pc.hxt);
      const T hyk = lane_down1(pc.hy, pc.hyt);
      T vx = pc.ex, vy = pc.ey, vz = pc.ez;
      const bool srcp = srcE && i >= sx0 && i < sx1;
      if (iin1 && jin) {
        const T cb = cb_of(cx0, cx1, cx2, cx3, pc.m, 0);
        const T dHz = pc.hzj - pc.hz;
        const T dHy = hyk - pc.hy;
        T v = vx + cb * (dHz * iky - dHy * ikz);
        if (sj >= 0) { const T q = bej * pc.qxy + cej * dHz; if (st && kin) stg<T>(psiB[psi_id(F_EX, 1, sj)], iy, q); v += cb * q; }
        { const T q = bek * pc.qxz + cek * dHy; if (st && sk >= 0) stg<T>(psiB[psi_id(F_EX, 2, 0)], iz, q); v -= cb * q; }
        vx = kin ? v : vx;
      }
      if (iin && jin1) {
        const T cb = cb_of(cy0, cy1, cy2, cy3, pc.m, 1);
        const T dHx = hxk - pc.hx;
        const T dHz = pc.hz1 - pc.hz;
        T v = vy + cb * (dHx * ikz - dHz * ikx);
        v = srcp ? v - jv : v;
        if (si >= 0) { const T q = bex * pc.qyx + cex * dHz; if (st && kin) stg<T>(psiB[psi_id(F_EY, 0, 0)], ix, q); v -= cb * q; }
        { const T q = bek * pc.qyz + cek * dHx; if (st && sk >= 0) stg<T>(psiB[psi_id(F_EY, 2, 0)], iz, q); v += cb * q; }
        vy = kin ? v : (srcp ? vy - jv : vy);
      } else {
        vy = srcp ? vy - jv : vy;
      }
      if (iin && jin) {
        const T cb = cb_of(cz0, cz1, cz2, cz3, pc.m, 2);
        const T dHy = pc.hy1 - pc.hy;
        const T dHx = pc.hxj - pc.hx;
        T v = vz + cb * (dHy * ikx - dHx * iky);
        if (si >= 0) { const T q = bex * pc.qzx + cex * dHy; if (st && kin1) stg<T>(psiB[psi_id(F_EZ, 0, 0)], ix, q); v += cb * q; }
        if (sj >= 0) { const T q = bej * pc.qzy + cej * dHx; if (st && kin1) stg<T>(psiB[psi_id(F_EZ, 1, sj)], iy, q); v -= cb * q; }
        vz = kin1 ? v : vz;
      }
      if (own) {
        // ---- new Ex, Ez of row j-1 (its owner's per-cell code, not stored)
        T vxm = pc.ex_, vzm = pc.ez_;
        const T hykm = lane_down1(pc.hy_, pc.hyt_);
        if (iin1 && jinm) {
          const T cb = cb_of(cx0, cx1, cx2, cx3, pc.m_, 0);
          const T dHz = pc.hz - pc.hz_;
          const T dHy = hykm - pc.hy_;
          T v = vxm + cb * (dHz * ikym - dHy * ikz);
          if (sjm >= 0) { const T q = bejm * pc.pxy_ + cejm * dHz; v += cb * q; }
          { const T q = bek * pc.pxz_ + cek * dHy; v -= cb * q; }
          vxm = kin ? v : vxm;
        }
        if (iin && jinm) {
          const T cb = cb_of(cz0, cz1, cz2, cz3, pc.m_, 2);
          const T dHy = pc.hy1_ - pc.hy_;
          const T dHx = pc.hx - pc.hx_;
          T v = vzm + cb * (dHy * ikx - dHx * ikym);
          if (si >= 0) { const T q = bex * pc.pzx_ + cex * dHy; v += cb * q; }
          if (sjm >= 0) { const T q = bejm * pc.pzy_ + cejm * dHx; v -= cb * q; }
          vzm = kin1 ? v : vzm;
        }
        // ---- new H, row j (k_3d_h per cell)
        const T exk = lane_up1(vx, vx), eyk = lane_up1(vy, vy);
        const T ikxh = lane_get(lx_ikh, ux), bhx = lane_get(lx_bh, ux), chx = lane_get(lx_ch, ux);
        int dih = 0, cih = 0;
        const int sih = slab_h(i, ib, ie, N, dih, cih);
        const unsigned ixh = (((unsigned)dih * ny + j) * nz + kc) * es + (sih == 1 ? xhi : 0u);
        const unsigned iyh = (((unsigned)i * N + djh) * nz + kc) * es;
        const unsigned izh = (((unsigned)i * ny + j) * N + dkh) * es + zsideh;
        const bool srch = srcH && i >= sx0 && i < sx1;
        T hx = pc.hx;
        if (iin && jin1) {
          const T dEy = vy - eyk;
          const T dEz = vz - vzm;
          T v = hx + db * (dEy * ikzh - dEz * ikyh);
          if (sjh >= 0) { const T q = bhj * pc.rxy + chj * dEz; if (owner && kin1) stg<T>(psiB[psi_id(F_HX, 1, sjh)], iyh, q); v -= db * q; }
          { const T q = bhk * pc.rxz + chk * dEy; if (owner && skh >= 0) stg<T>(psiB[psi_id(F_HX, 2, 0)], izh, q); v += db * q; }
          hx = kin1 ? v : hx;
        }
        T hy = srch ? pc.hy - jv : pc.hy;
        if (iin1 && jin) {
          const T dEz = vz - ezp;
          const T dEx = vx - exk;
          T v = hy + db * (dEz * ikxh - dEx * ikzh);
          if (sih >= 0) { const T q = bhx * pc.ryx + chx * dEz; if (owner && kin1) stg<T>(psiB[psi_id(F_HY, 0, 0)], ixh, q); v += db * q; }
          { const T q = bhk * pc.ryz + chk * dEx; if (owner && skh >= 0) stg<T>(psiB[psi_id(F_HY, 2, 0)], izh, q); v -= db * q; }
          hy = kin1 ? v : hy;
        }
        T hz = pc.hz;
        if (iin1 && jin1) {
          const T dEx = vx - vxm;
          const T dEy = vy - eyp;
          T v = hz + db * (dEx * ikyh - dEy * ikxh);
          if (sih >= 0) { const T q = bhx * pc.rzx + chx * dEy; if (owner && kin) stg<T>(psiB[psi_id(F_HZ, 0, 0)], ixh, q); v -= db * q; }
          if (sjh >= 0) { const T q = bhj * pc.rzy + chj * dEx; if (owner && kin) stg<T>(psiB[psi_id(F_HZ, 1, sjh)], iyh, q); v += db * q; }
          hz = kin ? v : hz;
        }
        if (owner) {
          stg_nt<T>(ExB, bo, vx);
          stg_nt<T>(EyB, bo, vy);
          stg_nt<T>(EzB, bo, vz);
          stg_nt<T>(HxB, bo, hx);
          stg_nt<T>(HyB, bo, hy);
          stg_nt<T>(HzB, bo, hz);
          if (jcap && i >= wx0 && i <= wx1) {
            const size_t c = (size_t)((i - wx0) * prow + (j - wy0));
            if (capE >= 0) {
              ring[(rbase + capE * 6 + 0) * pcells + c] = vx;
              ring[(rbase + capE * 6 + 1) * pcells + c] = vy;
            }
            if (capE2) {
              ring[(rbase + 6) * pcells + c] = vx;
              ring[(rbase + 7) * pcells + c] = vy;
            }
            if (capA) { ring[(rbase + 2) * pcells + c] = hx; ring[(rbase + 4) * pcells + c] = hy; }
            if (capB) { ring[(rbase + 3) * pcells + c] = hx; ring[(rbase + 5) * pcells + c] = hy; }
            if (capC) { ring[(rbase + 8) * pcells + c] = hx; ring[(rbase + 10) * pcells + c] = hy; }
            if (capD) { ring[(rbase + 9) * pcells + c] = hx; ring[(rbase + 11) * pcells + c] = hy; }
          }
        }
      }
      eyp = vy;
      ezp = vz;
      if (u < XM && i + 1 < iend) pc = pn;
    }
  }
}

// ------------------------------------------------- temporally blocked 3D ---
// TB = 2 Yee steps per launch, streamed along x.  A block of W waves takes W
// rows (y, one per wave) by 64 z-cells (one per lane) and marches them over a
// segment of planes; its own cells are the inner (W - 2 TB) rows by
// (64 - 2 TB) z-cells of the kTb3Seg planes from x0.  Step 1 runs one plane
// behind step 0: at position p, step 0 updates plane p - 2 (its E reads H of
// plane p - 1, loaded one position earlier) and step 1 plane p - 3 (its E
// reads H of plane p - 2, which step 0 has just produced; its H reads E of
// plane p - 4, which it produced itself one position earlier), while plane
// p loads.  Every field and psi value crosses HBM once per launch
// (with the TB ghost planes at both ends of the segment and the TB-cell
// ghost frame of rows and z-cells), instead of twice per step.  Neighbours:
// z by DPP lane shifts, y through LDS (old Hx and Hz of the row above for
// the E update, new Ex and Ez of the row below for the H update), x in
// registers.  Each step moves the cells that depend on values outside the
// tile one cell inwards from every side (E reads H at +1 and H reads E at -1
// along each axis), so after TB steps the inner cells are exact.  The
// per-cell arithmetic and update ranges are k_3d_e's and k_3d_h's.  The
// launch reads field/psi set ((n0 / TB) & 1) and writes the inner cells'
// state after its steps to the other set.
constexpr int kTb3Seg = 32;   // planes of a 3D blocked tile

// One cell's state: fields, psi and material byte.
template <typename T>
struct Cell3 {
  T ex, ey, ez, hx, hy, hz;
  T qxy, qzy, qxz, qyz, qyx, qzx;   // E psi: Ex_y, Ez_y, Ex_z, Ey_z, Ey_x, Ez_x
  T rxy, rzy, rxz, ryz, ryx, rzx;   // H psi: Hx_y, Hz_y, Hx_z, Hy_z, Hy_x, Hz_x
  unsigned m;
};

// Per-lane conditions of the blocked 3D step, one bit each.
enum : unsigned {
  L3_KIN = 1u, L3_KIN1 = 2u, L3_SRCE = 4u, L3_SRCH = 8u, L3_OWN = 16u, L3_SK = 32u, L3_SKH = 64u,
  L3_CAPE0 = 128u,    // Ex, Ey at port 0 (capture planes 0, 1)
  L3_CAPE1 = 256u,    // Ex, Ey at port 1 (planes 6, 7), or at port 0 when the ports coincide
  L3_CAPA = 512u, L3_CAPB = 1024u, L3_CAPC = 2048u, L3_CAPD = 4096u,   // Hx, Hy at port0, +1, port1, +1
  L3_CAP = L3_CAPE0 | L3_CAPE1 | L3_CAPA | L3_CAPB | L3_CAPC | L3_CAPD
};

// A copy of a per-lane value that the compiler cannot see through.  Taken
// once per position of the march, it keeps what the loop derives from the
// value (per-lane flag masks, readlanes of a lane table) where it is used,
// instead of hoisted out of the loop into scalar registers for the whole
// march, which does not have enough of them.
template <typename T>
__device__ __forceinline__ T opaque_vgpr(T v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename T, int TB, int W>
__global__ __launch_bounds__(64 * W) void k_3d_tb(const Slot<T>* __restrict__ slots,
                                                  const int4* __restrict__ tiles,
                                                  const int* __restrict__ counts, int tstep) {
  constexpr int TY = W - 2 * TB, TZ = 64 - 2 * TB;   // inner rows and z-cells of a tile
  static_assert(TB == 2 && TY > 0 && kTb3Seg + 2 * TB <= 64, "two steps per launch");
  constexpr unsigned es = sizeof(T);
  __shared__ T xh[TB][2][W][64];   // old Hx, Hz of each row at each step's plane
  __shared__ T xe[2][2][W][64];    // new Ex, Ez of each row (alternate steps alternate halves)
  const int ntiles = counts[0];
  const int lane = threadIdx.x;
  const int w = wave_id();
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n0 = s.base + tstep;
    if (!s.active || n0 >= s.nt) continue;      // uniform over the block
    const int nsteps = min(TB, s.nt - n0);
    const bool odd = ((n0 / TB) & 1) != 0;
    const int nx = s.nx, ny = s.ny, nz = s.nz, N = s.N;
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    const int x0 = tl.y, j0 = tl.z, k0 = tl.w;
    const int j = j0 - TB + w, k = k0 - TB + lane;
    const int jc = min(max(j, 0), ny - 1), kc = min(max(k, 0), nz - 1);
    const bool own = j >= j0 && j < j0 + TY && j < ny && k >= k0 && k < k0 + TZ && k < nz;
    const bool jgrid = j >= 0 && j < ny;
    const int xo1 = min(x0 + kTb3Seg, nx);                       // inner planes [x0, xo1)
    const int xf = max(x0 - TB, 0), xl = min(xo1 - 1 + TB, nx - 1);   // loaded planes [xf, xl]
    const unsigned pl = (unsigned)ny * nz;
    const unsigned jk = (unsigned)jc * nz + kc;
    // ---- row (wave-uniform) constants
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    int dj = 0, cj = 0, djh = 0, cjh = 0;
    const int sj = jin ? slab_e(j, jb, je, N, dj, cj) : -1;
    const int sjh = jin1 ? slab_h(j, jb, je, N, djh, cjh) : -1;
    const T iky = jin ? s.ike[1][j] : T(0), ikyh = jin1 ? s.ikh[1][j] : T(0);
    const T bej = sj >= 0 ? s.be[cj] : T(0), cej = sj >= 0 ? s.ceo[cj] : T(0);
    const T bhj = sjh >= 0 ? s.bh[cjh] : T(0), chj = sjh >= 0 ? s.cho[cjh] : T(0);
    // ---- lane (z) constants and conditions; zs: some z-cell of the tile
    // lies in a z slab
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dk = 0, ck = 0, dkh = 0, ckh = 0;
    const int sk = kin ? slab_e(k, kb, ke, N, dk, ck) : -1;
    const int skh = kin1 ? slab_h(k, kb, ke, N, dkh, ckh) : -1;
    const T ikz = kin ? s.ike[2][kc] : T(0), ikzh = kin1 ? s.ikh[2][kc] : T(0);
    const T bek = sk >= 0 ? s.be[ck] : T(0), cek = sk >= 0 ? s.ceo[ck] : T(0);
    const T bhk = skh >= 0 ? s.bh[ckh] : T(0), chk = skh >= 0 ? s.cho[ckh] : T(0);
    const bool zs = k0 - TB < kb + N || k0 - TB + 64 > ke - N;
    const unsigned xhi = s.psi_hi[0];
    // z-slab psi offsets within a plane row (the row part is added per plane)
    const unsigned izk = (unsigned)dk * es + (sk == 1 ? s.psi_hi[2] : 0u);
    const unsigned izkh = (unsigned)dkh * es + (skh == 1 ? s.psi_hi[2] : 0u);
    // sources: TE on Ey (E half-step), TM on Hy (start of the H half-step);
    // lane t holds the waveform of step n0 + t
    const bool ksrc = k >= s.sz0 && k < s.sz1 && j >= s.sy0 && j < s.sy1;
    const T lk_jv = s.src_field >= 0 && lane < nsteps ? s.jsrc[n0 + lane] : T(0);
    // capture: E at k = port (planes cap*6 + 0/1; 6/7 when the ports coincide),
    // Hx/Hy at k = port (2/4, 8/10) and port + 1 (3/5, 9/11)
    const bool jcap = own && j >= s.wy0 && j <= s.wy1;
    const int port0 = s.port[0], port1 = s.port[1];
    const unsigned lflags =
        (kin ? L3_KIN : 0u) | (kin1 ? L3_KIN1 : 0u) | (s.src_field == F_EY && ksrc ? L3_SRCE : 0u) |
        (s.src_field == F_HY && ksrc ? L3_SRCH : 0u) | (own ? L3_OWN : 0u) | (sk >= 0 ? L3_SK : 0u) |
        (skh >= 0 ? L3_SKH : 0u) | (jcap && k == port0 ? L3_CAPE0 | L3_CAPA : 0u) |
        (jcap && k == port1 ? L3_CAPE1 | L3_CAPC : 0u) | (jcap && k == port0 + 1 ? L3_CAPB : 0u) |
        (jcap && k == port1 + 1 ? L3_CAPD : 0u);
    // some lane of the tile captures (wave-uniform)
    const bool tcap = __builtin_amdgcn_readfirstlane(__any((lflags & L3_CAP) != 0u)) != 0;
    const bool tsrc = s.src_field >= 0;
    // ---- x-only coefficients: lane u holds plane xf + u
    T lx_ik = T(0), lx_be = T(0), lx_ce = T(0), lx_ikh = T(0), lx_bh = T(0), lx_ch = T(0);
    {
      const int ip = xf + lane;
      if (ip <= xl) {
        int d = 0, ci = 0;
        lx_ik = s.ike[0][ip];
        if (slab_e(ip, ib, ie, N, d, ci) >= 0) { lx_be = s.be[ci]; lx_ce = s.ceo[ci]; }
        lx_ikh = s.ikh[0][ip];
        if (slab_h(ip, ib, ie, N, d, ci) >= 0) { lx_bh = s.bh[ci]; lx_ch = s.cho[ci]; }
      }
    }

    // Plane i's fields and material byte, and its psi values, from set A
    // (rows and z-cells outside the grid load a clamped neighbour: no cell of
    // the grid that a step updates reads them).
    // The run's arrays as 32-bit byte offsets from its first array (field Ex
    // of set A; the allocation is below 4 GiB), one per lane of otab: 0-5
    // fields read (set A), 6-11 fields written (set B), 12 material, 13
    // capture ring, 14-25 psi read, 26-37 psi written (x slabs Ey Ez Hy Hz,
    // y slabs of this row's side Ex Ez Hx Hz, z slabs Ex Ey Hx Hy).  A
    // readlane turns one into a scalar operand where it is used: no array
    // pointer is held in (or reloaded into) scalar registers over the march.
    const char* const rb = reinterpret_cast<const char*>(s.f[F_EX]);
    unsigned otab = 0u;
    {
      OIDT_GLOBAL T* const* const fa = odd ? s.f2 : s.f;
      OIDT_GLOBAL T* const* const fb = odd ? s.f : s.f2;
      const void* ptr = nullptr;
      if (lane < 6) {
        ptr = fa[lane];
      } else if (lane < 12) {
        ptr = fb[lane - 6];
      } else if (lane == 12) {
        ptr = s.mat;
      } else if (lane == 13) {
        ptr = s.ring;
      } else if (lane < 38) {
        const int t = (lane - 14) % 12;
        OIDT_GLOBAL T* const* const pp = (lane < 26) == odd ? s.psi2 : s.psi;
        const int ys = max(sj, 0), ysh = max(sjh, 0);
        const int id = t == 0 ? psi_id(F_EY, 0, 0) : t == 1 ? psi_id(F_EZ, 0, 0) : t == 2 ? psi_id(F_HY, 0, 0)
                     : t == 3 ? psi_id(F_HZ, 0, 0) : t == 4 ? psi_id(F_EX, 1, ys) : t == 5 ? psi_id(F_EZ, 1, ys)
                     : t == 6 ? psi_id(F_HX, 1, ysh) : t == 7 ? psi_id(F_HZ, 1, ysh) : t == 8 ? psi_id(F_EX, 2, 0)
                     : t == 9 ? psi_id(F_EY, 2, 0) : t == 10 ? psi_id(F_HX, 2, 0) : psi_id(F_HY, 2, 0);
        ptr = pp[id];
      }
      if (ptr) otab = (unsigned)(reinterpret_cast<const char*>(ptr) - rb);
      // 40-51: the cb values (component c, index q at 40 + 4c + q), FP32
      // bits (FP64 runs keep them in scalar registers); 52-59: source and
      // capture constants
      if (sizeof(T) == 4 && lane >= 40 && lane < 52) otab = __float_as_uint((float)s.cbv[(lane - 40) >> 2][(lane - 40) & 3]);
      if (lane == 52) otab = (unsigned)s.sx0;
      if (lane == 53) otab = (unsigned)s.sx1;
      if (lane == 54) otab = (unsigned)s.wx0;
      if (lane == 55) otab = (unsigned)s.wx1;
      if (lane == 56) otab = (unsigned)s.wy0;
      if (lane == 57) otab = (unsigned)s.pcells;
      if (lane == 58) otab = (unsigned)s.prow;
      if (lane == 59) otab = (unsigned)s.nplanes;
    }
    unsigned ot = otab;   // the copy the accesses read (made opaque once per position)
    auto O = [&](int c) { return (unsigned)__builtin_amdgcn_readlane((int)ot, c); };
    const T* const rbT = reinterpret_cast<const T*>(rb);
    T* const rbW = const_cast<T*>(rbT);

    auto load_f = [&](int i, Cell3<T>& c) {
      const unsigned idx = (unsigned)i * pl + jk, bo = idx * es;
      c.hy = ldg<T>(rbT, O(F_HY) + bo);   // first: the step one plane behind needs them a position earlier
      c.hz = ldg<T>(rbT, O(F_HZ) + bo);
      c.ex = ldg<T>(rbT, O(F_EX) + bo);
      c.ey = ldg<T>(rbT, O(F_EY) + bo);
      c.ez = ldg<T>(rbT, O(F_EZ) + bo);
      c.hx = ldg<T>(rbT, O(F_HX) + bo);
      c.m = ldg<uint8_t>(reinterpret_cast<const uint8_t*>(rb), O(12) + idx);
    };
    auto load_q = [&](int i, unsigned lf, Cell3<T>& c) {
      int di = 0, ci = 0, dih = 0, cih = 0;
      const int si = slab_e(i, ib, ie, N, di, ci);
      const int sih = slab_h(i, ib, ie, N, dih, cih);
      // outside its slab a z lane reads element 0 (one line for the wave)
      const unsigned ix = si >= 0 ? ((unsigned)di * pl + jk) * es + (si == 1 ? xhi : 0u) : 0u;
      const unsigned ixh = sih >= 0 ? ((unsigned)dih * pl + jk) * es + (sih == 1 ? xhi : 0u) : 0u;
      const unsigned iy = sj >= 0 ? (((unsigned)i * N + dj) * nz + kc) * es : 0u;
      const unsigned iyh = sjh >= 0 ? (((unsigned)i * N + djh) * nz + kc) * es : 0u;
      const unsigned zrow = ((unsigned)i * ny + jc) * N * es;
      const unsigned iz = (lf & L3_SK) ? zrow + izk : 0u;
      const unsigned izh = (lf & L3_SKH) ? zrow + izkh : 0u;
      // psi loads only where a slab can hold the cell (the plane, the row or
      // the tile is uniform over the wave); elsewhere the values are zero
      c.qyx = c.qzx = c.ryx = c.rzx = c.qxy = c.qzy = c.rxy = c.rzy = c.qxz = c.qyz = c.rxz = c.ryz = T(0);
      if (si >= 0) {
        c.qyx = ldg<T>(rbT, O(14) + ix);
        c.qzx = ldg<T>(rbT, O(15) + ix);
      }
      if (sih >= 0) {
        c.ryx = ldg<T>(rbT, O(16) + ixh);
        c.rzx = ldg<T>(rbT, O(17) + ixh);
      }
      if (sj >= 0) {
        c.qxy = ldg<T>(rbT, O(18) + iy);
        c.qzy = ldg<T>(rbT, O(19) + iy);
      }
      if (sjh >= 0) {
        c.rxy = ldg<T>(rbT, O(20) + iyh);
        c.rzy = ldg<T>(rbT, O(21) + iyh);
      }
      if (zs) {
        c.qxz = ldg<T>(rbT, O(22) + iz);
        c.qyz = ldg<T>(rbT, O(23) + iz);
        c.rxz = ldg<T>(rbT, O(24) + izh);
        c.ryz = ldg<T>(rbT, O(25) + izh);
      }
    };

    // One Yee step of cell c (plane i, step n0 + st): E, then H.  hyn, hzn:
    // old Hy, Hz of plane i + 1; hxu, hzu: old Hx, Hz of row j + 1; pey,
    // pez: new Ey, Ez of plane i - 1 (replaced by this plane's).
    const T db = s.db;
    // cb of component c for cell material byte m: FP32 from the lane table
    // (a cross-lane read, no scalar registers; called only under wave-uniform
    // branches, so the table's lanes are active), FP64 from scalar registers
    const T cbd[3][4] = {{s.cbv[0][0], s.cbv[0][1], s.cbv[0][2], s.cbv[0][3]},
                         {s.cbv[1][0], s.cbv[1][1], s.cbv[1][2], s.cbv[1][3]},
                         {s.cbv[2][0], s.cbv[2][1], s.cbv[2][2], s.cbv[2][3]}};
    auto cbc = [&](unsigned m, int c) -> T {
      if constexpr (sizeof(T) == 4) {
        const int q = (m >> (2 * c)) & 3;
        return __int_as_float(__builtin_amdgcn_ds_bpermute((40 + 4 * c + q) * 4, (int)ot));
      } else {
        return cb_of(cbd[c][0], cbd[c][1], cbd[c][2], cbd[c][3], m, c);
      }
    };
    auto step = [&](Cell3<T>& c, int i, int st, unsigned lf, T hyn, T hzn, T hxu, T hzu, T& pey, T& pez) {
      const int u = i - xf;
      const T ikx = lane_get(lx_ik, u), bex = lane_get(lx_be, u), cex = lane_get(lx_ce, u);
      const T ikxh = lane_get(lx_ikh, u), bhx = lane_get(lx_bh, u), chx = lane_get(lx_ch, u);
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool si = i >= ib && (i < ib + N || i >= ie - N) && i < ie;          // in an E x slab
      const bool sih = i > ib && (i < ib + N || i > ie - N) && i < ie;           // in an H x slab
      const bool lkin = (lf & L3_KIN) != 0u, lkin1 = (lf & L3_KIN1) != 0u;
      // ---- E (k_3d_e per cell)
      const T hxk = lane_down1(c.hx, c.hx);
      const T hyk = lane_down1(c.hy, c.hy);
      T vx = c.ex, vy = c.ey, vz = c.ez;
      if (iin1 && jin) {
        const T cb = cbc(c.m, 0);
        const T dHz = hzu - c.hz;
        const T dHy = hyk - c.hy;
        T v = vx + cb * (dHz * iky - dHy * ikz);
        if (sj >= 0) { const T qv = bej * c.qxy + cej * dHz; c.qxy = qv; v += cb * qv; }
        if (zs) { const T qv = bek * c.qxz + cek * dHy; c.qxz = qv; v -= cb * qv; }
        vx = lkin ? v : vx;
      }
      const bool srcx = tsrc && i >= (int)O(52) && i < (int)O(53);
      const bool srcp = srcx && (lf & L3_SRCE) != 0u;
      const T jv = tsrc ? lane_get(lk_jv, st) : T(0);
      if (iin && jin1) {
        const T cb = cbc(c.m, 1);
        const T dHx = hxk - c.hx;
        const T dHz = hzn - c.hz;
        T v = vy + cb * (dHx * ikz - dHz * ikx);
        v = srcp ? v - jv : v;
        if (si) { const T qv = bex * c.qyx + cex * dHz; c.qyx = qv; v -= cb * qv; }
        if (zs) { const T qv = bek * c.qyz + cek * dHx; c.qyz = qv; v += cb * qv; }
        vy = lkin ? v : (srcp ? vy - jv : vy);
      } else {
        vy = srcp ? vy - jv : vy;
      }
      if (iin && jin) {
        const T cb = cbc(c.m, 2);
        const T dHy = hyn - c.hy;
        const T dHx = hxu - c.hx;
        T v = vz + cb * (dHy * ikx - dHx * iky);
        if (si) { const T qv = bex * c.qzx + cex * dHy; c.qzx = qv; v += cb * qv; }
        if (sj >= 0) { const T qv = bej * c.qzy + cej * dHx; c.qzy = qv; v -= cb * qv; }
        vz = lkin1 ? v : vz;
      }
      c.ex = vx;
      c.ey = vy;
      c.ez = vz;
      const bool capx = tcap && i >= x0 && i < xo1 && i >= (int)O(54) && i <= (int)O(55);
      auto put = [&](T v, unsigned plane) {
        const unsigned rbase = (unsigned)(tstep + st) * O(59);
        const unsigned cw = (unsigned)(i - (int)O(54)) * O(58) + (unsigned)(j - (int)O(56));
        stg<T>(rbW, O(13) + ((rbase + plane) * O(57) + cw) * es, v);
      };
      if (capx) {
        if (lf & L3_CAPE0) { put(vx, 0); put(vy, 1); }
        if (lf & L3_CAPE1) { put(vx, 6); put(vy, 7); }
      }
      // ---- new Ex, Ez of row j - 1 from the wave below
      T* const xa = &xe[st & 1][0][0][0];
      T* const xb = &xe[st & 1][1][0][0];
      xa[w * 64 + lane] = vx;
      xb[w * 64 + lane] = vz;
      lds_barrier();
      const T exm = w > 0 ? xa[(w - 1) * 64 + lane] : vx;
      const T ezm = w > 0 ? xb[(w - 1) * 64 + lane] : vz;
      // ---- H (k_3d_h per cell)
      const T exk = lane_up1(vx, vx), eyk = lane_up1(vy, vy);
      T hx = c.hx;
      if (iin && jin1) {
        const T dEy = vy - eyk;
        const T dEz = vz - ezm;
        T v = hx + db * (dEy * ikzh - dEz * ikyh);
        if (sjh >= 0) { const T qv = bhj * c.rxy + chj * dEz; c.rxy = qv; v -= db * qv; }
        if (zs) { const T qv = bhk * c.rxz + chk * dEy; c.rxz = qv; v += db * qv; }
        hx = lkin1 ? v : hx;
      }
      const bool srch = srcx && (lf & L3_SRCH) != 0u;
      T hy = srch ? c.hy - jv : c.hy;
      if (iin1 && jin) {
        const T dEz = vz - pez;
        const T dEx = vx - exk;
        T v = hy + db * (dEz * ikxh - dEx * ikzh);
        if (sih) { const T qv = bhx * c.ryx + chx * dEz; c.ryx = qv; v += db * qv; }
        if (zs) { const T qv = bhk * c.ryz + chk * dEx; c.ryz = qv; v -= db * qv; }
        hy = lkin1 ? v : hy;
      }
      T hz = c.hz;
      if (iin1 && jin1) {
        const T dEx = vx - exm;
        const T dEy = vy - pey;
        T v = hz + db * (dEx * ikyh - dEy * ikxh);
        if (sih) { const T qv = bhx * c.rzx + chx * dEy; c.rzx = qv; v -= db * qv; }
        if (sjh >= 0) { const T qv = bhj * c.rzy + chj * dEx; c.rzy = qv; v += db * qv; }
        hz = lkin ? v : hz;
      }
      c.hx = hx;
      c.hy = hy;
      c.hz = hz;
      pey = vy;
      pez = vz;
      if (capx) {
        if (lf & L3_CAPA) { put(hx, 2); put(hy, 4); }
        if (lf & L3_CAPB) { put(hx, 3); put(hy, 5); }
        if (lf & L3_CAPC) { put(hx, 8); put(hy, 10); }
        if (lf & L3_CAPD) { put(hx, 9); put(hy, 11); }
      }
    };

    // plane i's inner cells to set B (fields unconditionally: cells no step
    // updates hold the same value in both sets; psi where the steps update it)
    auto store = [&](const Cell3<T>& c, int i, unsigned lf) {
      const unsigned idx = (unsigned)i * pl + jk, bo = idx * es;
      stg_nt<T>(rbW, O(6 + F_EX) + bo, c.ex);
      stg_nt<T>(rbW, O(6 + F_EY) + bo, c.ey);
      stg_nt<T>(rbW, O(6 + F_EZ) + bo, c.ez);
      stg_nt<T>(rbW, O(6 + F_HX) + bo, c.hx);
      stg_nt<T>(rbW, O(6 + F_HY) + bo, c.hy);
      stg_nt<T>(rbW, O(6 + F_HZ) + bo, c.hz);
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool lkin = (lf & L3_KIN) != 0u, lkin1 = (lf & L3_KIN1) != 0u;
      int di = 0, ci = 0, dih = 0, cih = 0;
      const int si = slab_e(i, ib, ie, N, di, ci);
      const int sih = slab_h(i, ib, ie, N, dih, cih);
      if (si >= 0) {
        const unsigned ix = ((unsigned)di * pl + jk) * es + (si == 1 ? xhi : 0u);
        if (iin && jin1 && lkin) stg<T>(rbW, O(26) + ix, c.qyx);
        if (iin && jin && lkin1) stg<T>(rbW, O(27) + ix, c.qzx);
      }
      if (sih >= 0) {
        const unsigned ixh = ((unsigned)dih * pl + jk) * es + (sih == 1 ? xhi : 0u);
        if (iin1 && jin && lkin1) stg<T>(rbW, O(28) + ixh, c.ryx);
        if (iin1 && jin1 && lkin) stg<T>(rbW, O(29) + ixh, c.rzx);
      }
      if (sj >= 0) {
        const unsigned iy = (((unsigned)i * N + dj) * nz + kc) * es;
        if (iin1 && jin && lkin) stg<T>(rbW, O(30) + iy, c.qxy);
        if (iin && jin && lkin1) stg<T>(rbW, O(31) + iy, c.qzy);
      }
      if (sjh >= 0) {
        const unsigned iyh = (((unsigned)i * N + djh) * nz + kc) * es;
        if (iin && jin1 && lkin1) stg<T>(rbW, O(32) + iyh, c.rxy);
        if (iin1 && jin1 && lkin) stg<T>(rbW, O(33) + iyh, c.rzy);
      }
      if (zs) {
        const unsigned zrow = ((unsigned)i * ny + jc) * N * es;
        if (lf & L3_SK) {
          if (iin1 && jin) stg<T>(rbW, O(34) + zrow + izk, c.qxz);
          if (iin && jin1) stg<T>(rbW, O(35) + zrow + izk, c.qyz);
        }
        if (lf & L3_SKH) {
          if (iin && jin1) stg<T>(rbW, O(36) + zrow + izkh, c.rxz);
          if (iin1 && jin) stg<T>(rbW, O(37) + zrow + izkh, c.ryz);
        }
      }
    };

    // ---- the march.  At position p step 0 updates plane p - 2 and step 1
    // plane p - 3, while plane p loads: each plane's cells load two positions
    // before step 0 updates them (its Hy, Hz one position before, for step
    // 0's update of the plane below).  Four cell buffers take the roles in
    // turn (the loop is unrolled by four, so no buffer is copied, and no load
    // in flight waits).
    T pey[TB], pez[TB];
#pragma unroll
    for (int st = 0; st < TB; ++st) pey[st] = pez[st] = T(0);
    auto position = [&](int p, Cell3<T>& ld, const Cell3<T>& n1, Cell3<T>& c0, Cell3<T>& c1) {
      const unsigned lf = opaque_vgpr(lflags);
      ot = opaque_vgpr(otab);
      // rows outside the grid (ghost rows of the edge tiles) load nothing:
      // no updated cell reads them (E rows end below ny - 1, H rows start
      // above row 0), and their stale values only reach such cells
      if (p <= xl && jgrid) {
        load_f(p, ld);
        load_q(p, lf, ld);
      }
      // old Hx, Hz of both steps' planes for the wave below
      xh[0][0][w][lane] = c0.hx;
      xh[0][1][w][lane] = c0.hz;
      xh[1][0][w][lane] = c1.hx;
      xh[1][1][w][lane] = c1.hz;
      lds_barrier();
      // uniform over the block
      if (p - 2 >= xf && p - 2 <= xl) {
        const T hxu = w + 1 < W ? xh[0][0][w + 1][lane] : c0.hx;
        const T hzu = w + 1 < W ? xh[0][1][w + 1][lane] : c0.hz;
        step(c0, p - 2, 0, lf, n1.hy, n1.hz, hxu, hzu, pey[0], pez[0]);
      }
      if (nsteps > 1 && p - 3 >= xf && p - 3 <= xl) {
        const T hxu = w + 1 < W ? xh[1][0][w + 1][lane] : c1.hx;
        const T hzu = w + 1 < W ? xh[1][1][w + 1][lane] : c1.hz;
        step(c1, p - 3, 1, lf, c0.hy, c0.hz, hxu, hzu, pey[1], pez[1]);
      }
      // plane p - 3 is final
      if (p - 3 >= x0 && p - 3 < xo1 && (lf & L3_OWN)) store(c1, p - 3, lf);
    };
    Cell3<T> b0{}, b1{}, b2{}, b3{};
    const int pend = xo1 + TB;
    for (int p = xf; p <= pend; p += 4) {
      position(p, b0, b3, b2, b1);
      if (p + 1 <= pend) position(p + 1, b1, b0, b3, b2);
      if (p + 2 <= pend) position(p + 2, b2, b1, b0, b3);
      if (p + 3 <= pend) position(p + 3, b3, b2, b1, b0);
    }
    lds_barrier();   // the next tile reuses the exchange planes
  }
}

// ------------------------------------------------------------------- DFT ---
// Work item (slot, q, cell block, first frequency | frequencies << 16, at most
// kFB).  Folds the kChunk samples
// of the chunk in flight into acc[q][f][cell] (re += s cos, im -= s sin, as
// simul_fft_* of fdtd_funcs.py:128-144 and FDTD_3D :270-292) in FP64.  The
// sample is formed from the raw planes in the run precision with the
// reference's averaging expression.

// The sample of component q at window cell c is K values of one step's
// captured planes at fixed offsets o[] (from that step's plane 0), summed in
// the order ((o0 + o1) + o2) + o3 and, for K > 1, scaled by 1/K; returns K.
template <typename T>
__device__ __forceinline__ int sample_taps(const Slot<T>& s, int q, int c, size_t (&o)[4]) {
  const size_t P = s.pcells;
  if (s.scheme == SCHEME_3D) {
    const int port = q >> 2, comp = q & 3;
    const int wyn = s.wy1 - s.wy0;
    const int ii = c / wyn, jj = c - ii * wyn;
    const size_t b = (size_t)port * 6 * P + (size_t)ii * s.prow + jj;
    switch (comp) {
      case 0: o[0] = b; o[1] = b + s.prow; return 2;                                 // Ex, over i
      case 1: o[0] = P + b; o[1] = P + b + 1; return 2;                              // Ey, over j
      case 2: o[0] = 2 * P + b; o[1] = 2 * P + b + 1; o[2] = 3 * P + b; o[3] = 3 * P + b + 1; return 4;
      default:
        o[0] = 4 * P + b; o[1] = 4 * P + b + s.prow; o[2] = 5 * P + b; o[3] = 5 * P + b + s.prow;
        return 4;
    }
  }
  const int port = q >> 1, comp = q & 1;
  const size_t b = (size_t)port * 3 * P + c;
  if (s.scheme == SCHEME_TMZ) {
    if (comp == 0) { o[0] = b; return 1; }                   // Ez[ic, j]
    o[0] = P + b; o[1] = 2 * P + b; return 2;                // 0.5 (Hy[ic,j] + Hy[ic+1,j])
  }
  if (comp == 0) { o[0] = b; o[1] = b + 1; return 2; }      // 0.5 (Ey[ic,j] + Ey[ic,j+1])
  o[0] = P + b; o[1] = 2 * P + b; o[2] = P + b + 1; o[3] = 2 * P + b + 1;
  return 4;
}

// The chunk's samples smp[t], t < tv, of one cell (zero from tv on).  Every
// step's values load without a branch (steps past tv read step tv - 1 and are
// discarded), so all of them are in flight together.
template <typename T, int K>
__device__ __forceinline__ void gather_samples(const T* __restrict__ ring, size_t stride, const size_t (&o)[4],
                                               int tv, T (&smp)[kChunk]) {
#pragma unroll
  for (int t = 0; t < kChunk; ++t) {
    const T* r = ring + (size_t)min(t, tv - 1) * stride;
    T v = r[o[0]];
    if constexpr (K > 1) v = v + r[o[1]];
    if constexpr (K > 2) v = (v + r[o[2]]) + r[o[3]];
    if constexpr (K == 2) v = T(0.5) * v;
    if constexpr (K == 4) v = T(0.25) * v;
    smp[t] = t < tv ? v : T(0);
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_dft(const Slot<T>* __restrict__ slots,
                                                  const int4* __restrict__ work,
                                                  const int* __restrict__ counts) {
  __shared__ double cs[kFB][kChunk];
  __shared__ double sn[kFB][kChunk];
  const int nwork = counts[1];
  for (int ww = xcd_block(); ww < nwork; ww += gridDim.x) {
    const int4 wk = work[ww];
    const Slot<T>& s = slots[wk.x];
    if (!s.active) continue;
    const int n0 = s.base;
    const int tv = min(kChunk, s.nt - n0);
    if (tv <= 0) continue;
    const int q = wk.y;
    const bool hq = (s.scheme == SCHEME_3D) ? ((q & 3) >= 2) : ((q & 1) == 1);
    const double phi = hq ? 0.5 : 0.0;
    const int f0 = wk.w & 0xffff, fbn = wk.w >> 16;
    // the samples load first (cells past the window read its last cell and
    // fold nothing), in flight while the phase table is formed
    const int c = wk.z * kThreads + threadIdx.x;
    size_t o[4];
    const int k = sample_taps(s, q, min(c, s.wcells - 1), o);
    const size_t stride = (size_t)s.nplanes * s.pcells;
    T smp[kChunk];
    if (k == 1)
      gather_samples<T, 1>(s.ring, stride, o, tv, smp);
    else if (k == 2)
      gather_samples<T, 2>(s.ring, stride, o, tv, smp);
    else
      gather_samples<T, 4>(s.ring, stride, o, tv, smp);
    __syncthreads();
    for (int e = threadIdx.x; e < fbn * kChunk; e += kThreads) {
      const int fl = e / kChunk, t = e - fl * kChunk;
      const int f = f0 + fl;
      double cv = 0.0, sv = 0.0;
      if (f < s.nf && t < tv) {
        const double ph = (s.w[f] * ((double)(n0 + t) + phi)) * s.dt;
        sincos(ph, &sv, &cv);
      }
      cs[fl][t] = cv;
      sn[fl][t] = sv;
    }
    __syncthreads();
    if (c >= s.wcells) continue;
    const int fend = min(fbn, s.nf - f0);
    double* a = s.acc + (((size_t)q * s.nf + f0) * s.wcells + c) * 2;
    const size_t fs = (size_t)s.wcells * 2;   // between frequencies
    // four frequencies at a time: their accumulators load before the sums
    int fl = 0;
    for (; fl + 4 <= fend; fl += 4, a += 4 * fs) {
      double ar[4], ai[4], re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ar[u] = a[u * fs];
        ai[u] = a[u * fs + 1];
      }
#pragma unroll
      for (int t = 0; t < kChunk; ++t) {
        const double v = (double)smp[t];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          re[u] += v * cs[fl + u][t];
          im[u] += v * sn[fl + u][t];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u * fs] = ar[u] + re[u];
        a[u * fs + 1] = ai[u] - im[u];
      }
    }
    for (; fl < fend; ++fl, a += fs) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int t = 0; t < kChunk; ++t) {
        const double v = (double)smp[t];
        re += v * cs[fl][t];
        im += v * sn[fl][t];
      }
      a[0] += re;
      a[1] -= im;
    }
  }
}

// ------------------------------------------------------------ port power ---
// Time-averaged normal Poynting power of one port plane per frequency, summed
// over window cells [c0, c1) (the port-window sums of extract_sparams.py:88-104,
// FDTD_3D :114-131).  a / sub: the FP64 accumulators of the run and of the
// run whose phasors are subtracted first (sub may be null; the incident-field
// correction of S_ii).  q0 / qs: first accumulator row (port * components) of
// each.  One block per frequency; wave reductions with cross-lane shuffles,
// then LDS across the four waves.  out[f] = (re, im).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ __launch_bounds__(kThreads) void k_port_power(const double* __restrict__ a,
                                                         const double* __restrict__ sub, int scheme, int q0,
                                                         int qs, int nf, int wcells, int c0, int c1,
                                                         double* __restrict__ out) {
  __shared__ double part[2][kThreads / 64];
  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  auto ph = [&](const double* base, int q, int c, double& re, double& im) {
    const double* p = base + (((size_t)q * nf + f) * wcells + c) * 2;
    re = p[0];
    im = p[1];
  };
  double sr = 0.0, si = 0.0;
  for (int c = c0 + tid; c < c1; c += kThreads) {
    if (scheme == SCHEME_3D) {
      // components ex, ey, hx, hy: S = 0.5 (Ex Hy* - Ey Hx*)
      double v[4][2];
      for (int k = 0; k < 4; ++k) {
        ph(a, q0 + k, c, v[k][0], v[k][1]);
        if (sub) {
          double r, i;
          ph(sub, qs + k, c, r, i);
          v[k][0] -= r;
          v[k][1] -= i;
        }
      }
      sr += 0.5 * ((v[0][0] * v[3][0] + v[0][1] * v[3][1]) - (v[1][0] * v[2][0] + v[1][1] * v[2][1]));
      si += 0.5 * ((v[0][1] * v[3][0] - v[0][0] * v[3][1]) - (v[1][1] * v[2][0] - v[1][0] * v[2][1]));
    } else {
      double er, ei, hr, hi;
      ph(a, q0, c, er, ei);
      ph(a, q0 + 1, c, hr, hi);
      if (sub) {
        double r, i;
        ph(sub, qs, c, r, i);
        er -= r;
        ei -= i;
        ph(sub, qs + 1, c, r, i);
        hr -= r;
        hi -= i;
      }
      // TMz: -0.5 Ez Hy*; TEz: 0.5 Ey Hz*
      const double sgn = scheme == SCHEME_TMZ ? -0.5 : 0.5;
      sr += sgn * (er * hr + ei * hi);
      si += sgn * (ei * hr - er * hi);
    }
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  const int w = tid >> 6;
  if ((tid & 63) == 0) {
    part[0][w] = sr;
    part[1][w] = si;
  }
  __syncthreads();
  if (tid == 0) {
    double r = 0.0, i = 0.0;
    for (int k = 0; k < kThreads / 64; ++k) {
      r += part[0][k];
      i += part[1][k];
    }
    out[2 * f] = r;
    out[2 * f + 1] = i;
  }
}

// End-of-chunk: every resident run moves to its next chunk.  (A separate
// launch: counting the DFT launch's blocks with one atomic each, so that its
// last block could do this, measured slower on every batch but the smallest.)
template <typename T>
__global__ void k_advance(Slot<T>* __restrict__ slots, int nslots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots && slots[i].active) slots[i].base += kChunk;
}

// ------------------------------------------------ 2D TMz, lossy study form ---
// E half-step of the TM loss and wave-impedance study drivers
// (FDTD/GPU/FDTD_2D_TM/scattering_loss/fdtd_funcs.py:77-83, 94-114, source
// :147-152), on 0 < i < nx-1, 0 < j < ny-1:
//   Ez = MOD_E Ez + CURL_H ((Hy[i]-Hy[i-1]) / DEN_EX[i] + (Hx[j-1]-Hx[j]) / DEN_EY[j]),
// then the source, then on the CPML cells
//   psi = BE[q] psi + CE[q]/dx (same difference),  Ez += CURL_H dx psi,
// q = distance from the grid edge, 1 <= q < R (R = cpml_range = s.N), with
// its own psi array on each side.  CURL_H / MOD_E are the material values of
// components 2 / 0; DEN_* arrive as reciprocals in ike / ikh; ie = nx-1, je = ny-1.
template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tmzl_e(const Slot<T>* __restrict__ slots,
                                                     const int4* __restrict__ tiles,
                                                     const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, R = s.N, ie = s.ie, je = s.je;
    const int j = tl.z + threadIdx.x;
    if (j < 1 || j >= je) continue;
    const int i0 = max(tl.y + (int)threadIdx.y * XC, 1);
    const int i1 = min(tl.y + (int)threadIdx.y * XC + XC, ie);
    T* __restrict__ Ez = s.f[F_EZ];
    const T* __restrict__ Hx = s.f[F_HX];
    const T* __restrict__ Hy = s.f[F_HY];
    const T rdy = s.ike[1][j];
    const int qy = j < R ? j : (je - j < R ? je - j : -1);
    T* __restrict__ py = qy >= 0 ? s.psi[psi_id(F_EZ, 1, j < R ? 0 : 1)] : nullptr;
    const bool jsrc = s.src_field == F_EZ && j >= s.sy0 && j < s.sy1;
    const T jv = jsrc ? s.jsrc[n] : T(0);
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * ny + j;
      const uint8_t m = s.mat[idx];
      const T ch = cbsel(s, m, 2);
      const T me = cbsel(s, m, 0);
      const T dHy = Hy[idx] - Hy[idx - ny];
      const T dHx = Hx[idx - 1] - Hx[idx];
      T v = me * Ez[idx] + ch * (dHy * s.ike[0][i] + dHx * rdy);
      if (jsrc && i >= s.sx0 && i < s.sx1) v -= jv;
      const T chdx = ch * s.dxT;
      const int qx = i < R ? i : (ie - i < R ? ie - i : -1);
      if (qx >= 0)
        v += chdx * psi_upd(s.psi[psi_id(F_EZ, 0, i < R ? 0 : 1)] + (size_t)qx * ny + j, s.be[qx], s.ceo[qx], dHy);
      if (qy >= 0) v += chdx * psi_upd(py + (size_t)i * R + qy, s.be[qy], s.ceo[qy], dHx);
      Ez[idx] = v;
    }
  }
}

// H half-step of the same form (fdtd_funcs.py:85-92, 116-136), on i < nx-1,
// j < ny-1:
//   Hx += CURL_E (Ez[j]-Ez[j+1]) / DEN_HY[j],  Hy += CURL_E (Ez[i+1]-Ez[i]) / DEN_HX[i],
// then psi = BH[q] psi + CH[q]/dx (same difference), H += (dt/mu) psi on the
// CPML cells: low q = i, high q = nx-2-i (likewise in y), 1 <= q < R.
template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_tmzl_h(const Slot<T>* __restrict__ slots,
                                                     const int4* __restrict__ tiles,
                                                     const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, R = s.N, ie = s.ie, je = s.je;
    const int j = tl.z + threadIdx.x;
    if (j >= je) continue;
    const int i0 = tl.y + (int)threadIdx.y * XC;
    const int i1 = min(i0 + XC, ie);
    const T* __restrict__ Ez = s.f[F_EZ];
    T* __restrict__ Hx = s.f[F_HX];
    T* __restrict__ Hy = s.f[F_HY];
    const T ce = s.db, dtmu = s.dtmu;
    const T rdy = s.ikh[1][j];
    const int qy = (j >= 1 && j < R) ? j : (je - 1 - j >= 1 && je - 1 - j < R ? je - 1 - j : -1);
    T* __restrict__ py = qy >= 0 ? s.psi[psi_id(F_HX, 1, j < R ? 0 : 1)] : nullptr;
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * ny + j;
      const T ez = Ez[idx];
      const T dEy = ez - Ez[idx + 1];
      const T dEx = Ez[idx + ny] - ez;
      T hx = Hx[idx] + ce * dEy * rdy;
      T hy = Hy[idx] + ce * dEx * s.ikh[0][i];
      const int qx = (i >= 1 && i < R) ? i : (ie - 1 - i >= 1 && ie - 1 - i < R ? ie - 1 - i : -1);
      if (qx >= 0)
        hy += dtmu * psi_upd(s.psi[psi_id(F_HY, 0, i < R ? 0 : 1)] + (size_t)qx * ny + j, s.bh[qx], s.cho[qx], dEx);
      if (qy >= 0) hx += dtmu * psi_upd(py + (size_t)i * R + qy, s.bh[qy], s.cho[qy], dEy);
      Hx[idx] = hx;
      Hy[idx] = hy;
    }
  }
}

// Field monitors taken after one half-step: one work item = (monitor, first
// of kMonCells consecutive box cells, -, -); each lane takes kMonCells / 64
// cells kThreads apart, loads all of them (and the running sums) first, then
// folds and stores.
// Reads the current field set (the two-kernel forms keep the state in set A).
// DFT: the stencil sum is formed in T point by point, scaled and weighted in
// double, and the running sum is rounded back to T each step, as the
// reference's T accumulators are.
constexpr int kMonPer = 4;
constexpr int kMonCells = kThreads * kMonPer;

template <typename T>
__global__ __launch_bounds__(kThreads) void k_monitor(const Slot<T>* __restrict__ slots,
                                                      const Monitor<T>* __restrict__ mons,
                                                      const int4* __restrict__ work, int nwork, int tstep) {
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int4 it = work[w];
    const Monitor<T>& m = mons[it.x];
    const Slot<T>& s = slots[m.slot];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int r = n + m.shift;
    if (r < 0 || r >= m.nrec) continue;
    const int cells = m.ex * m.ey * m.ez;
    const T* __restrict__ f = s.f[m.field];
    const bool is3d = s.scheme == SCHEME_3D;
    // running sums first: their loads overlap the stencil's
    T re[kMonPer], im[kMonPer];
    if (m.mode == 0) {
#pragma unroll
      for (int k = 0; k < kMonPer; ++k) {
        const int c = it.y + (int)threadIdx.x + k * kThreads;
        re[k] = c < cells ? m.re[c] : T(0);
        im[k] = c < cells ? m.im[c] : T(0);
      }
    }
    T sum[kMonPer];
#pragma unroll
    for (int k = 0; k < kMonPer; ++k) {
      const int c = it.y + (int)threadIdx.x + k * kThreads;
      sum[k] = T(0);
      if (c >= cells) continue;
      const int cz = m.ez == 1 ? 0 : c % m.ez, cxy = m.ez == 1 ? c : c / m.ez;   // 2D boxes: no z division
      const int x = m.x0 + cxy / m.ey, y = m.y0 + cxy % m.ey, z = m.z0 + cz;
      for (int q = 0; q < m.npts; ++q) {
        const int xx = x + m.ox[q], yy = y + m.oy[q], zz = z + m.oz[q];
        T v = T(0);
        if (xx >= 0 && xx < m.gx && yy >= 0 && yy < m.gy && zz >= 0 && zz < m.gz)
          v = f[is3d ? ((size_t)xx * s.ny + yy) * s.nz + zz : (size_t)xx * s.ny + yy];
        sum[k] = q == 0 ? v : sum[k] + v;
      }
    }
    if (m.mode == 0) {
      const double cr = m.cs[2 * r], ci = m.cs[2 * r + 1];
#pragma unroll
      for (int k = 0; k < kMonPer; ++k) {
        const int c = it.y + (int)threadIdx.x + k * kThreads;
        if (c >= cells) continue;
        const double v = m.scale * (double)sum[k];
        m.re[c] = (T)((double)re[k] + v * cr);
        m.im[c] = (T)((double)im[k] + v * ci);
      }
    } else {
#pragma unroll
      for (int k = 0; k < kMonPer; ++k) {
        const int c = it.y + (int)threadIdx.x + k * kThreads;
        if (c < cells) m.re[(size_t)r * cells + c] = (T)(m.scale * (double)sum[k]);
      }
    }
  }
}

// Plain streaming copy (16 B per lane): calibrates the HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ c, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) c[i] = a[i];
}

}  // namespace oidt
