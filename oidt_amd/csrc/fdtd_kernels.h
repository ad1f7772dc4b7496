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
#include "fdtd_slot.h"

namespace oidt {

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

template <typename T>
__device__ __forceinline__ T psi_upd(T* p, T b, T c, T diff) {
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
        const T cb = s.cbt[s.mat[idx] * 3 + 2];
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
        const T cb = s.cbt[m * 3 + 0];
        const T dHz = Hz[idx + 1] - hz0;
        T v = Ex[idx] + cb * dHz * iky;
        if (sy >= 0) v += cb * psi_upd(s.psi[psi_id(F_EX, 1, sy)] + (size_t)i * N + dy, s.be[cy], s.ceo[cy], dHz);
        Ex[idx] = v;
      }
      // update_ey: i in [ib,ie), j in (jb,je); psi ey_x (-)
      T v = Ey[idx];
      if (by) {
        const T cb = s.cbt[m * 3 + 1];
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
// Block: threadIdx.x -> k (64 lanes, unit stride), threadIdx.y -> j (4 rows);
// each thread marches XC cells along x.

template <typename T, int XC>
__global__ __launch_bounds__(kThreads) void k_3d_e(const Slot<T>* __restrict__ slots,
                                                   const int4* __restrict__ tiles,
                                                   const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, nz = s.nz, N = s.N;
    const int k = tl.w + threadIdx.x, j = tl.z + threadIdx.y;
    if (j >= ny || k >= nz) continue;
    const int i0 = tl.y, i1 = min(tl.y + XC, s.nx);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    T* __restrict__ Ex = s.f[F_EX];
    T* __restrict__ Ey = s.f[F_EY];
    T* __restrict__ Ez = s.f[F_EZ];
    const T* __restrict__ Hx = s.f[F_HX];
    const T* __restrict__ Hy = s.f[F_HY];
    const T* __restrict__ Hz = s.f[F_HZ];
    const size_t pl = (size_t)ny * nz;
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dj, cj, dk, ck;
    const int sj = jin ? slab_e(j, jb, je, N, dj, cj) : -1;
    const int sk = kin ? slab_e(k, kb, ke, N, dk, ck) : -1;
    const T iky = jin ? s.ike[1][j] : T(0);
    const T ikz = kin ? s.ike[2][k] : T(0);
    const bool jksrc = s.src_field == F_EY && j >= s.sy0 && j < s.sy1 && k >= s.sz0 && k < s.sz1;
    const T jv = jksrc ? s.jsrc[n] : T(0);
    const bool jcap = j >= s.wy0 && j <= s.wy1;
    const int cap = !jcap ? -1 : (k == s.port[0] ? 0 : (k == s.port[1] ? 1 : -1));
    const int cap2 = (jcap && s.port[0] == s.port[1] && k == s.port[0]) ? 1 : -1;
    const size_t jk = (size_t)j * nz + k;
    T hz_next = (i0 < i1) ? Hz[(size_t)i0 * pl + jk] : T(0);
    T hy_next = (i0 < i1) ? Hy[(size_t)i0 * pl + jk] : T(0);
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * pl + jk;
      const T hz0 = hz_next, hy0 = hy_next;
      const bool more = i + 1 < s.nx;
      hz_next = more ? Hz[idx + pl] : T(0);
      hy_next = more ? Hy[idx + pl] : T(0);
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool bx = iin1 && jin && kin;
      const bool by = iin && jin1 && kin;
      const bool bz = iin && jin && kin1;
      const bool src = jksrc && i >= s.sx0 && i < s.sx1;
      const bool icap = cap >= 0 && i >= s.wx0 && i <= s.wx1;
      if (!(bx || by || bz || src || icap)) continue;
      const uint8_t m = (bx || by || bz) ? s.mat[idx] : 0;
      int di, ci;
      const int si = slab_e(i, ib, ie, N, di, ci);
      T vx = T(0), vy = T(0);
      const bool needx = bx || icap;
      // update_ex: i (ib,ie), j [jb,je), k [kb,ke); psi ex_y (+), ex_z (-)
      if (needx) vx = Ex[idx];
      if (bx) {
        const T cb = s.cbt[m * 3 + 0];
        const T dHz = Hz[idx + nz] - hz0;
        const T dHy = Hy[idx + 1] - hy0;
        vx = vx + cb * (dHz * iky - dHy * ikz);
        if (sj >= 0) vx += cb * psi_upd(s.psi[psi_id(F_EX, 1, sj)] + ((size_t)i * N + dj) * nz + k, s.be[cj], s.ceo[cj], dHz);
        if (sk >= 0) vx -= cb * psi_upd(s.psi[psi_id(F_EX, 2, sk)] + ((size_t)i * ny + j) * N + dk, s.be[ck], s.ceo[ck], dHy);
        Ex[idx] = vx;
      }
      // update_ey: i [ib,ie), j (jb,je), k [kb,ke); source (TE); psi ey_x (-), ey_z (+)
      const bool needy = by || src || icap;
      if (needy) vy = Ey[idx];
      if (by) {
        const T cb = s.cbt[m * 3 + 1];
        const T dHx = Hx[idx + 1] - Hx[idx];
        const T dHz = hz_next - hz0;
        vy = vy + cb * (dHx * ikz - dHz * s.ike[0][i]);
        if (src) vy -= jv;
        if (si >= 0) vy -= cb * psi_upd(s.psi[psi_id(F_EY, 0, si)] + ((size_t)di * ny + j) * nz + k, s.be[ci], s.ceo[ci], dHz);
        if (sk >= 0) vy += cb * psi_upd(s.psi[psi_id(F_EY, 2, sk)] + ((size_t)i * ny + j) * N + dk, s.be[ck], s.ceo[ck], dHx);
        Ey[idx] = vy;
      } else if (src) {
        vy -= jv;
        Ey[idx] = vy;
      }
      // update_ez: i [ib,ie), j [jb,je), k (kb,ke); psi ez_x (+), ez_y (-)
      if (bz) {
        const T cb = s.cbt[m * 3 + 2];
        const T dHy = hy_next - hy0;
        const T dHx = Hx[idx + nz] - Hx[idx];
        T v = Ez[idx] + cb * (dHy * s.ike[0][i] - dHx * iky);
        if (si >= 0) v += cb * psi_upd(s.psi[psi_id(F_EZ, 0, si)] + ((size_t)di * ny + j) * nz + k, s.be[ci], s.ceo[ci], dHy);
        if (sj >= 0) v -= cb * psi_upd(s.psi[psi_id(F_EZ, 1, sj)] + ((size_t)i * N + dj) * nz + k, s.be[cj], s.ceo[cj], dHx);
        Ez[idx] = v;
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
__global__ __launch_bounds__(kThreads) void k_3d_h(const Slot<T>* __restrict__ slots,
                                                   const int4* __restrict__ tiles,
                                                   const int* __restrict__ counts, int tstep) {
  const int ntiles = counts[0];
  for (int tt = xcd_block(); tt < ntiles; tt += gridDim.x) {
    const int4 tl = tiles[tt];
    const Slot<T>& s = slots[tl.x];
    const int n = s.base + tstep;
    if (!s.active || n >= s.nt) continue;
    const int ny = s.ny, nz = s.nz, N = s.N;
    const int k = tl.w + threadIdx.x, j = tl.z + threadIdx.y;
    if (j >= ny || k >= nz) continue;
    const int i0 = tl.y, i1 = min(tl.y + XC, s.nx);
    const int ib = s.ib, ie = s.ie, jb = s.jb, je = s.je, kb = s.kb, ke = s.ke;
    const T* __restrict__ Ex = s.f[F_EX];
    const T* __restrict__ Ey = s.f[F_EY];
    const T* __restrict__ Ez = s.f[F_EZ];
    T* __restrict__ Hx = s.f[F_HX];
    T* __restrict__ Hy = s.f[F_HY];
    T* __restrict__ Hz = s.f[F_HZ];
    const T db = s.db;
    const size_t pl = (size_t)ny * nz;
    const bool jin = j >= jb && j < je, jin1 = j > jb && j < je;
    const bool kin = k >= kb && k < ke, kin1 = k > kb && k < ke;
    int dj, cj, dk, ck;
    const int sj = jin1 ? slab_h(j, jb, je, N, dj, cj) : -1;
    const int sk = kin1 ? slab_h(k, kb, ke, N, dk, ck) : -1;
    const T iky = jin1 ? s.ikh[1][j] : T(0);
    const T ikz = kin1 ? s.ikh[2][k] : T(0);
    // 3D TM source on Hy (fdtd_main.py:369-370), applied before the H curl
    const bool jksrc = s.src_field == F_HY && j >= s.sy0 && j < s.sy1 && k >= s.sz0 && k < s.sz1;
    const T jv = jksrc ? s.jsrc[n] : T(0);
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
    const size_t jk = (size_t)j * nz + k;
    T ez_prev = (i0 > 0 && i0 < i1) ? Ez[(size_t)(i0 - 1) * pl + jk] : T(0);
    T ey_prev = (i0 > 0 && i0 < i1) ? Ey[(size_t)(i0 - 1) * pl + jk] : T(0);
    for (int i = i0; i < i1; ++i) {
      const size_t idx = (size_t)i * pl + jk;
      const T ez = Ez[idx], ey = Ey[idx];
      const T ezm = ez_prev, eym = ey_prev;
      ez_prev = ez;
      ey_prev = ey;
      const bool iin = i >= ib && i < ie, iin1 = i > ib && i < ie;
      const bool bx = iin && jin1 && kin1;
      const bool by = iin1 && jin && kin1;
      const bool bz = iin1 && jin1 && kin;
      const bool src = jksrc && i >= s.sx0 && i < s.sx1;
      const bool icap = anycap && i >= s.wx0 && i <= s.wx1;
      if (!(bx || by || bz || src || icap)) continue;
      int di, ci;
      const int si = slab_h(i, ib, ie, N, di, ci);
      // update_hx: i [ib,ie), j (jb,je), k (kb,ke); psi hx_y (-), hx_z (+)
      T vx = (bx || icap) ? Hx[idx] : T(0);
      if (bx) {
        const T dEy = ey - Ey[idx - 1];
        const T dEz = ez - Ez[idx - nz];
        vx = vx + db * (dEy * ikz - dEz * iky);
        if (sj >= 0) vx -= db * psi_upd(s.psi[psi_id(F_HX, 1, sj)] + ((size_t)i * N + dj) * nz + k, s.bh[cj], s.cho[cj], dEz);
        if (sk >= 0) vx += db * psi_upd(s.psi[psi_id(F_HX, 2, sk)] + ((size_t)i * ny + j) * N + dk, s.bh[ck], s.cho[ck], dEy);
        Hx[idx] = vx;
      }
      // update_hy: i (ib,ie), j [jb,je), k (kb,ke); psi hy_x (+), hy_z (-)
      T vy = (by || src || icap) ? Hy[idx] : T(0);
      if (src) vy -= jv;
      if (by) {
        const T dEz = ez - ezm;
        const T dEx = Ex[idx] - Ex[idx - 1];
        vy = vy + db * (dEz * s.ikh[0][i] - dEx * ikz);
        if (si >= 0) vy += db * psi_upd(s.psi[psi_id(F_HY, 0, si)] + ((size_t)di * ny + j) * nz + k, s.bh[ci], s.cho[ci], dEz);
        if (sk >= 0) vy -= db * psi_upd(s.psi[psi_id(F_HY, 2, sk)] + ((size_t)i * ny + j) * N + dk, s.bh[ck], s.cho[ck], dEx);
      }
      if (by || src) Hy[idx] = vy;
      // update_hz: i (ib,ie), j (jb,je), k [kb,ke); psi hz_x (-), hz_y (+)
      if (bz) {
        const T dEx = Ex[idx] - Ex[idx - nz];
        const T dEy = ey - eym;
        T v = Hz[idx] + db * (dEx * iky - dEy * s.ikh[0][i]);
        if (si >= 0) v -= db * psi_upd(s.psi[psi_id(F_HZ, 0, si)] + ((size_t)di * ny + j) * nz + k, s.bh[ci], s.cho[ci], dEy);
        if (sj >= 0) v += db * psi_upd(s.psi[psi_id(F_HZ, 1, sj)] + ((size_t)i * N + dj) * nz + k, s.bh[cj], s.cho[cj], dEx);
        Hz[idx] = v;
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

// ------------------------------------------------------------------- DFT ---
// Work item (slot, q, cell block, frequency block).  Folds the kChunk samples
// of the chunk in flight into acc[q][f][cell] (re += s cos, im -= s sin, as
// simul_fft_* of fdtd_funcs.py:128-144 and FDTD_3D :270-292) in FP64.  The
// sample is formed from the raw planes in the run precision with the
// reference's averaging expression.

template <typename T>
__device__ __forceinline__ T sample_at(const Slot<T>& s, const T* __restrict__ r, int q, int c) {
  // r points at plane 0 of this step
  const size_t P = s.pcells;
  if (s.scheme == SCHEME_3D) {
    const int port = q >> 2, comp = q & 3;
    const int wyn = s.wy1 - s.wy0;
    const int ii = c / wyn, jj = c - ii * wyn;
    const int rc = ii * s.prow + jj;
    const T* b = r + (size_t)port * 6 * P;
    switch (comp) {
      case 0: return T(0.5) * (b[rc] + b[rc + s.prow]);                              // Ex, over i
      case 1: return T(0.5) * (b[P + rc] + b[P + rc + 1]);                           // Ey, over j
      case 2: return T(0.25) * (b[2 * P + rc] + b[2 * P + rc + 1] + b[3 * P + rc] + b[3 * P + rc + 1]);
      default: return T(0.25) * (b[4 * P + rc] + b[4 * P + rc + s.prow] + b[5 * P + rc] + b[5 * P + rc + s.prow]);
    }
  }
  const int port = q >> 1, comp = q & 1;
  const T* b = r + (size_t)port * 3 * P;
  if (s.scheme == SCHEME_TMZ) {
    if (comp == 0) return b[c];                         // Ez[ic, j]
    return T(0.5) * (b[P + c] + b[2 * P + c]);          // 0.5 (Hy[ic,j] + Hy[ic+1,j])
  }
  if (comp == 0) return T(0.5) * (b[c] + b[c + 1]);     // 0.5 (Ey[ic,j] + Ey[ic,j+1])
  return T(0.25) * (b[P + c] + b[2 * P + c] + b[P + c + 1] + b[2 * P + c + 1]);
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
    const int f0 = wk.w * kFB;
    __syncthreads();
    for (int e = threadIdx.x; e < kFB * kChunk; e += kThreads) {
      const int fl = e / kChunk, t = e - fl * kChunk;
      const int f = f0 + fl;
      double c = 0.0, sv = 0.0;
      if (f < s.nf && t < tv) {
        const double ph = (s.w[f] * ((double)(n0 + t) + phi)) * s.dt;
        sincos(ph, &sv, &c);
      }
      cs[fl][t] = c;
      sn[fl][t] = sv;
    }
    __syncthreads();
    const int c = wk.z * kThreads + threadIdx.x;
    if (c >= s.wcells) continue;
    T smp[kChunk];
#pragma unroll
    for (int t = 0; t < kChunk; ++t)
      smp[t] = (t < tv) ? sample_at(s, s.ring + (size_t)t * s.nplanes * s.pcells, q, c) : T(0);
    const int fend = min(kFB, s.nf - f0);
    for (int fl = 0; fl < fend; ++fl) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int t = 0; t < kChunk; ++t) {
        const double v = (double)smp[t];
        re += v * cs[fl][t];
        im += v * sn[fl][t];
      }
      double* a = s.acc + (((size_t)q * s.nf + f0 + fl) * s.wcells + c) * 2;
      a[0] += re;
      a[1] -= im;
    }
  }
}

// End-of-chunk: every resident run moves to its next chunk.
template <typename T>
__global__ void k_advance(Slot<T>* __restrict__ slots, int nslots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots && slots[i].active) slots[i].base += kChunk;
}

}  // namespace oidt
