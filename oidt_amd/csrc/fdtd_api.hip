// C ABI of liboidt_fdtd.so (declared in include/oidt_fdtd.h).
//
// A batch owns a HIP stream (default priority), a device table of run
// descriptors (Slot<T>), a tile list for the stencil kernels and a work list
// for the DFT kernel.  One chunk = kChunk x (E launch, H launch) + DFT launch +
// advance launch, captured once per batch composition as a hipGraph and
// replayed.  A single run (oidt_solver) is a batch with one slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/oidt_fdtd.h"
#include "fdtd_kernels.h"

using namespace oidt;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(OIDT_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));       \
  } while (0)

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Pinned staging buffers for asynchronous uploads; a buffer is reused once the
// event recorded after its copy has completed.
struct Staging {
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  std::vector<Buf> bufs;
  int get(size_t bytes, void** out) {
    for (size_t i = 0; i < bufs.size(); ++i) {
      Buf& b = bufs[i];
      if (b.pending) {
        if (hipEventQuery(b.ev) == hipSuccess) b.pending = false;
        else continue;
      }
      if (b.cap >= bytes) { *out = b.p; return (int)i; }
    }
    Buf b;
    b.cap = std::max(bytes, (size_t)1 << 20);
    if (hipHostMalloc(&b.p, b.cap, hipHostMallocDefault) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) return -1;
    bufs.push_back(b);
    *out = b.p;
    return (int)bufs.size() - 1;
  }
  int submit(int id, hipStream_t st) {
    Buf& b = bufs[id];
    if (hipEventRecord(b.ev, st) != hipSuccess) return -1;
    b.pending = true;
    return 0;
  }
  void release() {
    for (auto& b : bufs) {
      if (b.ev) { (void)hipEventSynchronize(b.ev); (void)hipEventDestroy(b.ev); }
      if (b.p) (void)hipHostFree(b.p);
    }
    bufs.clear();
  }
};

// A field monitor of one run: its description and device buffers.
struct MonHost {
  oidt_monitor_desc d{};
  void* re = nullptr;
  void* im = nullptr;
  double* cs = nullptr;
  size_t cells = 0;
};

struct SlotHost {
  bool used = false;
  oidt_run_desc desc{};
  void* dmem = nullptr;
  size_t bytes = 0;
  std::vector<unsigned char> slot;  // host image of Slot<T>
  int64_t admit_chunk = 0;          // batch chunk counter when admitted
  // Steps between swaps of the run's two field sets: 0 for the two-kernel
  // forms (one set), 1 for the fused step, the launch depth for the
  // temporally blocked step.  After `steps` completed steps the state is in
  // set B when ceil(steps / tb) is odd.
  int tb = 0;
  int nplanes = 0, nq = 0, wcells = 0, pcells = 0, wyn = 0;
  std::vector<MonHost> mons;
};

bool field_used(int scheme, int f) {
  if (scheme == SCHEME_3D) return true;
  if (scheme == SCHEME_TMZ || scheme == SCHEME_TMZ_LOSSY) return f == F_EZ || f == F_HX || f == F_HY;
  return f == F_EX || f == F_EY || f == F_HZ;
}

// psi (component, axis) pairs present per scheme, from the reference allocations.
std::vector<std::pair<int, int>> psi_pairs(int scheme) {
  if (scheme == SCHEME_TMZ || scheme == SCHEME_TMZ_LOSSY) return {{F_EZ, 0}, {F_EZ, 1}, {F_HX, 1}, {F_HY, 0}};
  if (scheme == SCHEME_TEZ) return {{F_EX, 1}, {F_EY, 0}, {F_HZ, 0}, {F_HZ, 1}};
  return {{F_EX, 1}, {F_EX, 2}, {F_EY, 0}, {F_EY, 2}, {F_EZ, 0}, {F_EZ, 1},
          {F_HX, 1}, {F_HX, 2}, {F_HY, 0}, {F_HY, 2}, {F_HZ, 0}, {F_HZ, 1}};
}

size_t psi_elems(const oidt_run_desc& d, int axis) {
  const size_t nx = d.nx, ny = d.ny, nz = d.nz, N = d.num_cpml;
  if (axis == 0) return N * ny * nz;
  if (axis == 1) return nx * N * nz;
  return nx * ny * N;
}

int validate(const oidt_run_desc* d) {
  if (!d) return fail(OIDT_EINVAL, "null run description");
  if (d->abi_version != OIDT_ABI_VERSION) return fail(OIDT_EINVAL, "abi_version mismatch");
  if (d->scheme < 0 || d->scheme > 3) return fail(OIDT_EINVAL, "unknown scheme");
  if (d->precision != OIDT_F32 && d->precision != OIDT_F64) return fail(OIDT_EINVAL, "unknown precision");
  const bool is3d = d->scheme == SCHEME_3D;
  if (!is3d && d->nz != 1) return fail(OIDT_EINVAL, "2D runs need nz == 1");
  if (d->nf < 0 || d->nf > 65535) return fail(OIDT_EINVAL, "nf must be in [0, 65535]");
  if (d->scheme == SCHEME_TMZ_LOSSY && (d->offset != 1 || d->nf != 0))
    return fail(OIDT_EINVAL, "the lossy TMz form updates from offset 1 and has no port DFT (nf = 0)");
  if (d->scheme == SCHEME_TMZ_LOSSY && d->src_field >= 0 &&
      (d->src_box[0] < 1 || d->src_box[1] > d->nx - 1 || d->src_box[2] < 1 || d->src_box[3] > d->ny - 1))
    return fail(OIDT_EINVAL, "lossy TMz source box must lie inside the update range");
  const int N = d->num_cpml, o = d->offset;
  if (N < 2 || o < 1 || N < o) return fail(OIDT_EINVAL, "need num_cpml >= offset >= 1 and num_cpml >= 2");
  const int dims[3] = {d->nx, d->ny, d->nz};
  for (int a = 0; a < (is3d ? 3 : 2); ++a) {
    // both CPML slabs must fit inside the update range without overlapping
    if (dims[a] - 2 * o < 2 * N + 1) return fail(OIDT_EINVAL, "grid too small for the CPML thickness");
  }
  // Kernels address a field with a 32-bit byte offset into the padded layout
  // (the unit-stride axis rounded up to 128 bytes): each field array must
  // stay below 4 GiB.
  const int es = d->precision == OIDT_F64 ? 8 : 4;
  const int q = 128 / es;
  const double fast = is3d ? (double)((d->nz + q - 1) / q * q) : (double)((d->ny + q - 1) / q * q);
  const double slow = is3d ? (double)d->nx * d->ny : (double)d->nx;
  if (slow * fast * es > 4294967295.0)
    return fail(OIDT_EINVAL, "a field array would exceed 4 GiB (after padding the unit-stride axis to 128 bytes)");
  // The 3D half-step kernels address a high-side psi slab from the low-side
  // array of the same component (both sides and the fused step's second set
  // lie in between): three slabs must stay below 4 GiB.
  for (int a = 0; is3d && a < 3; ++a) {
    double cells = N;
    for (int b = 0; b < 3; ++b)
      if (b != a) cells *= b == 2 ? fast : (double)dims[b];
    if (3.0 * (cells * es + 256.0) > 4294967295.0)
      return fail(OIDT_EINVAL, "a CPML psi slab pair would exceed the 32-bit offset range");
  }
  if (d->nt < 0 || d->nf < 0) return fail(OIDT_EINVAL, "negative nt or nf");
  const int plim = is3d ? d->nz : d->nx;
  for (int p = 0; p < 2; ++p)
    if (d->port[p] < 0 || d->port[p] + 1 >= plim) return fail(OIDT_EINVAL, "port plane outside the grid");
  if (d->src_field < -1 || d->src_field > 5) return fail(OIDT_EINVAL, "bad src_field");
  if (d->src_field >= 0) {
    const int allowed = (d->scheme == SCHEME_TMZ || d->scheme == SCHEME_TMZ_LOSSY) ? F_EZ
                        : (d->scheme == SCHEME_TEZ ? F_HZ : -2);
    if (is3d ? (d->src_field != F_EY && d->src_field != F_HY) : d->src_field != allowed)
      return fail(OIDT_EINVAL, "src_field not driven by this scheme in the reference");
    for (int a = 0; a < 3; ++a)
      if (d->src_box[2 * a] > d->src_box[2 * a + 1]) return fail(OIDT_EINVAL, "inverted source box");
    if (is3d && d->src_field == F_HY) {
      // The fused H half-step applies the Hy source before the H curl.  The
      // reference applies it before the E-CPML, which reads Hy at
      // x in [ib, ib+N] u [ie-N, ie] and z in [kb, kb+N] u [ke-N, ke]; the two
      // orders agree only if the box avoids those cells.
      const int ib = o, ie = d->nx - o, kb = o, ke = d->nz - o;
      auto hits = [](int a0, int a1, int b0, int b1) { return a0 <= b1 && b0 < a1; };
      const int* s = d->src_box;
      if (hits(s[0], s[1], ib, ib + N) || hits(s[0], s[1], ie - N, ie) ||
          hits(s[4], s[5], kb, kb + N) || hits(s[4], s[5], ke - N, ke))
        return fail(OIDT_EINVAL, "3D TM source box overlaps the E-CPML stencil");
    }
  }
  return OIDT_OK;
}

// The device layout pads the unit-stride axis (y in 2D, z in 3D) to a
// multiple of 128 bytes, so every 64-lane row segment of every plane starts
// on a cache line (rows of 240 FP32 cells otherwise straddle lines and the
// kernels fetch and write partial lines).  Kernels index with the padded
// extent; every update range, CPML slab, source box, port and DFT window is
// taken from the real grid, so padding cells are never updated, and no
// updated cell reads one (the last updated cell along the axis is offset
// cells inside the real grid).
template <typename T>
oidt_run_desc padded(const oidt_run_desc& d) {
  oidt_run_desc p = d;
  constexpr int q = 128 / (int)sizeof(T);
  if (d.scheme == SCHEME_3D)
    p.nz = (d.nz + q - 1) / q * q;
  else
    p.ny = (d.ny + q - 1) / q * q;
  return p;
}

template <typename T>
struct Layout {
  size_t off_f[6], off_psi[36], off_f2[6], off_psi2[36], off_mat, off_ike[3], off_ikh[3], off_be, off_ceo, off_bh, off_cho,
      off_jsrc, off_ring, off_acc, off_w, total;
  int nplanes, nq, wcells, pcells, prow, wx0, wx1, wy0, wy1;
};

template <typename T>
Layout<T> layout(const oidt_run_desc& d) {
  Layout<T> L{};
  const bool is3d = d.scheme == SCHEME_3D;
  const oidt_run_desc pd = padded<T>(d);   // allocation extents
  const size_t cells = (size_t)pd.nx * pd.ny * pd.nz;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  // Every run holds a second set of fields and psi: the fused step form
  // reads one set and writes the other.
  const bool two = true;
  for (int f = 0; f < 6; ++f) L.off_f[f] = field_used(d.scheme, f) ? take(cells * sizeof(T)) : (size_t)-1;
  for (int f = 0; f < 6; ++f) L.off_f2[f] = two && field_used(d.scheme, f) ? take(cells * sizeof(T)) : (size_t)-1;
  for (int i = 0; i < 36; ++i) L.off_psi[i] = L.off_psi2[i] = (size_t)-1;
  for (auto pr : psi_pairs(d.scheme))
    for (int side = 0; side < 2; ++side) {
      L.off_psi[psi_id(pr.first, pr.second, side)] = take(psi_elems(pd, pr.second) * sizeof(T));
      if (two) L.off_psi2[psi_id(pr.first, pr.second, side)] = take(psi_elems(pd, pr.second) * sizeof(T));
    }
  L.off_mat = take(cells);
  for (int a = 0; a < 3; ++a) {
    const int n = a == 0 ? pd.nx : (a == 1 ? pd.ny : pd.nz);
    L.off_ike[a] = take(n * sizeof(T));
    L.off_ikh[a] = take(n * sizeof(T));
  }
  L.off_be = take(d.num_cpml * sizeof(T));
  L.off_ceo = take(d.num_cpml * sizeof(T));
  L.off_bh = take(d.num_cpml * sizeof(T));
  L.off_cho = take(d.num_cpml * sizeof(T));
  L.off_jsrc = take(std::max(d.nt, 1) * sizeof(T));
  const int N = d.num_cpml;
  L.wy0 = N;
  L.wy1 = d.ny - N;
  L.prow = L.wy1 - L.wy0 + 1;
  if (is3d) {
    L.wx0 = N;
    L.wx1 = d.nx - N;
    L.pcells = (L.wx1 - L.wx0 + 1) * L.prow;
    L.wcells = (L.wx1 - L.wx0) * (L.wy1 - L.wy0);
    L.nplanes = 12;
    L.nq = 8;
  } else {
    L.wx0 = L.wx1 = 0;
    L.pcells = L.prow;
    L.wcells = L.wy1 - L.wy0;
    L.nplanes = 6;
    L.nq = 4;
  }
  L.off_ring = take((size_t)kChunk * L.nplanes * L.pcells * sizeof(T));
  L.off_acc = take((size_t)L.nq * std::max(d.nf, 1) * L.wcells * 2 * sizeof(double));
  L.off_w = take(std::max(d.nf, 1) * sizeof(double));
  L.total = off;
  return L;
}

uint64_t footprint(const oidt_run_desc& d) {
  return d.precision == OIDT_F64 ? layout<double>(d).total : layout<float>(d).total;
}

}  // namespace

struct oidt_batch {
  int scheme = 0, precision = 0, max_slots = 0;
  int device = 0;              // device current at creation; the batch's stream and memory live there
  hipStream_t stream = nullptr;
  void* d_slots = nullptr;
  int* d_counts = nullptr;  // [0] stencil tiles, [1] DFT work items
  int4* d_tiles = nullptr;
  size_t tiles_cap = 0;
  int4* d_work = nullptr;
  size_t work_cap = 0;
  std::vector<SlotHost> slots;
  Staging staging;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool dirty = true;
  int ntiles = 0, nwork = 0;
  int xchunk = 4;
  int grid_override = 0;
  // the 3D march loads its read-once values non-temporally: set when the
  // resident runs' fields exceed the last-level cache (see rebuild_lists)
  bool stream_loads = false;
  int variant = 2;
  int64_t chunks_enqueued = 0;
  int64_t chunks_done = 0;
  std::deque<std::pair<hipEvent_t, int64_t>> marks;
  std::vector<hipEvent_t> free_events;
  hipStream_t aux = nullptr;   // reductions over finished runs, off the stepping stream
  double* d_pow = nullptr;     // [nf][2] port power output
  size_t pow_cap = 0;
  void* d_mons = nullptr;      // Monitor<T> table of all resident runs
  size_t mons_cap = 0;         // bytes
  int4* d_mwork[2] = {nullptr, nullptr};   // monitor work after the E / H half-step
  size_t mwork_cap[2] = {0, 0};
  int nmwork[2] = {0, 0};
  bool has_monitors() const {
    for (const auto& sh : slots)
      if (sh.used && !sh.mons.empty()) return true;
    return false;
  }
  size_t slot_bytes() const { return precision == OIDT_F64 ? sizeof(Slot<double>) : sizeof(Slot<float>); }
};

struct oidt_solver {
  oidt_batch* b = nullptr;
  int slot = -1;
};

namespace {

// Device-memory image of one run, built on the host.
template <typename T>
int build_slot(oidt_batch* b, SlotHost& sh, const oidt_run_desc& d, const oidt_run_inputs* in) {
  const Layout<T> L = layout<T>(d);
  const size_t cells = (size_t)d.nx * d.ny * d.nz;
  // maps per scheme and the material-byte component each is indexed under;
  // lossy TMz: {CURL_H, MOD_E}
  const int ncb = d.scheme == SCHEME_TMZ ? 1 : (d.scheme == SCHEME_3D ? 3 : 2);
  const int cb_comp[4][3] = {{2, -1, -1}, {0, 1, -1}, {0, 1, 2}, {2, 0, -1}};
  for (int c = 0; c < ncb; ++c)
    if (!in->cb[c]) return fail(OIDT_EINVAL, "missing cb map");
  const int naxes = d.scheme == SCHEME_3D ? 3 : 2;
  for (int a = 0; a < naxes; ++a)
    if (!in->kappa_e[a] || !in->kappa_h[a]) return fail(OIDT_EINVAL, "missing kappa array");
  if (!in->be || !in->ce || !in->bh || !in->ch) return fail(OIDT_EINVAL, "missing CPML coefficients");
  if (d.nt > 0 && !in->jsrc) return fail(OIDT_EINVAL, "missing J_SRC");
  if (d.nf > 0 && !in->freqs) return fail(OIDT_EINVAL, "missing freqs");

  // Material compression: per E component, the distinct cb values (at most
  // kMatMax) and a 2-bit index per cell, packed as bits 2c..2c+1 of one byte.
  const oidt_run_desc pd = padded<T>(d);
  const size_t row = d.scheme == SCHEME_3D ? (size_t)d.nz : (size_t)d.ny;            // real cells per row
  const size_t prow = d.scheme == SCHEME_3D ? (size_t)pd.nz : (size_t)pd.ny;         // padded
  const size_t pcells = (size_t)pd.nx * pd.ny * pd.nz;
  std::vector<uint8_t> mat(pcells, 0);
  T cbv[3][kMatMax] = {};
  for (int c = 0; c < ncb; ++c) {
    const int comp = cb_comp[d.scheme][c];
    const T* src = static_cast<const T*>(in->cb[c]);
    int nv = 0;
    T vals[kMatMax] = {};
    int last = -1;
    for (size_t e = 0; e < cells; ++e) {
      const T v = src[e];
      int id = -1;
      if (last >= 0 && vals[last] == v) id = last;
      for (int q = 0; id < 0 && q < nv; ++q)
        if (vals[q] == v) id = q;
      if (id < 0) {
        if (nv == kMatMax)
          return fail(OIDT_EINVAL, "cb map of component " + std::to_string(comp) + " has more than 4 distinct values");
        vals[nv] = v;
        id = nv++;
      }
      last = id;
      mat[e / row * prow + e % row] |= (uint8_t)(id << (2 * comp));
    }
    for (int q = 0; q < kMatMax; ++q) cbv[comp][q] = vals[q];
  }
  // Reciprocal kappa arrays and folded CPML c / dx.
  auto recip = [](const void* src, int n, std::vector<T>& out) {
    out.assign(n, T(0));
    if (!src) return;
    const T* s = static_cast<const T*>(src);
    for (int i = 0; i < n; ++i) out[i] = (T)(1.0 / (double)s[i]);
  };
  const int dims[3] = {d.nx, d.ny, d.nz};
  std::vector<T> ike[3], ikh[3];
  for (int a = 0; a < 3; ++a) {
    recip(a < naxes ? in->kappa_e[a] : nullptr, dims[a], ike[a]);
    recip(a < naxes ? in->kappa_h[a] : nullptr, dims[a], ikh[a]);
  }
  const int N = d.num_cpml;
  std::vector<T> be(N), ceo(N), bh(N), cho(N);
  for (int i = 0; i < N; ++i) {
    be[i] = static_cast<const T*>(in->be)[i];
    bh[i] = static_cast<const T*>(in->bh)[i];
    ceo[i] = (T)((double)static_cast<const T*>(in->ce)[i] / d.dx);
    cho[i] = (T)((double)static_cast<const T*>(in->ch)[i] / d.dx);
  }
  std::vector<double> w(std::max(d.nf, 1), 0.0);
  for (int f = 0; f < d.nf; ++f) w[f] = 2.0 * M_PI * in->freqs[f];

  // Device allocation (stream ordered) and zeroing of state.
  void* dm = nullptr;
  if (hipMallocAsync(&dm, L.total, b->stream) != hipSuccess)
    return fail(OIDT_ENOMEM, "device allocation of " + std::to_string(L.total) + " bytes failed");
  sh.dmem = dm;
  sh.bytes = L.total;
  char* base = static_cast<char*>(dm);
  HIP_TRY(hipMemsetAsync(dm, 0, L.total, b->stream));

  // Stage the constant arrays in pinned memory and upload.
  struct Piece { size_t off; const void* src; size_t bytes; };
  std::vector<Piece> pieces = {
      {L.off_mat, mat.data(), pcells},
      {L.off_be, be.data(), N * sizeof(T)},
      {L.off_ceo, ceo.data(), N * sizeof(T)},
      {L.off_bh, bh.data(), N * sizeof(T)},
      {L.off_cho, cho.data(), N * sizeof(T)},
      {L.off_w, w.data(), w.size() * sizeof(double)},
  };
  if (d.nt > 0) pieces.push_back({L.off_jsrc, in->jsrc, (size_t)d.nt * sizeof(T)});
  for (int a = 0; a < 3; ++a) {
    pieces.push_back({L.off_ike[a], ike[a].data(), ike[a].size() * sizeof(T)});
    pieces.push_back({L.off_ikh[a], ikh[a].data(), ikh[a].size() * sizeof(T)});
  }
  size_t total = 0;
  for (auto& p : pieces) total += align_up(p.bytes, 64);
  void* stage = nullptr;
  const int sid = b->staging.get(total, &stage);
  if (sid < 0) return fail(OIDT_ENOMEM, "pinned staging allocation failed");
  size_t so = 0;
  for (auto& p : pieces) {
    std::memcpy(static_cast<char*>(stage) + so, p.src, p.bytes);
    HIP_TRY(hipMemcpyAsync(base + p.off, static_cast<char*>(stage) + so, p.bytes, hipMemcpyHostToDevice, b->stream));
    so += align_up(p.bytes, 64);
  }
  if (b->staging.submit(sid, b->stream) != 0) return fail(OIDT_EHIP, "event record failed");

  // Descriptor.
  Slot<T> s{};
  s.nx = pd.nx; s.ny = pd.ny; s.nz = pd.nz; s.N = N;   // strides: padded extents
  s.ib = s.jb = d.offset; s.kb = d.scheme == SCHEME_3D ? d.offset : 0;
  s.ie = d.nx - d.offset; s.je = d.ny - d.offset; s.ke = d.scheme == SCHEME_3D ? d.nz - d.offset : 1;
  s.nt = d.nt; s.nf = d.nf; s.base = 0; s.active = 1; s.scheme = d.scheme;
  s.src_field = d.src_field;
  s.sx0 = d.src_box[0]; s.sx1 = d.src_box[1]; s.sy0 = d.src_box[2]; s.sy1 = d.src_box[3];
  s.sz0 = d.src_box[4]; s.sz1 = d.src_box[5];
  s.port[0] = d.port[0]; s.port[1] = d.port[1];
  s.wx0 = L.wx0; s.wx1 = L.wx1; s.wy0 = L.wy0; s.wy1 = L.wy1;
  s.prow = L.prow; s.pcells = L.pcells; s.wcells = L.wcells; s.nplanes = L.nplanes; s.nq = L.nq;
  s.db = (T)d.db;
  s.dxT = (T)d.dx;
  s.dtmu = (T)d.dt / (T)kMu0;
  s.dt = d.dt;
  for (int f = 0; f < 6; ++f) s.f[f] = L.off_f[f] == (size_t)-1 ? nullptr : (OIDT_GLOBAL T*)(base + L.off_f[f]);
  for (int i = 0; i < 36; ++i) s.psi[i] = L.off_psi[i] == (size_t)-1 ? nullptr : (OIDT_GLOBAL T*)(base + L.off_psi[i]);
  for (int f = 0; f < 6; ++f) s.f2[f] = L.off_f2[f] == (size_t)-1 ? nullptr : (OIDT_GLOBAL T*)(base + L.off_f2[f]);
  for (int i = 0; i < 36; ++i) s.psi2[i] = L.off_psi2[i] == (size_t)-1 ? nullptr : (OIDT_GLOBAL T*)(base + L.off_psi2[i]);
  for (int a = 0; a < 3; ++a) s.psi_hi[a] = 0;
  for (auto pr : psi_pairs(d.scheme)) {
    const size_t lo = L.off_psi[psi_id(pr.first, pr.second, 0)], hi = L.off_psi[psi_id(pr.first, pr.second, 1)];
    s.psi_hi[pr.second] = (unsigned)(hi - lo);
  }
  s.mat = (OIDT_GLOBAL const uint8_t*)(base + L.off_mat);
  for (int c = 0; c < 3; ++c)
    for (int q = 0; q < kMatMax; ++q) s.cbv[c][q] = cbv[c][q];
  for (int a = 0; a < 3; ++a) {
    s.ike[a] = (OIDT_GLOBAL const T*)(base + L.off_ike[a]);
    s.ikh[a] = (OIDT_GLOBAL const T*)(base + L.off_ikh[a]);
  }
  s.be = (OIDT_GLOBAL const T*)(base + L.off_be);
  s.ceo = (OIDT_GLOBAL const T*)(base + L.off_ceo);
  s.bh = (OIDT_GLOBAL const T*)(base + L.off_bh);
  s.cho = (OIDT_GLOBAL const T*)(base + L.off_cho);
  s.jsrc = (OIDT_GLOBAL const T*)(base + L.off_jsrc);
  s.ring = (OIDT_GLOBAL T*)(base + L.off_ring);
  s.acc = (OIDT_GLOBAL double*)(base + L.off_acc);
  s.w = (OIDT_GLOBAL const double*)(base + L.off_w);
  sh.slot.assign(reinterpret_cast<unsigned char*>(&s), reinterpret_cast<unsigned char*>(&s) + sizeof(s));
  sh.nplanes = L.nplanes; sh.nq = L.nq; sh.wcells = L.wcells; sh.pcells = L.pcells; sh.wyn = L.wy1 - L.wy0;
  return OIDT_OK;
}

int upload_bytes(oidt_batch* b, void* dst, const void* src, size_t bytes) {
  void* stage = nullptr;
  const int sid = b->staging.get(bytes, &stage);
  if (sid < 0) return fail(OIDT_ENOMEM, "pinned staging allocation failed");
  std::memcpy(stage, src, bytes);
  HIP_TRY(hipMemcpyAsync(dst, stage, bytes, hipMemcpyHostToDevice, b->stream));
  if (b->staging.submit(sid, b->stream) != 0) return fail(OIDT_EHIP, "event record failed");
  return OIDT_OK;
}

int grow(oidt_batch* b, int4** buf, size_t* cap, size_t need) {
  if (need <= *cap) return OIDT_OK;
  size_t nc = std::max(need, *cap * 2);
  int4* nb = nullptr;
  if (hipMallocAsync(reinterpret_cast<void**>(&nb), nc * sizeof(int4), b->stream) != hipSuccess)
    return fail(OIDT_ENOMEM, "tile list allocation failed");
  if (*buf) HIP_TRY(hipFreeAsync(*buf, b->stream));
  *buf = nb;
  *cap = nc;
  return OIDT_OK;
}

// SlotHost::tb of a run stepped by a launch form.
int run_tb(int variant, int xchunk) { return variant == 4 ? xchunk : (variant == 3 ? 1 : 0); }

// Kernels whose tiles are per wavefront (four tiles per block): the fused
// 2D steps.
bool march_tiles(const oidt_batch* b) { return b->variant == 3; }

// z-cells a wave of the fused 3D step owns (its lane 0 is the cell below).
constexpr int kHaloZ = kBX - 1;

// Waves (rows) of a blocked 3D step's block: 16 in FP32; 8 in FP64, whose
// cells take twice the registers.
int tb3_waves(int precision) { return precision == OIDT_F64 ? 8 : 16; }

// Monitor table and the two work lists (after the E and after the H
// half-step): one item per kMonCells consecutive box cells.
template <typename T>
int rebuild_monitors(oidt_batch* b) {
  std::vector<Monitor<T>> table;
  std::vector<int4> work[2];
  for (int si = 0; si < (int)b->slots.size(); ++si) {
    const SlotHost& sh = b->slots[si];
    if (!sh.used) continue;
    for (const MonHost& mh : sh.mons) {
      const oidt_monitor_desc& d = mh.d;
      Monitor<T> m{};
      m.re = (OIDT_GLOBAL T*)mh.re;
      m.im = (OIDT_GLOBAL T*)mh.im;
      m.cs = (OIDT_GLOBAL const double*)mh.cs;
      m.scale = d.scale;
      m.slot = si; m.field = d.field; m.mode = d.mode; m.npts = d.npts;
      for (int q = 0; q < 4; ++q) { m.ox[q] = d.off[q][0]; m.oy[q] = d.off[q][1]; m.oz[q] = d.off[q][2]; }
      m.x0 = d.box[0]; m.y0 = d.box[2]; m.z0 = d.box[4];
      m.ex = d.box[1] - d.box[0]; m.ey = d.box[3] - d.box[2]; m.ez = d.box[5] - d.box[4];
      m.gx = sh.desc.nx; m.gy = sh.desc.ny; m.gz = sh.desc.nz;
      m.shift = d.shift; m.nrec = d.nrec;
      const int id = (int)table.size();
      table.push_back(m);
      for (size_t c0 = 0; c0 < mh.cells; c0 += kMonCells) work[d.after_h ? 1 : 0].push_back(make_int4(id, (int)c0, 0, 0));
    }
  }
  int rc;
  const size_t tbytes = std::max<size_t>(table.size(), 1) * sizeof(Monitor<T>);
  if (tbytes > b->mons_cap) {
    void* nb = nullptr;
    if (hipMallocAsync(&nb, tbytes, b->stream) != hipSuccess) return fail(OIDT_ENOMEM, "monitor table allocation failed");
    if (b->d_mons) HIP_TRY(hipFreeAsync(b->d_mons, b->stream));
    b->d_mons = nb;
    b->mons_cap = tbytes;
  }
  if (!table.empty() && (rc = upload_bytes(b, b->d_mons, table.data(), table.size() * sizeof(Monitor<T>))) != OIDT_OK)
    return rc;
  for (int h = 0; h < 2; ++h) {
    if ((rc = grow(b, &b->d_mwork[h], &b->mwork_cap[h], std::max<size_t>(work[h].size(), 1))) != OIDT_OK) return rc;
    if (!work[h].empty() && (rc = upload_bytes(b, b->d_mwork[h], work[h].data(), work[h].size() * sizeof(int4))) != OIDT_OK)
      return rc;
    b->nmwork[h] = (int)work[h].size();
  }
  return OIDT_OK;
}

// Rebuild the tile and DFT work lists for the resident slots and upload them.
int rebuild_lists(oidt_batch* b) {
  std::vector<int4> tiles, work;
  const int xc = b->xchunk;
  for (int si = 0; si < (int)b->slots.size(); ++si) {
    const SlotHost& sh = b->slots[si];
    if (!sh.used) continue;
    const oidt_run_desc& d = sh.desc;
    if (d.scheme == SCHEME_3D && b->variant == 3) {
      // fused step: one wavefront per (row, 63 z-cells) x xc planes
      for (int x0 = 0; x0 < d.nx; x0 += xc)
        for (int y0 = 0; y0 < d.ny; ++y0)
          for (int z0 = 0; z0 < d.nz; z0 += kHaloZ) tiles.push_back(make_int4(si, x0, y0, z0));
    } else if (d.scheme == SCHEME_3D && b->variant == 4) {
      // blocked 3D step: one block per segment of kTb3Seg planes by its inner
      // rows and z-cells
      const int ty = tb3_waves(b->precision) - 2 * xc, tz = kBX - 2 * xc;
      for (int x0 = 0; x0 < d.nx; x0 += kTb3Seg)
        for (int y0 = 0; y0 < d.ny; y0 += ty)
          for (int z0 = 0; z0 < d.nz; z0 += tz) tiles.push_back(make_int4(si, x0, y0, z0));
    } else if (b->variant == 4) {
      // blocked step: one block per tile of its inner cells.  w = 1 marks an
      // interior tile: its extended tile lies inside every update range and
      // outside the CPML slabs and the source box, and no capture plane
      // falls among its inner planes
      const int tx = kTbPlanes - 2 * xc, ty = 64 - 2 * xc;
      const int xlo = d.offset + d.num_cpml, xhi = d.nx - d.offset - d.num_cpml;
      const int ylo = d.offset + d.num_cpml, yhi = d.ny - d.offset - d.num_cpml;
      for (int x0 = 0; x0 < d.nx; x0 += tx)
        for (int y0 = 0; y0 < d.ny; y0 += ty) {
          const int ex0 = x0 - xc, ex1 = ex0 + kTbPlanes, ey0 = y0 - xc, ey1 = ey0 + 64;
          bool in = ex0 >= xlo && ex1 <= xhi && ey0 >= ylo && ey1 <= yhi;
          if (d.src_field >= 0 && ex0 < d.src_box[1] && ex1 > d.src_box[0] && ey0 < d.src_box[3] && ey1 > d.src_box[2])
            in = false;
          for (const int pl : {d.port[0], d.port[0] + 1, d.port[1], d.port[1] + 1})
            if (pl >= x0 && pl < x0 + tx) in = false;
          tiles.push_back(make_int4(si, x0, y0, in ? 1 : 0));
        }
    } else if (d.scheme == SCHEME_3D) {
      for (int x0 = 0; x0 < d.nx; x0 += xc)
        for (int y0 = 0; y0 < d.ny; y0 += kBY)
          for (int z0 = 0; z0 < d.nz; z0 += kBX) tiles.push_back(make_int4(si, x0, y0, z0));
    } else {
      // the fused 2D steps give each wavefront its own tile of xc planes;
      // the other 2D forms give a block 4 waves x xc planes
      const int xstep = march_tiles(b) ? xc : xc * kBY;
      for (int x0 = 0; x0 < d.nx; x0 += xstep)
        for (int y0 = 0; y0 < d.ny; y0 += kBX) tiles.push_back(make_int4(si, x0, y0, 0));
    }
  }
  // DFT work items: kFB frequencies each, fewer when the whole batch would
  // otherwise have too few items to fill the GPU (a small single run: its
  // fold is a serial chain per item).  Each frequency's sum is formed the
  // same way whatever the grouping.
  int fb = kFB;
  auto items = [&](int f) {
    size_t n = 0;
    for (const SlotHost& sh : b->slots)
      if (sh.used) n += (size_t)sh.nq * ((sh.wcells + kThreads - 1) / kThreads) * ((sh.desc.nf + f - 1) / f);
    return n;
  };
  while (fb > 1 && items(fb) < 1024) fb >>= 1;
  for (int si = 0; si < (int)b->slots.size(); ++si) {
    const SlotHost& sh = b->slots[si];
    if (!sh.used) continue;
    const int cblocks = (sh.wcells + kThreads - 1) / kThreads;
    for (int q = 0; q < sh.nq; ++q)
      for (int cb = 0; cb < cblocks; ++cb)
        for (int f0 = 0; f0 < sh.desc.nf; f0 += fb) work.push_back(make_int4(si, q, cb, f0 | (fb << 16)));
  }
  // Field bytes of one step of the resident runs against the 256 MB
  // last-level cache: above 192 MB the march's read-once loads bypass it
  // (measured: 16 runs of 96 x 64 x nz, 856 MB, 7 % faster; one run of
  // 128 x 96 x 320, 94 MB, 12 % slower)
  size_t field_bytes = 0;
  for (const SlotHost& sh : b->slots)
    if (sh.used)
      field_bytes += (size_t)sh.desc.nx * sh.desc.ny * sh.desc.nz * 6 * (b->precision == OIDT_F64 ? 8 : 4);
  b->stream_loads = field_bytes > ((size_t)192 << 20);
  int rc;
  if ((rc = grow(b, &b->d_tiles, &b->tiles_cap, std::max<size_t>(tiles.size(), 1))) != OIDT_OK) return rc;
  if ((rc = grow(b, &b->d_work, &b->work_cap, std::max<size_t>(work.size(), 1))) != OIDT_OK) return rc;
  if (!tiles.empty() && (rc = upload_bytes(b, b->d_tiles, tiles.data(), tiles.size() * sizeof(int4))) != OIDT_OK) return rc;
  if (!work.empty() && (rc = upload_bytes(b, b->d_work, work.data(), work.size() * sizeof(int4))) != OIDT_OK) return rc;
  int counts[2] = {(int)tiles.size(), (int)work.size()};
  if ((rc = upload_bytes(b, b->d_counts, counts, sizeof(counts))) != OIDT_OK) return rc;
  b->ntiles = counts[0];
  b->nwork = counts[1];
  return b->precision == OIDT_F64 ? rebuild_monitors<double>(b) : rebuild_monitors<float>(b);
}

int grid_for(int items, int cap) {
  int g = std::min(std::max(items, 1), cap);
  return (g + 7) / 8 * 8;
}

template <typename T, int XC>
void launch_half(oidt_batch* b, int gs, int tstep, bool e_half) {
  const Slot<T>* sl = static_cast<const Slot<T>*>(b->d_slots);
  const dim3 blk(kBX, kBY), grd(gs);
  hipStream_t st = b->stream;
  const bool v1 = b->variant == 1;
  switch (b->scheme) {
    case SCHEME_TMZ:
      if (e_half) {
        if (v1) hipLaunchKernelGGL((k_tmz_e<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else hipLaunchKernelGGL((k_tmz_e2<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      } else {
        if (v1) hipLaunchKernelGGL((k_tmz_h<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else hipLaunchKernelGGL((k_tmz_h2<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      }
      break;
    case SCHEME_TMZ_LOSSY:
      if (e_half) hipLaunchKernelGGL((k_tmzl_e<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      else hipLaunchKernelGGL((k_tmzl_h<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      break;
    case SCHEME_TEZ:
      if (e_half) {
        if (v1) hipLaunchKernelGGL((k_tez_e<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else hipLaunchKernelGGL((k_tez_e2<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      } else {
        if (v1) hipLaunchKernelGGL((k_tez_h<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else hipLaunchKernelGGL((k_tez_h2<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      }
      break;
    default:
      if (e_half) {
        if (v1 && b->stream_loads) hipLaunchKernelGGL((k_3d_e<T, XC, true>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else if (v1) hipLaunchKernelGGL((k_3d_e<T, XC, false>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else hipLaunchKernelGGL((k_3d_e2<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      } else {
        if (v1 && b->stream_loads) hipLaunchKernelGGL((k_3d_h<T, XC, true>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else if (v1) hipLaunchKernelGGL((k_3d_h<T, XC, false>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
        else hipLaunchKernelGGL((k_3d_h2<T, XC>), grd, blk, 0, st, sl, b->d_tiles, b->d_counts, tstep);
      }
      break;
  }
}

// The fused step: the whole time step in the "E" launch; its tile depth
// (planes per wave) has more choices than the half-step forms' x-chunk.
template <typename T, int XM>
void launch_fused(oidt_batch* b, int gs, int tstep) {
  const Slot<T>* sl = static_cast<const Slot<T>*>(b->d_slots);
  const dim3 blk(kBX, kBY), grd(gs);
  if (b->scheme == SCHEME_TMZ) {
    hipLaunchKernelGGL((k_tmz_step<T, XM>), grd, blk, 0, b->stream, sl, b->d_tiles, b->d_counts, tstep);
  } else if (b->scheme == SCHEME_TEZ) {
    hipLaunchKernelGGL((k_tez_step<T, XM>), grd, blk, 0, b->stream, sl, b->d_tiles, b->d_counts, tstep);
  } else if constexpr (XM == 1 || XM == 2 || XM == 4 || XM == 8 || XM == 16) {
    hipLaunchKernelGGL((k_3d_halo<T, XM>), grd, blk, 0, b->stream, sl, b->d_tiles, b->d_counts, tstep);
  }
}

// The temporally blocked 2D step (TMz, TEz): xchunk steps per launch.  Blocks of 16
// waves x 2 planes in FP64 (half the registers per wave of 8 x 4, and the
// FP64 workloads are single small runs that need the extra waves per tile to
// hide latency), 8 x 4 in FP32 (better on large batches).  The shape depends
// only on the precision, so a run gives the same bits alone and in a batch.
template <typename T, int TB, int W>
void launch_blocked_w(oidt_batch* b, int gs, int tstep) {
  const Slot<T>* sl = static_cast<const Slot<T>*>(b->d_slots);
  if (b->scheme == SCHEME_TEZ)
    hipLaunchKernelGGL((k_tez_tb<T, TB, W>), dim3(gs), dim3(64, W), 0, b->stream, sl, b->d_tiles, b->d_counts, tstep);
  else
    hipLaunchKernelGGL((k_tmz_tb<T, TB, W>), dim3(gs), dim3(64, W), 0, b->stream, sl, b->d_tiles, b->d_counts, tstep);
}
template <typename T, int W>
void launch_blocked_tb(oidt_batch* b, int gs, int tstep) {
  switch (b->xchunk) {
    case 2: launch_blocked_w<T, 2, W>(b, gs, tstep); break;
    case 4: launch_blocked_w<T, 4, W>(b, gs, tstep); break;
    default: launch_blocked_w<T, 8, W>(b, gs, tstep); break;
  }
}
// The temporally blocked 3D step: 2 steps per launch, blocks of
// tb3_waves() rows.
template <typename T>
void launch_blocked3d(oidt_batch* b, int gs, int tstep) {
  const Slot<T>* sl = static_cast<const Slot<T>*>(b->d_slots);
  constexpr int W = sizeof(T) == 8 ? 8 : 16;
  hipLaunchKernelGGL((k_3d_tb<T, 2, W>), dim3(gs), dim3(64, W), 0, b->stream, sl, b->d_tiles, b->d_counts, tstep);
}

template <typename T>
void launch_blocked(oidt_batch* b, int gs, int tstep) {
  if (b->scheme == SCHEME_3D) launch_blocked3d<T>(b, gs, tstep);
  else if (sizeof(T) == 8) launch_blocked_tb<T, 16>(b, gs, tstep);
  else launch_blocked_tb<T, 8>(b, gs, tstep);
}

// Steps one stencil launch advances: the blocked step's depth, else 1.
int launch_steps(const oidt_batch* b) { return b->variant == 4 ? b->xchunk : 1; }

template <typename T>
void launch_stencil(oidt_batch* b, int gs, int tstep, bool e_half) {
  if (b->variant == 4) {
    if (e_half) launch_blocked<T>(b, gs, tstep);
    return;
  }
  if (b->variant == 3) {
    if (!e_half) return;
    switch (b->xchunk) {
      case 1: launch_fused<T, 1>(b, gs, tstep); break;
      case 2: launch_fused<T, 2>(b, gs, tstep); break;
      case 4: launch_fused<T, 4>(b, gs, tstep); break;
      case 8: launch_fused<T, 8>(b, gs, tstep); break;
      case 12: launch_fused<T, 12>(b, gs, tstep); break;
      case 16: launch_fused<T, 16>(b, gs, tstep); break;
      case 20: launch_fused<T, 20>(b, gs, tstep); break;
      case 24: launch_fused<T, 24>(b, gs, tstep); break;
      default: launch_fused<T, 32>(b, gs, tstep); break;
    }
    return;
  }
  switch (b->xchunk) {
    case 1: launch_half<T, 1>(b, gs, tstep, e_half); break;
    case 2: launch_half<T, 2>(b, gs, tstep, e_half); break;
    case 4: launch_half<T, 4>(b, gs, tstep, e_half); break;
    case 16: launch_half<T, 16>(b, gs, tstep, e_half); break;
    case 32: launch_half<T, 32>(b, gs, tstep, e_half); break;
    default: launch_half<T, 8>(b, gs, tstep, e_half); break;
  }
}

template <typename T>
void launch_dft(oidt_batch* b, int gd) {
  hipLaunchKernelGGL((k_dft<T>), dim3(gd), dim3(kThreads), 0, b->stream,
                     static_cast<const Slot<T>*>(b->d_slots), b->d_work, b->d_counts);
}

template <typename T>
void launch_advance(oidt_batch* b) {
  hipLaunchKernelGGL((k_advance<T>), dim3((b->max_slots + 255) / 256), dim3(256), 0, b->stream,
                     static_cast<Slot<T>*>(b->d_slots), b->max_slots);
}

// Grid of the stencil launches.  The marching step gives every wave exactly
// one tile (no cap): a wave holding a second tile at the end of the launch
// leaves a tail that measured 15-30 % of the step.  The other forms keep a
// persistent grid of at most 8192 blocks unless overridden.
int stencil_grid(const oidt_batch* b) {
  const bool m = march_tiles(b);
  const int items = m ? (b->ntiles + kBY - 1) / kBY : b->ntiles;
  return grid_for(items, b->grid_override > 0 ? b->grid_override : (m ? (1 << 24) : 8192));
}
int dft_grid(const oidt_batch* b) { return grid_for(b->nwork, 8192); }

template <typename T>
void launch_monitors(oidt_batch* b, int t, int h) {
  if (b->nmwork[h] == 0) return;
  hipLaunchKernelGGL((k_monitor<T>), dim3(grid_for(b->nmwork[h], 8192)), dim3(kThreads), 0, b->stream,
                     static_cast<const Slot<T>*>(b->d_slots), static_cast<const Monitor<T>*>(b->d_mons),
                     b->d_mwork[h], b->nmwork[h], t);
}

template <typename T>
int enqueue_chunk(oidt_batch* b) {
  const int gs = stencil_grid(b), gd = dft_grid(b);
  for (int t = 0; t < kChunk; t += launch_steps(b)) {
    launch_stencil<T>(b, gs, t, true);
    launch_monitors<T>(b, t, 0);
    if (b->variant < 3) launch_stencil<T>(b, gs, t, false);
    launch_monitors<T>(b, t, 1);
  }
  launch_dft<T>(b, gd);
  launch_advance<T>(b);
  return hipGetLastError() == hipSuccess ? OIDT_OK : fail(OIDT_EHIP, "kernel launch failed");
}

// Direct (non-graph) launches with an event pair around every kernel.
template <typename T>
int profile_chunks(oidt_batch* b, int nchunks, double* out) {
  const int gs = stencil_grid(b), gd = dft_grid(b);
  // kinds: 0 E half-step, 1 H half-step, 2 DFT fold, 3 advance, 4 / 5 monitors after E / H.
  // The one-launch forms (fused, blocked) have a single stencil kernel per
  // step: the chunk's stencil launches are timed back to back between one
  // pair of events (per-launch events would add their own gaps to launches
  // of a few microseconds); the two-kernel forms time each launch.
  struct Timed { int kind, launches; hipEvent_t a, z; };
  std::vector<Timed> ev;
  auto timed = [&](int kind, int launches, auto&& launch) -> int {
    Timed t{kind, launches, nullptr, nullptr};
    HIP_TRY(hipEventCreate(&t.a));
    HIP_TRY(hipEventCreate(&t.z));
    ev.push_back(t);
    HIP_TRY(hipEventRecord(t.a, b->stream));
    launch();
    HIP_TRY(hipEventRecord(t.z, b->stream));
    return OIDT_OK;
  };
  int rc;
  const int per = launch_steps(b);
  for (int c = 0; c < nchunks; ++c) {
    if (b->variant >= 3) {
      if ((rc = timed(0, kChunk / per, [&] {
             for (int t = 0; t < kChunk; t += per) launch_stencil<T>(b, gs, t, true);
           })) != OIDT_OK)
        return rc;
    } else {
      for (int t = 0; t < kChunk; ++t)
        for (int h = 0; h < 2; ++h) {
          if ((rc = timed(h, 1, [&] { launch_stencil<T>(b, gs, t, h == 0); })) != OIDT_OK) return rc;
          if (b->nmwork[h] > 0 && (rc = timed(4 + h, 1, [&] { launch_monitors<T>(b, t, h); })) != OIDT_OK) return rc;
        }
    }
    if ((rc = timed(2, 1, [&] { launch_dft<T>(b, gd); })) != OIDT_OK) return rc;
    if ((rc = timed(3, 1, [&] { launch_advance<T>(b); })) != OIDT_OK) return rc;
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(b->stream));
  double sum[6] = {0, 0, 0, 0, 0, 0};
  int cnt[6] = {0, 0, 0, 0, 0, 0};
  for (const Timed& t : ev) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, t.a, t.z));
    sum[t.kind] += ms;
    cnt[t.kind] += t.launches;
  }
  for (auto& t : ev) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.z);
  }
  for (int i = 0; i < 6; ++i) out[i] = cnt[i] ? sum[i] / cnt[i] : 0.0;
  b->chunks_enqueued += nchunks;
  b->chunks_done = b->chunks_enqueued;
  return OIDT_OK;
}

int capture_graph(oidt_batch* b) {
  if (b->gexec) {
    HIP_TRY(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  if (b->graph) {
    HIP_TRY(hipGraphDestroy(b->graph));
    b->graph = nullptr;
  }
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = b->precision == OIDT_F64 ? enqueue_chunk<double>(b) : enqueue_chunk<float>(b);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != OIDT_OK) return rc;
  if (e != hipSuccess) return fail(OIDT_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  b->dirty = false;
  return OIDT_OK;
}

void reap_marks(oidt_batch* b) {
  while (!b->marks.empty()) {
    auto& m = b->marks.front();
    if (hipEventQuery(m.first) != hipSuccess) break;
    b->chunks_done = m.second;
    b->free_events.push_back(m.first);
    b->marks.pop_front();
  }
}

// True once the GPU has completed every chunk the run needs (later chunks skip it).
bool slot_finished(oidt_batch* b, const SlotHost& sh) {
  reap_marks(b);
  const int64_t need = (sh.desc.nt + kChunk - 1) / kChunk;
  return b->chunks_done - sh.admit_chunk >= need;
}

int slot_ok(oidt_batch* b, int slot) {
  if (!b) return fail(OIDT_EINVAL, "null batch");
  if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return fail(OIDT_EINVAL, "no run in slot");
  return OIDT_OK;
}

// Fixed device array for oidt_batch_port_power_noalloc: the per-frequency
// (re, im) port powers of one call.  Each device holds its own copy for the
// life of the process; calls on one device share it.
constexpr int kNoallocMaxNf = OIDT_NOALLOC_MAX_NF;
__device__ __align__(256) double scratch[2 * kNoallocMaxNf];

// Device address of `scratch` per device, taken once with that device
// current (the host-side name is not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(OIDT_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess) return fail(OIDT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return OIDT_OK;
}

}  // namespace

extern "C" {

int32_t oidt_abi_version(void) { return OIDT_ABI_VERSION; }

const char* oidt_last_error(void) { return g_err.c_str(); }

int32_t oidt_chunk_steps(void) { return kChunk; }

int oidt_device_info_get(oidt_device_info* out) {
  if (!out) return fail(OIDT_EINVAL, "null output");
  std::memset(out, 0, sizeof(*out));
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  std::snprintf(out->name, sizeof(out->name), "%s", p.name);
  std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
  out->compute_units = p.multiProcessorCount;
  out->device_id = dev;
  size_t fr = 0, tot = 0;
  HIP_TRY(hipMemGetInfo(&fr, &tot));
  out->hbm_free_bytes = fr;
  out->hbm_total_bytes = tot;
  return OIDT_OK;
}

int oidt_run_footprint(const oidt_run_desc* desc, uint64_t* bytes) {
  int rc = validate(desc);
  if (rc != OIDT_OK) return rc;
  if (!bytes) return fail(OIDT_EINVAL, "null output");
  *bytes = footprint(*desc);
  return OIDT_OK;
}

int oidt_batch_create(int32_t scheme, int32_t precision, int32_t max_slots, oidt_batch** out) {
  if (!out) return fail(OIDT_EINVAL, "null output");
  *out = nullptr;
  if (scheme < 0 || scheme > 3) return fail(OIDT_EINVAL, "unknown scheme");
  if (precision != OIDT_F32 && precision != OIDT_F64) return fail(OIDT_EINVAL, "unknown precision");
  if (max_slots < 1 || max_slots > 4096) return fail(OIDT_EINVAL, "max_slots out of range");
  oidt_batch* b = new (std::nothrow) oidt_batch();
  if (!b) return fail(OIDT_ENOMEM, "host allocation failed");
  b->scheme = scheme;
  b->precision = precision;
  b->max_slots = max_slots;
  if (hipGetDevice(&b->device) != hipSuccess) {
    delete b;
    return fail(OIDT_EHIP, "hipGetDevice failed");
  }

  // measured defaults on MI355X: the prefetching march over 4-plane tiles in
  // 3D, the batched-load form with four cells per thread in 2D
  b->variant = scheme == SCHEME_3D ? 1 : 2;
  b->xchunk = 4;
  b->slots.resize(max_slots);
  auto bad = [&](hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    oidt_batch_destroy(b);
    return OIDT_EHIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e, "hipStreamCreate");
  if ((e = hipMalloc(&b->d_slots, (size_t)max_slots * b->slot_bytes())) != hipSuccess) return bad(e, "hipMalloc slots");
  if ((e = hipMemsetAsync(b->d_slots, 0, (size_t)max_slots * b->slot_bytes(), b->stream)) != hipSuccess) return bad(e, "hipMemset slots");
  if ((e = hipMalloc(reinterpret_cast<void**>(&b->d_counts), 2 * sizeof(int))) != hipSuccess) return bad(e, "hipMalloc counts");
  if ((e = hipMemsetAsync(b->d_counts, 0, 2 * sizeof(int), b->stream)) != hipSuccess) return bad(e, "hipMemset counts");
  *out = b;
  return OIDT_OK;
}

int oidt_batch_add(oidt_batch* b, const oidt_run_desc* desc, const oidt_run_inputs* in, int32_t* slot) {
  if (!b || !in || !slot) return fail(OIDT_EINVAL, "null argument");
  int rc = validate(desc);
  if (rc != OIDT_OK) return rc;
  if (desc->scheme != b->scheme || desc->precision != b->precision)
    return fail(OIDT_EINVAL, "run scheme/precision differs from the batch");
  int si = -1;
  for (int i = 0; i < b->max_slots; ++i)
    if (!b->slots[i].used) { si = i; break; }
  if (si < 0) return fail(OIDT_EFULL, "batch is full");
  SlotHost& sh = b->slots[si];
  sh = SlotHost();
  sh.desc = *desc;
  rc = b->precision == OIDT_F64 ? build_slot<double>(b, sh, *desc, in) : build_slot<float>(b, sh, *desc, in);
  if (rc != OIDT_OK) {
    if (sh.dmem) (void)hipFreeAsync(sh.dmem, b->stream);
    sh = SlotHost();
    return rc;
  }
  rc = upload_bytes(b, static_cast<char*>(b->d_slots) + (size_t)si * b->slot_bytes(), sh.slot.data(), sh.slot.size());
  if (rc != OIDT_OK) return rc;
  sh.used = true;
  sh.admit_chunk = b->chunks_enqueued;
  sh.tb = run_tb(b->variant, b->xchunk);
  b->dirty = true;
  *slot = si;
  return OIDT_OK;
}

int oidt_batch_set_launch(oidt_batch* b, int32_t xchunk, int32_t grid_blocks, int32_t variant) {
  if (!b) return fail(OIDT_EINVAL, "null batch");
  if (variant < 0 || variant > 4) return fail(OIDT_EINVAL, "variant must be 0 (keep), 1, 2, 3 or 4");
  if (variant == 0) variant = b->variant;
  const int xc = xchunk ? xchunk : b->xchunk;   // the depth this launch form will use
  if (variant == 4) {
    if (b->scheme == SCHEME_3D) {
      if (xc != 2) return fail(OIDT_EINVAL, "the blocked 3D step takes 2 steps per launch");
    } else if (b->scheme != SCHEME_TMZ && b->scheme != SCHEME_TEZ) {
      return fail(OIDT_EINVAL, "the blocked step (variant 4) is for TMz, TEz and 3D runs");
    } else if (xc != 2 && xc != 4 && xc != 8) {
      return fail(OIDT_EINVAL, "the blocked step takes 2, 4 or 8 steps per launch");
    }
  }
  if (variant == 3 && b->scheme == SCHEME_3D && xc != 1 && xc != 2 && xc != 4 && xc != 8 && xc != 16)
    return fail(OIDT_EINVAL, "the fused 3D step takes 1, 2, 4, 8 or 16 planes per wave");
  for (const auto& sh : b->slots)
    if (sh.used && b->chunks_enqueued > sh.admit_chunk && sh.tb != run_tb(variant, xc))
      return fail(OIDT_ESTATE, "cannot change how a run in progress swaps field sets (two-kernel, fused or blocked form and depth)");
  const bool pow2 = xc == 1 || xc == 2 || xc == 4 || xc == 8 || xc == 16 || xc == 32;
  const bool fused_only = xc == 12 || xc == 20 || xc == 24;
  if (!pow2 && !(fused_only && variant == 3))
    return fail(OIDT_EINVAL, "xchunk must be 0 (keep), 1, 2, 4, 8, 16 or 32 (also 12, 20 or 24 for the fused step)");
  if (grid_blocks < 0) return fail(OIDT_EINVAL, "negative grid size");
  if (variant == 3 && b->scheme == SCHEME_TMZ_LOSSY)
    return fail(OIDT_EINVAL, "the lossy TMz form has no fused step (variant 1 or 2)");
  if (variant >= 3 && b->has_monitors())
    return fail(OIDT_ESTATE, "field monitors need a two-kernel form (variant 1 or 2)");
  b->variant = variant;
  for (auto& sh : b->slots)
    if (sh.used && b->chunks_enqueued == sh.admit_chunk) sh.tb = run_tb(variant, xc);
  if (xchunk) b->xchunk = xchunk;
  b->grid_override = grid_blocks;
  b->dirty = true;
  return OIDT_OK;
}

int oidt_batch_get_launch(const oidt_batch* b, int32_t* xchunk, int32_t* grid_blocks, int32_t* variant) {
  if (!b) return fail(OIDT_EINVAL, "null batch");
  if (xchunk) *xchunk = b->xchunk;
  if (grid_blocks) *grid_blocks = b->grid_override;
  if (variant) *variant = b->variant;
  return OIDT_OK;
}

int oidt_batch_advance(oidt_batch* b, int32_t nchunks) {
  if (!b) return fail(OIDT_EINVAL, "null batch");
  if (nchunks < 0) return fail(OIDT_EINVAL, "negative chunk count");
  if (nchunks == 0) return OIDT_OK;
  int rc;
  if (b->dirty) {
    if ((rc = rebuild_lists(b)) != OIDT_OK) return rc;
    if ((rc = capture_graph(b)) != OIDT_OK) return rc;
  }
  for (int c = 0; c < nchunks; ++c) HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
  b->chunks_enqueued += nchunks;
  hipEvent_t ev;
  if (!b->free_events.empty()) {
    ev = b->free_events.back();
    b->free_events.pop_back();
  } else {
    HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  HIP_TRY(hipEventRecord(ev, b->stream));
  b->marks.emplace_back(ev, b->chunks_enqueued);
  return OIDT_OK;
}

int oidt_batch_prepare(oidt_batch* b) {
  if (!b) return fail(OIDT_EINVAL, "null batch");
  if (!b->dirty) return OIDT_OK;
  int rc;
  if ((rc = rebuild_lists(b)) != OIDT_OK) return rc;
  return capture_graph(b);
}

int oidt_batch_sync(oidt_batch* b) {
  if (!b) return fail(OIDT_EINVAL, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  reap_marks(b);
  return OIDT_OK;
}

int oidt_batch_poll(oidt_batch* b, int32_t slot, int32_t* steps_done) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  if (!steps_done) return fail(OIDT_EINVAL, "null output");
  reap_marks(b);
  const SlotHost& sh = b->slots[slot];
  const int64_t ch = std::max<int64_t>(0, b->chunks_done - sh.admit_chunk);
  *steps_done = (int32_t)std::min<int64_t>(sh.desc.nt, ch * kChunk);
  return OIDT_OK;
}

int oidt_batch_progress(oidt_batch* b, int64_t* chunks_enqueued, int64_t* chunks_done) {
  if (!b || !chunks_enqueued || !chunks_done) return fail(OIDT_EINVAL, "null argument");
  reap_marks(b);
  *chunks_enqueued = b->chunks_enqueued;
  *chunks_done = b->chunks_done;
  return OIDT_OK;
}

int oidt_batch_read_dft(oidt_batch* b, int32_t slot, int32_t port, int32_t comp, double* re, double* im) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  const SlotHost& sh = b->slots[slot];
  const int ncomp = sh.nq / 2;
  if (port < 0 || port > 1 || comp < 0 || comp >= ncomp) return fail(OIDT_EINVAL, "bad port or component");
  if (!re || !im) return fail(OIDT_EINVAL, "null output");
  const oidt_run_desc& d = sh.desc;
  const size_t n = (size_t)d.nf * sh.wcells;
  if (n == 0) return OIDT_OK;
  const size_t off = d.precision == OIDT_F64 ? layout<double>(d).off_acc : layout<float>(d).off_acc;
  const int q = port * ncomp + comp;
  std::vector<double> buf(2 * n);
  // A finished run's accumulators are final while later chunks run for the
  // others: read without waiting for the batch stream.
  if (!slot_finished(b, sh)) HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(buf.data(), static_cast<char*>(sh.dmem) + off + (size_t)q * n * 2 * sizeof(double),
                    2 * n * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t e = 0; e < n; ++e) {
    re[e] = buf[2 * e];
    im[e] = buf[2 * e + 1];
  }
  return OIDT_OK;
}

int oidt_batch_port_power(oidt_batch* b, int32_t slot, int32_t port, int32_t sub_slot, int32_t sub_port,
                          int32_t row0, int32_t nrows, double* re, double* im) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  const SlotHost& sh = b->slots[slot];
  const oidt_run_desc& d = sh.desc;
  if (port < 0 || port > 1) return fail(OIDT_EINVAL, "bad port");
  if (!re || !im) return fail(OIDT_EINVAL, "null output");
  if (d.scheme == SCHEME_TMZ_LOSSY) return fail(OIDT_EINVAL, "the lossy TMz form has no port DFT");
  const SlotHost* ss = nullptr;
  if (sub_slot >= 0) {
    if ((rc = slot_ok(b, sub_slot)) != OIDT_OK) return rc;
    ss = &b->slots[sub_slot];
    if (sub_port < 0 || sub_port > 1) return fail(OIDT_EINVAL, "bad port of the subtracted run");
    if (ss->desc.nf != d.nf || ss->wcells != sh.wcells || ss->wyn != sh.wyn)
      return fail(OIDT_EINVAL, "the subtracted run has other port planes or frequencies");
  }
  const int rowlen = d.scheme == SCHEME_3D ? sh.wyn : 1;
  const int rows = rowlen > 0 ? sh.wcells / rowlen : 0;
  if (row0 < 0 || nrows < 0 || row0 + nrows > rows) return fail(OIDT_EINVAL, "port window outside the DFT window");
  if (d.nf == 0) return OIDT_OK;
  // Accumulators of finished runs are final; otherwise wait for the batch.
  if (!slot_finished(b, sh) || (ss && !slot_finished(b, *ss))) HIP_TRY(hipStreamSynchronize(b->stream));
  if (!b->aux) HIP_TRY(hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking));
  if (b->pow_cap < (size_t)d.nf) {
    if (b->d_pow) HIP_TRY(hipFree(b->d_pow));
    b->d_pow = nullptr;
    b->pow_cap = 0;
    HIP_TRY(hipMalloc(&b->d_pow, (size_t)d.nf * 2 * sizeof(double)));
    b->pow_cap = d.nf;
  }
  const size_t off = d.precision == OIDT_F64 ? layout<double>(d).off_acc : layout<float>(d).off_acc;
  const double* a = reinterpret_cast<const double*>(static_cast<char*>(sh.dmem) + off);
  const double* sa = nullptr;
  if (ss) {
    const size_t soff = ss->desc.precision == OIDT_F64 ? layout<double>(ss->desc).off_acc
                                                      : layout<float>(ss->desc).off_acc;
    sa = reinterpret_cast<const double*>(static_cast<char*>(ss->dmem) + soff);
  }
  const int ncomp = sh.nq / 2;
  hipLaunchKernelGGL(k_port_power, dim3(d.nf), dim3(kThreads), 0, b->aux, a, sa, d.scheme, port * ncomp,
                     ss ? sub_port * ncomp : 0, d.nf, sh.wcells, row0 * rowlen, (row0 + nrows) * rowlen, b->d_pow);
  HIP_TRY(hipGetLastError());
  std::vector<double> buf(2 * (size_t)d.nf);
  HIP_TRY(hipMemcpyAsync(buf.data(), b->d_pow, buf.size() * sizeof(double), hipMemcpyDeviceToHost, b->aux));
  HIP_TRY(hipStreamSynchronize(b->aux));
  for (int f = 0; f < d.nf; ++f) {
    re[f] = buf[2 * f];
    im[f] = buf[2 * f + 1];
  }
  return OIDT_OK;
}

int oidt_batch_port_power_noalloc(oidt_batch* b, int32_t slot, int32_t port, int32_t sub_slot, int32_t sub_port,
                                  int32_t row0, int32_t nrows, double* re, double* im) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  const SlotHost& sh = b->slots[slot];
  const oidt_run_desc& d = sh.desc;
  if (port < 0 || port > 1) return fail(OIDT_EINVAL, "bad port");
  if (!re || !im) return fail(OIDT_EINVAL, "null output");
  if (d.scheme == SCHEME_TMZ_LOSSY) return fail(OIDT_EINVAL, "the lossy TMz form has no port DFT");
  if (d.nf > kNoallocMaxNf)
    return fail(OIDT_EINVAL, "oidt_batch_port_power_noalloc takes at most " + std::to_string(kNoallocMaxNf) +
                                 " frequencies; the run has " + std::to_string(d.nf));
  const SlotHost* ss = nullptr;
  if (sub_slot >= 0) {
    if ((rc = slot_ok(b, sub_slot)) != OIDT_OK) return rc;
    ss = &b->slots[sub_slot];
    if (sub_port < 0 || sub_port > 1) return fail(OIDT_EINVAL, "bad port of the subtracted run");
    if (ss->desc.nf != d.nf || ss->wcells != sh.wcells || ss->wyn != sh.wyn)
      return fail(OIDT_EINVAL, "the subtracted run has other port planes or frequencies");
  }
  const int rowlen = d.scheme == SCHEME_3D ? sh.wyn : 1;
  const int rows = rowlen > 0 ? sh.wcells / rowlen : 0;
  if (row0 < 0 || nrows < 0 || row0 + nrows > rows) return fail(OIDT_EINVAL, "port window outside the DFT window");
  if (d.nf == 0) return OIDT_OK;
  double* out = nullptr;
  if ((rc = scratch_for(b->device, &out)) != OIDT_OK) return rc;
  const size_t off = d.precision == OIDT_F64 ? layout<double>(d).off_acc : layout<float>(d).off_acc;
  const double* a = reinterpret_cast<const double*>(static_cast<char*>(sh.dmem) + off);
  const double* sa = nullptr;
  if (ss) {
    const size_t soff = ss->desc.precision == OIDT_F64 ? layout<double>(ss->desc).off_acc
                                                      : layout<float>(ss->desc).off_acc;
    sa = reinterpret_cast<const double*>(static_cast<char*>(ss->dmem) + soff);
  }
  const int ncomp = sh.nq / 2;
  // On the batch stream, after the work enqueued there: the accumulators are
  // final once the runs' chunks have completed.  k_port_power writes every
  // (re, im) pair of frequencies [0, nf) (one block per frequency), so
  // nothing in `scratch` needs clearing first.
  hipLaunchKernelGGL(k_port_power, dim3(d.nf), dim3(kThreads), 0, b->stream, a, sa, d.scheme, port * ncomp,
                     ss ? sub_port * ncomp : 0, d.nf, sh.wcells, row0 * rowlen, (row0 + nrows) * rowlen, out);
  HIP_TRY(hipGetLastError());
  // split the interleaved (re, im) pairs straight into the caller's arrays
  HIP_TRY(hipMemcpy2DAsync(re, sizeof(double), out, 2 * sizeof(double), sizeof(double), d.nf,
                           hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipMemcpy2DAsync(im, sizeof(double), out + 1, 2 * sizeof(double), sizeof(double), d.nf,
                           hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return OIDT_OK;
}

int oidt_batch_read_field(oidt_batch* b, int32_t slot, int32_t field, void* host, uint64_t nbytes) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  const SlotHost& sh = b->slots[slot];
  const oidt_run_desc& d = sh.desc;
  if (field < 0 || field > 5 || !field_used(d.scheme, field)) return fail(OIDT_EINVAL, "field not part of this scheme");
  const size_t es = d.precision == OIDT_F64 ? 8 : 4;
  const size_t bytes = (size_t)d.nx * d.ny * d.nz * es;
  if (!host || nbytes != bytes) return fail(OIDT_EINVAL, "host buffer size mismatch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  reap_marks(b);
  // The fused and blocked steps alternate field sets (see SlotHost::tb).
  const int64_t steps = std::min<int64_t>(d.nt, std::max<int64_t>(0, b->chunks_done - sh.admit_chunk) * kChunk);
  const bool in_b = sh.tb > 0 && (((steps + sh.tb - 1) / sh.tb) & 1);
  const size_t off = d.precision == OIDT_F64 ? (in_b ? layout<double>(d).off_f2[field] : layout<double>(d).off_f[field])
                                             : (in_b ? layout<float>(d).off_f2[field] : layout<float>(d).off_f[field]);
  const bool is3d = d.scheme == SCHEME_3D;
  const size_t row = (is3d ? (size_t)d.nz : (size_t)d.ny) * es;
  const size_t prow = (is3d ? (size_t)(d.precision == OIDT_F64 ? padded<double>(d).nz : padded<float>(d).nz)
                            : (size_t)(d.precision == OIDT_F64 ? padded<double>(d).ny : padded<float>(d).ny)) * es;
  const size_t rows = bytes / row;
  HIP_TRY(hipMemcpy2D(host, row, static_cast<char*>(sh.dmem) + off, prow, row, rows, hipMemcpyDeviceToHost));
  return OIDT_OK;
}

int oidt_batch_remove(oidt_batch* b, int32_t slot) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  SlotHost& sh = b->slots[slot];
  // Deactivate in stream order, then free after all enqueued work on it.
  std::vector<unsigned char> zero(b->slot_bytes(), 0);
  rc = upload_bytes(b, static_cast<char*>(b->d_slots) + (size_t)slot * b->slot_bytes(), zero.data(), zero.size());
  if (rc != OIDT_OK) return rc;
  if (sh.dmem) HIP_TRY(hipFreeAsync(sh.dmem, b->stream));
  for (auto& mh : sh.mons)
    for (void* p : {mh.re, mh.im, static_cast<void*>(mh.cs)})
      if (p) HIP_TRY(hipFreeAsync(p, b->stream));
  sh = SlotHost();
  b->dirty = true;
  return OIDT_OK;
}

int oidt_batch_add_monitor(oidt_batch* b, int32_t slot, const oidt_monitor_desc* m, int32_t* monitor) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  if (!m || !monitor) return fail(OIDT_EINVAL, "null argument");
  SlotHost& sh = b->slots[slot];
  const oidt_run_desc& d = sh.desc;
  if (b->chunks_enqueued != sh.admit_chunk) return fail(OIDT_ESTATE, "monitors are attached before the run starts");
  if (b->variant >= 3) return fail(OIDT_ESTATE, "field monitors need a two-kernel form (variant 1 or 2)");
  if (m->field < 0 || m->field > 5 || !field_used(d.scheme, m->field))
    return fail(OIDT_EINVAL, "monitor field not present in this scheme");
  if ((m->after_h != 0 && m->after_h != 1) || (m->mode != OIDT_MON_DFT && m->mode != OIDT_MON_CAPTURE))
    return fail(OIDT_EINVAL, "monitor after_h must be 0 or 1 and mode OIDT_MON_DFT or OIDT_MON_CAPTURE");
  if (m->npts < 1 || m->npts > 4) return fail(OIDT_EINVAL, "monitor stencil takes 1 to 4 points");
  for (int q = 0; q < m->npts; ++q)
    for (int a = 0; a < 3; ++a)
      if (m->off[q][a] < -1 || m->off[q][a] > 1) return fail(OIDT_EINVAL, "monitor stencil offsets must be -1, 0 or 1");
  const int dims[3] = {d.nx, d.ny, d.nz};
  for (int a = 0; a < 3; ++a)
    if (m->box[2 * a] < 0 || m->box[2 * a] >= m->box[2 * a + 1] || m->box[2 * a + 1] > dims[a])
      return fail(OIDT_EINVAL, "monitor box empty or outside the grid");
  if (m->nrec < 1) return fail(OIDT_EINVAL, "monitor needs nrec >= 1");
  if (!std::isfinite(m->scale)) return fail(OIDT_EINVAL, "monitor scale not finite");
  if (m->mode == OIDT_MON_DFT && !m->cs) return fail(OIDT_EINVAL, "DFT monitor needs its cos/sin table");
  const size_t cells = (size_t)(m->box[1] - m->box[0]) * (m->box[3] - m->box[2]) * (m->box[5] - m->box[4]);
  if (cells > (size_t)1 << 30) return fail(OIDT_EINVAL, "monitor box larger than 2^30 cells");
  const size_t esz = b->precision == OIDT_F64 ? sizeof(double) : sizeof(float);
  MonHost mh;
  mh.d = *m;
  mh.d.cs = nullptr;
  mh.cells = cells;
  const size_t rbytes = (m->mode == OIDT_MON_DFT ? cells : cells * (size_t)m->nrec) * esz;
  auto undo = [&]() {
    for (void* p : {mh.re, mh.im, static_cast<void*>(mh.cs)})
      if (p) (void)hipFreeAsync(p, b->stream);
  };
  if (hipMallocAsync(&mh.re, rbytes, b->stream) != hipSuccess) {
    undo();
    return fail(OIDT_ENOMEM, "monitor allocation of " + std::to_string(rbytes) + " bytes failed");
  }
  if (m->mode == OIDT_MON_DFT) {
    if (hipMallocAsync(&mh.im, cells * esz, b->stream) != hipSuccess ||
        hipMallocAsync(reinterpret_cast<void**>(&mh.cs), (size_t)m->nrec * 2 * sizeof(double), b->stream) != hipSuccess) {
      undo();
      return fail(OIDT_ENOMEM, "monitor allocation failed");
    }
    if ((rc = upload_bytes(b, mh.cs, m->cs, (size_t)m->nrec * 2 * sizeof(double))) != OIDT_OK) {
      undo();
      return rc;
    }
    if (hipMemsetAsync(mh.im, 0, cells * esz, b->stream) != hipSuccess) {
      undo();
      return fail(OIDT_EHIP, "monitor clear failed");
    }
  }
  if (hipMemsetAsync(mh.re, 0, rbytes, b->stream) != hipSuccess) {
    undo();
    return fail(OIDT_EHIP, "monitor clear failed");
  }
  sh.mons.push_back(mh);
  b->dirty = true;
  *monitor = (int32_t)sh.mons.size() - 1;
  return OIDT_OK;
}

int oidt_batch_read_monitor(oidt_batch* b, int32_t slot, int32_t monitor, void* re, void* im, uint64_t nbytes) {
  int rc = slot_ok(b, slot);
  if (rc != OIDT_OK) return rc;
  const SlotHost& sh = b->slots[slot];
  if (monitor < 0 || monitor >= (int)sh.mons.size()) return fail(OIDT_EINVAL, "no such monitor");
  const MonHost& mh = sh.mons[monitor];
  const size_t esz = b->precision == OIDT_F64 ? sizeof(double) : sizeof(float);
  const bool dft = mh.d.mode == OIDT_MON_DFT;
  const size_t want = (dft ? mh.cells : mh.cells * (size_t)mh.d.nrec) * esz;
  if (nbytes != want) return fail(OIDT_EINVAL, "buffer size must be " + std::to_string(want) + " bytes");
  if (!re || (dft && !im)) return fail(OIDT_EINVAL, "null output");
  HIP_TRY(hipStreamSynchronize(b->stream));
  reap_marks(b);
  HIP_TRY(hipMemcpy(re, mh.re, want, hipMemcpyDeviceToHost));
  if (dft) HIP_TRY(hipMemcpy(im, mh.im, want, hipMemcpyDeviceToHost));
  return OIDT_OK;
}

int oidt_batch_profile_all(oidt_batch* b, int32_t nchunks, double* out_ms6) {
  if (!b || !out_ms6) return fail(OIDT_EINVAL, "null argument");
  if (nchunks < 1) return fail(OIDT_EINVAL, "need at least one chunk");
  int rc;
  if (b->dirty) {
    if ((rc = rebuild_lists(b)) != OIDT_OK) return rc;
    if ((rc = capture_graph(b)) != OIDT_OK) return rc;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  reap_marks(b);
  return b->precision == OIDT_F64 ? profile_chunks<double>(b, nchunks, out_ms6)
                                  : profile_chunks<float>(b, nchunks, out_ms6);
}

int oidt_batch_profile(oidt_batch* b, int32_t nchunks, double* out_ms4) {
  if (!out_ms4) return fail(OIDT_EINVAL, "null argument");
  double all[6];
  const int rc = oidt_batch_profile_all(b, nchunks, all);
  if (rc == OIDT_OK)
    for (int i = 0; i < 4; ++i) out_ms4[i] = all[i];
  return rc;
}

int oidt_copy_probe(uint64_t nbytes, int32_t reps, double* ms) {
  if (!ms || nbytes < 16 || reps < 1) return fail(OIDT_EINVAL, "bad copy probe arguments");
  const size_t n4 = nbytes / 16;
  float4* a = nullptr;
  float4* c = nullptr;
  HIP_TRY(hipMalloc(&a, n4 * 16));
  if (hipMalloc(&c, n4 * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(OIDT_ENOMEM, "copy probe allocation failed");
  }
  HIP_TRY(hipMemset(a, 0, n4 * 16));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const int blocks = (int)std::min<size_t>((n4 + 255) / 256, 65536);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, nullptr, a, c, n4);
  HIP_TRY(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, nullptr, a, c, n4);
  HIP_TRY(hipEventRecord(e1, nullptr));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(c);
  return OIDT_OK;
}

int oidt_solver_profile(oidt_solver* s, int32_t nchunks, double* out_ms4) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_profile(s->b, nchunks, out_ms4);
}

int oidt_solver_profile_all(oidt_solver* s, int32_t nchunks, double* out_ms6) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_profile_all(s->b, nchunks, out_ms6);
}

int oidt_batch_destroy(oidt_batch* b) {
  if (!b) return OIDT_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& sh : b->slots) {
    if (sh.dmem) (void)hipFree(sh.dmem);
    for (auto& mh : sh.mons)
      for (void* p : {mh.re, mh.im, static_cast<void*>(mh.cs)})
        if (p) (void)hipFree(p);
  }
  if (b->d_mons) (void)hipFree(b->d_mons);
  for (int h = 0; h < 2; ++h)
    if (b->d_mwork[h]) (void)hipFree(b->d_mwork[h]);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  if (b->d_work) (void)hipFree(b->d_work);
  if (b->d_counts) (void)hipFree(b->d_counts);
  if (b->d_slots) (void)hipFree(b->d_slots);
  for (auto& m : b->marks) (void)hipEventDestroy(m.first);
  for (auto e : b->free_events) (void)hipEventDestroy(e);
  b->staging.release();
  if (b->aux) (void)hipStreamSynchronize(b->aux);
  if (b->d_pow) (void)hipFree(b->d_pow);
  if (b->aux) (void)hipStreamDestroy(b->aux);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return OIDT_OK;
}

// ---- single run = batch of one -------------------------------------------

int oidt_solver_create(const oidt_run_desc* desc, const oidt_run_inputs* in, oidt_solver** out) {
  if (!out) return fail(OIDT_EINVAL, "null output");
  *out = nullptr;
  int rc = validate(desc);
  if (rc != OIDT_OK) return rc;
  oidt_solver* s = new (std::nothrow) oidt_solver();
  if (!s) return fail(OIDT_ENOMEM, "host allocation failed");
  if ((rc = oidt_batch_create(desc->scheme, desc->precision, 1, &s->b)) != OIDT_OK) {
    delete s;
    return rc;
  }
  if ((rc = oidt_batch_add(s->b, desc, in, &s->slot)) != OIDT_OK) {
    oidt_batch_destroy(s->b);
    delete s;
    return rc;
  }
  *out = s;
  return OIDT_OK;
}

int oidt_solver_set_launch(oidt_solver* s, int32_t xchunk, int32_t grid_blocks, int32_t variant) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_set_launch(s->b, xchunk, grid_blocks, variant);
}

int oidt_solver_advance(oidt_solver* s, int32_t nsteps) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  if (nsteps < 0) return fail(OIDT_EINVAL, "negative step count");
  const SlotHost& sh = s->b->slots[s->slot];
  const int64_t total = (sh.desc.nt + kChunk - 1) / kChunk;
  const int64_t done = s->b->chunks_enqueued - sh.admit_chunk;
  const int64_t want = std::min<int64_t>(total - done, (nsteps + kChunk - 1) / kChunk);
  return want > 0 ? oidt_batch_advance(s->b, (int32_t)want) : OIDT_OK;
}

int oidt_solver_sync(oidt_solver* s) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_sync(s->b);
}

int oidt_solver_steps(oidt_solver* s, int32_t* steps) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_poll(s->b, s->slot, steps);
}

int oidt_solver_read_dft(oidt_solver* s, int32_t port, int32_t comp, double* re, double* im) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_read_dft(s->b, s->slot, port, comp, re, im);
}

int oidt_solver_read_field(oidt_solver* s, int32_t field, void* host, uint64_t nbytes) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_read_field(s->b, s->slot, field, host, nbytes);
}

int oidt_solver_add_monitor(oidt_solver* s, const oidt_monitor_desc* m, int32_t* monitor) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_add_monitor(s->b, s->slot, m, monitor);
}

int oidt_solver_read_monitor(oidt_solver* s, int32_t monitor, void* re, void* im, uint64_t nbytes) {
  if (!s) return fail(OIDT_EINVAL, "null solver");
  return oidt_batch_read_monitor(s->b, s->slot, monitor, re, im, nbytes);
}

int oidt_solver_destroy(oidt_solver* s) {
  if (!s) return OIDT_OK;
  const int rc = oidt_batch_destroy(s->b);
  delete s;
  return rc;
}

}  // extern "C"
