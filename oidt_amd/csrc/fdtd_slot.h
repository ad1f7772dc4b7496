// Device-resident description of one run ("slot") and the launch constants
// shared by the kernels and the host library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oidt {

// Every pointer in a run descriptor points into device-global memory.  The
// device pass says so in the type (address space 1): loads and stores through
// them, and through the locals initialised from them, compile to global
// instructions rather than flat ones (which take 64-bit per-lane addresses
// and count on both the vector-memory and the LDS counter).  The host pass
// sees plain pointers; the layout is the same (8-byte pointers).
#if defined(__HIP_DEVICE_COMPILE__)
#define OIDT_GLOBAL __attribute__((address_space(1)))
#else
#define OIDT_GLOBAL
#endif

// Steps per DFT chunk: the stencil kernels record port-plane samples into a
// ring of this many steps; the DFT kernel folds a whole chunk into the FP64
// accumulators at once.
constexpr int kChunk = 32;
// Stencil block: 64 lanes along the fastest array axis (z in 3D, y in 2D)
// times 4 rows.
constexpr int kBX = 64;
constexpr int kBY = 4;
constexpr int kThreads = kBX * kBY;
// Frequencies per DFT work item.
constexpr int kFB = 32;
// Distinct cb values per E component (2-bit index per component in the
// material byte).
constexpr int kMatMax = 4;
// Vacuum permeability of the reference configurations (mu = 4e-7 pi).
constexpr double kMu0 = 4.0e-7 * 3.14159265358979323846;

// SCHEME_TMZ_LOSSY: the 2D TMz update form of the loss and wave-impedance
// study drivers (lossy E update, CPML over the whole border).
enum : int { SCHEME_TMZ = 0, SCHEME_TEZ = 1, SCHEME_3D = 2, SCHEME_TMZ_LOSSY = 3 };
enum : int { F_EX = 0, F_EY = 1, F_EZ = 2, F_HX = 3, F_HY = 4, F_HZ = 5 };

// psi arrays are addressed as psi[(component * 3 + axis) * 2 + side]
// (component 0..5 = Ex..Hz, axis 0..2 = x..z, side 0 = low, 1 = high).
// Layouts follow the reference allocation (fdtd_main.py:128-137, 3D :147-175):
//   x-slab [N][ny][nz], y-slab [nx][N][nz], z-slab [nx][ny][N].
__host__ __device__ constexpr int psi_id(int comp, int axis, int side) {
  return (comp * 3 + axis) * 2 + side;
}

template <typename T>
struct Slot {
  int nx, ny, nz, N;
  int ib, jb, kb, ie, je, ke;
  int nt, nf;
  int base;    // first step of the chunk in flight (advanced on the device)
  int active;
  int scheme;
  int src_field;  // -1: none
  int sx0, sx1, sy0, sy1, sz0, sz1;
  int port[2];
  int wx0, wx1, wy0, wy1;  // DFT window (half-open); wx unused in 2D
  int prow;                // raw capture plane row length (wy1 - wy0 + 1)
  int pcells;              // raw capture plane cells
  int wcells;              // window cells
  int nplanes;             // raw planes per step
  int nq;                  // DFT sample components
  // Byte distance from a component's low-side psi array to its high-side one
  // along the same axis (the same for every component of an axis): a kernel
  // picks the side of a cell by adding it to a 32-bit offset from the
  // low-side array.
  unsigned psi_hi[3];
  T db;
  T dxT;    // delta_x in the run precision (lossy TMz: curl_h * dx)
  T dtmu;   // delta_t / mu in the run precision (lossy TMz H CPML)
  double dt;
  OIDT_GLOBAL T* f[6];       // field set A (current state of the two-kernel forms)
  OIDT_GLOBAL T* psi[36];
  OIDT_GLOBAL T* f2[6];      // field set B: the fused step reads one set and writes the other
  OIDT_GLOBAL T* psi2[36];
  T cbv[3][kMatMax];    // distinct cb values of Ex, Ey, Ez
  OIDT_GLOBAL const uint8_t* mat;   // per cell: bits 2c..2c+1 index cbv[c]
  OIDT_GLOBAL const T* ike[3];      // 1 / (kappa_e * dx) per axis
  OIDT_GLOBAL const T* ikh[3];      // 1 / (kappa_h * dx) per axis
  OIDT_GLOBAL const T* be;
  OIDT_GLOBAL const T* ceo;         // ce / dx
  OIDT_GLOBAL const T* bh;
  OIDT_GLOBAL const T* cho;         // ch / dx
  OIDT_GLOBAL const T* jsrc;        // [nt]
  OIDT_GLOBAL T* ring;              // [kChunk][nplanes][pcells]
  OIDT_GLOBAL double* acc;          // [nq][nf][wcells][2]
  OIDT_GLOBAL const double* w;      // [nf] angular frequency 2*pi*f
};

// Field monitor (loss and impedance study drivers): the scaled stencil sum of
// one field component at every cell of a box, folded into a single-frequency
// running DFT (mode 0) or stored per step (mode 1).
template <typename T>
struct Monitor {
  OIDT_GLOBAL T* re;              // mode 0: real part [cells]; mode 1: records [nrec][cells]
  OIDT_GLOBAL T* im;              // mode 0: imaginary part [cells]
  OIDT_GLOBAL const double* cs;   // mode 0: (cos, sin) weight of each record [nrec][2]
  double scale;
  int slot, field, mode, npts;
  int ox[4], oy[4], oz[4];
  int x0, y0, z0, ex, ey, ez;   // box origin and extent
  int gx, gy, gz;               // real grid extents: stencil points outside read 0
  int shift, nrec;
};

}  // namespace oidt
