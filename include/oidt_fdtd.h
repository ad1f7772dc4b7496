/*
 * oidt_fdtd.h -- C ABI of the MI355X FDTD time-stepping library (liboidt_fdtd.so).
 *
 * The library replaces the device half of the reference driver
 * FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_main.py (and its TE / 3D
 * twins): the host->device copies of :216-261, the per-step kernel launches of
 * the time loop :275-304 (TE :278-307, 3D FDTD_3D/.../fdtd_main.py:357-406) and
 * the copy-back of the running-DFT accumulators :349-357.  Everything else --
 * parameter derivation, geometry, CPML grading, waveform, extraction, file
 * output -- stays in the Python host code, which hands the library exactly the
 * arrays the reference hands its Numba kernels (fdtd_funcs.py signatures).
 *
 * Conventions
 *   - All functions return 0 (OIDT_OK) on success or a negative oidt_status;
 *     oidt_last_error() then returns a thread-local message.
 *   - Arrays are C-order ([x][y] or [x][y][z], z fastest), in the run's
 *     precision (float or double) unless stated otherwise.
 *   - The library owns all device memory; host buffers are read or written only
 *     during the call that receives them.
 *   - A handle is used from one host thread at a time; distinct handles are
 *     independent.  No C++ exception crosses this boundary.
 */
#ifndef OIDT_FDTD_H
#define OIDT_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OIDT_ABI_VERSION 1

typedef enum {
  OIDT_OK = 0,
  OIDT_EINVAL = -1,   /* invalid argument or inconsistent description */
  OIDT_EHIP = -2,     /* HIP runtime error */
  OIDT_ENOMEM = -3,   /* device allocation failed */
  OIDT_ESTATE = -4,   /* call not valid in the handle's current state */
  OIDT_EFULL = -5     /* batch has no free slot */
} oidt_status;

/* Update schemes: FDTD_2D_TM, FDTD_2D_TE, FDTD_3D of the reference, and the
 * lossy 2D TMz form of the TM loss and wave-impedance study drivers
 * (FDTD_2D_TM/scattering_loss/fdtd_funcs.py:77-152): offset 1, num_cpml =
 * cpml_range, nf = 0, cb = {CURL_H, MOD_E} (at most 4 distinct values each),
 * kappa_e / kappa_h = {DEN_EX, DEN_EY} / {DEN_HX, DEN_HY} (no dx factor),
 * db = CURL_E; the H CPML scale is dt / mu with mu = 4e-7 pi; two-kernel forms
 * only. */
typedef enum { OIDT_TMZ = 0, OIDT_TEZ = 1, OIDT_3D = 2, OIDT_TMZ_LOSSY = 3 } oidt_scheme;

/* Precision of fields, psi and coefficients ("precision" of user_config.py:65).
 * DFT accumulators are always double. */
typedef enum { OIDT_F32 = 0, OIDT_F64 = 1 } oidt_precision;

/* Field components (index into the library's field table). */
typedef enum {
  OIDT_EX = 0, OIDT_EY = 1, OIDT_EZ = 2, OIDT_HX = 3, OIDT_HY = 4, OIDT_HZ = 5
} oidt_field;

/* One simulation run: the scalars of one `python fdtd_main.py <line> <f|b> <i|r>`
 * invocation, in grid units.  Update bounds follow fdtd_main.py:113-117:
 * ib = jb = kb = offset, ie = nx - offset, ... */
typedef struct oidt_run_desc {
  int32_t abi_version;  /* must equal OIDT_ABI_VERSION */
  int32_t scheme;       /* oidt_scheme */
  int32_t precision;    /* oidt_precision */
  int32_t nx, ny, nz;   /* nz = 1 for the 2D schemes */
  int32_t num_cpml;     /* CPML thickness N (user_config num_cpml) */
  int32_t offset;       /* PEC border, 2 in the reference */
  int32_t nt;           /* number of time steps */
  int32_t nf;           /* number of DFT frequencies (sparam_num_freqs), at most 65535 */
  int32_t src_field;    /* oidt_field driven by J_SRC: TMz Ez, TEz Hz, 3D Ey (TE) or Hy (TM) */
  int32_t src_box[6];   /* x0,x1,y0,y1,z0,z1: half-open box of update_source (z0=0,z1=1 in 2D) */
  int32_t port[2];      /* port planes: p1x,p2x (2D) or p1z,p2z (3D) */
  double dx;            /* delta_x, already rounded to the run precision */
  double dt;            /* delta_t, already rounded to the run precision */
  double db;            /* DB = delta_t / mu as passed to the H kernels */
} oidt_run_desc;

/* Arrays of one run, as the reference hands them to its kernels. */
typedef struct oidt_run_inputs {
  const void* cb[3];       /* E-component update maps dt/eps, [nx][ny][nz]:
                              TMz {CBZ}, TEz {CBX, CBY}, 3D {CBX, CBY, CBZ} */
  const void* kappa_e[3];  /* KXE, KYE, KZE: kappa*dx per axis (KZE NULL in 2D) */
  const void* kappa_h[3];  /* KXH, KYH, KZH */
  const void* be;          /* BE[num_cpml] */
  const void* ce;          /* CE[num_cpml] */
  const void* bh;          /* BH[num_cpml] */
  const void* ch;          /* CH[num_cpml] */
  const void* jsrc;        /* J_SRC[nt] */
  const double* freqs;     /* sparam_freqs[nf], Hz */
} oidt_run_inputs;

/* Field monitor of the loss and impedance study drivers: at every cell of a
 * box, the sum over 1-4 stencil points of one field component (points outside
 * the grid read 0), times `scale`, taken once per step right after the E or
 * the H half-step, and either
 *   mode OIDT_MON_DFT:     re += v * cs[r][0], im += v * cs[r][1]   (single-
 *                          frequency running DFT; the product and sum are
 *                          formed in double and rounded to the run precision,
 *                          as the reference's float32 accumulators are), or
 *   mode OIDT_MON_CAPTURE: out[r][cell] = v                          (time record),
 * where r = step + shift, and steps with r outside [0, nrec) are skipped.
 * Cells are ordered x-major, z fastest, over the box.
 * Replaces simul_fft (FDTD_2D_TM/scattering_loss/fdtd_funcs.py:138-144, TE
 * :113-120), simul_fft_efield / simul_fft_hfield (FDTD_3D/scattering_loss/
 * fdtd_funcs.py:217-242) and map_efield_zwave / map_hfield_zwave
 * (FDTD_2D_TM/wave_impedance/fdtd_funcs.py:155-168, FDTD_2D_TE/wave_impedance/
 * fdtd_funcs.py:169-181, FDTD_3D/wave_impedance/fdtd_funcs.py:321-335). */
typedef enum { OIDT_MON_DFT = 0, OIDT_MON_CAPTURE = 1 } oidt_monitor_mode;

typedef struct oidt_monitor_desc {
  int32_t field;        /* oidt_field */
  int32_t after_h;      /* 0: taken after the E half-step, 1: after the H half-step */
  int32_t mode;         /* oidt_monitor_mode */
  int32_t npts;         /* stencil points, 1..4 */
  int32_t off[4][3];    /* stencil offsets (x, y, z) of each point, each in -1..1 */
  double scale;         /* factor on the stencil sum: 1, 0.5, 0.25 in the reference */
  int32_t box[6];       /* x0,x1,y0,y1,z0,z1: half-open box of monitored cells */
  int32_t shift;        /* record index r = step + shift */
  int32_t nrec;         /* records kept, r in [0, nrec) */
  const double* cs;     /* OIDT_MON_DFT: [nrec][2] cos and sin weight of each record */
} oidt_monitor_desc;

typedef struct oidt_device_info {
  char name[64];
  char arch[32];
  int32_t compute_units;
  int32_t device_id;
  uint64_t hbm_total_bytes;
  uint64_t hbm_free_bytes;
} oidt_device_info;

typedef struct oidt_solver oidt_solver;
typedef struct oidt_batch oidt_batch;

/* ---- library ---------------------------------------------------------- */

/* ABI version compiled into the library. */
int32_t oidt_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* oidt_last_error(void);

/* Device 0 properties and free/total HBM (admission by footprint).
 * Replaces cuda.current_context().get_memory_info(), fdtd_main.py:315. */
int oidt_device_info_get(oidt_device_info* out);

/* Device bytes a run will hold while resident (fields, psi, material index,
 * coefficient tables, DFT ring and accumulators).  Replaces the estimate of
 * fdtd_main.py:207-213. */
int oidt_run_footprint(const oidt_run_desc* desc, uint64_t* bytes);

/* Samples per DFT chunk: steps advance in chunks of this many. */
int32_t oidt_chunk_steps(void);

/* ---- single run ---------------------------------------------------------- */

/* Allocate and upload one run (fields and psi zeroed).
 * Replaces fdtd_main.py:216-261 (cuda.to_device of every array). */
int oidt_solver_create(const oidt_run_desc* desc, const oidt_run_inputs* in, oidt_solver** out);

/* Launch knobs of a single run, as oidt_batch_set_launch; call before the
 * first oidt_solver_advance to change between fused and two-kernel forms.
 * No counterpart in the reference (fixed launch shapes, fdtd_main.py:263-273). */
int oidt_solver_set_launch(oidt_solver* s, int32_t xchunk, int32_t grid_blocks, int32_t variant);

/* Enqueue up to nsteps further steps (rounded up to whole chunks, never past
 * nt).  Asynchronous.  Replaces the body of the time loop fdtd_main.py:275-304. */
int oidt_solver_advance(oidt_solver* s, int32_t nsteps);

/* Wait for enqueued work. */
int oidt_solver_sync(oidt_solver* s);

/* Steps completed once the enqueued work has finished. */
int oidt_solver_steps(oidt_solver* s, int32_t* steps);

/* Copy the running DFT of one port component to the host: re/im are
 * [nf][cells] doubles over the reference window ([N, ny-N) in 2D,
 * [N, nx-N) x [N, ny-N) in 3D).  comp: 0 = E sample, 1 = H sample (2D);
 * 0 = Ex, 1 = Ey, 2 = Hx, 3 = Hy (3D).  port: 0 or 1.
 * Replaces the copy_to_host calls of fdtd_main.py:349-357. */
int oidt_solver_read_dft(oidt_solver* s, int32_t port, int32_t comp, double* re, double* im);

/* Copy one field component ([nx][ny][nz], run precision) to the host. */
int oidt_solver_read_field(oidt_solver* s, int32_t field, void* host, uint64_t nbytes);

/* Field monitors of a single run; as oidt_batch_add_monitor /
 * oidt_batch_read_monitor below. */
int oidt_solver_add_monitor(oidt_solver* s, const oidt_monitor_desc* m, int32_t* monitor);
int oidt_solver_read_monitor(oidt_solver* s, int32_t monitor, void* re, void* im, uint64_t nbytes);

/* Free the run's device memory.  Replaces cuda.close(), fdtd_main.py:388. */
int oidt_solver_destroy(oidt_solver* s);

/* ---- batch: many runs, one launch per half-step -------------------------- */
/* No counterpart in the reference (which runs one simulation per process,
 * simulation_launcher.py:26-35).  All runs of a batch share scheme and
 * precision; grid sizes, step counts and frequency counts may differ.  Runs can
 * be added and removed between advance calls (continuous batching). */

int oidt_batch_create(int32_t scheme, int32_t precision, int32_t max_slots, oidt_batch** out);

/* Upload a run into a free slot; it starts at step 0 with the next advance. */
int oidt_batch_add(oidt_batch* b, const oidt_run_desc* desc, const oidt_run_inputs* in, int32_t* slot);

/* Enqueue nchunks chunks for every resident unfinished run.  Asynchronous. */
int oidt_batch_advance(oidt_batch* b, int32_t nchunks);

/* Build the launch lists and capture the chunk graph of the batch's current
 * composition without enqueuing any step (the first advance after an add,
 * remove or launch change does this itself).  Lets a caller keep graph
 * capture and list uploads out of a timed or latency-critical region.
 * No counterpart in the reference. */
int oidt_batch_prepare(oidt_batch* b);

int oidt_batch_sync(oidt_batch* b);

/* Non-blocking progress: steps completed by the GPU for a slot (finished
 * when equal to the run's nt). */
int oidt_batch_poll(oidt_batch* b, int32_t slot, int32_t* steps_done);

/* Chunks enqueued so far and chunks the GPU has completed (non-blocking). */
int oidt_batch_progress(oidt_batch* b, int64_t* chunks_enqueued, int64_t* chunks_done);

/* DFT of one slot (layout as oidt_solver_read_dft).  A run whose last chunk
 * has completed is read without waiting for the rest of the batch. */
int oidt_batch_read_dft(oidt_batch* b, int32_t slot, int32_t port, int32_t comp, double* re, double* im);
int oidt_batch_read_field(oidt_batch* b, int32_t slot, int32_t field, void* host, uint64_t nbytes);

/* Port power on the device, per frequency: the sum over rows [row0, row0 +
 * nrows) of the DFT window of the time-averaged normal Poynting density of
 * port `port` of run `slot` (2D: one row = one cell along y; 3D: one row =
 * one x index with every y of the window), after subtracting the phasors of
 * port `sub_port` of run `sub_slot` (sub_slot = -1: nothing subtracted).
 * TMz -0.5 Ez Hy*, TEz 0.5 Ey Hz*, 3D 0.5 (Ex Hy* - Ey Hx*), in FP64.
 * Replaces the host sums of extract_sparams.py:88-104 (TM, TE) and
 * FDTD_3D/.../extract_sparams.py:114-131.  re/im receive nf values. */
int oidt_batch_port_power(oidt_batch* b, int32_t slot, int32_t port, int32_t sub_slot, int32_t sub_port,
                          int32_t row0, int32_t nrows, double* re, double* im);

/* Largest frequency count oidt_batch_port_power_noalloc takes. */
#define OIDT_NOALLOC_MAX_NF 16384

/* oidt_batch_port_power without any allocation: no hipMalloc, hipFree,
 * stream creation or host heap use on its success path.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  The
 * per-frequency sums pass through a fixed device array of the library
 * (2 x OIDT_NOALLOC_MAX_NF doubles = 256 KiB per device, held until the
 * process ends); runs with more frequencies are refused (OIDT_EINVAL).  The
 * reduction is enqueued on the batch's own stream, after the work already
 * enqueued there, and the call returns once re / im are written.  Every call
 * on one device shares that array: calls that overlap on one device (from
 * several host threads) race and must be serialised by the caller.
 * Arguments and results as oidt_batch_port_power (identical values). */
int oidt_batch_port_power_noalloc(oidt_batch* b, int32_t slot, int32_t port, int32_t sub_slot, int32_t sub_port,
                                  int32_t row0, int32_t nrows, double* re, double* im);

/* Attach a field monitor (oidt_monitor_desc) to a run that has not started;
 * *monitor receives its index within the run.  Monitors need a two-kernel
 * form (set_launch variant 1 or 2): the fused step has no point between the
 * half-steps (OIDT_ESTATE).  Replaces the monitor arrays' allocation and
 * cuda.to_device of FDTD_2D_TM/scattering_loss/fdtd_main.py:51-56 and their
 * launches in the time loop (:216-218; TE :222-223; 3D :341-344; wave
 * impedance 2D TM :210-211, 2D TE :220-221, 3D :298-299). */
int oidt_batch_add_monitor(oidt_batch* b, int32_t slot, const oidt_monitor_desc* m, int32_t* monitor);

/* Copy a monitor to the host (waits for the batch stream): OIDT_MON_DFT fills
 * re and im with [cells] values, OIDT_MON_CAPTURE fills re with [nrec][cells]
 * (im unused); run precision; nbytes is the size of each buffer.  Replaces
 * the copy_to_host calls of FDTD_2D_TM/scattering_loss/fdtd_main.py:260-262
 * (3D :386-409; wave impedance 2D TM :253-254, 3D :342-346). */
int oidt_batch_read_monitor(oidt_batch* b, int32_t slot, int32_t monitor, void* re, void* im, uint64_t nbytes);

/* Release a slot (its device memory is returned in stream order). */
int oidt_batch_remove(oidt_batch* b, int32_t slot);

/* Launch knobs for the autotuner: x-chunk depth of the stencil tiles (cells per
 * thread along x: 1, 2, 4, 8, 16 or 32; 0 keeps the current value; default
 * 4), the persistent grid size (0 = automatic) and the stencil kernel form
 * (0 keeps the current one; 1 = per-plane march, the 3D default, which in 3D
 * issues the next plane's loads before updating the current one; 2 = batched
 * loads per x-chunk, the 2D default; 3 = fused Yee step, one launch per step,
 * all schemes, xchunk also 12, 20 or 24 in 2D, 1, 2, 4, 8 or 16 in 3D; 4 = temporally
 * blocked step, TMz, TEz and 3D, xchunk = steps per launch, 2, 4 or 8 in 2D,
 * 2 in 3D).  Once a run
 * has started, its form cannot change between two-kernel, fused and blocked,
 * nor the blocked step's depth (OIDT_ESTATE).
 * No counterpart in the reference (fixed launch shapes, fdtd_main.py:263-273). */
int oidt_batch_set_launch(oidt_batch* b, int32_t xchunk, int32_t grid_blocks, int32_t variant);

/* Current launch knobs (any pointer may be NULL). */
int oidt_batch_get_launch(const oidt_batch* b, int32_t* xchunk, int32_t* grid_blocks, int32_t* variant);

int oidt_batch_destroy(oidt_batch* b);

/* ---- measurement -------------------------------------------------------- */
/* No counterpart in the reference, which times the whole loop with time()
 * (fdtd_main.py:271-341).  These run real work and advance the runs. */

/* Enqueue nchunks chunks as individual launches with HIP event pairs on the
 * batch stream, wait, and return the mean duration (ms) per launch of
 * out_ms4[0] E half-step (the whole step in the fused and blocked forms),
 * [1] H half-step, [2] DFT fold, [3] chunk advance.  Two-kernel forms get an
 * event pair around every kernel; the fused and blocked forms one pair around
 * each chunk's back-to-back stencil launches. */
int oidt_batch_profile(oidt_batch* b, int32_t nchunks, double* out_ms4);
int oidt_solver_profile(oidt_solver* s, int32_t nchunks, double* out_ms4);

/* As oidt_batch_profile, plus out_ms6[4] / [5]: mean duration of the field
 * monitor launch after the E / H half-step (0 when there is none). */
int oidt_batch_profile_all(oidt_batch* b, int32_t nchunks, double* out_ms6);
int oidt_solver_profile_all(oidt_solver* s, int32_t nchunks, double* out_ms6);

/* Mean duration (ms) of a plain device copy kernel moving nbytes (read nbytes,
 * write nbytes), over reps launches: the reference point for HBM counters. */
int oidt_copy_probe(uint64_t nbytes, int32_t reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* OIDT_FDTD_H */
