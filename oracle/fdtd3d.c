/*
 * oracle/fdtd3d.c -- FP64 C restatement of the reference 3D FDTD time step.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle): it is the checker for the
 * HIP path at sizes the NumPy oracle cannot reach in seconds, and the timed
 * CPU baseline.  The product library never links it.
 *
 * It restates, statement group by statement group and in the reference launch
 * order (FDTD/GPU/FDTD_3D/s-parameter_extraction/fdtd_main.py:357-406), the
 * kernels of FDTD/GPU/FDTD_3D/s-parameter_extraction/fdtd_funcs.py:
 *   update_ex / update_ey / update_ez          :44-70
 *   update_source                              :104-110 (called :364-379)
 *   the six E CPML kernels (ex_y .. ez_y)      :114-186
 *   update_hx / update_hy / update_hz          :74-100
 *   the six H CPML kernels (hx_y .. hz_y)      :190-266
 *   simul_fft_efield / simul_fft_hfield        :270-292
 * with the same expressions, operand order and update ranges as the NumPy
 * oracle (oracle/fdtd.py, class Yee3D), which the physics pins check.  Built
 * with -ffp-contract=off so no multiply-add is fused: every operation rounds
 * as the NumPy oracle's does, and the two agree to the last bits (tested).
 *
 * Arrays are C-order [x][y][z] (z fastest), FP64.  psi arrays: x slabs
 * [N][ny][nz], y slabs [nx][N][nz], z slabs [nx][ny][N], low side then high
 * side.  DFT accumulators: complex (re, im interleaved) [nf][nx][ny] per
 * (component, port), components ex, ey, hx, hy, ports 1, 2.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

typedef struct o3d_state {
  int32_t nx, ny, nz, N, ib;    /* grid, CPML thickness, PEC offset (ib = jb = kb) */
  int32_t nf, src_is_hy;        /* frequencies; source on Hy (TM) instead of Ey (TE) */
  int32_t src[6];               /* x0,x1,y0,y1,z0,z1 (already clipped to >= 0) */
  int32_t port[2];              /* p1z, p2z */
  double dx, db, dt;
  const double *cbx, *cby, *cbz;                   /* [nx][ny][nz] */
  const double *kxe, *kye, *kze, *kxh, *kyh, *kzh; /* kappa * dx per axis */
  const double *be, *ce, *bh, *ch;                 /* [N] */
  const double *J;                                 /* [nt] */
  const double *w;                                 /* [nf] angular frequencies */
  double *f[6];                                    /* Ex Ey Ez Hx Hy Hz */
  double *psi[12][2];  /* exy exz eyx eyz ezx ezy hxy hxz hyx hyz hzx hzy, [low, high] */
  double *acc[4][2];   /* ex ey hx hy x port 1, 2 */
} o3d_state;

#define IX(i, j, k) (((size_t)(i) * ny + (j)) * nz + (k))

static void e_half(o3d_state* s, int n) {
  const int nx = s->nx, ny = s->ny, nz = s->nz, N = s->N;
  const int ib = s->ib, jb = s->ib, kb = s->ib, ie = nx - s->ib, je = ny - s->ib, ke = nz - s->ib;
  double *Ex = s->f[0], *Ey = s->f[1], *Ez = s->f[2];
  const double *Hx = s->f[3], *Hy = s->f[4], *Hz = s->f[5];
  const double dx = s->dx;
  /* update_ex: i (ib,ie), j [jb,je), k [kb,ke) */
#pragma omp parallel for schedule(static)
  for (int i = ib + 1; i < ie; ++i)
    for (int j = jb; j < je; ++j)
      for (int k = kb; k < ke; ++k) {
        const size_t c = IX(i, j, k);
        Ex[c] += s->cbx[c] * ((Hz[IX(i, j + 1, k)] - Hz[c]) / s->kye[j] - (Hy[IX(i, j, k + 1)] - Hy[c]) / s->kze[k]);
      }
  /* update_ey: i [ib,ie), j (jb,je), k [kb,ke) */
#pragma omp parallel for schedule(static)
  for (int i = ib; i < ie; ++i)
    for (int j = jb + 1; j < je; ++j)
      for (int k = kb; k < ke; ++k) {
        const size_t c = IX(i, j, k);
        Ey[c] += s->cby[c] * ((Hx[IX(i, j, k + 1)] - Hx[c]) / s->kze[k] - (Hz[IX(i + 1, j, k)] - Hz[c]) / s->kxe[i]);
      }
  /* update_ez: i [ib,ie), j [jb,je), k (kb,ke) */
#pragma omp parallel for schedule(static)
  for (int i = ib; i < ie; ++i)
    for (int j = jb; j < je; ++j)
      for (int k = kb + 1; k < ke; ++k) {
        const size_t c = IX(i, j, k);
        Ez[c] += s->cbz[c] * ((Hy[IX(i + 1, j, k)] - Hy[c]) / s->kxe[i] - (Hx[IX(i, j + 1, k)] - Hx[c]) / s->kye[j]);
      }
  /* update_source on Ey (TE) or Hy (TM), right after the E updates */
  {
    double* F = s->src_is_hy ? s->f[4] : s->f[1];
    const double Jn = s->J[n];
    for (int i = s->src[0]; i < s->src[1]; ++i)
      for (int j = s->src[2]; j < s->src[3]; ++j)
        for (int k = s->src[4]; k < s->src[5]; ++k) F[IX(i, j, k)] -= Jn;
  }
  /* E CPML, each kernel low side (cells b + d, grading index N-1-d) then
     high side (cells e - N + d, grading index d) */
  for (int side = 0; side < 2; ++side) { /* ex_y (+): i (ib,ie), k [kb,ke); psi [nx][N][nz] */
    double* p = s->psi[0][side];
#pragma omp parallel for schedule(static)
    for (int i = ib + 1; i < ie; ++i)
      for (int d = 0; d < N; ++d) {
        const int g = side ? d : N - 1 - d, j = side ? je - N + d : jb + d;
        for (int k = kb; k < ke; ++k) {
          double* q = p + ((size_t)i * N + d) * nz + k;
          *q = s->be[g] * *q + s->ce[g] * (Hz[IX(i, j + 1, k)] - Hz[IX(i, j, k)]) / dx;
          Ex[IX(i, j, k)] += s->cbx[IX(i, j, k)] * *q;
        }
      }
  }
  for (int side = 0; side < 2; ++side) { /* ex_z (-): i (ib,ie), j [jb,je); psi [nx][ny][N] */
    double* p = s->psi[1][side];
#pragma omp parallel for schedule(static)
    for (int i = ib + 1; i < ie; ++i)
      for (int j = jb; j < je; ++j)
        for (int d = 0; d < N; ++d) {
          const int g = side ? d : N - 1 - d, k = side ? ke - N + d : kb + d;
          double* q = p + ((size_t)i * ny + j) * N + d;
          *q = s->be[g] * *q + s->ce[g] * (Hy[IX(i, j, k + 1)] - Hy[IX(i, j, k)]) / dx;
          Ex[IX(i, j, k)] -= s->cbx[IX(i, j, k)] * *q;
        }
  }
  for (int side = 0; side < 2; ++side) { /* ey_x (-): j (jb,je), k [kb,ke); psi [N][ny][nz] */
    double* p = s->psi[2][side];
#pragma omp parallel for schedule(static)
    for (int d = 0; d < N; ++d) {
      const int g = side ? d : N - 1 - d, i = side ? ie - N + d : ib + d;
      for (int j = jb + 1; j < je; ++j)
        for (int k = kb; k < ke; ++k) {
          double* q = p + ((size_t)d * ny + j) * nz + k;
          *q = s->be[g] * *q + s->ce[g] * (Hz[IX(i + 1, j, k)] - Hz[IX(i, j, k)]) / dx;
          Ey[IX(i, j, k)] -= s->cby[IX(i, j, k)] * *q;
        }
    }
  }
  for (int side = 0; side < 2; ++side) { /* ey_z (+): i [ib,ie), j (jb,je); psi [nx][ny][N] */
    double* p = s->psi[3][side];
#pragma omp parallel for schedule(static)
    for (int i = ib; i < ie; ++i)
      for (int j = jb + 1; j < je; ++j)
        for (int d = 0; d < N; ++d) {
          const int g = side ? d : N - 1 - d, k = side ? ke - N + d : kb + d;
          double* q = p + ((size_t)i * ny + j) * N + d;
          *q = s->be[g] * *q + s->ce[g] * (Hx[IX(i, j, k + 1)] - Hx[IX(i, j, k)]) / dx;
          Ey[IX(i, j, k)] += s->cby[IX(i, j, k)] * *q;
        }
  }
  for (int side = 0; side < 2; ++side) { /* ez_x (+): j [jb,je), k (kb,ke); psi [N][ny][nz] */
    double* p = s->psi[4][side];
#pragma omp parallel for schedule(static)
    for (int d = 0; d < N; ++d) {
      const int g = side ? d : N - 1 - d, i = side ? ie - N + d : ib + d;
      for (int j = jb; j < je; ++j)
        for (int k = kb + 1; k < ke; ++k) {
          double* q = p + ((size_t)d * ny + j) * nz + k;
          *q = s->be[g] * *q + s->ce[g] * (Hy[IX(i + 1, j, k)] - Hy[IX(i, j, k)]) / dx;
          Ez[IX(i, j, k)] += s->cbz[IX(i, j, k)] * *q;
        }
    }
  }
  for (int side = 0; side < 2; ++side) { /* ez_y (-): i [ib,ie), k (kb,ke); psi [nx][N][nz] */
    double* p = s->psi[5][side];
#pragma omp parallel for schedule(static)
    for (int i = ib; i < ie; ++i)
      for (int d = 0; d < N; ++d) {
        const int g = side ? d : N - 1 - d, j = side ? je - N + d : jb + d;
        for (int k = kb + 1; k < ke; ++k) {
          double* q = p + ((size_t)i * N + d) * nz + k;
          *q = s->be[g] * *q + s->ce[g] * (Hx[IX(i, j + 1, k)] - Hx[IX(i, j, k)]) / dx;
          Ez[IX(i, j, k)] -= s->cbz[IX(i, j, k)] * *q;
        }
      }
  }
}

static void h_half(o3d_state* s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz, N = s->N;
  const int ib = s->ib, jb = s->ib, kb = s->ib, ie = nx - s->ib, je = ny - s->ib, ke = nz - s->ib;
  const double *Ex = s->f[0], *Ey = s->f[1], *Ez = s->f[2];
  double *Hx = s->f[3], *Hy = s->f[4], *Hz = s->f[5];
  const double dx = s->dx, db = s->db;
  /* update_hx: i [ib,ie), j (jb,je), k (kb,ke) */
#pragma omp parallel for schedule(static)
  for (int i = ib; i < ie; ++i)
    for (int j = jb + 1; j < je; ++j)
      for (int k = kb + 1; k < ke; ++k) {
        const size_t c = IX(i, j, k);
        Hx[c] += db * ((Ey[c] - Ey[IX(i, j, k - 1)]) / s->kzh[k] - (Ez[c] - Ez[IX(i, j - 1, k)]) / s->kyh[j]);
      }
  /* update_hy: i (ib,ie), j [jb,je), k (kb,ke) */
#pragma omp parallel for schedule(static)
  for (int i = ib + 1; i < ie; ++i)
    for (int j = jb; j < je; ++j)
      for (int k = kb + 1; k < ke; ++k) {
        const size_t c = IX(i, j, k);
        Hy[c] += db * ((Ez[c] - Ez[IX(i - 1, j, k)]) / s->kxh[i] - (Ex[c] - Ex[IX(i, j, k - 1)]) / s->kzh[k]);
      }
  /* update_hz: i (ib,ie), j (jb,je), k [kb,ke) */
#pragma omp parallel for schedule(static)
  for (int i = ib + 1; i < ie; ++i)
    for (int j = jb + 1; j < je; ++j)
      for (int k = kb; k < ke; ++k) {
        const size_t c = IX(i, j, k);
        Hz[c] += db * ((Ex[c] - Ex[IX(i, j - 1, k)]) / s->kyh[j] - (Ey[c] - Ey[IX(i - 1, j, k)]) / s->kxh[i]);
      }
  /* H CPML, d = 1..N-1: low side cells b + d with grading index N - d, high
     side cells e - N + d with grading index d */
  for (int side = 0; side < 2; ++side) { /* hx_y (-): i [ib,ie), k (kb,ke); psi [nx][N][nz] */
    double* p = s->psi[6][side];
#pragma omp parallel for schedule(static)
    for (int i = ib; i < ie; ++i)
      for (int d = 1; d < N; ++d) {
        const int g = side ? d : N - d, j = side ? je - N + d : jb + d;
        for (int k = kb + 1; k < ke; ++k) {
          double* q = p + ((size_t)i * N + d) * nz + k;
          *q = s->bh[g] * *q + s->ch[g] * (Ez[IX(i, j, k)] - Ez[IX(i, j - 1, k)]) / dx;
          Hx[IX(i, j, k)] -= db * *q;
        }
      }
  }
  for (int side = 0; side < 2; ++side) { /* hx_z (+): i [ib,ie), j (jb,je); psi [nx][ny][N] */
    double* p = s->psi[7][side];
#pragma omp parallel for schedule(static)
    for (int i = ib; i < ie; ++i)
      for (int j = jb + 1; j < je; ++j)
        for (int d = 1; d < N; ++d) {
          const int g = side ? d : N - d, k = side ? ke - N + d : kb + d;
          double* q = p + ((size_t)i * ny + j) * N + d;
          *q = s->bh[g] * *q + s->ch[g] * (Ey[IX(i, j, k)] - Ey[IX(i, j, k - 1)]) / dx;
          Hx[IX(i, j, k)] += db * *q;
        }
  }
  for (int side = 0; side < 2; ++side) { /* hy_x (+): j [jb,je), k (kb,ke); psi [N][ny][nz] */
    double* p = s->psi[8][side];
#pragma omp parallel for schedule(static)
    for (int d = 1; d < N; ++d) {
      const int g = side ? d : N - d, i = side ? ie - N + d : ib + d;
      for (int j = jb; j < je; ++j)
        for (int k = kb + 1; k < ke; ++k) {
          double* q = p + ((size_t)d * ny + j) * nz + k;
          *q = s->bh[g] * *q + s->ch[g] * (Ez[IX(i, j, k)] - Ez[IX(i - 1, j, k)]) / dx;
          Hy[IX(i, j, k)] += db * *q;
        }
    }
  }
  for (int side = 0; side < 2; ++side) { /* hy_z (-): i (ib,ie), j [jb,je); psi [nx][ny][N] */
    double* p = s->psi[9][side];
#pragma omp parallel for schedule(static)
    for (int i = ib + 1; i < ie; ++i)
      for (int j = jb; j < je; ++j)
        for (int d = 1; d < N; ++d) {
          const int g = side ? d : N - d, k = side ? ke - N + d : kb + d;
          double* q = p + ((size_t)i * ny + j) * N + d;
          *q = s->bh[g] * *q + s->ch[g] * (Ex[IX(i, j, k)] - Ex[IX(i, j, k - 1)]) / dx;
          Hy[IX(i, j, k)] -= db * *q;
        }
  }
  for (int side = 0; side < 2; ++side) { /* hz_x (-): j (jb,je), k [kb,ke); psi [N][ny][nz] */
    double* p = s->psi[10][side];
#pragma omp parallel for schedule(static)
    for (int d = 1; d < N; ++d) {
      const int g = side ? d : N - d, i = side ? ie - N + d : ib + d;
      for (int j = jb + 1; j < je; ++j)
        for (int k = kb; k < ke; ++k) {
          double* q = p + ((size_t)d * ny + j) * nz + k;
          *q = s->bh[g] * *q + s->ch[g] * (Ey[IX(i, j, k)] - Ey[IX(i - 1, j, k)]) / dx;
          Hz[IX(i, j, k)] -= db * *q;
        }
    }
  }
  for (int side = 0; side < 2; ++side) { /* hz_y (+): i (ib,ie), k [kb,ke); psi [nx][N][nz] */
    double* p = s->psi[11][side];
#pragma omp parallel for schedule(static)
    for (int i = ib + 1; i < ie; ++i)
      for (int d = 1; d < N; ++d) {
        const int g = side ? d : N - d, j = side ? je - N + d : jb + d;
        for (int k = kb; k < ke; ++k) {
          double* q = p + ((size_t)i * N + d) * nz + k;
          *q = s->bh[g] * *q + s->ch[g] * (Ex[IX(i, j, k)] - Ex[IX(i, j - 1, k)]) / dx;
          Hz[IX(i, j, k)] += db * *q;
        }
      }
  }
}

/* Port DFT at the two planes k = p1z, p2z over i in [N, nx-N), j in [N, ny-N):
   Ex averaged over i, Ey over j, Hx over (j, k), Hy over (i, k); E samples at
   phase w n dt, H samples at w (n + 1/2) dt; re += v cos, im -= v sin. */
static void dft(o3d_state* s, int n) {
  const int nx = s->nx, ny = s->ny, nz = s->nz, N = s->N, nf = s->nf;
  const double *Ex = s->f[0], *Ey = s->f[1], *Hx = s->f[3], *Hy = s->f[4];
#pragma omp parallel for schedule(static)
  for (int f = 0; f < nf; ++f) {
    const double wn = s->w[f] * (double)n, wh = s->w[f] * ((double)n + 0.5);
    const double cE = cos(wn * s->dt), sE = sin(wn * s->dt), cH = cos(wh * s->dt), sH = sin(wh * s->dt);
    for (int p = 0; p < 2; ++p) {
      const int kc = s->port[p];
      double *aex = s->acc[0][p], *aey = s->acc[1][p], *ahx = s->acc[2][p], *ahy = s->acc[3][p];
      for (int i = N; i < nx - N; ++i)
        for (int j = N; j < ny - N; ++j) {
          const double sx = 0.5 * (Ex[IX(i, j, kc)] + Ex[IX(i + 1, j, kc)]);
          const double sy = 0.5 * (Ey[IX(i, j, kc)] + Ey[IX(i, j + 1, kc)]);
          const double hx = 0.25 * (Hx[IX(i, j, kc)] + Hx[IX(i, j + 1, kc)] + Hx[IX(i, j, kc + 1)] +
                                    Hx[IX(i, j + 1, kc + 1)]);
          const double hy = 0.25 * (Hy[IX(i, j, kc)] + Hy[IX(i + 1, j, kc)] + Hy[IX(i, j, kc + 1)] +
                                    Hy[IX(i + 1, j, kc + 1)]);
          const size_t a = (((size_t)f * nx + i) * ny + j) * 2;
          aex[a] += sx * cE;
          aex[a + 1] += sx * -sE;
          aey[a] += sy * cE;
          aey[a + 1] += sy * -sE;
          ahx[a] += hx * cH;
          ahx[a + 1] += hx * -sH;
          ahy[a] += hy * cH;
          ahy[a + 1] += hy * -sH;
        }
    }
  }
}

/* Advance steps n0 .. n0 + nsteps - 1.  Returns 0. */
int o3d_run(o3d_state* s, int32_t n0, int32_t nsteps) {
  for (int n = n0; n < n0 + nsteps; ++n) {
    e_half(s, n);
    h_half(s);
    dft(s, n);
  }
  return 0;
}

int o3d_state_size(void) { return (int)sizeof(o3d_state); }
