/*
 * CPU oracle for the Poisson image-editing solver cores.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path never calls it.
 *
 * Every function restates one loop of the reference solver cores
 * (Trinkle23897/Fast-Poisson-Image-Editing, fpie 0.3.2) in plain C, with the
 * same arithmetic types and the same summation order, so that the float32
 * state after n iterations is bit-identical to the reference's:
 *
 *   reference loop                                  oracle function
 *   fpie/core/gcc/equ.cc:14-32    partition         po_partition_rowmajor
 *   fpie/core/openmp/equ.cc:22-56 partition         po_partition_redblack
 *   fpie/core/gcc/equ.cc:41-54,74-80                po_equ_gs
 *   fpie/core/gcc/equ.cc:56-72    calc_error        po_equ_err_gcc
 *   fpie/core/openmp/equ.cc:67-80,107-118           po_equ_redblack
 *   fpie/core/openmp/equ.cc:82-105 calc_error       po_equ_err_openmp
 *   fpie/np_solver.py:31-41       Jacobi            po_equ_jacobi
 *   fpie/core/gcc/grid.cc:21-36,57-75               po_grid_gs
 *   fpie/core/gcc/grid.cc:38-55   calc_error        po_grid_err_gcc
 *   fpie/core/openmp/grid.cc:35-43,75-103 Jacobi    po_grid_jacobi_ulrd
 *   fpie/core/openmp/grid.cc:46-73 calc_error       po_grid_err_openmp
 *   fpie/np_solver.py:79-88       Jacobi            po_grid_jacobi_udlr
 *   fpie/core/gcc/equ.cc:82-84, grid.cc:76-80 clamp po_clamp_u8
 *
 * Two further functions are not restatements but CPU models of the GPU
 * schedules, used by the tests to show on the CPU that the schedules give the
 * same bits as the sequential loops above:
 *   po_grid_wavefront  - skewed two-colour schedule of lexicographic GS
 *   po_grid_colour     - red-black schedule on the grid layout
 *
 * Layouts follow the reference: Equ uses A int32[N][4] (ids of the up, down,
 * left, right neighbours), X and B float[N][3]; Grid uses mask int32[h][w],
 * tgt and grad float[h][w][3] (interleaved RGB).
 *
 * Build: gcc -O2 -std=c99 -ffp-contract=off -fopenmp -shared -fPIC (see
 * oracle/Makefile).  -ffp-contract=off keeps every add and multiply rounded
 * on its own, as the reference's x86-64 build does.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- ids --- */

/* fpie/core/gcc/equ.cc:14-32: ids 1.. in row-major order over mask > 0. */
int po_partition_rowmajor(const int32_t* mask, int h, int w, int32_t* ids) {
  int cnt = 0;
  for (int i = 0; i < h; ++i)
    for (int j = 0; j < w; ++j)
      ids[i * w + j] = mask[i * w + j] > 0 ? ++cnt : 0;
  return cnt;
}

/* fpie/core/openmp/equ.cc:22-56: (i+j) odd pixels first, then even;
 * returns n_mid = first id of the even colour. */
int po_partition_redblack(const int32_t* mask, int h, int w, int32_t* ids) {
  int cnt = 0;
  for (int i = 0; i < h; ++i)
    for (int j = 0; j < w; ++j)
      if ((i + j) % 2 == 1) ids[i * w + j] = mask[i * w + j] > 0 ? ++cnt : 0;
  int n_mid = cnt + 1;
  for (int i = 0; i < h; ++i)
    for (int j = 0; j < w; ++j)
      if ((i + j) % 2 == 0) ids[i * w + j] = mask[i * w + j] > 0 ? ++cnt : 0;
  return n_mid;
}

/* -------------------------------------------------------------- Equ ---- */

/* One equation update, fpie/core/gcc/equ.cc:41-54: the float sum
 * B + X[U] + X[D] + X[L] + X[R] left to right, then divided by the double
 * 4.0 and stored as float. */
static inline void equ_update(const int32_t* A, float* X, const float* B,
                              int i) {
  const int32_t* a = A + 4 * i;
  for (int c = 0; c < 3; ++c) {
    float s = B[3 * i + c];
    s = s + X[3 * a[0] + c];
    s = s + X[3 * a[1] + c];
    s = s + X[3 * a[2] + c];
    s = s + X[3 * a[3] + c];
    X[3 * i + c] = (float)((double)s / 4.0);
  }
}

/* fpie/core/gcc/equ.cc:74-80: in place, ids 1..N-1 in increasing order. */
void po_equ_gs(int n, const int32_t* A, float* X, const float* B, int iters) {
  for (int k = 0; k < iters; ++k)
    for (int i = 1; i < n; ++i) equ_update(A, X, B, i);
}

/* fpie/core/openmp/equ.cc:107-118: ids [1, n_mid) then [n_mid, N).  Each
 * phase only reads the other colour, so the loop order inside a phase does
 * not change the bits; the parallel loop gives the same result. */
void po_equ_redblack(int n, int n_mid, const int32_t* A, float* X,
                     const float* B, int iters, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  for (int k = 0; k < iters; ++k) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i < n_mid; ++i) equ_update(A, X, B, i);
#pragma omp parallel for schedule(static)
    for (int i = n_mid; i < n; ++i) equ_update(A, X, B, i);
  }
}

/* fpie/np_solver.py:31-41: X' = (B + X[A0] + X[A1] + X[A2] + X[A3]) / 4.0
 * over every row (row 0 included), all float32; tmp holds N*3 floats. */
void po_equ_jacobi(int n, const int32_t* A, float* X, const float* B,
                   int iters, float* tmp) {
  for (int k = 0; k < iters; ++k) {
    for (int i = 0; i < n; ++i) {
      const int32_t* a = A + 4 * i;
      for (int c = 0; c < 3; ++c) {
        float s = B[3 * i + c];
        s = s + X[3 * a[0] + c];
        s = s + X[3 * a[1] + c];
        s = s + X[3 * a[2] + c];
        s = s + X[3 * a[3] + c];
        tmp[3 * i + c] = s / 4.0f;
      }
    }
    memcpy(X, tmp, sizeof(float) * 3 * (size_t)n);
  }
}

/* The float sum of fpie/core/gcc/equ.cc:56-72 and the residual term in
 * double: |(B + X[U] + X[D] + X[L] + X[R]) - X * 4.0|. */
static inline double equ_term(const int32_t* A, const float* X,
                              const float* B, int i, int c) {
  const int32_t* a = A + 4 * i;
  float s = B[3 * i + c];
  s = s + X[3 * a[0] + c];
  s = s + X[3 * a[1] + c];
  s = s + X[3 * a[2] + c];
  s = s + X[3 * a[3] + c];
  return fabs((double)s - (double)X[3 * i + c] * 4.0);
}

/* fpie/core/gcc/equ.cc:56-72: double terms accumulated into a float. */
void po_equ_err_gcc(int n, const int32_t* A, const float* X, const float* B,
                    float* err) {
  err[0] = err[1] = err[2] = 0.0f;
  for (int i = 1; i < n; ++i)
    for (int c = 0; c < 3; ++c)
      err[c] = (float)((double)err[c] + equ_term(A, X, B, i, c));
}

/* fpie/core/openmp/equ.cc:82-105: each term rounded to float, then a
 * sequential float sum. */
void po_equ_err_openmp(int n, const int32_t* A, const float* X,
                       const float* B, float* err) {
  err[0] = err[1] = err[2] = 0.0f;
  for (int i = 1; i < n; ++i)
    for (int c = 0; c < 3; ++c) err[c] += (float)equ_term(A, X, B, i, c);
}

/* The exact (double) residual terms of every equation, for tests that check
 * a differently ordered sum against math.fsum: out double[N][3], row 0 = 0. */
void po_equ_terms(int n, const int32_t* A, const float* X, const float* B,
                  double* out) {
  out[0] = out[1] = out[2] = 0.0;
  for (int i = 1; i < n; ++i)
    for (int c = 0; c < 3; ++c) out[3 * i + c] = equ_term(A, X, B, i, c);
}

/* -------------------------------------------------------------- Grid --- */

/* fpie/core/gcc/grid.cc:21-36: grad + up + left + right + down (float),
 * divided by the double 4.0. */
static inline float grid_sum_ulrd(const float* t, const float* g, int off3,
                                  int m3, int c) {
  float s = g[off3 + c];
  s = s + t[off3 - m3 + c];
  s = s + t[off3 - 3 + c];
  s = s + t[off3 + 3 + c];
  s = s + t[off3 + m3 + c];
  return s;
}

/* fpie/np_solver.py:81-88: grad + up + down + left + right (float). */
static inline float grid_sum_udlr(const float* t, const float* g, int off3,
                                  int m3, int c) {
  float s = g[off3 + c];
  s = s + t[off3 - m3 + c];
  s = s + t[off3 + m3 + c];
  s = s + t[off3 - 3 + c];
  s = s + t[off3 + 3 + c];
  return s;
}

/* fpie/core/gcc/grid.cc:57-75: tiles of gx rows x gy columns in row-major
 * tile order, pixels row-major inside a tile, masked pixels updated in
 * place. */
void po_grid_gs(int h, int w, const int32_t* mask, float* tgt,
                const float* grad, int iters, int gx, int gy) {
  int m3 = 3 * w;
  int bx = (h + gx - 1) / gx, by = (w + gy - 1) / gy;
  for (int k = 0; k < iters; ++k)
    for (int b = 0; b < bx * by; ++b) {
      int x0 = b / by * gx, y0 = b % by * gy;
      for (int x = x0; x < x0 + gx && x < h; ++x)
        for (int y = y0; y < y0 + gy && y < w; ++y) {
          int id = x * w + y;
          if (!mask[id]) continue;
          for (int c = 0; c < 3; ++c)
            tgt[3 * id + c] =
                (float)((double)grid_sum_ulrd(tgt, grad, 3 * id, m3, c) / 4.0);
        }
    }
}

/* fpie/core/openmp/grid.cc:75-103 (update :35-43): Jacobi through a
 * scratch copy with the g+U+L+R+D order; tmp holds h*w*3 floats. */
void po_grid_jacobi_ulrd(int h, int w, const int32_t* mask, float* tgt,
                         const float* grad, int iters, float* tmp,
                         int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  int m3 = 3 * w;
  long hw = (long)h * w;
  for (int k = 0; k < iters; ++k) {
#pragma omp parallel for schedule(static)
    for (long id = 0; id < hw; ++id)
      if (mask[id])
        for (int c = 0; c < 3; ++c)
          tmp[3 * id + c] = (float)(
              (double)grid_sum_ulrd(tgt, grad, 3 * (int)id, m3, c) / 4.0);
#pragma omp parallel for schedule(static)
    for (long id = 0; id < hw; ++id)
      if (mask[id])
        for (int c = 0; c < 3; ++c) tgt[3 * id + c] = tmp[3 * id + c];
  }
}

/* fpie/np_solver.py:79-88: Jacobi with the g+U+D+L+R order, float32
 * division by 4.0. */
void po_grid_jacobi_udlr(int h, int w, const int32_t* mask, float* tgt,
                         const float* grad, int iters, float* tmp) {
  int m3 = 3 * w;
  long hw = (long)h * w;
  for (int k = 0; k < iters; ++k) {
    for (long id = 0; id < hw; ++id)
      if (mask[id])
        for (int c = 0; c < 3; ++c)
          tmp[3 * id + c] = grid_sum_udlr(tgt, grad, 3 * (int)id, m3, c) / 4.0f;
    for (long id = 0; id < hw; ++id)
      if (mask[id])
        for (int c = 0; c < 3; ++c) tgt[3 * id + c] = tmp[3 * id + c];
  }
}

static inline double grid_term(const float* t, const float* g, int id, int m3,
                               int c) {
  float s = grid_sum_ulrd(t, g, 3 * id, m3, c);
  return fabs((double)s - (double)t[3 * id + c] * 4.0);
}

/* fpie/core/gcc/grid.cc:38-55. */
void po_grid_err_gcc(int h, int w, const int32_t* mask, const float* tgt,
                     const float* grad, float* err) {
  int m3 = 3 * w;
  err[0] = err[1] = err[2] = 0.0f;
  for (int id = 0; id < h * w; ++id)
    if (mask[id])
      for (int c = 0; c < 3; ++c)
        err[c] = (float)((double)err[c] + grid_term(tgt, grad, id, m3, c));
}

/* fpie/core/openmp/grid.cc:46-73: float terms, sequential float sum. */
void po_grid_err_openmp(int h, int w, const int32_t* mask, const float* tgt,
                        const float* grad, float* err) {
  int m3 = 3 * w;
  err[0] = err[1] = err[2] = 0.0f;
  for (int id = 0; id < h * w; ++id)
    if (mask[id])
      for (int c = 0; c < 3; ++c) err[c] += (float)grid_term(tgt, grad, id, m3, c);
}

/* Exact double residual terms of the grid system (g+U+L+R+D float sum when
 * order == 1, g+U+D+L+R when order == 0); out double[h][w][3], 0 off mask. */
void po_grid_terms(int h, int w, const int32_t* mask, const float* tgt,
                   const float* grad, int order, double* out) {
  int m3 = 3 * w;
  for (int id = 0; id < h * w; ++id)
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      if (mask[id]) {
        float s = order ? grid_sum_ulrd(tgt, grad, 3 * id, m3, c)
                        : grid_sum_udlr(tgt, grad, 3 * id, m3, c);
        v = fabs((double)s - (double)tgt[3 * id + c] * 4.0);
      }
      out[3 * id + c] = v;
    }
}

/* fpie/core/gcc/equ.cc:82-84 and grid.cc:76-80: clamp to [0, 255] and
 * convert (truncate) to unsigned char. */
void po_clamp_u8(const float* x, long n, uint8_t* out) {
  for (long i = 0; i < n; ++i)
    out[i] = (uint8_t)(x[i] < 0 ? 0 : x[i] > 255 ? 255 : x[i]);
}

/* ------------------------------------------------ schedule models ---- */

/* Skewed two-colour schedule of lexicographic Gauss-Seidel on the grid
 * layout: step t updates, in place, every masked pixel (i, j) with
 * d = t - i - j even and 0 <= d < 2 * iters (sweep k = d / 2).  At step t
 * the up and left neighbours hold sweep k and the down and right
 * neighbours sweep k - 1, exactly as in the sequential row-major loop, and
 * one step only writes one colour, so the result equals po_grid_gs.
 * order: 1 = g+U+L+R+D (gcc grid), 0 = g+U+D+L+R (Equ in grid form). */
void po_grid_wavefront(int h, int w, const int32_t* mask, float* tgt,
                       const float* grad, int iters, int order) {
  int m3 = 3 * w;
  int steps = 2 * iters + (h - 1) + (w - 1) - 1;
  for (int t = 0; t < steps; ++t)
    for (int i = 0; i < h; ++i)
      for (int j = 0; j < w; ++j) {
        int id = i * w + j, d = t - i - j;
        if (!mask[id] || d < 0 || d >= 2 * iters || (d & 1)) continue;
        for (int c = 0; c < 3; ++c) {
          float s = order ? grid_sum_ulrd(tgt, grad, 3 * id, m3, c)
                          : grid_sum_udlr(tgt, grad, 3 * id, m3, c);
          tgt[3 * id + c] = s * 0.25f;
        }
      }
}

/* Red-black schedule on the grid layout: per iteration the (i+j)-odd
 * masked pixels, then the even ones, in place (the colour order of
 * fpie/core/openmp/equ.cc:109-118). */
void po_grid_colour(int h, int w, const int32_t* mask, float* tgt,
                    const float* grad, int iters, int order) {
  int m3 = 3 * w;
  for (int k = 0; k < iters; ++k)
    for (int colour = 1; colour >= 0; --colour)
      for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
          int id = i * w + j;
          if (!mask[id] || ((i + j) & 1) != colour) continue;
          for (int c = 0; c < 3; ++c) {
            float s = order ? grid_sum_ulrd(tgt, grad, 3 * id, m3, c)
                            : grid_sum_udlr(tgt, grad, 3 * id, m3, c);
            tgt[3 * id + c] = s * 0.25f;
          }
        }
}
