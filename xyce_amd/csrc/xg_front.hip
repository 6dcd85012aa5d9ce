// Multifrontal LU of the components too large for the task lists of
// xg_lu.hip (the 20k-node random logic, the 550k-unknown power grid): the
// refactor-and-solve that replaces N_LAS_AmesosSolver::solve (KLU,
// N_LAS_AmesosSolver.C:231-470) for them.
//
// The plan (xg_lu_setup.cpp) gives every component a tree of dense fronts:
// front s owns k consecutive pivot positions (its fully summed columns)
// and has m >= k rows; its m x m block is row-major in one value buffer.
// Per Newton iteration:
//  1. every front is cleared and the Jacobian's entries are scattered into
//     the fronts that own them;
//  2. level by level from the leaves (a level's fronts are independent),
//     one workgroup per front adds its children's contribution blocks (in
//     child order), then factors its k fully summed columns: blocked
//     right-looking LU, pivots chosen by magnitude among the fully summed
//     rows only, panels in LDS, the trailing update leaving the Schur
//     complement (the contribution block) in the bottom-right corner;
//  3. the forward solve runs the same levels: a front gathers its
//     right-hand side, adds its children's remainders, applies its row
//     swaps, solves with its unit lower block and leaves the remainder for
//     its parent; the backward solve walks the levels back down;
//  4. one workgroup per component checks the normwise backward error
//     against the unfactored rows, applies x += dx and forms the weighted
//     RMS norm, as the task-list kernel does.
// Every sum is in a fixed order, so the results do not depend on the
// launch.  A pivot column with no nonzero candidate, or a backward error
// above the bound, hands the component to the dense partial-pivoting
// kernel (as the task lists do); beyond that kernel's size it stays
// singular (stats[2] = 1) with its iterate unchanged.
#include <algorithm>
#include <cfloat>

#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kFB = 256;           // threads per front workgroup
constexpr int kFinishBlock = 1024;  // threads of the per-component check
constexpr int kSolveBlock = 512;    // threads of a solve workgroup for fronts above 512 rows

// (value, smallest index) maximum over the block: per wavefront, then
// across the B / 64 wavefronts.
template <int B>
__device__ __forceinline__ void blk_argmax(double v, int idx, double *sv, int *si,
                                           double &best, int &bi) {
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(idx, off);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    sv[w] = v;
    si[w] = idx;
  }
  __syncthreads();
  best = sv[0];
  bi = si[0];
  for (int q = 1; q < B / 64; ++q)
    if (sv[q] > best || (sv[q] == best && si[q] < bi)) {
      best = sv[q];
      bi = si[q];
    }
  __syncthreads();
}

template <int B>
__device__ __forceinline__ double blk_reduce(double v, double *sv, bool is_max) {
  const int t = threadIdx.x;
  sv[t] = v;
  __syncthreads();
  for (int s = B / 2; s > 0; s >>= 1) {
    if (t < s) sv[t] = is_max ? fmax(sv[t], sv[t + s]) : sv[t] + sv[t + s];
    __syncthreads();
  }
  const double r = sv[0];
  __syncthreads();
  return r;
}

// Partial LU of the m x m row-major front A: its first k columns, pivots
// among rows [j, k) for column j.  Panels of NB columns factor in the LDS
// buffer P (rows kb..m-1, stride NB + 1); the panel's row swaps are then
// applied to the other columns, U12 is solved column by column in
// registers and the trailing block takes the panel's rank-NB update (each
// element's updates in ascending pivot order, as in the column-by-column
// algorithm).  piv[j] = the row swapped with row j.  False when a pivot
// column has no nonzero candidate.
//
// Columns [k0, k1) only, with the row swaps applied to columns [swap_lo, cl)
// and the updates confined to columns below cl: one panel of the
// many-workgroup form below, which applies the swaps left of swap_lo
// afterwards (k0 = 0, k1 = k, cl = m, swap_lo = 0 is the whole
// factorization).
template <int NB, int BS = kFB>
__device__ bool front_factor(double *A, int m, int k, double *P, double *sv, int *si,
                             int *spiv, int32_t *piv, int k0, int k1, int cl,
                             int swap_lo = 0) {
  const int t = threadIdx.x;
  constexpr int PS = NB + 1;
  for (int kb = k0; kb < k1; kb += NB) {
    const int nb = min(NB, k1 - kb);
    const int rows = m - kb;      // panel rows
    const int cand = k - kb;      // rows eligible as pivots
    for (int e = t; e < rows * NB; e += BS) {
      const int r = e / NB, c = e % NB;
      P[r * PS + c] = c < nb ? A[(int64_t)(kb + r) * m + kb + c] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
      double v = -1.0;
      int idx = 0x7fffffff;
      for (int r = c + t; r < cand; r += BS) {
        const double a = fabs(P[r * PS + c]);
        if (a > v) {
          v = a;
          idx = r;
        }
      }
      double best;
      int br;
      blk_argmax<BS>(v, idx, sv, si, best, br);
      if (!(best > 0.0) || !(best <= DBL_MAX)) return false;  // same on every thread
      if (t == 0) spiv[c] = kb + br;
      if (br != c)
        for (int cc = t; cc < NB; cc += BS) {
          const double tmp = P[c * PS + cc];
          P[c * PS + cc] = P[br * PS + cc];
          P[br * PS + cc] = tmp;
        }
      __syncthreads();
      const double pv = P[c * PS + c];
      for (int r = c + 1 + t; r < rows; r += BS) {
        const double l = P[r * PS + c] / pv;
        P[r * PS + c] = l;
        for (int cc = c + 1; cc < nb; ++cc) P[r * PS + cc] -= l * P[c * PS + cc];
      }
      __syncthreads();
    }
    for (int e = t; e < rows * NB; e += BS) {
      const int r = e / NB, c = e % NB;
      if (c < nb) A[(int64_t)(kb + r) * m + kb + c] = P[r * PS + c];
    }
    if (t < nb) piv[kb + t] = spiv[t];
    // the panel's row swaps on every other column, in pivot order
    for (int j = swap_lo + t; j < cl; j += BS) {
      if (j >= kb && j < kb + nb) continue;
      for (int c = 0; c < nb; ++c) {
        const int pr = spiv[c];
        if (pr != kb + c) {
          const double tmp = A[(int64_t)(kb + c) * m + j];
          A[(int64_t)(kb + c) * m + j] = A[(int64_t)pr * m + j];
          A[(int64_t)pr * m + j] = tmp;
        }
      }
    }
    __syncthreads();
    const int j0 = kb + nb, ncols = cl - j0;
    // U12 of column j (rows kb..kb+nb-1) by the unit lower panel block
    auto solve_u = [&](int j, double *u) {
#pragma unroll
      for (int c = 0; c < NB; ++c) u[c] = c < nb ? A[(int64_t)(kb + c) * m + j] : 0.0;
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int i = c + 1; i < NB; ++i)
          if (i < nb) u[i] -= P[i * PS + c] * u[c];
    };
    // rows j0+g, j0+g+G, ... of column j take the panel's updates
    auto sweep = [&](int j, int g, int G, const double *u) {
      for (int i = j0 + g; i < m; i += 4 * G) {
        double x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ii = i + q * G;
          x[q] = ii < m ? A[(int64_t)ii * m + j] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ii = i + q * G;
          if (ii < m) {
            const double *l = P + (ii - kb) * PS;
            double v = x[q];
#pragma unroll
            for (int c = 0; c < NB; ++c) v -= l[c] * u[c];
            A[(int64_t)ii * m + j] = v;
          }
        }
      }
    };
    if (ncols >= BS) {
      for (int j = j0 + t; j < cl; j += BS) {
        double u[NB];
        solve_u(j, u);
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < nb) A[(int64_t)(kb + c) * m + j] = u[c];
        sweep(j, 0, 1, u);
      }
    } else if (ncols > 0) {
      const int G = min(BS / ncols, 16);
      const bool mine = t < ncols * G;
      const int j = j0 + t % ncols, g = t / ncols;
      double u[NB];
      if (mine) solve_u(j, u);
      __syncthreads();
      if (mine) {
        if (g == 0)
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c < nb) A[(int64_t)(kb + c) * m + j] = u[c];
        sweep(j, g, G, u);
      }
    }
    __syncthreads();
  }
  return true;
}

__device__ __forceinline__ bool front_skip(const FrontArgs &f, int c) {
  return (f.gate && *f.gate) || (f.active && !f.active[c]) || f.bad[c];
}

// Clears the fronts and the failure flags, then (second kernel) scatters
// the Jacobian's entries into them.
__global__ void __launch_bounds__(kFB) front_clear_kernel(FrontArgs f) {
  if (f.gate && *f.gate) return;
  const int64_t stride = (int64_t)gridDim.x * kFB;
  for (int64_t i = (int64_t)blockIdx.x * kFB + threadIdx.x; i < f.n_vals; i += stride)
    f.vals[i] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kFB + threadIdx.x; i < f.n_comp; i += stride)
    f.bad[i] = 0;
}

__global__ void __launch_bounds__(kFB) front_scatter_kernel(FrontArgs f, const double *J) {
  if (f.gate && *f.gate) return;
  const int64_t stride = (int64_t)gridDim.x * kFB;
  for (int64_t e = (int64_t)blockIdx.x * kFB + threadIdx.x; e < f.n_ent; e += stride)
    f.vals[f.ent_fpos[e]] = J[f.ent_k[e]];
}

template <int NB, int ROWS, int BS>
__global__ void __launch_bounds__(BS) front_factor_kernel(FrontArgs f, int lv0) {
  __shared__ double P[ROWS * (NB + 1)];
  __shared__ double sv[BS / 64];
  __shared__ int si[BS / 64];
  __shared__ int spiv[NB];
  const int s = f.lv[lv0 + blockIdx.x];
  const int c = f.comp[s];
  if (front_skip(f, c)) return;
  const int t = threadIdx.x;
  const int m = f.m[s], k = f.k[s];
  if (m > f.big_m) return;  // factored by many workgroups (launch_big_front)
  double *A = f.vals + f.fo[s];
  // the children's contribution blocks, in child order
  for (int q = f.ch_ptr[s]; q < f.ch_ptr[s + 1]; ++q) {
    const int ch = f.ch[q];
    const int mc = f.m[ch], kc = f.k[ch], w = mc - kc;
    const double *C = f.vals + f.fo[ch] + (int64_t)kc * mc + kc;
    const int32_t *map = f.cbmap + f.ro[ch] + kc;
    for (int64_t e = t; e < (int64_t)w * w; e += BS) {
      const int i = (int)(e / w), j = (int)(e % w);
      A[(int64_t)map[i] * m + map[j]] += C[(int64_t)i * mc + j];
    }
    __syncthreads();
  }
  if (!front_factor<NB, BS>(A, m, k, P, sv, si, spiv, f.piv + f.ro[s], 0, k, m))
    if (t == 0) f.bad[c] = 1;
}

// Fronts of up to kTinyM rows, one thread each (a level of the tree's
// bottom holds up to ~10^5 of them): the children's contribution blocks in
// child order, then the right-looking LU with whole-row swaps, pivots among
// the fully summed rows (largest magnitude, smallest index on ties).  Every
// element takes the same operations in the same order as in the blocked
// workgroup form, so both give the same bits.
constexpr int kTinyM = 16;

__global__ void __launch_bounds__(64) front_tiny_kernel(FrontArgs f, int lv0, int nfr) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q >= nfr) return;
  const int s = f.lv[lv0 + q];
  const int c = f.comp[s];
  if (front_skip(f, c)) return;
  const int m = f.m[s], k = f.k[s];
  if (m > f.big_m) return;
  double *A = f.vals + f.fo[s];
  for (int qq = f.ch_ptr[s]; qq < f.ch_ptr[s + 1]; ++qq) {
    const int ch = f.ch[qq];
    const int mc = f.m[ch], kc = f.k[ch], w = mc - kc;
    const double *C = f.vals + f.fo[ch] + (int64_t)kc * mc + kc;
    const int32_t *map = f.cbmap + f.ro[ch] + kc;
    for (int i = 0; i < w; ++i)
      for (int j = 0; j < w; ++j) A[map[i] * m + map[j]] += C[i * mc + j];
  }
  int32_t *piv = f.piv + f.ro[s];
  for (int col = 0; col < k; ++col) {
    double best = -1.0;
    int br = col;
    for (int r = col; r < k; ++r) {
      const double a = fabs(A[r * m + col]);
      if (a > best) {
        best = a;
        br = r;
      }
    }
    if (!(best > 0.0) || !(best <= DBL_MAX)) {
      f.bad[c] = 1;
      return;
    }
    piv[col] = br;
    if (br != col)
      for (int j = 0; j < m; ++j) {
        const double tmp = A[col * m + j];
        A[col * m + j] = A[br * m + j];
        A[br * m + j] = tmp;
      }
    const double pv = A[col * m + col];
    for (int r = col + 1; r < m; ++r) {
      const double l = A[r * m + col] / pv;
      A[r * m + col] = l;
      for (int j = col + 1; j < m; ++j) A[r * m + j] -= l * A[col * m + j];
    }
  }
}

// ---- many-workgroup form of the large fronts (rows m > big_m), all of a
// level's at once (blockIdx.y or .z = the front among the level's large
// ones, f.big[lb + that]): per panel of kBigW fully summed columns, one
// workgroup per front factors the panel (swaps on the columns left of the
// panel's end), then the columns right of it take the panel's row swaps
// and their U12 rows (one thread per column), then the trailing block
// takes the rank-kBigW update in 64 x 64 tiles.  A front whose columns
// are all factored leaves the later panel launches at once.  Every element
// sees the same operations in the same order as in the one-workgroup
// form, so both give the same bits.
constexpr int kBigW = 32;   // panel width
// threads of a panel workgroup (per logic factorisation, 256: 5.8 ms of
// panels, 512: 4.2 ms, 1024: 5.9 ms)
constexpr int kPanelBlock = 512;
constexpr int kTile = 64;   // trailing-update tile

// The children's contribution blocks, one workgroup per receiving row of
// a front (blockIdx.x = the row among the front's receiving rows): the
// child rows that land on it, in child order, each added across its
// columns, so every element takes its children's terms in child order.
__global__ void __launch_bounds__(kFB) front_extend_kernel(FrontArgs f, int lb) {
  const int s = f.big[lb + blockIdx.y];
  if (front_skip(f, f.comp[s])) return;
  const int64_t q = f.xr_off[s] + blockIdx.x;
  if (q >= f.xr_off[s + 1]) return;
  const int m = f.m[s];
  double *Ar = f.vals + f.fo[s] + (int64_t)f.xr_row[q] * m;
  for (int64_t e = f.ext_ptr[q]; e < f.ext_ptr[q + 1]; ++e) {
    const int ch = f.ext_ch[e], i = f.ext_i[e];
    const int mc = f.m[ch], kc = f.k[ch], w = mc - kc;
    const double *C = f.vals + f.fo[ch] + (int64_t)(kc + i) * mc + kc;
    const int32_t *map = f.cbmap + f.ro[ch] + kc;
    for (int j = threadIdx.x; j < w; j += kFB) Ar[map[j]] += C[j];
    __syncthreads();
  }
}

template <int NB, int ROWS, int BS>
__global__ void __launch_bounds__(BS) front_panel_kernel(FrontArgs f, int lb, int kb) {
  __shared__ double P[ROWS * (NB + 1)];
  __shared__ double sv[BS / 64];
  __shared__ int si[BS / 64];
  __shared__ int spiv[NB];
  const int s = f.big[lb + blockIdx.x];
  const int c = f.comp[s];
  if (front_skip(f, c)) return;
  const int m = f.m[s], k = f.k[s];
  if (kb >= k) return;
  const int cl = min(kb + kBigW, k);
  if (!front_factor<NB, BS>(f.vals + f.fo[s], m, k, P, sv, si, spiv, f.piv + f.ro[s], kb, cl,
                            cl, kb))
    if (threadIdx.x == 0) f.bad[c] = 1;
}

// Columns j >= cl: the panel's row swaps in pivot order, then U12 = L11^-1
// (rows kb..cl-1) with the unit lower panel block, each u[i] updated in
// ascending pivot order.
// One wavefront per 64 columns: the work is a few dependent loads and a
// 32-term solve per column, so more, smaller workgroups spread it over
// more CUs.
constexpr int kSwapBlock = 64;

__global__ void __launch_bounds__(kSwapBlock) front_swap_trsm_kernel(FrontArgs f, int lb, int kb) {
  __shared__ double L11[kBigW][kBigW + 1];
  __shared__ int32_t pv[kBigW];
  const int s = f.big[lb + blockIdx.y];
  if (front_skip(f, f.comp[s])) return;
  const int m = f.m[s], k = f.k[s];
  if (kb >= k) return;
  const int cl = min(kb + kBigW, k), nw = cl - kb;
  if (cl + (int)(blockIdx.x * kSwapBlock) >= m) return;
  double *A = f.vals + f.fo[s];
  const int32_t *piv = f.piv + f.ro[s];
  for (int e = threadIdx.x; e < kBigW * kBigW; e += kSwapBlock) {
    const int i = e / kBigW, cc = e % kBigW;
    L11[i][cc] = (i < nw && cc < nw) ? A[(int64_t)(kb + i) * m + kb + cc] : 0.0;
  }
  if (threadIdx.x < kBigW) pv[threadIdx.x] = threadIdx.x < nw ? piv[kb + threadIdx.x] : 0;
  __syncthreads();
  const int j = cl + blockIdx.x * kSwapBlock + threadIdx.x;
  if (j >= m) return;
  for (int cc = 0; cc < nw; ++cc) {
    const int pr = pv[cc];
    if (pr != kb + cc) {
      const double tmp = A[(int64_t)(kb + cc) * m + j];
      A[(int64_t)(kb + cc) * m + j] = A[(int64_t)pr * m + j];
      A[(int64_t)pr * m + j] = tmp;
    }
  }
  double u[kBigW];
#pragma unroll
  for (int cc = 0; cc < kBigW; ++cc) u[cc] = cc < nw ? A[(int64_t)(kb + cc) * m + j] : 0.0;
#pragma unroll
  for (int cc = 0; cc < kBigW; ++cc)
#pragma unroll
    for (int i = cc + 1; i < kBigW; ++i)
      if (i < nw) u[i] -= L11[i][cc] * u[cc];
#pragma unroll
  for (int cc = 0; cc < kBigW; ++cc)
    if (cc < nw) A[(int64_t)(kb + cc) * m + j] = u[cc];
}

// Rows and columns >= cl: A -= L21 U12, each element's terms in ascending
// pivot order.  Grid: (tiles over columns, tiles over rows, fronts).
__global__ void __launch_bounds__(kFB) front_update_kernel(FrontArgs f, int lb, int kb) {
  __shared__ double Lt[kTile][kBigW + 1];
  __shared__ double Ut[kBigW][kTile + 1];
  const int s = f.big[lb + blockIdx.z];
  if (front_skip(f, f.comp[s])) return;
  const int m = f.m[s], k = f.k[s];
  if (kb >= k) return;
  const int cl = min(kb + kBigW, k), nw = cl - kb;
  double *A = f.vals + f.fo[s];
  const int r0 = cl + blockIdx.y * kTile, c0 = cl + blockIdx.x * kTile;
  if (r0 >= m || c0 >= m) return;
  for (int e = threadIdx.x; e < kTile * kBigW; e += kFB) {
    const int r = e / kBigW, cc = e % kBigW;
    Lt[r][cc] = (r0 + r < m && cc < nw) ? A[(int64_t)(r0 + r) * m + kb + cc] : 0.0;
  }
  for (int e = threadIdx.x; e < kBigW * kTile; e += kFB) {
    const int cc = e / kTile, q = e % kTile;
    Ut[cc][q] = (cc < nw && c0 + q < m) ? A[(int64_t)(kb + cc) * m + c0 + q] : 0.0;
  }
  __syncthreads();
  // thread: one column of the tile, rows r, r + 4, ... (16 rows)
  const int q = threadIdx.x % kTile, rg = threadIdx.x / kTile;
  const int j = c0 + q;
  if (j >= m) return;
  for (int r = rg; r < kTile; r += kFB / kTile) {
    const int i = r0 + r;
    if (i >= m) break;
    double v = A[(int64_t)i * m + j];
    for (int cc = 0; cc < nw; ++cc) v -= Lt[r][cc] * Ut[cc][q];
    A[(int64_t)i * m + j] = v;
  }
}

// The row swaps of the later panels on the factored columns j < k (the L
// part), deferred from the panels: column j takes every pivot after its
// own panel, in pivot order, as the one-workgroup form applies them.
__global__ void __launch_bounds__(kFB) front_left_swaps_kernel(FrontArgs f, int lb) {
  const int s = f.big[lb + blockIdx.y];
  if (front_skip(f, f.comp[s])) return;
  const int m = f.m[s], k = f.k[s];
  const int j = blockIdx.x * kFB + threadIdx.x;
  if (j >= k) return;
  double *A = f.vals + f.fo[s];
  const int32_t *piv = f.piv + f.ro[s];
  for (int c = (j / kBigW + 1) * kBigW; c < k; ++c) {
    const int pr = piv[c];
    if (pr != c) {
      const double tmp = A[(int64_t)c * m + j];
      A[(int64_t)c * m + j] = A[(int64_t)pr * m + j];
      A[(int64_t)pr * m + j] = tmp;
    }
  }
}

// Forward solve of one level's fronts: the front's right-hand side (its
// fully summed rows' entries, zero below), its children's remainders,
// its row swaps, then L y = b with the unit lower columns; y goes to the
// position vector, the remainder below stays in wrhs for the parent.
// Blocked by 64 columns: the diagonal block on one wavefront (the
// solution broadcast by lane shuffles), then every row below takes the
// block's 64 terms.  Each element takes its terms in ascending column
// order, as the column-by-column algorithm gives them.
template <int ROWS, int BS>
__global__ void __launch_bounds__(BS) front_fwd_kernel(FrontArgs f, const double *bp,
                                                        int lv0) {
  __shared__ double ws[ROWS];
  const int s = f.lv[lv0 + blockIdx.x];
  const int c = f.comp[s];
  if (front_skip(f, c)) return;
  const int t = threadIdx.x;
  const int m = f.m[s], k = f.k[s];
  const double *A = f.vals + f.fo[s];
  const int32_t *rows = f.rows + f.ro[s];
  double *w = f.wrhs + f.ro[s];
  for (int i = t; i < m; i += BS) ws[i] = i < k ? bp[rows[i]] : 0.0;
  __syncthreads();
  for (int q = f.ch_ptr[s]; q < f.ch_ptr[s + 1]; ++q) {
    const int ch = f.ch[q];
    const int kc = f.k[ch], wc = f.m[ch] - kc;
    const double *r = f.wrhs + f.ro[ch] + kc;
    const int32_t *map = f.cbmap + f.ro[ch] + kc;
    for (int i = t; i < wc; i += BS) ws[map[i]] += r[i];
    __syncthreads();
  }
  // the row swaps in pivot order: one wavefront finds the pivots that
  // move a row (64 at a time) and its first lane applies just those
  const int32_t *pv = f.piv + f.ro[s];
  if (t < 64) {
    for (int j0 = 0; j0 < k; j0 += 64) {
      const int j = j0 + t;
      const int p = j < k ? pv[j] : j;
      unsigned long long mv = __ballot(p != j);
      while (mv) {
        const int b = __ffsll(mv) - 1;
        const int pb = __shfl(p, b);
        if (t == 0) {
          const double tmp = ws[j0 + b];
          ws[j0 + b] = ws[pb];
          ws[pb] = tmp;
        }
        mv &= mv - 1;
      }
    }
  }
  __syncthreads();
  for (int jb = 0; jb < k; jb += 64) {
    const int je = min(jb + 64, k);
    if (t < 64) {
      const int i = jb + t;
      double wi = i < je ? ws[i] : 0.0;
      for (int j = jb; j < je; ++j) {
        const double yj = __shfl(wi, j - jb);
        if (i > j && i < je) wi -= A[(int64_t)i * m + j] * yj;
      }
      if (i < je) ws[i] = wi;
    }
    __syncthreads();
    for (int i = je + t; i < m; i += BS) {
      double r = ws[i];
      const double *a = A + (int64_t)i * m;
      for (int j = jb; j < je; ++j) r -= a[j] * ws[j];
      ws[i] = r;
    }
    __syncthreads();
  }
  for (int i = t; i < m; i += BS) {
    w[i] = ws[i];
    if (i < k) f.v[rows[i]] = ws[i];
  }
}

// Backward solve of one level's fronts (levels from the root down): the
// rows below the fully summed block are solved already (by ancestors).
// Blocked by 64 columns from the last: the diagonal block on one
// wavefront, then every row above takes the block's terms; each element
// takes its terms in descending column order, as the column-by-column
// algorithm gives them.
template <int ROWS, int BS>
__global__ void __launch_bounds__(BS) front_bwd_kernel(FrontArgs f, int lv0) {
  __shared__ double tb[ROWS], xs[ROWS];
  const int s = f.lv[lv0 + blockIdx.x];
  const int c = f.comp[s];
  if (front_skip(f, c)) return;
  const int t = threadIdx.x;
  const int m = f.m[s], k = f.k[s];
  const double *A = f.vals + f.fo[s];
  const int32_t *rows = f.rows + f.ro[s];
  // the right-hand side less the solved rows below: a wavefront per row,
  // its lanes strided over the row, summed by a fixed shuffle tree
  for (int j = k + t; j < m; j += BS) xs[j - k] = f.v[rows[j]];
  __syncthreads();
  const int w = t >> 6, lane = t & 63;
  for (int i = w; i < k; i += BS / 64) {
    const double *a = A + (int64_t)i * m;
    double p = 0.0;
    for (int j = k + lane; j < m; j += 64) p += a[j] * xs[j - k];
    for (int off = 32; off >= 1; off >>= 1) p += __shfl_xor(p, off);
    if (lane == 0) tb[i] = f.v[rows[i]] - p;
  }
  __syncthreads();
  for (int jb = ((k - 1) / 64) * 64; jb >= 0 && k > 0; jb -= 64) {
    const int je = min(jb + 64, k);
    if (t < 64) {
      const int i = jb + t;
      double ti = i < je ? tb[i] : 0.0;
      for (int j = je - 1; j >= jb; --j) {
        double xj = 0.0;
        if (i == j) xj = ti / A[(int64_t)j * m + j];
        xj = __shfl(xj, j - jb);
        if (i == j)
          ti = xj;
        else if (i < j)
          ti -= A[(int64_t)i * m + j] * xj;
      }
      if (i < je) tb[i] = ti;
    }
    __syncthreads();
    for (int i = t; i < jb; i += BS) {
      double r = tb[i];
      const double *a = A + (int64_t)i * m;
      for (int j = je - 1; j >= jb; --j) r -= a[j] * tb[j];
      tb[i] = r;
    }
    __syncthreads();
  }
  for (int i = t; i < k; i += BS) f.v[rows[i]] = tb[i];
}

// The residual and |J| row-sum maxima of a component's long rows in
// [r0, r1) (rows of more than kLongRow entries, listed per component):
// each row by the whole block, strided partial sums then a fixed tree.
template <int B>
__device__ void long_rows_check(const FrontArgs &f, const SparseLuArgs &a, int c, int r0,
                                int r1, double *sv, double &rm, double &am) {
  const int beg = a.cstart[c];
  const int eb = a.ent_off[c];
  const double *jp = a.Jp + eb;
  const int32_t *ecol = a.ent_col + eb, *ers = a.ent_rs + beg + c;
  const double *v = f.v + beg;
  for (int q = f.long_ptr[c]; q < f.long_ptr[c + 1]; ++q) {
    const int i = f.long_rows[q];
    if (i < r0 || i >= r1) continue;
    double s = 0.0, as = 0.0;
    for (int j = ers[i] + threadIdx.x; j < ers[i + 1]; j += B) {
      const double qv = jp[j];
      s = s + qv * v[ecol[j]];
      as = as + fabs(qv);
    }
    s = blk_reduce<B>(s, sv, false);
    as = blk_reduce<B>(as, sv, false);
    rm = fmax(rm, fabs(a.bp[beg + i] - s));
    am = fmax(am, as);
  }
}

// Per front component: backward error of the solution against the
// unfactored rows (entry-ordered J and position-ordered rhs of the permute
// kernel), then x += dx and the weighted RMS norm of dx
// (N_NLS_DampedNewton.C:311-335), as the task-list kernel does.
__global__ void __launch_bounds__(kFinishBlock) front_finish_kernel(FrontArgs f,
                                                                    SparseLuArgs a) {
  __shared__ double sv[kFinishBlock];
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  if (f.gate && *f.gate) return;
  if (!f.front_comp[c]) return;
  if (f.fin_split && f.fin_split[c]) return;  // the chunked kernels below
  if (a.active && !a.active[c]) return;  // stats written by the task kernel
  const int beg = a.cstart[c], n = a.cstart[c + 1] - beg;
  double *stats = a.stats + 4 * (size_t)c;
  double rmax = 0.0;
  for (int i = t; i < n; i += kFinishBlock) rmax = fmax(rmax, fabs(a.bp[beg + i]));
  const double max_rhs = blk_reduce<kFinishBlock>(rmax, sv, true);
  bool good = !f.bad[c];
  if (good) {
    const int eb = a.ent_off[c];
    const double *jp = a.Jp + eb;
    const int32_t *ecol = a.ent_col + eb, *ers = a.ent_rs + beg + c;
    const double *v = f.v + beg;
    double rm = 0.0, am = 0.0, xm = 0.0;
    for (int i = t; i < n; i += kFinishBlock) {
      xm = fmax(xm, fabs(v[i]));
      if (ers[i + 1] - ers[i] > kLongRow) continue;  // below, by the block
      double r = a.bp[beg + i], as = 0.0;
      for (int j = ers[i]; j < ers[i + 1]; ++j) {
        const double q = jp[j];
        r = r - q * v[ecol[j]];
        as = as + fabs(q);
      }
      rm = fmax(rm, fabs(r));
      am = fmax(am, as);
    }
    long_rows_check<kFinishBlock>(f, a, c, 0, n, sv, rm, am);
    rm = blk_reduce<kFinishBlock>(rm, sv, true);
    am = blk_reduce<kFinishBlock>(am, sv, true);
    xm = blk_reduce<kFinishBlock>(xm, sv, true);
    good = rm <= a.max_berr * (am * xm + max_rhs);  // false for NaN / inf too
  }
  if (!good) {
    // singular for the fronts' restricted pivoting: the dense partial-
    // pivoting kernel launched next refactors the component when it has a
    // dense form, and reports it singular otherwise
    if (t == 0) {
      stats[0] = max_rhs;
      stats[1] = 0.0;
      stats[2] = 1.0;
      stats[3] = (double)n;
      a.fallback[c] = 1;
    }
    return;
  }
  const int32_t *qcol = a.qcol + beg;
  double acc = 0.0;
  for (int i = t; i < n; i += kFinishBlock) {
    const int g = qcol[i];
    const double dx = f.v[beg + i];
    const double w = a.reltol * fmax(fabs(a.wa[g]), fabs(a.wb[g])) + a.abstol;
    const double q = dx / w;
    acc = acc + q * q;
    a.x[g] = a.x[g] + dx;
  }
  const double ss = blk_reduce<kFinishBlock>(acc, sv, false);
  if (t == 0) {
    stats[0] = max_rhs;
    stats[1] = sqrt(ss / n);
    stats[2] = 0.0;
    stats[3] = (double)n;
  }
}

// ---- the same finish for components of more than kFinishSplit unknowns,
// in chunks of kFinishChunk rows: (1) per chunk the maxima of |rhs|, of
// the residual, of the row sums of |J| and of |dx|; (2) per component
// their maxima (exact in any order) and the verdict; (3) per chunk
// x += dx and the chunk's sum of squared weighted steps; (4) per component
// those sums in chunk order, the weighted RMS norm.
__global__ void __launch_bounds__(kFB) finish_part_kernel(FrontArgs f, SparseLuArgs a) {
  __shared__ double sv[kFB];
  if (f.gate && *f.gate) return;
  const int ch = blockIdx.x, t = threadIdx.x;
  const int c = f.fin_chunk_comp[ch];
  if (a.active && !a.active[c]) return;
  const int beg = a.cstart[c], n = a.cstart[c + 1] - beg;
  const int r0 = f.fin_chunk_beg[ch], r1 = min(n, r0 + kFinishChunk);
  double rmax = 0.0, rm = 0.0, am = 0.0, xm = 0.0;
  const bool bad = f.bad[c] != 0;
  const int eb = a.ent_off[c];
  const double *jp = a.Jp + eb;
  const int32_t *ecol = a.ent_col + eb, *ers = a.ent_rs + beg + c;
  const double *v = f.v + beg;
  for (int i = r0 + t; i < r1; i += kFB) {
    const double b = a.bp[beg + i];
    rmax = fmax(rmax, fabs(b));
    if (!bad) {
      xm = fmax(xm, fabs(v[i]));
      if (ers[i + 1] - ers[i] > kLongRow) continue;  // below, by the block
      double r = b, as = 0.0;
      for (int j = ers[i]; j < ers[i + 1]; ++j) {
        const double q = jp[j];
        r = r - q * v[ecol[j]];
        as = as + fabs(q);
      }
      rm = fmax(rm, fabs(r));
      am = fmax(am, as);
    }
  }
  if (!bad) long_rows_check<kFB>(f, a, c, r0, r1, sv, rm, am);
  double *p = f.fin_part + 4 * (size_t)ch;
  rmax = blk_reduce<kFB>(rmax, sv, true);
  rm = blk_reduce<kFB>(rm, sv, true);
  am = blk_reduce<kFB>(am, sv, true);
  xm = blk_reduce<kFB>(xm, sv, true);
  if (t == 0) {
    p[0] = rmax;
    p[1] = rm;
    p[2] = am;
    p[3] = xm;
  }
}

__global__ void __launch_bounds__(kFB) finish_decide_kernel(FrontArgs f, SparseLuArgs a) {
  __shared__ double sv[kFB];
  if (f.gate && *f.gate) return;
  const int c = f.fin_comps[blockIdx.x], t = threadIdx.x;
  if (a.active && !a.active[c]) return;
  const int n = a.cstart[c + 1] - a.cstart[c];
  double m4[4] = {0.0, 0.0, 0.0, 0.0};
  for (int ch = f.fin_cptr[blockIdx.x] + t; ch < f.fin_cptr[blockIdx.x + 1]; ch += kFB)
    for (int q = 0; q < 4; ++q) m4[q] = fmax(m4[q], f.fin_part[4 * (size_t)ch + q]);
  for (int q = 0; q < 4; ++q) m4[q] = blk_reduce<kFB>(m4[q], sv, true);
  const bool good = !f.bad[c] && m4[1] <= a.max_berr * (m4[2] * m4[3] + m4[0]);
  if (t == 0) {
    f.fin_good[c] = good ? 1 : 0;
    double *stats = a.stats + 4 * (size_t)c;
    stats[0] = m4[0];
    if (!good) {
      stats[1] = 0.0;
      stats[2] = 1.0;
      stats[3] = (double)n;
      a.fallback[c] = 1;
    }
  }
}

__global__ void __launch_bounds__(kFB) finish_update_kernel(FrontArgs f, SparseLuArgs a) {
  __shared__ double sv[kFB];
  if (f.gate && *f.gate) return;
  const int ch = blockIdx.x, t = threadIdx.x;
  const int c = f.fin_chunk_comp[ch];
  if (a.active && !a.active[c]) return;
  if (!f.fin_good[c]) return;
  const int beg = a.cstart[c], n = a.cstart[c + 1] - beg;
  const int r0 = f.fin_chunk_beg[ch], r1 = min(n, r0 + kFinishChunk);
  const int32_t *qcol = a.qcol + beg;
  double acc = 0.0;
  for (int i = r0 + t; i < r1; i += kFB) {
    const int g = qcol[i];
    const double dx = f.v[beg + i];
    const double w = a.reltol * fmax(fabs(a.wa[g]), fabs(a.wb[g])) + a.abstol;
    const double q = dx / w;
    acc = acc + q * q;
    a.x[g] = a.x[g] + dx;
  }
  acc = blk_reduce<kFB>(acc, sv, false);
  if (t == 0) f.fin_part[4 * (size_t)ch] = acc;
}

__global__ void __launch_bounds__(64) finish_sum_kernel(FrontArgs f, SparseLuArgs a) {
  if (f.gate && *f.gate) return;
  const int c = f.fin_comps[blockIdx.x];
  if (a.active && !a.active[c]) return;
  if (!f.fin_good[c] || threadIdx.x != 0) return;
  const int n = a.cstart[c + 1] - a.cstart[c];
  double ss = 0.0;
  for (int ch = f.fin_cptr[blockIdx.x]; ch < f.fin_cptr[blockIdx.x + 1]; ++ch)
    ss = ss + f.fin_part[4 * (size_t)ch];
  double *stats = a.stats + 4 * (size_t)c;
  stats[1] = sqrt(ss / n);
  stats[2] = 0.0;
  stats[3] = (double)n;
}

// One workgroup per front of the level: one wavefront for fronts of up
// to 64 rows (cheap barriers, four times the workgroups per CU), four for
// larger ones.  The block size changes which thread does an operation,
// never an element's operations or their order.
template <int NB, int ROWS>
hipError_t launch_level(const FrontArgs &f, int lv0, int nfr, hipStream_t s) {
  if (ROWS <= 64)
    hipLaunchKernelGGL((front_factor_kernel<NB, ROWS, 64>), dim3(nfr), dim3(64), 0, s, f, lv0);
  else
    hipLaunchKernelGGL((front_factor_kernel<NB, ROWS, kFB>), dim3(nfr), dim3(kFB), 0, s, f,
                       lv0);
  return hipGetLastError();
}

// A level's fronts of more than big_m rows (f.big[lb, lb + nb)) by many
// workgroups: their children's contribution blocks (one launch, a
// workgroup per receiving row), then per panel position the panel,
// swap / U12 and trailing-update launches over all of them, then the
// deferred left swaps.
hipError_t launch_big_fronts(const FrontArgs &f, int lb, int nb, hipStream_t s) {
  int max_m = 0, max_k = 0;
  int64_t max_rows = 0;
  for (int q = lb; q < lb + nb; ++q) {
    const int sf = f.big_host[q];
    max_m = std::max(max_m, f.m_host[sf]);
    max_k = std::max(max_k, f.k_host[sf]);
    max_rows = std::max(max_rows, f.xr_off_host[sf + 1] - f.xr_off_host[sf]);
  }
  if (max_rows > 0)
    hipLaunchKernelGGL(front_extend_kernel, dim3((unsigned)max_rows, nb), dim3(kFB), 0, s, f,
                       lb);
  for (int kb = 0; kb < max_k; kb += kBigW) {
    if (max_m <= 1934)
      hipLaunchKernelGGL((front_panel_kernel<8, 1934, kPanelBlock>), dim3(nb), dim3(kPanelBlock),
                         0, s, f, lb, kb);
    else
      hipLaunchKernelGGL((front_panel_kernel<4, kFrontMaxM, kPanelBlock>), dim3(nb),
                         dim3(kPanelBlock), 0, s, f, lb,
                         kb);
    // the largest trailing block at this panel position
    int rest = 0;
    for (int q = lb; q < lb + nb; ++q) {
      const int sf = f.big_host[q];
      if (kb < f.k_host[sf])
        rest = std::max(rest, f.m_host[sf] - std::min(kb + kBigW, f.k_host[sf]));
    }
    if (rest <= 0) continue;
    hipLaunchKernelGGL(front_swap_trsm_kernel, dim3((rest + kSwapBlock - 1) / kSwapBlock, nb),
                       dim3(kSwapBlock), 0, s, f, lb, kb);
    const unsigned tiles = (unsigned)((rest + kTile - 1) / kTile);
    hipLaunchKernelGGL(front_update_kernel, dim3(tiles, tiles, nb), dim3(kFB), 0, s, f, lb,
                       kb);
  }
  if (max_k > kBigW)
    hipLaunchKernelGGL(front_left_swaps_kernel, dim3((max_k + kFB - 1) / kFB, nb), dim3(kFB),
                       0, s, f, lb);
  return hipGetLastError();
}

}  // namespace

// The panel form for fronts of up to `max_m` rows: 0 when none fits.
int front_form(int32_t max_m) {
  if (max_m <= 64) return 1;
  if (max_m <= 256) return 2;
  if (max_m <= 1024) return 3;
  if (max_m <= 1934) return 4;
  if (max_m <= kFrontMaxM) return 5;
  return 0;
}

hipError_t launch_fronts(const FrontArgs &f, const SparseLuArgs &a, hipStream_t s) {
  if (f.n_fronts == 0) return hipSuccess;
  const unsigned gclear =
      (unsigned)std::min<int64_t>((std::max(f.n_vals, (int64_t)f.n_comp) + kFB - 1) / kFB, 8192);
  hipLaunchKernelGGL(front_clear_kernel, dim3(gclear), dim3(kFB), 0, s, f);
  const unsigned gsc = (unsigned)std::min<int64_t>((f.n_ent + kFB - 1) / kFB + 1, 8192);
  hipLaunchKernelGGL(front_scatter_kernel, dim3(gsc), dim3(kFB), 0, s, f, a.J);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (int L = 0; L < f.n_levels; ++L) {
    const int lv0 = f.lv_ptr_host[L], nfr = f.lv_ptr_host[L + 1] - lv0;
    // the level's fronts up to big_m rows, one workgroup each
    int32_t small_max = 0;
    bool any_big = false;
    for (int q = lv0; q < lv0 + nfr; ++q) {
      const int32_t mq = f.m_host[f.lv_host[q]];
      if (mq > f.big_m) any_big = true;
      else small_max = std::max(small_max, mq);
    }
    if (small_max > 0 && small_max <= kTinyM) {
      hipLaunchKernelGGL(front_tiny_kernel, dim3((nfr + 63) / 64), dim3(64), 0, s, f, lv0, nfr);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    } else if (small_max > 0) {
      switch (front_form(small_max)) {
        case 1: e = launch_level<16, 64>(f, lv0, nfr, s); break;
        case 2: e = launch_level<16, 256>(f, lv0, nfr, s); break;
        case 3: e = launch_level<16, 1024>(f, lv0, nfr, s); break;
        case 4: e = launch_level<8, 1934>(f, lv0, nfr, s); break;
        default: e = launch_level<4, kFrontMaxM>(f, lv0, nfr, s); break;
      }
      if (e != hipSuccess) return e;
    }
    // the level's larger fronts, at most 65535 per launch (grid y / z)
    if (any_big)
      for (int lb = f.big_ptr_host[L]; lb < f.big_ptr_host[L + 1]; lb += 65535)
        if ((e = launch_big_fronts(f, lb, std::min(65535, f.big_ptr_host[L + 1] - lb), s)) !=
            hipSuccess)
          return e;
  }
  // the solves' LDS holds a front's rows: the level's largest front picks
  // the instance (the largest fronts with 512 threads)
  for (int L = 0; L < f.n_levels; ++L) {
    const int lv0 = f.lv_ptr_host[L], nfr = f.lv_ptr_host[L + 1] - lv0;
    const int mm = f.lv_maxm_host[L];
    const double *bp = a.bp;
    if (mm <= 64)
      hipLaunchKernelGGL((front_fwd_kernel<64, kFB>), dim3(nfr), dim3(kFB), 0, s, f, bp, lv0);
    else if (mm <= 512)
      hipLaunchKernelGGL((front_fwd_kernel<512, kFB>), dim3(nfr), dim3(kFB), 0, s, f, bp, lv0);
    else
      hipLaunchKernelGGL((front_fwd_kernel<kFrontMaxM, kSolveBlock>), dim3(nfr), dim3(kSolveBlock),
                         0, s, f, bp, lv0);
  }
  for (int L = f.n_levels - 1; L >= 0; --L) {
    const int lv0 = f.lv_ptr_host[L], nfr = f.lv_ptr_host[L + 1] - lv0;
    const int mm = f.lv_maxm_host[L];
    if (mm <= 64)
      hipLaunchKernelGGL((front_bwd_kernel<64, kFB>), dim3(nfr), dim3(kFB), 0, s, f, lv0);
    else if (mm <= 512)
      hipLaunchKernelGGL((front_bwd_kernel<512, kFB>), dim3(nfr), dim3(kFB), 0, s, f, lv0);
    else
      hipLaunchKernelGGL((front_bwd_kernel<kFrontMaxM, kSolveBlock>), dim3(nfr),
                         dim3(kSolveBlock), 0, s, f, lv0);
  }
  hipLaunchKernelGGL(front_finish_kernel, dim3(f.n_comp), dim3(kFinishBlock), 0, s, f, a);
  if (f.n_fin_chunks > 0) {
    hipLaunchKernelGGL(finish_part_kernel, dim3(f.n_fin_chunks), dim3(kFB), 0, s, f, a);
    hipLaunchKernelGGL(finish_decide_kernel, dim3(f.n_fin_comps), dim3(kFB), 0, s, f, a);
    hipLaunchKernelGGL(finish_update_kernel, dim3(f.n_fin_chunks), dim3(kFB), 0, s, f, a);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(f.n_fin_comps), dim3(64), 0, s, f, a);
  }
  return hipGetLastError();
}

}  // namespace xg
