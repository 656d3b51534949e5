// ba_dense.hip — the direct route: the reduced camera system of a small bundle assembled densely,
// factored by Cholesky and solved, with a matrix-free certificate. It is the reference's opt-in
// `schurCache` route (src/sfm/bundle/gpu/joint-focal.shade.ts:1206-1383, admitted by
// joint-focal.ts:102-108), which caches S densely and then runs its PCG on the cached matrix; here S
// is factored instead, so one solve replaces the PCG loop of an LM iteration. Both routes end with
// the same check: the true residual b - S x from the matrix-free operator, PCG restarted from x when
// it misses the tolerance (certifyCachedBundleCg j:1330-1383).
//
// The dense array of a bundle (m_b = 6 C_b [+ 1 focal] unknowns padded to a multiple u of 32 with
// identity rows, n = u + 1 rows, row-major): the lower triangle of rows 0..u-1 holds S and is factored
// in place into L (S = L L^T); row u holds the reduced right-hand side. Camera c of the bundle owns
// rows 6c'..6c'+5 (c' = c - first camera), the focal unknown row 6 C_b.
//
//   k_dense_rows     one workgroup per camera c of a direct bundle writes its row block of the lower
//                    triangle. Block (c, d), d < c: -sum, over the points both cameras observe, of
//                    W_c V_p^-1 W_d^T = G_c G_d^T, with G = W L_p (V_p^-1 = L_p L_p^T) written per
//                    camera-major slot by k_camera_pass as 18 contiguous words, so both operands are
//                    whole 144-byte loads (the blocked W layout would scatter a gathered slot over 18
//                    cache lines). Camera c's points go into an LDS table indexed by point id, and
//                    camera d's list is streamed against it (both lists ascend by point). Diagonal
//                    block: S_cc from k_camera_finish. Focal border: H_cf - sum_i G_i L_p^T fc_p.
//                    Each entry has one owner thread that adds the workgroup sums in a fixed order.
//   k_dense_factor   one workgroup per direct bundle: right-looking Cholesky in 32-column panels (the
//                    diagonal block factored by one wave in LDS, the panel by forward substitution,
//                    the trailing lower triangle updated from LDS-staged panel rows), then the forward
//                    and back substitution of the right-hand side into x.
//   k_certify        the matrix-free residual r = b - S x of the solution (after the implicit
//                    sweeps of x): certified, corrected from the factor, or handed to PCG.
//   k_dense_correct  x += L^-T L^-1 r: one step of iterative refinement from the stored factor (the
//                    mixed form factors the FP32 cross blocks' operator and certifies in FP64).
#include "ba_kernels.h"

namespace tcvba {

constexpr uint32_t DENSE_NONE = 0xffffffffu;

// First slot in [lo, hi) of an ascending point list whose point is >= p.
__device__ __forceinline__ uint32_t first_slot(const uint32_t* cp, uint32_t lo, uint32_t hi,
                                               uint32_t p) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cp[mid] < p)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The sum over the workgroup of each of acc[0..N): thread t < N receives entry t (others 0).
// Butterfly within each wave, then the waves in order. Call from uniform control flow.
template <int N>
__device__ __forceinline__ double dense_block_sum(double acc[N], double (*red)[36]) {
#pragma unroll
  for (int k = 0; k < N; k++)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < N; k++) red[wave][k] = acc[k];
  __syncthreads();
  double v = 0;
  if (threadIdx.x < (uint32_t)N)
    for (int w = 0; w < DENSE_THREADS / 64; w++) v += red[w][threadIdx.x];
  return v;
}

// acc += G_a G_b^T = W_a V_p^-1 W_b^T for the two camera-major slots a (row camera) and b of one
// point p (G = W L_p, 18 contiguous words per slot).
__device__ __forceinline__ void pair_block(const double* camG, uint32_t a, uint32_t b,
                                           double acc[36]) {
  const double* Ga = camG + (size_t)a * 18;
  const double* Gb = camG + (size_t)b * 18;
  double x[18], y[18];
#pragma unroll
  for (int k = 0; k < 18; k++) {
    x[k] = Ga[k];
    y[k] = Gb[k];
  }
#pragma unroll
  for (int j = 0; j < 6; j++)
#pragma unroll
    for (int l = 0; l < 6; l++)
      acc[j * 6 + l] += x[j * 3] * y[l * 3] + x[j * 3 + 1] * y[l * 3 + 1] + x[j * 3 + 2] * y[l * 3 + 2];
}

// Whether cameras lc and ld of bundle g (local indices) observe a common point: bit ld of row lc of
// the bundle's covisibility bitmap (k_covisibility).
__device__ __forceinline__ bool covisible(const Problem& q, const Bundle& g, uint32_t lc,
                                          uint32_t ld) {
  const uint32_t words = (g.cam1 - g.cam0 + 31) / 32;
  return (q.covis[g.covis_off + (size_t)lc * words + ld / 32] >> (ld % 32)) & 1u;
}

// One thread per point of a direct bundle: mark every pair of its observing cameras in the bundle's
// covisibility bitmap (integer OR: the result does not depend on the order). Once per solve.
__global__ void k_covisibility(Problem q) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= q.P) return;
  const Bundle& g = q.bundles[q.point_bundle[p]];
  if (!g.direct) return;
  const uint32_t words = (g.cam1 - g.cam0 + 31) / 32;
  const uint32_t* ps = point_start(q);
  for (uint32_t i = ps[p]; i < ps[p + 1]; i++) {
    const uint32_t a = obs_camera(q)[i] - g.cam0;
    for (uint32_t j = ps[p]; j < ps[p + 1]; j++) {
      const uint32_t d = obs_camera(q)[j] - g.cam0;
      if (d < a) atomicOr(&q.covis[g.covis_off + (size_t)a * words + d / 32], 1u << (d % 32));
    }
  }
}

template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_dense_rows(Problem q) {
  // camera c's slots of one chunk: a table indexed by point id when their points span at most
  // DENSE_TABLE ids, else their (ascending) point ids, searched by bisection
  __shared__ uint32_t table[DENSE_TABLE];
  __shared__ uint32_t span[2 * DENSE_LARGE_CAMERAS];
  __shared__ double red[DENSE_THREADS / 64][36];
  const uint32_t c = q.direct_cams[xcd_contiguous(blockIdx.x, gridDim.x)];
  const uint32_t b = q.camera_bundle[c];
  const double* sb = bstate(q, b);
  // only a bundle whose PCG set-up succeeded with a non-zero right-hand side is solved
  if (sb[S_ACTIVE] == 0 || sb[S_PCG_ACTIVE] == 0) return;
  const Bundle& g = q.bundles[b];
  const uint32_t tid = threadIdx.x;
  const uint32_t n = g.dense_n, u = n - 1, lc = c - g.cam0, r0 = 6 * lc;
  double* D = q.dense + g.dense_off;
  const bool var = variable_camera(q, c);
  const uint32_t* cs = camera_start(q);
  const uint32_t* cp = q.cam_point;
  // owner thread t < 36 of entry (t / 6, t % 6) of each 6x6 block of the row block: blocks d < c
  // start at zero, the diagonal block is S_cc (the identity for a fixed camera)
  if (tid < 36) {
    const uint32_t i = tid / 6, j = tid % 6;
    double* row = D + (size_t)(r0 + i) * n;
    for (uint32_t ld = 0; ld < lc; ld++) row[6 * ld + j] = 0;
    if (j <= i) row[r0 + j] = q.camS[(size_t)c * 36 + tid];
  }
  if (tid < 6) D[(size_t)u * n + r0 + tid] = q.camRhs[(size_t)c * 6 + tid];
  if (SHARED) {
    double acc6[6] = {0, 0, 0, 0, 0, 0};
    if (var)
      for (uint32_t s = cs[c] + tid; s < cs[c + 1]; s += DENSE_THREADS) {
        // W_s V_p^-1 fc_p = G_s (L_p^T fc_p)
        const double* pf = q.pfc + (size_t)cp[s] * 3;
        const double* G = q.camG + (size_t)s * 18;
        const double f0 = pf[0], f1 = pf[1], f2 = pf[2];
#pragma unroll
        for (int j = 0; j < 6; j++) acc6[j] += G[j * 3] * f0 + G[j * 3 + 1] * f1 + G[j * 3 + 2] * f2;
      }
    const double v = dense_block_sum<6>(acc6, red);
    const uint32_t f = 6 * (g.cam1 - g.cam0);
    if (tid < 6) D[(size_t)f * n + r0 + tid] = var ? q.camHcf[(size_t)c * 6 + tid] - v : 0.0;
    if (lc == 0 && tid == 0) {
      D[(size_t)f * n + f] = sb[S_FOCAL_SCHUR];
      D[(size_t)u * n + f] = sb[S_FOCAL_RHS];
    }
  }
  if (!var || lc == 0) return;
  const uint32_t cend = cs[c + 1];
  for (uint32_t cc = cs[c]; cc < cend;) {
    // the chunk [cc, ce) of camera c's list and the point ids [p_lo, p_hi) it spans
    const uint32_t p_lo = cp[cc];
    uint32_t ce = first_slot(cp, cc, cend, p_lo + DENSE_TABLE);
    const bool tabled = ce - cc >= DENSE_CHUNK_MIN || ce == cend;
    if (!tabled) ce = cend - cc > DENSE_TABLE ? cc + DENSE_TABLE : cend;
    const uint32_t p_hi = cp[ce - 1] + 1;
    __syncthreads();  // the previous chunk's reads of the table and spans
    // each covisible list's slots whose points lie in [p_lo, p_hi)
    for (uint32_t ld = tid; ld < lc; ld += DENSE_THREADS) {
      const uint32_t d = g.cam0 + ld;
      uint32_t lo = 0, hi = 0;
      if (covisible(q, g, lc, ld) && variable_camera(q, d)) {
        lo = first_slot(cp, cs[d], cs[d + 1], p_lo);
        hi = first_slot(cp, lo, cs[d + 1], p_hi);
      }
      span[2 * ld] = lo;
      span[2 * ld + 1] = hi;
    }
    if (tabled) {
      for (uint32_t k = tid; k < p_hi - p_lo; k += DENSE_THREADS) table[k] = DENSE_NONE;
      __syncthreads();
      for (uint32_t s = cc + tid; s < ce; s += DENSE_THREADS) table[cp[s] - p_lo] = s;
    } else {
      for (uint32_t s = cc + tid; s < ce; s += DENSE_THREADS) table[s - cc] = cp[s];
    }
    __syncthreads();
    // wave w takes the earlier cameras ld = w, w + 4, ...: each block (c, d) is summed by one
    // wave (lane-strided matches, then a butterfly), with no workgroup barrier per block
    const uint32_t nc = ce - cc, wave = tid / 64, lane = tid % 64;
    for (uint32_t ld = wave; ld < lc; ld += DENSE_THREADS / 64) {
      const uint32_t dlo = span[2 * ld], dhi = span[2 * ld + 1];
      if (dlo == dhi) continue;
      double acc[36];
#pragma unroll
      for (int k = 0; k < 36; k++) acc[k] = 0;
      bool hit = false;
      for (uint32_t s = dlo + lane; s < dhi; s += 64) {
        const uint32_t p = cp[s];
        uint32_t sc = DENSE_NONE;
        if (tabled) {
          sc = table[p - p_lo];
        } else {
          uint32_t lo = 0, hi = nc;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (table[mid] < p)
              lo = mid + 1;
            else
              hi = mid;
          }
          if (lo < nc && table[lo] == p) sc = cc + lo;
        }
        if (sc == DENSE_NONE) continue;
        hit = true;
        pair_block(q.camG, sc, s, acc);
      }
      if (!__any(hit)) continue;
#pragma unroll
      for (int k = 0; k < 36; k++)
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
      double* blk = D + (size_t)r0 * n + 6 * ld;
#pragma unroll
      for (int k = 0; k < 36; k++)
        if (lane == (uint32_t)k) blk[(size_t)(k / 6) * n + k % 6] -= acc[k];
    }
    cc = ce;
  }
}

// Global row of unknown i < m_b of bundle g's dense system.
__device__ __forceinline__ uint32_t dense_row(const Bundle& g, uint32_t i) {
  const uint32_t cams = 6 * (g.cam1 - g.cam0);
  return i < cams ? 6 * g.cam0 + i : g.focal_row;
}

template <bool SHARED>
__device__ __forceinline__ uint32_t dense_unknowns(const Bundle& g) {
  return 6 * (g.cam1 - g.cam0) + (SHARED ? 1u : 0u);
}

constexpr int NB = DENSE_NB;
constexpr uint32_t DENSE_MAX_ROWS = (6 * DENSE_MAX_CAMERAS + 1 + NB - 1) / NB * NB;

// Stage the lower triangle of the diagonal block at (kb, kb) of A into L (zero above).
__device__ __forceinline__ void stage_diagonal(const double* A, uint32_t n, uint32_t kb,
                                               double (*L)[NB + 1]) {
  for (uint32_t e = threadIdx.x; e < NB * NB; e += DENSE_THREADS) {
    const uint32_t i = e / NB, j = e % NB;
    L[i][j] = j <= i ? A[(size_t)(kb + i) * n + kb + j] : 0.0;
  }
}

// y <- L^-T L^-1 y for the factor L in the lower triangle of A's first u rows (u a multiple of 32);
// y in LDS. Blocked: each 32-row diagonal block by one wave, the rest of the rows by the workgroup.
__device__ void dense_substitute(const double* A, uint32_t n, uint32_t u, double* y,
                                 double (*L)[NB + 1]) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t kb = 0; kb < u; kb += NB) {
    __syncthreads();
    stage_diagonal(A, n, kb, L);
    __syncthreads();
    if (tid < 64) {
      for (uint32_t j = 0; j < (uint32_t)NB; j++) {
        if (tid == j) y[kb + j] /= L[j][j];
        wave_sync();
        if (tid > j && tid < (uint32_t)NB) y[kb + tid] -= L[tid][j] * y[kb + j];
        wave_sync();
      }
    }
    __syncthreads();
    for (uint32_t i = kb + NB + tid; i < u; i += DENSE_THREADS) {
      const double* row = A + (size_t)i * n + kb;
      double sum = 0;
#pragma unroll 8
      for (uint32_t t = 0; t < (uint32_t)NB; t++) sum += row[t] * y[kb + t];
      y[i] -= sum;
    }
  }
  for (uint32_t k = u; k > 0; k -= NB) {
    const uint32_t kb = k - NB;
    __syncthreads();
    stage_diagonal(A, n, kb, L);
    __syncthreads();
    if (tid < 64) {
      for (int j = NB - 1; j >= 0; j--) {
        if (tid == (uint32_t)j) y[kb + j] /= L[j][j];
        wave_sync();
        if (tid < (uint32_t)j) y[kb + tid] -= L[j][tid] * y[kb + j];
        wave_sync();
      }
    }
    __syncthreads();
    for (uint32_t i = tid; i < kb; i += DENSE_THREADS) {
      double sum = 0;
#pragma unroll 8
      for (uint32_t t = 0; t < (uint32_t)NB; t++) sum += A[(size_t)(kb + t) * n + i] * y[kb + t];
      y[i] -= sum;
    }
  }
  __syncthreads();
}

template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_dense_factor(Problem q) {
  __shared__ double L[NB][NB + 1];
  __shared__ double Pi[NB][NB + 1];
  __shared__ double Pj[NB][NB + 1];
  __shared__ double y[DENSE_MAX_ROWS];
  __shared__ int bad;
  const uint32_t b = q.small_bundles[blockIdx.x];
  double* s = bstate(q, b);
  if (s[S_ACTIVE] == 0 || s[S_PCG_ACTIVE] == 0) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1, mb = dense_unknowns<SHARED>(g), tid = threadIdx.x;
  double* A = q.dense + g.dense_off;
  if (tid == 0) bad = 0;
  // the padding rows mb..u-1: identity, and a zero right-hand side
  for (uint32_t i = mb; i < u; i++)
    for (uint32_t j = tid; j <= i; j += DENSE_THREADS) A[(size_t)i * n + j] = j == i ? 1.0 : 0.0;
  for (uint32_t j = mb + tid; j < u; j += DENSE_THREADS) A[(size_t)u * n + j] = 0;
  for (uint32_t kb = 0; kb < u; kb += NB) {
    __syncthreads();
    stage_diagonal(A, n, kb, L);
    __syncthreads();
    if (tid < 64) {
      // one wave, lane i on row i of the diagonal block
      for (uint32_t k = 0; k < (uint32_t)NB; k++) {
        if (tid == k) {
          const double d = L[k][k];
          if (!(d > 0) || !(d < INFINITY)) bad = 1;
          L[k][k] = sqrt(fabs(d));
        }
        wave_sync();
        if (tid > k && tid < (uint32_t)NB) L[tid][k] /= L[k][k];
        wave_sync();
        if (tid > k && tid < (uint32_t)NB) {
          const double lik = L[tid][k];
          for (uint32_t j = k + 1; j <= tid; j++) L[tid][j] -= lik * L[j][k];
        }
        wave_sync();
      }
    }
    __syncthreads();
    if (bad) return;  // x stays zero and the bundle's PCG runs as set up
    for (uint32_t e = tid; e < NB * NB; e += DENSE_THREADS) {
      const uint32_t i = e / NB, j = e % NB;
      if (j <= i) A[(size_t)(kb + i) * n + kb + j] = L[i][j];
    }
    // panel: each row below the block solves x L_kk^T = a
    for (uint32_t i = kb + NB + tid; i < u; i += DENSE_THREADS) {
      double* row = A + (size_t)i * n + kb;
      double a[NB];
#pragma unroll
      for (int j = 0; j < NB; j++) a[j] = row[j];
#pragma unroll
      for (int j = 0; j < NB; j++) {
        double v = a[j];
#pragma unroll
        for (int t = 0; t < j; t++) v -= a[t] * L[j][t];
        a[j] = v / L[j][j];
        // keep the scheduler from hoisting the LDS reads of later rows (528 values in all)
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int j = 0; j < NB; j++) row[j] = a[j];
    }
    // trailing lower triangle: A_ij -= sum_t P_it P_jt for kb + 32 <= j <= i < u
    for (uint32_t ib = kb + NB; ib < u; ib += NB) {
      __syncthreads();  // the panel rows are written; the previous tile's reads of Pi are done
      for (uint32_t e = tid; e < NB * NB; e += DENSE_THREADS) {
        const uint32_t i = e / NB, t = e % NB;
        Pi[i][t] = A[(size_t)(ib + i) * n + kb + t];
      }
      for (uint32_t jb = kb + NB; jb <= ib; jb += NB) {
        __syncthreads();
        for (uint32_t e = tid; e < NB * NB; e += DENSE_THREADS) {
          const uint32_t j = e / NB, t = e % NB;
          Pj[j][t] = A[(size_t)(jb + j) * n + kb + t];
        }
        __syncthreads();
        for (uint32_t e = tid; e < NB * NB; e += DENSE_THREADS) {
          const uint32_t i = e / NB, j = e % NB;
          if (jb + j <= ib + i) {
            double sum = 0;
#pragma unroll 8
            for (int t = 0; t < NB; t++) sum += Pi[i][t] * Pj[j][t];
            A[(size_t)(ib + i) * n + jb + j] -= sum;
          }
        }
      }
    }
  }
  for (uint32_t i = tid; i < u; i += DENSE_THREADS) y[i] = A[(size_t)u * n + i];
  dense_substitute(A, n, u, y, L);
  double* x = vx(q);
  for (uint32_t i = tid; i < mb; i += DENSE_THREADS) x[dense_row(g, i)] = y[i];
  if (tid == 0) {
    s[S_PCG_ACTIVE] = 0;
    s[S_DIRECT] = 1;
  }
}

// One correction of iterative refinement: x += S^-1 r from the stored factor, for each bundle whose
// direct solution is not yet certified (r from k_certify).
template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_dense_correct(Problem q) {
  __shared__ double L[NB][NB + 1];
  __shared__ double y[DENSE_MAX_ROWS];
  const uint32_t b = q.small_bundles[blockIdx.x];
  const double* s = bstate(q, b);
  if (s[S_ACTIVE] == 0 || s[S_DIRECT] != 1) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1, mb = dense_unknowns<SHARED>(g), tid = threadIdx.x;
  const double* A = q.dense + g.dense_off;
  const double* r = vr(q);
  for (uint32_t i = tid; i < u; i += DENSE_THREADS) y[i] = i < mb ? r[dense_row(g, i)] : 0.0;
  dense_substitute(A, n, u, y, L);
  double* x = vx(q);
  for (uint32_t i = tid; i < mb; i += DENSE_THREADS) x[dense_row(g, i)] += y[i];
}

// ------------------------------------------------------------------ large bundles
//
// A bundle of more than DENSE_MAX_CAMERAS cameras is factored by a sequence of launches, one
// 64-column panel at a time:
//   k_lfact_diag    the diagonal block L_kk (one workgroup: a step per column, the block's entries
//                   held in registers by fixed owner threads, one barrier per step) and its inverse
//                   X_kk = L_kk^-1 (the same scheme, a step per row), stored beside the array;
//   k_lfact_panel   the rows below as a product with the inverse, P = A X_kk^T (a workgroup per 64
//                   rows, 4x4 outputs per thread), and with it the forward substitution of the
//                   right-hand side, which is the array's last row: y_kb = X_kk w_kb, then
//                   w_i -= P_i . y_kb for the workgroup's own rows;
//   k_lfact_update  the panel's outer products subtracted from the trailing lower triangle (a
//                   workgroup per 64x64 tile, 4x4 outputs per thread).
// The back substitution runs a launch per block (k_lsolve_bwd): x_kb = X_kk^T y_kb, then
// y_i -= L_kb,i^T x_kb for the rows above. A correction (iterative refinement) also runs the
// forward substitution block by block (k_lsolve_fwd). Beside a large bundle's (u + 1)^2 array the
// dense words hold the forward-substituted vector (u + 1 words) and the u / 64 inverses X_kk.

constexpr int LNB = DENSE_LNB;

// A large bundle's dense array as the multi-launch kernels address it: A (n x n, row-major), the
// forward vector after it (n words), then the u / 64 inverses X_kk (64 x 64, row-major), and the
// bundle's state words. The host passes the pointers, so that a kernel's first loads of the array
// go out together with its load of the state words.
struct LargeDense {
  double* A;
  double* state;
  uint32_t n;
};

__device__ __forceinline__ double* large_vector(const LargeDense& d) {
  return d.A + (size_t)d.n * d.n;
}
__device__ __forceinline__ double* large_inverse(const LargeDense& d, uint32_t kb) {
  return large_vector(d) + d.n + (size_t)kb * LNB;
}

// Word 41 (failed pivot) is set by k_lfact_diag; mode 0 is the first solve (the bundle's PCG set-up
// still pending), mode 1 a correction (an uncertified direct solution).
// The words are read together (no short-circuit), in one round trip.
__device__ __forceinline__ bool large_skip(const double* s, int mode) {
  const double active = s[S_ACTIVE], pcg = s[S_PCG_ACTIVE], bad = s[S_DENSE_BAD],
               direct = s[S_DIRECT];
  if (mode == 0) return (active == 0) | (pcg == 0) | (bad != 0);
  return (active == 0) | (direct != 1);
}

// Identity padding rows, a zero right-hand side beyond the unknowns, the failure word cleared.
template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_lfact_begin(Problem q, uint32_t b) {
  double* s = bstate(q, b);
  if (s[S_ACTIVE] == 0 || s[S_PCG_ACTIVE] == 0) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1, mb = dense_unknowns<SHARED>(g);
  double* A = q.dense + g.dense_off;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (uint32_t i = mb; i < u; i++)
    for (uint32_t j = gid; j <= i; j += stride) A[(size_t)i * n + j] = j == i ? 1.0 : 0.0;
  for (uint32_t j = mb + gid; j < u; j += stride) A[(size_t)u * n + j] = 0;
  if (gid == 0) s[S_DENSE_BAD] = 0;
}

// (row, column) of entry e of a lower triangle enumerated row by row (e = i (i + 1) / 2 + j).
__device__ __forceinline__ void lower_entry(uint32_t e, uint32_t& i, uint32_t& j) {
  uint32_t r = (uint32_t)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > e) r--;
  while ((r + 1) * (r + 2) / 2 <= e) r++;
  i = r;
  j = e - r * (r + 1) / 2;
}

// One doubling level of the inverse of the 64 x 64 lower-triangular L (in C, zero above the
// diagonal) into X (zero above the diagonal; its diagonal blocks of size H already inverted): for
// each pair of diagonal H x H blocks at p = 2 H b, X21 = -X22 (L21 X11). T = L21 X11 is kept in C's
// upper triangle (T_rc at C[p + c][p + H + r]), which nothing else reads. The 32 H outputs of a
// phase are spread over the 256 threads, H / 8 per thread from H = 8 up, their sums interleaved.
template <int H>
__device__ __forceinline__ void inverse_level(double (*C)[LNB + 1], double (*X)[LNB + 1]) {
  constexpr uint32_t OUT = 32 * H, Q = OUT > DENSE_THREADS ? OUT / DENSE_THREADS : 1;
  const uint32_t tid = threadIdx.x;
  const bool on = OUT >= DENSE_THREADS || tid < OUT;
  uint32_t r[Q], c[Q], p[Q];
#pragma unroll
  for (uint32_t q = 0; q < Q; q++) {
    const uint32_t o = tid + DENSE_THREADS * q;
    r[q] = o % (H * H) / H;
    c[q] = o % H;
    p[q] = 2 * H * (o / (H * H));
  }
  double sum[Q];
  if (on) {
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) sum[q] = 0;
#pragma unroll
    for (int m = 0; m < H; m++)
#pragma unroll
      for (uint32_t q = 0; q < Q; q++)
        sum[q] += C[p[q] + H + r[q]][p[q] + m] * X[p[q] + m][p[q] + c[q]];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) C[p[q] + c[q]][p[q] + H + r[q]] = sum[q];
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) sum[q] = 0;
#pragma unroll
    for (int m = 0; m < H; m++)
#pragma unroll
      for (uint32_t q = 0; q < Q; q++)
        sum[q] += X[p[q] + H + r[q]][p[q] + H + m] * C[p[q] + c[q]][p[q] + H + m];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) X[p[q] + H + r[q]][p[q] + c[q]] = -sum[q];
  }
  __syncthreads();
}

// 1 / v: the hardware estimate and one Newton step (relative error a few ulp).
__device__ __forceinline__ double recip(double v) {
  const double r = __builtin_amdgcn_rcp(v);
  return fma(r, fma(-v, r, 1.0), r);
}

// LDS words of the diagonal-block routine (two 64 x 65 arrays, the published columns of two block
// steps, their pivot coefficients, 1 / sqrt of the pivots); the update kernel's two panel tiles
// share the same buffer.
constexpr int DIAG_LDS = 2 * LNB * (LNB + 1) + 2 * 4 * LNB + 2 * 16 + LNB;

// The diagonal block of panel kb, L_kk and X_kk = L_kk^-1, by the 256 threads of one workgroup.
// Thread t holds the 4x4 block a (rows 4 (t / 16).., columns 4 (t % 16)..) of the whole symmetric
// 64 x 64 block in registers. Cholesky in 16 block steps of 4 columns: the threads owning columns
// k..k+3 publish them (zero above row k), the owner of the 4x4 pivot block publishes its
// elimination coefficients; after one barrier every thread eliminates its own rows' and columns'
// 4 values and subtracts the rank-4 product from its 16 entries. The eliminated columns scaled by
// 1 / sqrt(pivot) are L. The inverse is built by doubling (inverse_level), six levels of two
// products each.
__device__ void diag_block(const LargeDense& d, uint32_t kb, double a[4][4], double* smem) {
  auto C = reinterpret_cast<double(*)[LNB + 1]>(smem);  // row k: L's column k, then L
  auto X = reinterpret_cast<double(*)[LNB + 1]>(smem + LNB * (LNB + 1));
  double* pre = smem + 2 * LNB * (LNB + 1);  // [2][4][64]: the published columns of a block step
  double* coef = pre + 2 * 4 * LNB;          // [2][16]: f10 f20 f21 f30 f31 f32, then 1 / pivot
  double* rroot = coef + 2 * 16;
  const uint32_t n = d.n, tid = threadIdx.x, i0 = (tid / 16) * 4, j0 = (tid % 16) * 4;
  double* A = d.A + (size_t)kb * n + kb;
  for (uint32_t k = 0; k < (uint32_t)LNB; k += 4) {
    double* P = pre + ((k / 4) & 1) * 4 * LNB;
    double* F = coef + ((k / 4) & 1) * 16;
    if (j0 == k) {
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int c = 0; c < 4; c++) P[c * LNB + i0 + x] = i0 + x >= k ? a[x][c] : 0.0;
      if (i0 == k) {
        // the pivot block's elimination: e[m][l] = a_ml - sum_{l' < l} e[m][l'] e[l][l'] / e[l'][l']
        double e[4][4], g[4];
#pragma unroll
        for (int l = 0; l < 4; l++) {
#pragma unroll
          for (int m = l; m < 4; m++) {
            double v = a[m][l];
#pragma unroll
            for (int t = 0; t < l; t++) v -= e[m][t] * e[l][t] * g[t];
            e[m][l] = v;
          }
          g[l] = recip(e[l][l]);
        }
        F[0] = e[1][0] * g[0];
        F[1] = e[2][0] * g[0];
        F[2] = e[2][1] * g[1];
        F[3] = e[3][0] * g[0];
        F[4] = e[3][1] * g[1];
        F[5] = e[3][2] * g[2];
#pragma unroll
        for (int l = 0; l < 4; l++) F[6 + l] = g[l];
      }
    }
    __syncthreads();
    const double f10 = F[0], f20 = F[1], f21 = F[2], f30 = F[3], f31 = F[4], f32 = F[5];
    // the 8 vectors (4 rows, 4 columns) eliminated level by level, so that each level's 8 chains
    // interleave; then the rank-4 product, one column level at a time (16 independent updates)
    double v[8][4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int x = 0; x < 4; x++) {
        v[x][c] = P[c * LNB + i0 + x];
        v[4 + x][c] = P[c * LNB + j0 + x];
      }
#pragma unroll
    for (int e = 0; e < 8; e++) v[e][1] -= v[e][0] * f10;
#pragma unroll
    for (int e = 0; e < 8; e++) v[e][2] -= v[e][0] * f20 + v[e][1] * f21;
#pragma unroll
    for (int e = 0; e < 8; e++) v[e][3] -= v[e][0] * f30 + v[e][1] * f31 + v[e][2] * f32;
#pragma unroll
    for (int l = 0; l < 4; l++) {
      const double g = F[6 + l];
      double cj[4];
#pragma unroll
      for (int y = 0; y < 4; y++) cj[y] = v[4 + y][l] * g;
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) a[x][y] -= v[x][l] * cj[y];
    }
    if (j0 == k) {
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int l = 0; l < 4; l++) C[k + l][i0 + x] = i0 + x >= k + l ? v[x][l] : 0.0;
    }
  }
  __syncthreads();
  if (tid < (uint32_t)LNB) {
    const double piv = C[tid][tid];
    rroot[tid] = 1.0 / sqrt(fabs(piv));
    if (!(piv > 0) || !(piv < INFINITY)) d.state[S_DENSE_BAD] = 1;
  }
  __syncthreads();
  double l[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) l[x][y] = C[j0 + y][i0 + x] * rroot[j0 + y];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
      if (i0 + x >= j0 + y) A[(size_t)(i0 + x) * n + j0 + y] = l[x][y];
  __syncthreads();  // C is read: it takes L (row-major, zero above the diagonal)
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      C[i0 + x][j0 + y] = l[x][y];
      X[i0 + x][j0 + y] = i0 + x == j0 + y ? 1.0 / l[x][y] : 0.0;
    }
  __syncthreads();
  inverse_level<1>(C, X);
  inverse_level<2>(C, X);
  inverse_level<4>(C, X);
  inverse_level<8>(C, X);
  inverse_level<16>(C, X);
  inverse_level<32>(C, X);
  double* Xg = large_inverse(d, kb);
  for (uint32_t e = tid; e < (uint32_t)(LNB * LNB); e += DENSE_THREADS) {
    const uint32_t i = e / LNB, j = e % LNB;
    Xg[e] = j <= i ? X[i][j] : 0.0;
  }
}

// The first diagonal block (the later ones are factored by the update launch before them).
__global__ __launch_bounds__(DENSE_THREADS) void k_lfact_diag(LargeDense d, uint32_t kb) {
  __shared__ double smem[DIAG_LDS];
  const bool skip = large_skip(d.state, 0);
  const uint32_t n = d.n, tid = threadIdx.x, i0 = (tid / 16) * 4, j0 = (tid % 16) * 4;
  const double* A = d.A + (size_t)kb * n + kb;
  double a[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint32_t i = i0 + x, j = j0 + y;
      a[x][y] = A[(size_t)max(i, j) * n + min(i, j)];
    }
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) settle(a[x][y]);
  if (skip) return;
  diag_block(d, kb, a, smem);
}

// The 64 values X v (TRANSPOSE: X^T v) for a 64 x 64 X with row stride LD and v in LDS, into out
// (LDS): four threads per output, partial sums of 16 terms combined by shuffles. 256 threads.
template <bool TRANSPOSE, int LD>
__device__ __forceinline__ void block_matvec(const double* X, const double* v, double* out) {
  const uint32_t tid = threadIdx.x, o = tid / 4, part = tid % 4;
  double sum = 0;
#pragma unroll
  for (int k = 0; k < LNB / 4; k++) {
    const uint32_t c = part * (LNB / 4) + k;
    sum += (TRANSPOSE ? X[c * LD + o] : X[o * LD + c]) * v[c];
  }
  sum += __shfl_xor(sum, 1, 64);
  sum += __shfl_xor(sum, 2, 64);
  if (part == 0) out[o] = sum;
}

typedef double dense_d4 __attribute__((ext_vector_type(4)));

// acc[b] += the 16 x 16 block (rows 16 w.., columns 16 b..) of Am Bm^T over K = 64, for wave w of
// the workgroup, by FP64 matrix-core products (16x16x4): Am, Bm are 64 x 64 in LDS (row stride
// LNB + 1). Operand lane l holds row l % 16, column l / 16 of each 16 x 4 slice; register e of
// acc[b] in lane l is entry (16 w + l / 16 + 4 e, 16 b + l % 16).
__device__ __forceinline__ void strip_product(const double (*Am)[LNB + 1],
                                              const double (*Bm)[LNB + 1], dense_d4 acc[4]) {
  const uint32_t lane = threadIdx.x % 64, w = threadIdx.x / 64, r = lane % 16, kq = lane / 16;
#pragma unroll 4
  for (int kk = 0; kk < LNB / 4; kk++) {
    const double a = Am[16 * w + r][4 * kk + kq];
    double b[4];
#pragma unroll
    for (int c = 0; c < 4; c++) b[c] = Bm[16 * c + r][4 * kk + kq];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[c], acc[c], 0, 0, 0);
  }
}

// The panel below block kb as P = A X_kk^T (a workgroup per 64 rows, a 16-row strip per wave on
// the matrix cores), and
// the forward substitution of block kb: y_kb = X_kk w_kb (workgroup 0 stores it; the last block
// has no rows below and runs only this), w_i -= P_i . y_kb for the workgroup's rows.
__global__ __launch_bounds__(DENSE_THREADS) void k_lfact_panel(LargeDense d, uint32_t kb) {
  __shared__ double Xs[LNB][LNB + 1];
  __shared__ double Ps[LNB][LNB + 1];
  __shared__ double wk[LNB], y[LNB];
  const bool skip = large_skip(d.state, 0);
  const uint32_t n = d.n, u = n - 1, tid = threadIdx.x;
  double* A = d.A;
  double* w = A + (size_t)u * n;
  const double* Xg = large_inverse(d, kb);
  const uint32_t r0 = kb + LNB + LNB * blockIdx.x;
  const bool rows = r0 < u;
  // every load issued before the first wait: X_kk, w_kb and the 64 panel rows (thread tid: column
  // tid % 64 of rows tid / 64 + 4 r)
  const uint32_t ti = tid / LNB, tc = tid % LNB;
  double xv[LNB / 4], pv[LNB / 4];
#pragma unroll
  for (int r = 0; r < LNB / 4; r++) {
    xv[r] = Xg[(ti + 4 * r) * LNB + tc];
    pv[r] = rows ? A[(size_t)(r0 + ti + 4 * r) * n + kb + tc] : 0.0;
  }
  double wv = w[kb + tc];
#pragma unroll
  for (int r = 0; r < LNB / 4; r++) {
    settle(xv[r]);
    settle(pv[r]);
  }
  settle(wv);
  if (skip) return;
  if (tid < (uint32_t)LNB) wk[tid] = wv;
#pragma unroll
  for (int r = 0; r < LNB / 4; r++) {
    Xs[ti + 4 * r][tc] = xv[r];
    Ps[ti + 4 * r][tc] = pv[r];
  }
  __syncthreads();
  block_matvec<false, LNB + 1>(&Xs[0][0], wk, y);  // y_kb = X_kk w_kb
  __syncthreads();
  if (blockIdx.x == 0 && tid < (uint32_t)LNB) large_vector(d)[kb + tid] = y[tid];
  if (!rows) return;
  dense_d4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = dense_d4{0, 0, 0, 0};
  strip_product(Ps, Xs, acc);
  const uint32_t lane = tid % 64, row0 = 16 * (tid / 64) + lane / 16, col = lane % 16;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const uint32_t i = row0 + 4 * e;
    double part = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      A[(size_t)(r0 + i) * n + kb + 16 * c + col] = acc[c][e];
      part += acc[c][e] * y[16 * c + col];
    }
    // the 16 lanes of row i (one 16-lane group) hold its 64 columns
    for (int o = 8; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    if (col == 0) w[r0 + i] -= part;
  }
}

// Tile (ti, tj) of logical index L over the T x T lower triangle of tiles, in the order of 16 x 16
// super-tiles (row-major over the lower triangle of super-tiles, row-major inside each): a run of
// consecutive indices, which one XCD receives (xcd_contiguous), works through one super-tile at a
// time and finds its 16 + 16 panel tiles in that XCD's L2.
constexpr uint32_t SUPER_TILE = 16;
__device__ __forceinline__ void super_tile(uint32_t L, uint32_t T, uint32_t& ti, uint32_t& tj) {
  for (uint32_t I = 0;; I++) {
    const uint32_t r0 = I * SUPER_TILE, h = min(SUPER_TILE, T - r0);
    const uint32_t full = h * SUPER_TILE * I, row = full + h * (h + 1) / 2;
    if (L < row) {
      if (L < full) {
        const uint32_t rem = L % (h * SUPER_TILE);
        ti = r0 + rem / SUPER_TILE;
        tj = L / (h * SUPER_TILE) * SUPER_TILE + rem % SUPER_TILE;
      } else {
        uint32_t r, c;
        lower_entry(L - full, r, c);
        ti = r0 + r;
        tj = r0 + c;
      }
      return;
    }
    L -= row;
  }
}

// The trailing lower triangle from row and column base = kb + 64 panels less the outer products
// of the `panels` (1 or 2) solved panels at columns kb..base-1: a workgroup per 64x64 tile
// (super_tile order), a 16-row strip per wave on the matrix cores, the panels one after the
// other into the same accumulators, so that each output tile is read and written once for both.
// `column`: only the first tile column (rows base.., columns base..base+63), one panel. The
// workgroup of tile (0, 0), the next diagonal block (blockIdx 0, dispatched first), goes on to
// factor it.
__global__ __launch_bounds__(DENSE_THREADS) void k_lfact_update(LargeDense d, uint32_t kb,
                                                                 uint32_t panels, uint32_t column) {
  __shared__ double smem[DIAG_LDS];
  auto Pi = reinterpret_cast<double(*)[LNB + 1]>(smem);
  auto Pj = reinterpret_cast<double(*)[LNB + 1]>(smem + LNB * (LNB + 1));
  const bool skip = large_skip(d.state, 0);
  const uint32_t n = d.n, tid = threadIdx.x, base = kb + LNB * panels;
  double* A = d.A;
  uint32_t ti, tj;
  if (column) {
    ti = xcd_contiguous(blockIdx.x, gridDim.x);
    tj = 0;
  } else {
    super_tile(xcd_contiguous(blockIdx.x, gridDim.x), (n - 1 - base) / LNB, ti, tj);
  }
  const uint32_t ib = base + LNB * ti, jb = base + LNB * tj;
  // the thread's outputs in the matrix-core layout: out[c][e] is entry (row0 + 4 e, 16 c + col)
  const uint32_t lane = tid % 64, row0 = 16 * (tid / 64) + lane / 16, col = lane % 16;
  // all loads in flight at once: the first panel's two tiles (thread tid: column tid % 64 of rows
  // tid / 64 + 4 r) and the thread's 16 outputs (entries above the diagonal of a diagonal tile
  // are read and left unchanged)
  double pa[LNB / 4], pb[LNB / 4], out[4][4];
  const uint32_t sr = tid / LNB, sc = tid % LNB;
#pragma unroll
  for (int r = 0; r < LNB / 4; r++) {
    pa[r] = A[(size_t)(ib + sr + 4 * r) * n + kb + sc];
    pb[r] = A[(size_t)(jb + sr + 4 * r) * n + kb + sc];
  }
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int e = 0; e < 4; e++) out[c][e] = A[(size_t)(ib + row0 + 4 * e) * n + jb + 16 * c + col];
#pragma unroll
  for (int r = 0; r < LNB / 4; r++) {
    settle(pa[r]);
    settle(pb[r]);
  }
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int e = 0; e < 4; e++) settle(out[c][e]);
  if (skip) return;
  dense_d4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = dense_d4{0, 0, 0, 0};
  for (uint32_t p = 0; p < panels; p++) {
    if (p) __syncthreads();  // the previous panel's tiles are read
#pragma unroll
    for (int r = 0; r < LNB / 4; r++) {
      Pi[sr + 4 * r][sc] = pa[r];
      Pj[sr + 4 * r][sc] = pb[r];
    }
    if (p + 1 < panels) {
      // the next panel's tiles, in flight during this panel's products
      const uint32_t pc = kb + LNB * (p + 1);
#pragma unroll
      for (int r = 0; r < LNB / 4; r++) {
        pa[r] = A[(size_t)(ib + sr + 4 * r) * n + pc + sc];
        pb[r] = A[(size_t)(jb + sr + 4 * r) * n + pc + sc];
      }
    }
    __syncthreads();
    strip_product(Pi, Pj, acc);
  }
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const uint32_t i = row0 + 4 * e, j = 16 * c + col;
      out[c][e] -= acc[c][e];
      if (ti != tj || j <= i) A[(size_t)(ib + i) * n + jb + j] = out[c][e];
    }
  if (ti != 0 || tj != 0) return;
  // tile (0, 0) is the next panel's diagonal block, final now: this workgroup factors it, its lower
  // triangle mirrored through LDS into the whole symmetric block in the routine's layout
  __syncthreads();  // the panel tiles are read
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const uint32_t i = row0 + 4 * e, j = 16 * c + col;
      if (j <= i) Pi[i][j] = Pi[j][i] = out[c][e];
    }
  __syncthreads();
  const uint32_t i0 = (tid / 16) * 4, j0 = (tid % 16) * 4;
  double a4[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) a4[x][y] = Pi[i0 + x][j0 + y];
  __syncthreads();
  diag_block(d, base, a4, smem);
}

// Forward step of block kb of a correction: y_kb = X_kk w_kb into the forward vector (workgroup 0),
// and w_i -= L_i,kb y_kb for the rows i below, w the array's last row.
__global__ __launch_bounds__(DENSE_THREADS) void k_lsolve_fwd(LargeDense d, uint32_t kb) {
  __shared__ double wk[LNB], y[LNB];
  const bool skip = large_skip(d.state, 1);
  const uint32_t n = d.n, u = n - 1, tid = threadIdx.x;
  double* A = d.A;
  double* w = A + (size_t)u * n;
  double wv = w[kb + tid % LNB];
  settle(wv);
  if (skip) return;
  if (tid < (uint32_t)LNB) wk[tid] = wv;
  __syncthreads();
  block_matvec<false, LNB>(large_inverse(d, kb), wk, y);
  __syncthreads();
  if (blockIdx.x == 0 && tid < (uint32_t)LNB) large_vector(d)[kb + tid] = y[tid];
  const uint32_t i = kb + LNB + blockIdx.x * DENSE_THREADS + tid;
  if (i >= u) return;
  const double* row = A + (size_t)i * n + kb;
  double sum = 0;
#pragma unroll 8
  for (int t = 0; t < LNB; t++) sum += row[t] * y[t];
  w[i] -= sum;
}

// Backward step of block kb: x_kb = X_kk^T y_kb into w (workgroup 0), and y_i -= L_kb,i^T x_kb for
// the rows i above, y the forward vector.
__global__ __launch_bounds__(DENSE_THREADS) void k_lsolve_bwd(LargeDense d, uint32_t kb,
                                                               int mode) {
  __shared__ double yk[LNB], x[LNB];
  const bool skip = large_skip(d.state, mode);
  const uint32_t n = d.n, u = n - 1, tid = threadIdx.x;
  double* A = d.A;
  double* w = A + (size_t)u * n;
  double* yv = large_vector(d);
  // the thread's column of the block row, loaded with the rest before the first wait
  const uint32_t i = blockIdx.x * DENSE_THREADS + tid;
  const uint32_t ic = min(i, kb ? kb - 1 : 0u);
  double col[LNB];
#pragma unroll
  for (int t = 0; t < LNB; t++) col[t] = A[(size_t)(kb + t) * n + ic];
  double yvk = yv[kb + tid % LNB];
#pragma unroll
  for (int t = 0; t < LNB; t++) settle(col[t]);
  settle(yvk);
  if (skip) return;
  if (tid < (uint32_t)LNB) yk[tid] = yvk;
  __syncthreads();
  block_matvec<true, LNB>(large_inverse(d, kb), yk, x);
  __syncthreads();
  if (blockIdx.x == 0 && tid < (uint32_t)LNB) w[kb + tid] = x[tid];
  if (i >= kb) return;
  double sum = 0;
#pragma unroll
  for (int t = 0; t < LNB; t++) sum += col[t] * x[t];
  yv[i] -= sum;
}

// The residual r into w before a correction (zero in the padding).
template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_lsolve_in(Problem q, uint32_t b) {
  const double* s = bstate(q, b);
  if (large_skip(s, 1)) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1, mb = dense_unknowns<SHARED>(g);
  double* w = q.dense + g.dense_off + (size_t)u * n;
  const double* r = vr(q);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < u; i += gridDim.x * blockDim.x)
    w[i] = i < mb ? r[dense_row(g, i)] : 0.0;
}

// x = w (mode 0: the solution, the bundle's PCG set aside) or x += w (mode 1: a correction).
template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_lsolve_out(Problem q, uint32_t b, int mode) {
  double* s = bstate(q, b);
  if (large_skip(s, mode)) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1, mb = dense_unknowns<SHARED>(g);
  const double* w = q.dense + g.dense_off + (size_t)u * n;
  double* x = vx(q);
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t i = gid; i < mb; i += gridDim.x * blockDim.x) {
    if (mode == 0)
      x[dense_row(g, i)] = w[i];
    else
      x[dense_row(g, i)] += w[i];
  }
  if (mode == 0 && gid == 0) {
    s[S_PCG_ACTIVE] = 0;
    s[S_DIRECT] = 1;
  }
}

// The certificate of a direct solution, per bundle with word 38 = 1, after the implicit sweeps of
// x: r = b - S x with the matrix-free FP64 operator. |r|^2 <= tol^2 |r0|^2: certified (words 38 =
// 2, 24, 22, 39). Otherwise r stays in the residual vector for k_dense_correct, or on the final
// pass PCG restarts from x (z = M^-1 r, p = z; word 38 cleared). A non-finite residual restarts PCG
// from zero at once, as the reference restarts from its original zero iterate.
template <bool SHARED>
__global__ __launch_bounds__(CG_THREADS) void k_certify(Problem q, int final_pass, double tol) {
  __shared__ double lds[2 * (CG_THREADS / 64)];
  const uint32_t bi = q.direct_bundles[blockIdx.x];
  const Bundle& g = q.bundles[bi];
  double* s = bstate(q, bi);
  if (s[S_ACTIVE] == 0 || s[S_DIRECT] != 1) return;
  const uint32_t tid = threadIdx.x;
  const double r0 = s[S_R0];
  double* x = vx(q);
  double* r = vr(q);
  double* p = vp(q);
  double* S = vs(q);
  double* z = vz(q);
  (void)schur_finish<SHARED>(q, g, s, x, S, lds);
  __syncthreads();
  const uint32_t r_lo = 6 * g.cam0, r_hi = 6 * g.cam1;
  double res = 0;
  for (uint32_t i = r_lo + tid; i < r_hi; i += CG_THREADS) {
    const double ri = q.camRhs[i] - S[i];
    r[i] = ri;
    res += ri * ri;
  }
  if (SHARED && tid == 0) {
    const uint32_t f = g.focal_row;
    r[f] = s[S_FOCAL_RHS] - S[f];
    res += r[f] * r[f];
  }
  res = block_sum(res, lds);
  const bool finite = res < INFINITY;
  if (finite && res <= tol * tol * r0) {
    if (tid == 0) {
      s[S_R] = res;
      s[S_CONVERGED] = 1;
      s[S_DIRECT] = 2;
      s[S_DIRECT_SOLVES] += 1;
    }
    return;
  }
  if (finite && !final_pass) {
    if (tid == 0) s[S_R] = res;
    return;
  }
  __syncthreads();  // every row of r before it is replaced or read by the block products
  if (!finite) {
    for (uint32_t i = r_lo + tid; i < r_hi; i += CG_THREADS) {
      x[i] = 0;
      r[i] = q.camRhs[i];
    }
    if (SHARED && tid == 0) {
      x[g.focal_row] = 0;
      r[g.focal_row] = s[S_FOCAL_RHS];
    }
    res = r0;
    __syncthreads();
  }
  double rho = 0;
  for (uint32_t i = r_lo + tid; i < r_hi; i += CG_THREADS) {
    const double zi = block_row(q.camInv, i, r);
    z[i] = zi;
    p[i] = zi;
    rho += zi * r[i];
  }
  if (SHARED && tid == 0) {
    const uint32_t f = g.focal_row;
    z[f] = s[S_FOCAL_INV] * r[f];
    p[f] = z[f];
    rho += z[f] * r[f];
  }
  rho = block_sum(rho, lds);
  if (tid == 0) {
    s[S_R] = res;
    s[S_RHO] = rho;
    s[S_RHO_NEXT] = rho;
    s[S_DIRECT] = 0;
    s[S_STEP_LIMIT] = s[S_STEPS] + pcg_budget(q, g);  // PCG from x with the bundle's budget
    s[S_PASS_STOP] = q.tol * q.tol * s[S_R0];  // to the tolerance of the original system
    if (rho > 0)
      s[S_PCG_ACTIVE] = 1;
    else
      s[S_PCG_FAILED] = 1;
  }
}

}  // namespace tcvba
