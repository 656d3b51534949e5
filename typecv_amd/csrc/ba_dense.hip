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

// The workgroup's logical index such that each of the 8 XCDs (which receive workgroups round-robin
// in dispatch order) runs one contiguous range of the grid: the rows of one bundle then share that
// XCD's L2, where the bundle's G rows are read once per row camera.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t i, uint32_t n) {
#ifndef TCV_XCD_REMAP
#define TCV_XCD_REMAP 1
#endif
#if TCV_XCD_REMAP
  const uint32_t q = n / 8, r = n % 8, x = i % 8, k = i / 8;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
#else
  (void)n;
  return i;
#endif
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
    const uint32_t nc = ce - cc;
    for (uint32_t ld = 0; ld < lc; ld++) {
      const uint32_t dlo = span[2 * ld], dhi = span[2 * ld + 1];
      if (dlo == dhi) continue;
      double acc[36];
#pragma unroll
      for (int k = 0; k < 36; k++) acc[k] = 0;
      int hit = 0;
      for (uint32_t s = dlo + tid; s < dhi; s += DENSE_THREADS) {
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
        hit = 1;
        pair_block(q.camG, sc, s, acc);
      }
      if (!__syncthreads_or(hit)) continue;
      const double v = dense_block_sum<36>(acc, red);
      if (tid < 36) D[(size_t)(r0 + tid / 6) * n + 6 * ld + tid % 6] -= v;
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
// 64-column panel at a time: k_lfact_diag factors the diagonal block (one wave, lane per row),
// k_lfact_panel solves the rows below against it (a thread per row), k_lfact_update subtracts the
// panel's outer products from the trailing lower triangle (a workgroup per 64x64 tile, 4x4 outputs
// per thread). The substitution runs a launch per block and direction: every workgroup of a step
// solves the step's 64x64 triangular system itself (the same sums everywhere), workgroup 0 stores
// the result, and each workgroup then updates its own rows. The dense array keeps u + 1 more words
// after its (u + 1)^2 for the forward-substituted vector.

constexpr int LNB = DENSE_LNB;

// Word 41 (failed pivot) is set by k_lfact_diag; mode 0 is the first solve (the bundle's PCG set-up
// still pending), mode 1 a correction (an uncertified direct solution).
__device__ __forceinline__ bool large_skip(const double* s, int mode) {
  if (s[S_ACTIVE] == 0) return true;
  if (mode == 0) return s[S_PCG_ACTIVE] == 0 || s[S_DENSE_BAD] != 0;
  return s[S_DIRECT] != 1;
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

template <bool SHARED>
__global__ __launch_bounds__(64) void k_lfact_diag(Problem q, uint32_t b, uint32_t kb) {
  __shared__ double L[LNB][LNB + 1];
  double* s = bstate(q, b);
  if (large_skip(s, 0)) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, lane = threadIdx.x;
  double* A = q.dense + g.dense_off;
  for (uint32_t e = lane; e < LNB * LNB; e += 64) {
    const uint32_t i = e / LNB, j = e % LNB;
    L[i][j] = j <= i ? A[(size_t)(kb + i) * n + kb + j] : 0.0;
  }
  wave_sync();
  bool bad = false;
  for (uint32_t k = 0; k < (uint32_t)LNB; k++) {
    if (lane == k) {
      const double d = L[k][k];
      bad = !(d > 0) || !(d < INFINITY);
      L[k][k] = sqrt(fabs(d));
    }
    wave_sync();
    if (lane > k) L[lane][k] /= L[k][k];
    wave_sync();
    if (lane > k) {
      const double lik = L[lane][k];
      for (uint32_t j = k + 1; j <= lane; j++) L[lane][j] -= lik * L[j][k];
    }
    wave_sync();
  }
  if (__any(bad) && lane == 0) s[S_DENSE_BAD] = 1;
  for (uint32_t e = lane; e < LNB * LNB; e += 64) {
    const uint32_t i = e / LNB, j = e % LNB;
    if (j <= i) A[(size_t)(kb + i) * n + kb + j] = L[i][j];
  }
}

template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_lfact_panel(Problem q, uint32_t b, uint32_t kb) {
  __shared__ double L[LNB][LNB + 1];
  const double* s = bstate(q, b);
  if (large_skip(s, 0)) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1;
  double* A = q.dense + g.dense_off;
  for (uint32_t e = threadIdx.x; e < LNB * LNB; e += DENSE_THREADS) {
    const uint32_t i = e / LNB, j = e % LNB;
    L[i][j] = j <= i ? A[(size_t)(kb + i) * n + kb + j] : 0.0;
  }
  __syncthreads();
  const uint32_t i = kb + LNB + blockIdx.x * DENSE_THREADS + threadIdx.x;
  if (i >= u) return;
  double* row = A + (size_t)i * n + kb;
  double a[LNB];
#pragma unroll
  for (int j = 0; j < LNB; j++) a[j] = row[j];
#pragma unroll
  for (int j = 0; j < LNB; j++) {
    double v = a[j];
#pragma unroll
    for (int t = 0; t < j; t++) v -= a[t] * L[j][t];
    a[j] = v / L[j][j];
    // keep the scheduler from hoisting the LDS reads of later rows
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int j = 0; j < LNB; j++) row[j] = a[j];
}

// Tile index t of the lower triangle of T x T tiles: (row, column) with column <= row.
__device__ __forceinline__ void lower_tile(uint32_t t, uint32_t& ti, uint32_t& tj) {
  uint32_t r = (uint32_t)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > t) r--;
  while ((r + 1) * (r + 2) / 2 <= t) r++;
  ti = r;
  tj = t - r * (r + 1) / 2;
}

template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_lfact_update(Problem q, uint32_t b,
                                                                 uint32_t kb) {
  __shared__ double Pi[LNB][LNB + 1];
  __shared__ double Pj[LNB][LNB + 1];
  const double* s = bstate(q, b);
  if (large_skip(s, 0)) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n;
  double* A = q.dense + g.dense_off;
  uint32_t ti, tj;
  lower_tile(blockIdx.x, ti, tj);
  const uint32_t ib = kb + LNB + LNB * ti, jb = kb + LNB + LNB * tj;
  for (uint32_t e = threadIdx.x; e < LNB * LNB; e += DENSE_THREADS) {
    const uint32_t i = e / LNB, t = e % LNB;
    Pi[i][t] = A[(size_t)(ib + i) * n + kb + t];
    Pj[i][t] = A[(size_t)(jb + i) * n + kb + t];
  }
  __syncthreads();
  const uint32_t i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[a][c] = 0;
  for (int t = 0; t < LNB; t++) {
    double x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      x[a] = Pi[i0 + a][t];
      y[a] = Pj[j0 + a][t];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[a][c] += x[a] * y[c];
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++)
      if (ti != tj || j0 + c <= i0 + a) A[(size_t)(ib + i0 + a) * n + jb + j0 + c] -= acc[a][c];
}

// Forward step of block kb: y_kb = L_kk^-1 w_kb into the vector after the array (workgroup 0),
// and w_i -= L_i,kb y_kb for the rows i below, w the array's last row.
template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_lsolve_fwd(Problem q, uint32_t b, uint32_t kb,
                                                               int mode) {
  __shared__ double L[LNB][LNB + 1];
  __shared__ double y[LNB];
  const double* s = bstate(q, b);
  if (large_skip(s, mode)) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1, tid = threadIdx.x;
  double* A = q.dense + g.dense_off;
  double* w = A + (size_t)u * n;
  double* yv = A + (size_t)n * n;
  for (uint32_t e = tid; e < LNB * LNB; e += DENSE_THREADS) {
    const uint32_t i = e / LNB, j = e % LNB;
    L[i][j] = j <= i ? A[(size_t)(kb + i) * n + kb + j] : 0.0;
  }
  if (tid < (uint32_t)LNB) y[tid] = w[kb + tid];
  __syncthreads();
  if (tid < 64) {
    for (uint32_t j = 0; j < (uint32_t)LNB; j++) {
      if (tid == j) y[j] /= L[j][j];
      wave_sync();
      if (tid > j) y[tid] -= L[tid][j] * y[j];
      wave_sync();
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid < (uint32_t)LNB) yv[kb + tid] = y[tid];
  const uint32_t i = kb + LNB + blockIdx.x * DENSE_THREADS + tid;
  if (i >= u) return;
  const double* row = A + (size_t)i * n + kb;
  double sum = 0;
#pragma unroll 8
  for (int t = 0; t < LNB; t++) sum += row[t] * y[t];
  w[i] -= sum;
}

// Backward step of block kb: x_kb = L_kk^-T y_kb into w (workgroup 0), and y_i -= L_kb,i^T x_kb for
// the rows i above.
template <bool SHARED>
__global__ __launch_bounds__(DENSE_THREADS) void k_lsolve_bwd(Problem q, uint32_t b, uint32_t kb,
                                                               int mode) {
  __shared__ double L[LNB][LNB + 1];
  __shared__ double x[LNB];
  const double* s = bstate(q, b);
  if (large_skip(s, mode)) return;
  const Bundle& g = q.bundles[b];
  const uint32_t n = g.dense_n, u = n - 1, tid = threadIdx.x;
  double* A = q.dense + g.dense_off;
  double* w = A + (size_t)u * n;
  double* yv = A + (size_t)n * n;
  for (uint32_t e = tid; e < LNB * LNB; e += DENSE_THREADS) {
    const uint32_t i = e / LNB, j = e % LNB;
    L[i][j] = j <= i ? A[(size_t)(kb + i) * n + kb + j] : 0.0;
  }
  if (tid < (uint32_t)LNB) x[tid] = yv[kb + tid];
  __syncthreads();
  if (tid < 64) {
    for (int j = LNB - 1; j >= 0; j--) {
      if (tid == (uint32_t)j) x[j] /= L[j][j];
      wave_sync();
      if (tid < (uint32_t)j) x[tid] -= L[j][tid] * x[j];
      wave_sync();
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid < (uint32_t)LNB) w[kb + tid] = x[tid];
  const uint32_t i = blockIdx.x * DENSE_THREADS + tid;
  if (i >= kb) return;
  double sum = 0;
#pragma unroll 8
  for (int t = 0; t < LNB; t++) sum += A[(size_t)(kb + t) * n + i] * x[t];
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
    if (rho > 0)
      s[S_PCG_ACTIVE] = 1;
    else
      s[S_PCG_FAILED] = 1;
  }
}

}  // namespace tcvba
