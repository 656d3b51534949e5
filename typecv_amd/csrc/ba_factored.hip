// ba_factored.hip — the Schur sweeps over Jacobian factors, FP32 (the mixed form) or FP64 (the FP64
// form); included by ba_engine.hip after ba_kernels.hip.
//
// An observation's cross block W = w Jc^T Jp (and the focal row w Jf^T Jp) is a product of its
// Jacobian rows, and those are fixed by three small pieces: the projection derivative
// D = d(u, v)/dq (2 x 3), the camera-frame point q and, on the shared route, Jf = d(u, v)/d log f.
// With s = sqrt(w), D' = s D and Jf' = s Jf:
//   Jp = D' R (R the camera's rotation),  Jc (omega, v) = D' [-[q]x | I]  (no v part for a
//   directional point, all zero for a fixed camera).
// The mixed form stores these KF = 9 (11 with the focal column) FP32 words per observation in place
// of the 18 (21) FP32 words of W, and forms every product in FP64 from the widened words and the
// camera's FP64 rotation (read from the scene):
//   point side  W^T v_c + wf v_f = R^T (D'^T g),  g = D' (v_omega x q + v_v) + Jf' v_f
//   camera side W y            : h = D' (R y);  omega: q x (D'^T h),  v: D'^T h,  focal: Jf' . h
// so a Schur product streams 2 (KF + 1) words per observation instead of 2 (K + 1). Sums keep the
// stored sweeps' orders (segment sums per point, slot order and wave butterflies per camera task).
// The FP64 form (but for the exact systems of the polish and the adjoint, whose exact-Hessian
// correction has no factored form) stores D' and Jf' in FP64, 64 (48) bytes per observation
// instead of the 168 (144) of its cross block, and recomputes q = R X + t from the pose and the
// point where the products use it (camera_point, the bits the passes compute): the products'
// gathers of the pose and the point come from cache, where stored q would be 24 more bytes of
// stream per observation.
#pragma once

namespace tcvba {

// The factors' camera-frame point q of a pose and a point (f[6..8]; the FP64 form's).
__device__ __forceinline__ void point_factor(const double* pose, const double* X, bool dir,
                                             double* f) {
  camera_point(pose, X[0], X[1], X[2], dir, f + 6);
}

// The factors of observation i (camera c, point p), widened: the stored words (FT: float or double;
// stored word k at w0 + k * stride, the point-major words: a tile's row length, the camera-major
// ones: 64), with q from the pose and the point in the FP64 form; RECOMPUTE: relinearised from the
// scene and rounded as the passes store them (the streamed route, the same bits; w0 unused).
template <class FT, bool SHARED, bool RECOMPUTE>
__device__ __forceinline__ void load_factors(const Problem& q, const FT* w0, uint32_t stride,
                                             uint32_t i, uint32_t c, uint32_t p, const Bundle& g,
                                             double fscale, double* f) {
  if (RECOMPUTE) {
    Lin L;
    const bool ok =
        linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L);
    double raw[11];
    cross_factors(L, ok, SHARED, raw);
#pragma unroll
    for (int k = 0; k < 11; k++) f[k] = (double)(FT)raw[k];
  } else {
#pragma unroll
    for (int k = 0; k < kf_words<FT>(SHARED); k++)
      f[kf_raw<FT>(k)] = (double)__builtin_nontemporal_load(w0 + k * stride);
    if (sizeof(FT) == 8)
      point_factor(q.scene + (size_t)c * 12, q.scene + (size_t)q.C * 12 + (size_t)p * 3,
                   directional_point(g, p), f);
  }
  if (!SHARED) f[9] = f[10] = 0;
}

// The point side of one observation: R^T D'^T (D' (v_omega x q [+ v_v]) [+ Jf' v_f]). Branch-free
// (a fixed camera's pose term is multiplied by 0), so the gathers of v, R and the camera's flag go
// out together instead of v waiting for the flag (146 against 151 us per sweep at 2,000 cameras).
__device__ __forceinline__ void factored_point_term(const double* f, const double* R,
                                                    const double* v, double vf, bool var,
                                                    bool dir, double t[3]) {
  const double vm = var ? 1.0 : 0.0, tm = dir ? 0.0 : 1.0;
  const double w0 = v[1] * f[8] - v[2] * f[7] + tm * v[3];
  const double w1 = v[2] * f[6] - v[0] * f[8] + tm * v[4];
  const double w2 = v[0] * f[7] - v[1] * f[6] + tm * v[5];
  const double g0 = vm * (f[0] * w0 + f[1] * w1 + f[2] * w2) + f[9] * vf;
  const double g1 = vm * (f[3] * w0 + f[4] * w1 + f[5] * w2) + f[10] * vf;
  const double a0 = f[0] * g0 + f[3] * g1, a1 = f[1] * g0 + f[4] * g1, a2 = f[2] * g0 + f[5] * g1;
#pragma unroll
  for (int j = 0; j < 3; j++) t[j] = R[j] * a0 + R[3 + j] * a1 + R[6 + j] * a2;
}

// The point sweep over the factors (the modes of sweep_tile: SCHUR on p, UPDATE on x).
template <class FT, bool SHARED, int MODE, bool RECOMPUTE>
__device__ __forceinline__ void sweep_tile_factored(const Problem& q, uint32_t t, uint32_t lane,
                                                    double (*st)[3]) {
  const uint32_t fp = q.tiles[t], ep = q.tiles[t + 1], b = q.tile_bundle[t];
  const double* sb = bstate(q, b);
  if (MODE == SWEEP_SCHUR && sb[S_PCG_ACTIVE] == 0) return;
  if (MODE == SWEEP_UPDATE && (sb[S_ACTIVE] == 0 || sb[S_PCG_FAILED] != 0)) return;
  const Bundle& g = q.bundles[b];
  if (q.n_unstored != 0 && !stores_w(g)) return;
  const double* v = MODE == SWEEP_SCHUR ? vp(q) : vx(q);
  const uint32_t* ps = point_start(q);
  const uint32_t s = ps[fp], e = ps[ep];
  const uint32_t my = fp + lane;
  const bool has = my < ep;
  const uint32_t mys = has ? ps[my] : 0, mye = has ? ps[my + 1] : 0;
  const double vf = SHARED ? v[6 * q.C + b] : 0.0;
  const double fscale = RECOMPUTE && SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  double acc[3] = {0, 0, 0};
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    if (i < e) {
      const uint32_t c = obs_camera(q)[i], pt = obs_point(q)[i];
      const bool var = variable_camera(q, c), dir = directional_point(g, pt);
      double f[11], R[9], vc[6];
      const FT* w0 = RECOMPUTE ? nullptr : tile_word<FT>(q.Wp, kf_words<FT>(SHARED), s, e - s, 0, i);
      load_factors<FT, SHARED, RECOMPUTE>(q, w0, e - s, i, c, pt, g, fscale, f);
      const double* pose = q.scene + (size_t)c * 12;
#pragma unroll
      for (int k = 0; k < 9; k++) R[k] = pose[k];
#pragma unroll
      for (int k = 0; k < 6; k++) vc[k] = v[c * 6 + k];
      double tj[3];
      factored_point_term(f, R, vc, vf, var, dir, tj);
      st[lane][0] = tj[0];
      st[lane][1] = tj[1];
      st[lane][2] = tj[2];
    }
    wave_sync();
    if (has) {
      const uint32_t lo = mys > base ? mys : base, hi = mye < base + TILE ? mye : base + TILE;
      segment_sum(st, base, lo, hi, acc);
    }
    wave_sync();
  }
  if (!has) return;
  const double* pb = q.pblk + (size_t)my * 12;
  if (MODE == SWEEP_SCHUR) {
    double* y = vy(q) + (size_t)my * 3;
    for (int j = 0; j < 3; j++) y[j] = pb[j * 3] * acc[0] + pb[j * 3 + 1] * acc[1] + pb[j * 3 + 2] * acc[2];
  } else {
    const double r0 = pb[9] - acc[0], r1 = pb[10] - acc[1], r2 = pb[11] - acc[2];
    const size_t at = (size_t)q.C * 12 + (size_t)my * 3;
    for (int j = 0; j < 3; j++)
      q.trial[at + j] = q.scene[at + j] + (pb[j * 3] * r0 + pb[j * 3 + 1] * r1 + pb[j * 3 + 2] * r2);
  }
}

template <bool SHARED, int MODE, bool RECOMPUTE = false, class FT = float>
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_point_sweep_factored(Problem q) {
  __shared__ double st_all[TILE_WAVES][TILE][3];
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const uint32_t t = point_tile(wave);
  if (t >= q.n_tiles) return;
  sweep_tile_factored<FT, SHARED, MODE, RECOMPUTE>(q, t, lane, st_all[wave]);
}

// The camera sweep over the factors (FT: float or double): one workgroup per camera-list task, the
// camera's rotation and flags uniform, k_camera_sweep's slot order and reduction.
template <bool SHARED, bool RECOMPUTE = false, class FT = float>
__global__ __launch_bounds__(TASK_THREADS) void k_camera_sweep_factored(Problem q) {
  constexpr int NV = SHARED ? 7 : 6;
  __shared__ double red[TASK_THREADS / 64][NV];
  const uint32_t t = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t c = q.tasks[4 * t], s = q.tasks[4 * t + 1], e = q.tasks[4 * t + 2],
                 b = q.tasks[4 * t + 3];
  constexpr int PER = TASK_OBS / TASK_THREADS;
  double state = bstate(q, b)[S_PCG_ACTIVE];
  uint32_t slot[PER], pt[PER];
  const uint32_t last = e > s ? e - 1 : 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    slot[k] = min(s + threadIdx.x + TASK_THREADS * k, last);
    pt[k] = e > s ? q.cam_point[slot[k]] : 0;
  }
  settle_uniform(state);
#pragma unroll
  for (int k = 0; k < PER; k++) settle(pt[k]);
  if (state == 0) return;
  const Bundle& g = q.bundles[b];
  if (q.n_unstored != 0 && !stores_w(g)) return;
  const bool var = variable_camera(q, c);
  const double* pose = q.scene + (size_t)c * 12;
  double R[9];
#pragma unroll
  for (int k = 0; k < 9; k++) R[k] = pose[k];
  const double fscale = RECOMPUTE && SHARED ? focal_scale(bstate(q, b)[S_LOG_FOCAL]) : 1.0;
  const double* y = vy(q);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0;
  if (e > s) {
    double yv[PER][3];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const double f = s + threadIdx.x + TASK_THREADS * k < e ? 1.0 : 0.0;
#pragma unroll
      for (int j = 0; j < 3; j++) yv[k][j] = y[(size_t)pt[k] * 3 + j] * f;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const uint32_t n = slot[k];
      double fw[11];
      load_factors<FT, SHARED, RECOMPUTE>(q, RECOMPUTE ? nullptr : fptr<SHARED, FT>(q.Wc, 0, n), TILE,
                                      RECOMPUTE ? camera_obs(q)[n] : n, c, pt[k], g, fscale, fw);
      const bool dir = directional_point(g, pt[k]);
      const double b0 = R[0] * yv[k][0] + R[1] * yv[k][1] + R[2] * yv[k][2];
      const double b1 = R[3] * yv[k][0] + R[4] * yv[k][1] + R[5] * yv[k][2];
      const double b2 = R[6] * yv[k][0] + R[7] * yv[k][1] + R[8] * yv[k][2];
      const double h0 = fw[0] * b0 + fw[1] * b1 + fw[2] * b2;
      const double h1 = fw[3] * b0 + fw[4] * b1 + fw[5] * b2;
      const double d0 = fw[0] * h0 + fw[3] * h1, d1 = fw[1] * h0 + fw[4] * h1,
                   d2 = fw[2] * h0 + fw[5] * h1;
      if (var) {
        acc[0] += fw[7] * d2 - fw[8] * d1;
        acc[1] += fw[8] * d0 - fw[6] * d2;
        acc[2] += fw[6] * d1 - fw[7] * d0;
        if (!dir) {
          acc[3] += d0;
          acc[4] += d1;
          acc[5] += d2;
        }
      }
      if (SHARED) acc[NV - 1] += fw[9] * h0 + fw[10] * h1;
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < NV; k++) red[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double v = 0;
    for (int w = 0; w < TASK_THREADS / 64; w++) v += red[w][threadIdx.x];
    q.task_part[(size_t)t * 8 + threadIdx.x] = v;
  }
}

}  // namespace tcvba
