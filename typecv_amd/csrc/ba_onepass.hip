// ba_onepass.hip — the Schur product in one pass over the stored cross blocks (FP64) or FP32
// Jacobian factors (mixed), included by ba_engine.hip after ba_factored.hip.
//
// The two sweeps read every observation's stored words twice: the point sweep for
// y_p = V_p^-1 sum_i W_i^T p_c(i), the camera sweep for sum_i W_i y_p(i) in camera order. Here one
// wave keeps a tile's words in registers across both halves: it forms y_p for the tile's points
// (the point sweep's terms and segment sums), then each observation's camera term W_i y_p, and
// adds those into per-wave camera partials in LDS. Points are sorted by home camera, so the
// cameras that a workgroup's (at most OP_TILES) consecutive tiles observe form a narrow window (at
// most OP_SPAN cameras, checked at set-up; wider problems keep the two sweeps). A workgroup's tiles
// belong to one bundle and are counted from the bundle's first tile, so a batch sums each bundle as
// its solve alone does. The workgroup writes one row of partials per camera of its window;
// k_onepass_reduce sums each camera's rows in workgroup order. Every sum has a fixed order:
//   * within a 64-observation pass, each camera row takes its lanes' terms one after another in
//     lane order (rounds of at most one lane per row, the lane's rank among the lanes of its row
//     found by ballots of the row's bits; the adds are LDS double additions, each row's in round
//     order);
//   * a wave adds its passes to its partials in tile order, the workgroup its four waves in wave
//     order, the reduction the workgroups in index order.
// So the product is deterministic (and the streamed route, which recomputes the factors, gives the
// same bits), though summed in another order than the two sweeps.
#pragma once

namespace tcvba {

// The camera side of one observation from its factors f, the camera's rotation R and the point's
// y: h = D' (R y); omega part q x (D'^T h), v part D'^T h (none for a directional point, neither
// for a fixed camera), focal part Jf' . h.
template <bool SHARED>
__device__ __forceinline__ void factored_camera_term(const double* f, const double* R,
                                                     const double* y, bool var, bool dir,
                                                     double* hv) {
  constexpr int NV = SHARED ? 7 : 6;
  const double b0 = R[0] * y[0] + R[1] * y[1] + R[2] * y[2];
  const double b1 = R[3] * y[0] + R[4] * y[1] + R[5] * y[2];
  const double b2 = R[6] * y[0] + R[7] * y[1] + R[8] * y[2];
  const double h0 = f[0] * b0 + f[1] * b1 + f[2] * b2;
  const double h1 = f[3] * b0 + f[4] * b1 + f[5] * b2;
  const double d0 = f[0] * h0 + f[3] * h1, d1 = f[1] * h0 + f[4] * h1, d2 = f[2] * h0 + f[5] * h1;
#pragma unroll
  for (int k = 0; k < NV; k++) hv[k] = 0;
  if (var) {
    hv[0] = f[7] * d2 - f[8] * d1;
    hv[1] = f[8] * d0 - f[6] * d2;
    hv[2] = f[6] * d1 - f[7] * d0;
    if (!dir) {
      hv[3] = d0;
      hv[4] = d1;
      hv[5] = d2;
    }
  }
  if (SHARED) hv[NV - 1] = f[9] * h0 + f[10] * h1;
}

// The factors of observation i as stored (FP32; RECOMPUTE: relinearised and rounded as the passes
// store them), kept in FP32 registers between the two halves.
template <bool SHARED, bool RECOMPUTE>
__device__ __forceinline__ void load_factors32(const Problem& q, uint32_t slot, uint32_t i,
                                               uint32_t c, uint32_t p, const Bundle& g,
                                               double fscale, float* f) {
  constexpr int KF = kf_words(SHARED);
  if (RECOMPUTE) {
    Lin L;
    const bool ok =
        linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L);
    double raw[11];
    cross_factors(L, ok, SHARED, raw);
#pragma unroll
    for (int k = 0; k < KF; k++) f[k] = (float)raw[k];
  } else {
#pragma unroll
    for (int k = 0; k < KF; k++) f[k] = __builtin_nontemporal_load(fptr<SHARED>(q.Wp, k, slot));
  }
#pragma unroll
  for (int k = KF; k < 11; k++) f[k] = 0;
}

__device__ __forceinline__ void widen11(const float* f, double* d) {
#pragma unroll
  for (int k = 0; k < 11; k++) d[k] = (double)f[k];
}

// A workgroup's camera window: rows lo .. lo + span - 1 of its bundle's cameras [cam0, cam0 + n),
// or, for a bundle whose cameras close a loop (a trajectory that goes round), of the cameras with
// the first half n / 2 moved past the end (camera c < cam0 + n / 2 taken as c + n), whichever is
// narrower. `op_lo` holds lo with bit 31 set for the second kind.
struct Window {
  uint32_t lo, cam0, n, half;
  bool wrapped;
  __device__ uint32_t row(uint32_t c) const {
    return (wrapped && c - cam0 < half ? c + n : c) - lo;
  }
  __device__ uint32_t camera(uint32_t r) const {
    const uint32_t v = lo + r;
    return wrapped && v >= cam0 + n ? v - n : v;
  }
};

// The cameras of a workgroup's window, staged in LDS once per product: each camera's PCG direction
// p_c, whether it is variable and, for the factors (NR = 9), its rotation. The tiles' gathers of
// them become LDS reads.
template <int NR>
struct WindowCams {
  double p[OP_SPAN][6];
  double R[OP_SPAN][NR];
  uint32_t var[OP_SPAN];
};

__device__ __forceinline__ Window window_of(const Problem& q, uint32_t w) {
  const uint32_t raw = q.op_lo[w];
  const Bundle& g = q.bundles[q.tile_bundle[q.op_tile[w]]];
  Window x;
  x.lo = raw & 0x7fffffffu;
  x.wrapped = (raw >> 31) != 0;
  x.cam0 = g.cam0;
  x.n = g.cam1 - g.cam0;
  x.half = x.n / 2;
  return x;
}

// The stored words of one observation that the two halves share. The mixed form keeps its FP32
// factors (the camera's rotation is read again, from L2, for the camera half); the FP64 form its
// cross block.
template <class WT, bool SHARED>
struct ObsWords;

template <bool SHARED>
struct ObsWords<float, SHARED> {
  float f[11];
  template <bool RECOMPUTE>
  __device__ __forceinline__ void load(const Problem& q, uint32_t slot, uint32_t i, uint32_t c,
                                       uint32_t p, const Bundle& g, double fscale) {
    load_factors32<SHARED, RECOMPUTE>(q, slot, i, c, p, g, fscale, f);
  }
  __device__ __forceinline__ void point_term(const double* R, const double* vc, double vf,
                                             bool var, bool dir, double t[3]) const {
    double d[11];
    widen11(f, d);
    factored_point_term(d, R, vc, vf, var, dir, t);
  }
  __device__ __forceinline__ void camera_term(const double* R, const double* y, bool var, bool dir,
                                              double* hv) const {
    double d[11];
    widen11(f, d);
    factored_camera_term<SHARED>(d, R, y, var, dir, hv);
  }
};

template <bool SHARED>
struct ObsWords<double, SHARED> {
  static constexpr int K = SHARED ? 21 : 18;
  double w[K];
  template <bool RECOMPUTE>
  __device__ __forceinline__ void load(const Problem& q, uint32_t slot, uint32_t i, uint32_t c,
                                       uint32_t p, const Bundle& g, double fscale) {
    if (RECOMPUTE) {
      relinearised_w<double, SHARED>(q, g, c, p, i, fscale, w);
    } else {
#pragma unroll
      for (int k = 0; k < K; k++) w[k] = wload(wptr<double, SHARED>(q.Wp, k, slot, q.M));
    }
  }
  // k_point_sweep's expression: t_j = sum_k W[k][j] v_c[k] (+ W_f[j] v_f)
  __device__ __forceinline__ void point_term(const double*, const double* vc, double vf, bool,
                                             bool, double t[3]) const {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double a = 0;
#pragma unroll
      for (int k = 0; k < 6; k++) a += w[k * 3 + j] * vc[k];
      if (SHARED) a += w[18 + j] * vf;
      t[j] = a;
    }
  }
  // W y: the camera rows and, on the shared route, the focal row
  __device__ __forceinline__ void camera_term(const double*, const double* y, bool, bool,
                                              double* hv) const {
#pragma unroll
    for (int k = 0; k < 6; k++) hv[k] = w[k * 3] * y[0] + w[k * 3 + 1] * y[1] + w[k * 3 + 2] * y[2];
    if (SHARED) hv[6] = w[18] * y[0] + w[19] * y[1] + w[20] * y[2];
  }
};

// One tile of the wave: the point half (y_p of its points, kept in `st`), then the camera half
// added into `part` (one row per camera of the window).
template <class WT, bool SHARED, bool RECOMPUTE>
__device__ __forceinline__ void onepass_tile(const Problem& q, uint32_t t, uint32_t lane,
                                             const Window& win,
                                             const WindowCams<sizeof(WT) == 4 ? 9 : 1>& wc,
                                             double (*st)[3], double (*part)[SHARED ? 7 : 6]) {
  constexpr int NV = SHARED ? 7 : 6;
  const uint32_t fp = q.tiles[t], ep = q.tiles[t + 1], b = q.tile_bundle[t];
  const double* sb = bstate(q, b);
  if (sb[S_PCG_ACTIVE] == 0) return;
  const Bundle& g = q.bundles[b];
  if (q.n_unstored != 0 && !stores_w(g)) return;
  const double* v = vp(q);
  const uint32_t* ps = point_start(q);
  const uint32_t s = ps[fp], e = ps[ep];
  const uint32_t my = fp + lane;
  const bool has = my < ep;
  const uint32_t mys = has ? ps[my] : 0, mye = has ? ps[my + 1] : 0;
  const double vf = SHARED ? v[6 * q.C + b] : 0.0;
  const double fscale = RECOMPUTE && SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0;
  const bool multi = e - s > (uint32_t)TILE;  // one point of more than 64 observations
  uint32_t c = 0, pt = 0, row = 0;
  bool var = false, dir = false;
  ObsWords<WT, SHARED> ow;
  double acc[3] = {0, 0, 0};
  // the point half (the point sweep's terms and sums)
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    if (i < e) {
      c = obs_camera(q)[i];
      pt = obs_point(q)[i];
      row = win.row(c);
      var = wc.var[row] != 0;
      dir = directional_point(g, pt);
      ow.template load<RECOMPUTE>(q, q.tile_slot[t] + (i - s), i, c, pt, g, fscale);
      double tj[3];
      ow.point_term(wc.R[row], wc.p[row], vf, var, dir, tj);
      st[lane][0] = tj[0];
      st[lane][1] = tj[1];
      st[lane][2] = tj[2];
    }
    wave_sync();
    if (has) {
      const uint32_t l = mys > base ? mys : base, h = mye < base + TILE ? mye : base + TILE;
      segment_sum(st, base, l, h, acc);
    }
    wave_sync();
  }
  if (has) {
    const double* pb = q.pblk + (size_t)my * 12;
#pragma unroll
    for (int j = 0; j < 3; j++)
      st[lane][j] = pb[j * 3] * acc[0] + pb[j * 3 + 1] * acc[1] + pb[j * 3 + 2] * acc[2];
  }
  wave_sync();
  // the camera half, pass by pass (a point of more than 64 observations reloads its words)
  for (uint32_t base = s; base < e; base += TILE) {
    const uint32_t i = base + lane;
    const bool ok = i < e;
    if (multi && ok) {
      c = obs_camera(q)[i];
      pt = obs_point(q)[i];
      row = win.row(c);
      var = wc.var[row] != 0;
      dir = directional_point(g, pt);
      ow.template load<RECOMPUTE>(q, q.tile_slot[t] + (i - s), i, c, pt, g, fscale);
    }
    double hv[NV];
    if (ok) ow.camera_term(wc.R[row], st[pt - fp], var, dir, hv);
    if (!ok) row = 0;
    // each lane's rank among the lanes of the same camera row: the rows' bits by ballot
    unsigned long long same = __ballot(ok);
#pragma unroll
    for (int bit = 0; bit < 6; bit++) {
      const unsigned long long set = __ballot((row >> bit) & 1u);
      same &= (row >> bit) & 1u ? set : ~set;
    }
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(same >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
    // rounds of one lane per row, in rank (lane) order: each row gets its lanes' terms one after
    // another. No two lanes of one round add to the same row, and the wave's LDS operations are
    // applied in issue order, so each row's sum has a fixed order: the LDS adds (ds_add_f64, IEEE
    // double additions done by the LDS) need no wait between rounds, only before the rows are read
    for (uint32_t r = 0; __ballot(ok && rank >= r) != 0ull; r++) {
      if (ok && rank == r) {
#pragma unroll
        for (int k = 0; k < NV; k++)
          __hip_atomic_fetch_add(&part[row][k], hv[k], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
    }
    wave_sync();
  }
}

// One workgroup per op_tile range (OP_WAVE_TILES tiles per wave), XCD-contiguous order.
template <class WT, bool SHARED, bool RECOMPUTE = false>
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_schur_onepass(Problem q) {
  constexpr int NV = SHARED ? 7 : 6;
  __shared__ double st_all[TILE_WAVES][TILE][3];
  __shared__ double part_all[TILE_WAVES][OP_SPAN][NV];
  __shared__ WindowCams<sizeof(WT) == 4 ? 9 : 1> wc;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const uint32_t w = xcd_contiguous(blockIdx.x, gridDim.x);
  const Window win = window_of(q, w);
  const uint32_t row0 = q.op_off[w], span = q.op_off[w + 1] - row0;
  {
    // the window's cameras: thread r stages row r (rows past the span stay unread)
    const uint32_t r = threadIdx.x;
    if (r < span) {
      const uint32_t c = win.camera(r);
      const double* v = vp(q) + (size_t)c * 6;
#pragma unroll
      for (int k = 0; k < 6; k++) wc.p[r][k] = v[k];
      if constexpr (sizeof(WT) == 4) {
        const double* pose = q.scene + (size_t)c * 12;
#pragma unroll
        for (int k = 0; k < 9; k++) wc.R[r][k] = pose[k];
      }
      wc.var[r] = variable_camera(q, c) ? 1u : 0u;
    }
  }
  __syncthreads();
  static_assert(OP_SPAN == TILE, "one partial row per lane");
#pragma unroll
  for (int k = 0; k < NV; k++) part_all[wave][lane][k] = 0;
  wave_sync();
  const uint32_t t0 = q.op_tile[w] + wave * OP_WAVE_TILES;
  const uint32_t t1 = min(t0 + OP_WAVE_TILES, q.op_tile[w + 1]);
  for (uint32_t t = t0; t < t1; t++)
    onepass_tile<WT, SHARED, RECOMPUTE>(q, t, lane, win, wc, st_all[wave], part_all[wave]);
  __syncthreads();
  double* out = q.op_part + (size_t)row0 * 8;
  for (uint32_t x = threadIdx.x; x < span * 8; x += blockDim.x) {
    const uint32_t k = x >> 3, j = x & 7;
    double sum = 0;
    if (j < (uint32_t)NV) {
      sum = part_all[0][k][j];
#pragma unroll
      for (int u = 1; u < TILE_WAVES; u++) sum += part_all[u][k][j];
    }
    out[x] = sum;
  }
}

// Each camera's rows of partials summed in workgroup order (a wave per camera: lanes take the rows
// in turn, then a butterfly), written as the camera sweep's task partials: the whole sum in the
// camera's first task, zero in the others.
template <bool SHARED>
__global__ __launch_bounds__(256) void k_onepass_reduce(Problem q) {
  constexpr int NV = SHARED ? 7 : 6;
  const uint32_t c = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  const uint32_t lane = threadIdx.x & 63;
  if (c >= q.C) return;
  const uint32_t b = q.camera_bundle[c];
  if (bstate(q, b)[S_PCG_ACTIVE] == 0) return;
  if (q.n_unstored != 0 && !stores_w(q.bundles[b])) return;  // the unstored sweeps write those
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0;
  for (uint32_t r = q.op_cam[c] + lane; r < q.op_cam[c + 1]; r += 64) {
    const double* row = q.op_part + (size_t)q.op_row[r] * 8;
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] += row[k];
  }
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  const uint32_t t0 = q.task_start[c], t1 = q.task_start[c + 1];
  if (lane < (uint32_t)NV)
    for (uint32_t t = t0; t < t1; t++) q.task_part[(size_t)t * 8 + lane] = t == t0 ? acc[lane] : 0.0;
}

// Set-up: the camera window of each one-pass workgroup (Window): the narrower of the plain and the
// wrapped range of its observations' cameras (span 0 for a workgroup without observations).
__global__ __launch_bounds__(256) void k_onepass_windows(Problem q, uint32_t* lo, uint32_t* span) {
  __shared__ uint32_t red[4][4];
  const uint32_t w = blockIdx.x;
  const uint32_t t0 = q.op_tile[w], t1 = q.op_tile[w + 1];
  const Bundle& g = q.bundles[q.tile_bundle[t0]];
  const uint32_t n = g.cam1 - g.cam0, half = n / 2;
  const uint32_t* ps = point_start(q);
  const uint32_t s = ps[q.tiles[t0]], e = ps[q.tiles[t1]];
  uint32_t v[4] = {0xffffffffu, 0, 0xffffffffu, 0};  // plain min, max; wrapped min, max
  for (uint32_t i = s + threadIdx.x; i < e; i += blockDim.x) {
    const uint32_t c = obs_camera(q)[i], cw = c - g.cam0 < half ? c + n : c;
    v[0] = min(v[0], c);
    v[1] = max(v[1], c);
    v[2] = min(v[2], cw);
    v[3] = max(v[3], cw);
  }
  for (int o = 32; o > 0; o >>= 1) {
    v[0] = min(v[0], (uint32_t)__shfl_xor(v[0], o, 64));
    v[1] = max(v[1], (uint32_t)__shfl_xor(v[1], o, 64));
    v[2] = min(v[2], (uint32_t)__shfl_xor(v[2], o, 64));
    v[3] = max(v[3], (uint32_t)__shfl_xor(v[3], o, 64));
  }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 4; k++) red[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int u = 1; u < 4; u++) {
      v[0] = min(v[0], red[u][0]);
      v[1] = max(v[1], red[u][1]);
      v[2] = min(v[2], red[u][2]);
      v[3] = max(v[3], red[u][3]);
    }
    const uint32_t plain = v[1] - v[0] + 1, wrapped = v[3] - v[2] + 1;
    if (e <= s) {
      lo[w] = g.cam0;
      span[w] = 0;
    } else if (wrapped < plain) {
      lo[w] = v[2] | 0x80000000u;
      span[w] = wrapped;
    } else {
      lo[w] = v[0];
      span[w] = plain;
    }
  }
}

}  // namespace tcvba
