// ba_onepass.hip — the Schur product in one pass over each observation's stored Jacobian factors
// (ba_factored.hip; FP64 in the FP64 form, FP32 in the mixed), included by ba_engine.hip after
// ba_factored.hip.
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
//   * a point's terms (its observations are consecutive lanes of a pass) by a segmented inclusive
//     scan over DPP lane shifts (1, 2, 4 and 8 within each row of 16 lanes, then the last lanes of
//     rows 0 and 2 into rows 1 and 3 and that of row 1 into rows 2 and 3): the segment's last lane
//     holds its sum, with no LDS traffic;
//   * within a 64-observation pass, each camera row component takes its lanes' terms in an order
//     fixed by the lanes' ranks among the pass's lanes of that row (add_terms: in each LDS add of
//     the pass the lanes of one row add distinct components, so no two lanes meet);
//   * a wave adds its passes to its partials in tile order, the workgroup its four waves in wave
//     order, the reduction the workgroups in index order.
// So the product is deterministic (and the streamed route, which recomputes the words, gives the
// same bits), though summed in another order than the two sweeps.
//
// What fixes those orders is fixed for a whole solve, so set-up (k_onepass_codes) writes it down
// once: per stored slot a code word (the window row of its observation's camera, its rank, its
// point within the tile, its distance from the first lane of its point's segment in the pass and
// whether it is the segment's last), and per tile a record of eight words (first slot,
// observations, first point, points, first observation, rounds). A wave reads its eight tiles'
// records with one load and, in the mixed form, loads the next tile's codes, factors and V^-1
// before it works on the current one.
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
// p_c, whether it is variable and its rotation (NR = 9), which the factors' products need, or its
// whole pose, rotation then translation (NR = 12: the FP64 form, whose products recompute the
// camera-frame points). The tiles' gathers of them become LDS reads.
template <int NR>
struct WindowCams {
  static_assert(NR == 9 || NR == 12, "a rotation, or a whole pose");
  double p[OP_SPAN][6];
  double R[OP_SPAN][NR];
  uint32_t var[OP_SPAN];
};

// The pose words a window stages per camera for the products over WT factors.
template <class WT>
constexpr int window_words() {
  return sizeof(WT) == 8 ? 12 : 9;
}

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

// The record words of a tile (op_trec, OP_RECORD words per tile).
enum : uint32_t {
  TR_SLOT = 0,    // first slot of its code words (a multiple of 64)
  TR_OBS = 1,     // observations
  TR_POINT = 2,   // first point
  TR_POINTS = 3,  // points
  TR_FIRST = 4,   // first observation (point-major)
  TR_ROUNDS = 5,  // the largest rank of any of the tile's passes, plus one
};
static_assert(OP_RECORD * OP_WAVE_TILES == TILE, "a wave's records: one word per lane");

// A slot's code word (op_code): bits 0-5 the window row of its observation's camera, 6-11 its rank
// among the pass's observations of that row (lanes below it with the same row), 12-17 its point
// counted from the tile's first point, 18-23 its distance from the first lane of its point's
// observations in the pass, bit 24 set on the last of those lanes.
struct SlotCode {
  uint32_t row, rank, lp, dist;
  bool last;
  __device__ explicit SlotCode(uint32_t x)
      : row(x & 63u), rank((x >> 6) & 63u), lp((x >> 12) & 63u), dist((x >> 18) & 63u),
        last(((x >> 24) & 1u) != 0) {}
};

__device__ __forceinline__ uint32_t slot_code(uint32_t row, uint32_t rank, uint32_t lp,
                                              uint32_t dist, bool last) {
  return row | rank << 6 | lp << 12 | dist << 18 | (last ? 1u : 0u) << 24;
}

// DPP lane moves of a double (two 32-bit moves): lane l takes the value of lane l - K of its row of
// 16 lanes (0 where that lane is in another row), or, for the broadcasts, of the last lane of the
// row before (rows 1 and 3 from lanes 15 and 47; rows 2 and 3 from lane 31; 0 in the other rows).
enum : int { kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143 };

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}

// One step of the segmented scan: x += the moved x where `take`, x unchanged elsewhere.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void scan_step(double x[3], bool take) {
  double m[3];
#pragma unroll
  for (int c = 0; c < 3; c++) m[c] = dpp_move<CTRL, ROW_MASK>(x[c]);
#pragma unroll
  for (int c = 0; c < 3; c++) x[c] = take ? x[c] + m[c] : x[c];
}

// Segmented inclusive scan over the wave's 64 lanes: lane l, whose segment began `dist` lanes
// before it, ends with the sum of x over its segment's lanes up to l (Hillis-Steele within each
// row of 16 lanes, then the carries across rows). Every lane of the wave takes part.
__device__ __forceinline__ void segmented_scan3(double x[3], uint32_t dist, uint32_t lane) {
  const uint32_t in_row = lane & 15u, row = lane >> 4;
  scan_step<kDppRowShr + 1, 0xf>(x, dist >= 1 && in_row >= 1);
  scan_step<kDppRowShr + 2, 0xf>(x, dist >= 2 && in_row >= 2);
  scan_step<kDppRowShr + 4, 0xf>(x, dist >= 4 && in_row >= 4);
  scan_step<kDppRowShr + 8, 0xf>(x, dist >= 8 && in_row >= 8);
  // lanes 15 and 47 into rows 1 and 3 where the segment began in the row before
  scan_step<kDppRowBcast15, 0xa>(x, (row & 1u) != 0 && dist > in_row);
  // lane 31 into rows 2 and 3 where the segment began at or before lane 31
  scan_step<kDppRowBcast31, 0xc>(x, dist > (lane & 31u));
}

// The stored words of one observation that the two halves share. The mixed form keeps its FP32
// factors; the FP64 form its cross block. `load_stored` reads lane `lane` of a tile's (or a pass
// of a tile's) words (`w0`: their first word; word k at w0 + k * n, n the tile's observations),
// `load_recomputed` relinearises them from the scene.
template <class WT, bool SHARED>
struct ObsWords;

template <bool SHARED>
struct ObsWords<float, SHARED> {
  static constexpr int K = kf_words<float>(SHARED);
  float f[11];
  __device__ __forceinline__ void load_stored(const float* w0, uint32_t n, uint32_t lane) {
#pragma unroll
    for (int k = 0; k < K; k++) f[k] = __builtin_nontemporal_load(w0 + k * n + lane);
#pragma unroll
    for (int k = K; k < 11; k++) f[k] = 0;
  }
  // q is among the stored words
  __device__ __forceinline__ void set_point(const double*, const double*, bool) {}
  // relinearised and rounded as the passes store them
  __device__ __forceinline__ void load_recomputed(const Problem& q, uint32_t i, uint32_t c,
                                                  uint32_t p, const Bundle& g, double fscale) {
    Lin L;
    const bool ok =
        linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L);
    double raw[11];
    cross_factors(L, ok, SHARED, raw);
#pragma unroll
    for (int k = 0; k < 11; k++) f[k] = (float)raw[k];
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

// The FP64 form's words in this product: the same factors in FP64 (the passes store them so when
// the one-pass product runs), every product formed as for the FP32 ones.
template <bool SHARED>
struct ObsWords<double, SHARED> {
  static constexpr int K = kf_words<double>(SHARED);
  double f[11];
  __device__ __forceinline__ void load_stored(const double* w0, uint32_t n, uint32_t lane) {
#pragma unroll
    for (int k = 0; k < K; k++)
      f[kf_raw<double>(k)] = __builtin_nontemporal_load(w0 + k * n + lane);
    if (!SHARED) f[9] = f[10] = 0;
  }
  // q from the window camera's pose and the point, as the passes compute it
  __device__ __forceinline__ void set_point(const double* pose, const double* X, bool dir) {
    camera_point(pose, X[0], X[1], X[2], dir, f + 6);
  }
  // relinearised as the passes store them
  __device__ __forceinline__ void load_recomputed(const Problem& q, uint32_t i, uint32_t c,
                                                  uint32_t p, const Bundle& g, double fscale) {
    Lin L;
    const bool ok =
        linearize(q, c, p, i, fscale, variable_camera(q, c), directional_point(g, p), L);
    cross_factors(L, ok, SHARED, f);
  }
  __device__ __forceinline__ void point_term(const double* R, const double* vc, double vf,
                                             bool var, bool dir, double t[3]) const {
    factored_point_term(f, R, vc, vf, var, dir, t);
  }
  __device__ __forceinline__ void camera_term(const double* R, const double* y, bool var, bool dir,
                                              double* hv) const {
    factored_camera_term<SHARED>(f, R, y, var, dir, hv);
  }
};

// The first stored word of the tile whose observations begin at `first` (tile_word's layout).
// The scene coordinates of point p.
__device__ __forceinline__ const double* point_of(const Problem& q, uint32_t p) {
  return q.scene + (size_t)q.C * 12 + (size_t)p * 3;
}

template <class WT, bool SHARED>
__device__ __forceinline__ const WT* stored_tile(const Problem& q, uint32_t first) {
  return static_cast<const WT*>(q.Wp) + (size_t)first * ObsWords<WT, SHARED>::K;
}

// Word k (a wave-uniform index) of the wave's tile records.
__device__ __forceinline__ uint32_t record_word(uint32_t rec, uint32_t k) {
  return (uint32_t)__builtin_amdgcn_readlane((int)rec, (int)k);
}

// What one tile needs before its LDS work: each lane's code, its stored words (the FP32 factors,
// when AHEAD) and V^-1 of the tile's point numbered as the lane.
template <class WT, bool SHARED>
struct TileIn {
  uint32_t code;
  ObsWords<WT, SHARED> ow;
  double vinv[9];
};

template <class WT, bool SHARED, bool AHEAD>
__device__ __forceinline__ void load_tile(const Problem& q, uint32_t rec, uint32_t j,
                                          uint32_t lane, TileIn<WT, SHARED>& in) {
  const uint32_t slot = record_word(rec, OP_RECORD * j + TR_SLOT);
  const uint32_t n = record_word(rec, OP_RECORD * j + TR_OBS);
  const uint32_t fp = record_word(rec, OP_RECORD * j + TR_POINT);
  const uint32_t np = record_word(rec, OP_RECORD * j + TR_POINTS);
  in.code = n ? q.op_code[slot + lane] : 0u;
  if (AHEAD && lane < n && n <= (uint32_t)TILE)
    in.ow.load_stored(stored_tile<WT, SHARED>(q, record_word(rec, OP_RECORD * j + TR_FIRST)), n,
                      lane);
  if (lane < np) {
    const double* pb = q.pblk + (size_t)(fp + lane) * 12;
#pragma unroll
    for (int k = 0; k < 9; k++) in.vinv[k] = pb[k];
  }
}

// h[j] = h[(j + S) % NV] where `on` (a register rotation by S, as selects).
template <int NV, int S>
__device__ __forceinline__ void rotate_if(double* h, bool on) {
  double t[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) t[j] = h[(j + S) % NV];
#pragma unroll
  for (int j = 0; j < NV; j++) h[j] = on ? t[j] : h[j];
}

// The camera half's adds of one pass, conflict-free and dense: the lanes are taken in groups of NV
// ranks (rank / NV), and in step j of a group the lane of rank rho = rank % NV adds component
// (rho + j) % NV of its term to its row. Lanes of one row have distinct ranks, so within a step
// they add to distinct components, and NV steps (one LDS add each, every lane of the group active)
// add every term. Each row component takes its lanes' terms in an order fixed by their ranks
// (group by group; within a group rho = c, c - 1, ..., c + 1 mod NV for component c): the wave's
// LDS operations are applied in issue order, and the LDS adds (ds_add_f64, IEEE double additions
// done by the LDS) need no wait between steps, only before the rows are read.
template <int NV>
__device__ __forceinline__ void add_terms(double (*part)[NV], const SlotCode& k, bool ok,
                                          uint32_t rounds, const double* hv) {
  static_assert(NV <= 8, "the rotation covers ranks 0..7 of a group");
  const uint32_t rho = k.rank % NV, grp = k.rank / NV;
  double hr[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) hr[j] = hv[j];
  rotate_if<NV, 1>(hr, (rho & 1u) != 0);
  rotate_if<NV, 2>(hr, (rho & 2u) != 0);
  rotate_if<NV, 4>(hr, (rho & 4u) != 0);
  const uint32_t groups = (rounds + NV - 1) / NV;
  for (uint32_t gi = 0; gi < groups; gi++) {
    const bool mine = ok && grp == gi;
    uint32_t c = rho;
#pragma unroll
    for (int j = 0; j < NV; j++) {
      if (mine)
        __hip_atomic_fetch_add(&part[k.row][c], hr[j], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
      c = c + 1 == (uint32_t)NV ? 0u : c + 1;
    }
  }
}

// The per-workgroup constants of a product's tiles.
template <int NR>
struct OnePassCtx {
  const Problem& q;
  const Bundle& g;
  const Window& win;
  const WindowCams<NR>& wc;
  double vf, fscale;
};

// One tile of at most 64 observations (one pass): the point half (the point sweep's terms and
// segment sums, y_p kept in `st`), then the camera half added into `part`.
template <class WT, bool SHARED, bool RECOMPUTE, bool AHEAD, int NR>
__device__ __forceinline__ void onepass_tile(const OnePassCtx<NR>& x, uint32_t rec, uint32_t j,
                                             uint32_t lane, TileIn<WT, SHARED>& in,
                                             double (*st)[3], double (*part)[SHARED ? 7 : 6]) {
  constexpr int NV = SHARED ? 7 : 6;
  const uint32_t n = record_word(rec, OP_RECORD * j + TR_OBS);
  const uint32_t fp = record_word(rec, OP_RECORD * j + TR_POINT);
  const uint32_t np = record_word(rec, OP_RECORD * j + TR_POINTS);
  const uint32_t first = record_word(rec, OP_RECORD * j + TR_FIRST);
  const uint32_t rounds = record_word(rec, OP_RECORD * j + TR_ROUNDS);
  const SlotCode k(in.code);
  const bool ok = lane < n;
  const uint32_t pt = fp + k.lp;
  const bool var = x.wc.var[k.row] != 0, dir = directional_point(x.g, pt);
  if (!AHEAD && ok) {
    if (RECOMPUTE)
      in.ow.load_recomputed(x.q, first + lane, x.win.camera(k.row), pt, x.g, x.fscale);
    else
      in.ow.load_stored(stored_tile<WT, SHARED>(x.q, first), n, lane);
  }
  if (!RECOMPUTE && ok) in.ow.set_point(x.wc.R[k.row], point_of(x.q, pt), dir);
  // the point half: each point's terms summed by the scan; its last lane hands the sum to the
  // point's lane (st[lp]), which forms y_p = V_p^-1 sum
  double tj[3] = {0, 0, 0};
  if (ok) in.ow.point_term(x.wc.R[k.row], x.wc.p[k.row], x.vf, var, dir, tj);
  segmented_scan3(tj, k.dist, lane);
  if (ok && k.last) {
    st[k.lp][0] = tj[0];
    st[k.lp][1] = tj[1];
    st[k.lp][2] = tj[2];
  }
  wave_sync();
  if (lane < np) {
    const double acc[3] = {st[lane][0], st[lane][1], st[lane][2]};
#pragma unroll
    for (int c = 0; c < 3; c++)
      st[lane][c] = in.vinv[c * 3] * acc[0] + in.vinv[c * 3 + 1] * acc[1] + in.vinv[c * 3 + 2] * acc[2];
  }
  wave_sync();
  double hv[NV];
  if (ok) in.ow.camera_term(x.wc.R[k.row], st[k.lp], var, dir, hv);
  add_terms<NV>(part, k, ok, rounds, hv);
  wave_sync();
}

// A tile of one point with more than 64 observations: both halves pass by pass, each pass's codes
// and words loaded (or relinearised) where it is used.
template <class WT, bool SHARED, bool RECOMPUTE, int NR>
__device__ __forceinline__ void onepass_long_tile(const OnePassCtx<NR>& x, uint32_t rec, uint32_t j,
                                               uint32_t lane, const double* vinv,
                                               double (*st)[3], double (*part)[SHARED ? 7 : 6]) {
  constexpr int NV = SHARED ? 7 : 6;
  const uint32_t slot = record_word(rec, OP_RECORD * j + TR_SLOT);
  const uint32_t n = record_word(rec, OP_RECORD * j + TR_OBS);
  const uint32_t fp = record_word(rec, OP_RECORD * j + TR_POINT);
  const uint32_t np = record_word(rec, OP_RECORD * j + TR_POINTS);
  const uint32_t first = record_word(rec, OP_RECORD * j + TR_FIRST);
  const uint32_t rounds = record_word(rec, OP_RECORD * j + TR_ROUNDS);
  ObsWords<WT, SHARED> ow;
  auto load = [&](uint32_t base, const SlotCode& k) {
    if (RECOMPUTE) {
      ow.load_recomputed(x.q, first + base + lane, x.win.camera(k.row), fp + k.lp, x.g, x.fscale);
    } else {
      ow.load_stored(stored_tile<WT, SHARED>(x.q, first) + base, n, lane);
      ow.set_point(x.wc.R[k.row], point_of(x.q, fp + k.lp), directional_point(x.g, fp + k.lp));
    }
  };
  // the point's terms pass by pass: each pass's sum (the scan's last lane) added in pass order
  double acc[3] = {0, 0, 0};
  for (uint32_t base = 0; base < n; base += TILE) {
    const SlotCode k(x.q.op_code[slot + base + lane]);
    const bool ok = base + lane < n;
    double tj[3] = {0, 0, 0};
    if (ok) {
      load(base, k);
      ow.point_term(x.wc.R[k.row], x.wc.p[k.row], x.vf, x.wc.var[k.row] != 0,
                    directional_point(x.g, fp + k.lp), tj);
    }
    segmented_scan3(tj, k.dist, lane);
    if (ok && k.last) {
      st[0][0] = tj[0];
      st[0][1] = tj[1];
      st[0][2] = tj[2];
    }
    wave_sync();
    if (lane < np) {
      acc[0] += st[0][0];
      acc[1] += st[0][1];
      acc[2] += st[0][2];
    }
    wave_sync();
  }
  if (lane < np) {
#pragma unroll
    for (int c = 0; c < 3; c++)
      st[lane][c] = vinv[c * 3] * acc[0] + vinv[c * 3 + 1] * acc[1] + vinv[c * 3 + 2] * acc[2];
  }
  wave_sync();
  for (uint32_t base = 0; base < n; base += TILE) {
    const SlotCode k(x.q.op_code[slot + base + lane]);
    const bool ok = base + lane < n;
    double hv[NV];
    if (ok) {
      load(base, k);
      ow.camera_term(x.wc.R[k.row], st[k.lp], x.wc.var[k.row] != 0,
                     directional_point(x.g, fp + k.lp), hv);
    }
    add_terms<NV>(part, k, ok, rounds, hv);
    wave_sync();
  }
}

// The LDS of one one-pass workgroup: each wave's point terms and y (st) and partial rows, and the
// window's cameras.
template <int NV, int NR>
struct OnePassLds {
  double st[TILE_WAVES][TILE][3];
  double part[TILE_WAVES][OP_SPAN][NV];
  WindowCams<NR> wc;
};

// One-pass workgroup w (its op_tile range, OP_WAVE_TILES tiles per wave). Called by every thread of
// a workgroup of TILE * TILE_WAVES threads; returns with L free for reuse.
template <class WT, bool SHARED, bool RECOMPUTE>
__device__ __forceinline__ void onepass_workgroup(const Problem& q, uint32_t w,
                                                  OnePassLds<SHARED ? 7 : 6,
                                                             window_words<WT>()>& L) {
  constexpr int NV = SHARED ? 7 : 6;
  constexpr int NR = window_words<WT>();
  // words loaded one tile ahead: the FP32 factors (not the FP64 blocks, whose registers would cut
  // the waves a SIMD holds, nor relinearised words, which need the tile's codes first)
  constexpr bool AHEAD = sizeof(WT) == 4 && !RECOMPUTE;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const uint32_t tw0 = q.op_tile[w], tw1 = q.op_tile[w + 1];
  const uint32_t b = q.tile_bundle[tw0];
  const double* sb = bstate(q, b);
  // the workgroup's tiles are of one bundle: its PCG state decides for all of them
  if (sb[S_PCG_ACTIVE] == 0) return;
  const Bundle& g = q.bundles[b];
  if (q.n_unstored != 0 && !stores_w(g)) return;
  // the wave's tiles and their records (lane OP_RECORD j + k: word k of the wave's tile j)
  const uint32_t t0 = tw0 + wave * OP_WAVE_TILES;
  const uint32_t nt = t0 < tw1 ? min((uint32_t)OP_WAVE_TILES, tw1 - t0) : 0u;
  const uint32_t rec = lane < OP_RECORD * nt ? q.op_trec[(size_t)t0 * OP_RECORD + lane] : 0u;
  TileIn<WT, SHARED> cur;
  if (nt) load_tile<WT, SHARED, AHEAD>(q, rec, 0, lane, cur);
  const Window win = window_of(q, w);
  const uint32_t row0 = q.op_off[w], span = q.op_off[w + 1] - row0;
  {
    // the window's cameras: thread r stages row r (rows past the span stay unread)
    const uint32_t r = threadIdx.x;
    if (r < span) {
      const uint32_t c = win.camera(r);
      const double* v = vp(q) + (size_t)c * 6;
#pragma unroll
      for (int k = 0; k < 6; k++) L.wc.p[r][k] = v[k];
      const double* pose = q.scene + (size_t)c * 12;
#pragma unroll
      for (int k = 0; k < NR; k++) L.wc.R[r][k] = pose[k];
      L.wc.var[r] = variable_camera(q, c) ? 1u : 0u;
    }
  }
  __syncthreads();
  static_assert(OP_SPAN == TILE, "one partial row per lane");
#pragma unroll
  for (int k = 0; k < NV; k++) L.part[wave][lane][k] = 0;
  wave_sync();
  const OnePassCtx<NR> x{q, g, win, L.wc, SHARED ? vp(q)[6 * q.C + b] : 0.0,
                         RECOMPUTE && SHARED ? focal_scale(sb[S_LOG_FOCAL]) : 1.0};
  for (uint32_t j = 0; j < nt; j++) {
    TileIn<WT, SHARED> next;
    if (j + 1 < nt) load_tile<WT, SHARED, AHEAD>(q, rec, j + 1, lane, next);
    const uint32_t n = record_word(rec, OP_RECORD * j + TR_OBS);
    if (n > (uint32_t)TILE)
      onepass_long_tile<WT, SHARED, RECOMPUTE, NR>(x, rec, j, lane, cur.vinv, L.st[wave],
                                                   L.part[wave]);
    else if (n)
      onepass_tile<WT, SHARED, RECOMPUTE, AHEAD, NR>(x, rec, j, lane, cur, L.st[wave],
                                                     L.part[wave]);
    cur = next;
  }
  __syncthreads();
  double* out = q.op_part + (size_t)row0 * 8;
  for (uint32_t e = threadIdx.x; e < span * 8; e += blockDim.x) {
    const uint32_t k = e >> 3, c = e & 7;
    double sum = 0;
    if (c < (uint32_t)NV) {
      sum = L.part[0][k][c];
#pragma unroll
      for (int u = 1; u < TILE_WAVES; u++) sum += L.part[u][k][c];
    }
    out[e] = sum;
  }
  __syncthreads();
}

// One workgroup per op_tile range (or per entry of op_map), XCD-contiguous order.
template <class WT, bool SHARED, bool RECOMPUTE = false>
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_schur_onepass(Problem q) {
  __shared__ OnePassLds<SHARED ? 7 : 6, window_words<WT>()> L;
  const uint32_t i = xcd_contiguous(blockIdx.x, gridDim.x);
  onepass_workgroup<WT, SHARED, RECOMPUTE>(q, q.op_map ? q.op_map[i] : i, L);
}

// Each camera's rows of partials summed in workgroup order (a wave per camera: lanes take the rows
// in turn, then a butterfly), written as the camera sweep's task partials: the whole sum in the
// camera's first task, zero in the others.
template <bool SHARED>
__device__ __forceinline__ void onepass_reduce_camera(const Problem& q, uint32_t c, uint32_t lane) {
  constexpr int NV = SHARED ? 7 : 6;
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

template <bool SHARED>
__global__ __launch_bounds__(256) void k_onepass_reduce(Problem q) {
  const uint32_t k = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  if (k < (q.cam_map ? q.n_cam_map : q.C))
    onepass_reduce_camera<SHARED>(q, q.cam_map ? q.cam_map[k] : k, threadIdx.x & 63);
}

// In the PCG loop after the one-pass product: camera c's rows of partials summed as
// onepass_reduce_camera sums them (or, for a bundle whose cross blocks are not stored, the task
// partials of its own camera sweep), then its rows of S p, (S p)_c = A_c p_c [+ H_cf p_f] - the
// partials, by lanes 0..5 in k_cg_product's expressions, and the camera's p . S p and focal share
// H_cf . p_c - partial[6], summed over the six rows in row order, for k_cg_step (cam_dot). One
// launch with the reduction's parallelism, a wave per camera, in place of the reduction and
// k_cg_product.
template <bool SHARED>
__device__ __forceinline__ void onepass_product_camera(const Problem& q, uint32_t c,
                                                       uint32_t lane) {
  constexpr int NV = SHARED ? 7 : 6;
  const uint32_t b = q.camera_bundle[c];
  if (bstate(q, b)[S_PCG_ACTIVE] == 0) return;
  const double* p = vp(q);
  const uint32_t j = lane < 6 ? lane : 0, row = 6 * c + j;
  double A[6], pc[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    A[k] = q.camA[(size_t)c * 36 + j * 6 + k];
    pc[k] = p[6 * (size_t)c + k];
  }
  const double pr = p[row], pf = SHARED ? p[6 * q.C + b] : 0.0;
  const double hcf = SHARED ? q.camHcf[row] : 0.0;
  double part[NV];
  if (q.n_unstored != 0 && !stores_w(q.bundles[b])) {
    camera_sums<NV>(q, c, part);
  } else {
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0;
    for (uint32_t r = q.op_cam[c] + lane; r < q.op_cam[c + 1]; r += 64) {
      const double* rw = q.op_part + (size_t)q.op_row[r] * 8;
#pragma unroll
      for (int k = 0; k < NV; k++) acc[k] += rw[k];
    }
#pragma unroll
    for (int k = 0; k < NV; k++) {
      for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
      part[k] = 0.0 + acc[k];  // what camera_sums forms from the one task partial
    }
  }
  double a = SHARED ? hcf * pf : 0.0, av = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) av += A[k] * pc[k];
  a += av;
  double pj = part[0];
#pragma unroll
  for (int k = 1; k < 6; k++) pj = j == (uint32_t)k ? part[k] : pj;
  a -= pj;
  const double d = pr * a, fs = SHARED ? hcf * pr - (j == 0 ? part[NV - 1] : 0.0) : 0.0;
  if (lane < 6) vs(q)[row] = a;
  double dc = 0, fc = 0;
#pragma unroll
  for (int r = 0; r < 6; r++) {
    dc += __shfl(d, r, 64);
    fc += __shfl(fs, r, 64);
  }
  if (lane == 0) {
    q.cam_dot[2 * (size_t)c] = dc;
    q.cam_dot[2 * (size_t)c + 1] = fc;
  }
}

template <bool SHARED>
__global__ __launch_bounds__(256) void k_onepass_product(Problem q) {
  const uint32_t k = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  if (k < (q.cam_map ? q.n_cam_map : q.C))
    onepass_product_camera<SHARED>(q, q.cam_map ? q.cam_map[k] : k, threadIdx.x & 63);
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

// Set-up, after the windows: each slot's code word and each tile's record, in the product's
// workgroup and wave layout (wave v of workgroup w takes tiles op_tile[w] + 8 v ...). A lane's rank
// is the number of lanes below it, in its pass, whose observation has a camera of the same row
// (the rows' bits compared by ballots).
__global__ __launch_bounds__(TILE* TILE_WAVES) void k_onepass_codes(Problem q, uint32_t* code,
                                                                     uint32_t* trec) {
  const uint32_t wave = threadIdx.x / TILE, lane = threadIdx.x % TILE;
  const uint32_t w = blockIdx.x;
  const Window win = window_of(q, w);
  const uint32_t t0 = q.op_tile[w] + wave * OP_WAVE_TILES;
  const uint32_t t1 = min(t0 + OP_WAVE_TILES, q.op_tile[w + 1]);
  const uint32_t* ps = point_start(q);
  for (uint32_t t = t0; t < t1; t++) {
    const uint32_t fp = q.tiles[t], ep = q.tiles[t + 1], slot = q.tile_slot[t];
    const uint32_t s = ps[fp], e = ps[ep];
    uint32_t rounds = 0;
    for (uint32_t base = s; base < e; base += TILE) {
      const uint32_t i = base + lane;
      const bool ok = i < e;
      const uint32_t row = ok ? win.row(obs_camera(q)[i]) : 0u;
      const uint32_t pt = ok ? obs_point(q)[i] : fp;
      // the lanes of the point's observations in this pass: [first, end)
      const uint32_t first = ok ? max(ps[pt], base) : i, end = ok ? min(ps[pt + 1], base + TILE) : i + 1;
      unsigned long long same = __ballot(ok);
#pragma unroll
      for (int bit = 0; bit < 6; bit++) {
        const unsigned long long set = __ballot((row >> bit) & 1u);
        same &= (row >> bit) & 1u ? set : ~set;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(same >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
      code[slot + (base - s) + lane] =
          slot_code(ok ? row : 0u, ok ? rank : 0u, pt - fp, i - first, i + 1 == end);
      uint32_t r = ok ? rank + 1 : 0u;
      for (int o = 32; o > 0; o >>= 1) r = max(r, (uint32_t)__shfl_xor(r, o, 64));
      rounds = max(rounds, r);
    }
    if (lane < (uint32_t)OP_RECORD) {
      const uint32_t word = lane == TR_SLOT     ? slot
                            : lane == TR_OBS    ? e - s
                            : lane == TR_POINT  ? fp
                            : lane == TR_POINTS ? ep - fp
                            : lane == TR_FIRST  ? s
                            : lane == TR_ROUNDS ? rounds
                                                : 0u;
      trec[(size_t)t * OP_RECORD + lane] = word;
    }
  }
}

}  // namespace tcvba
