// geodesy.hip -- geodesic reconstruction, leveling and the h-extrema chain on gfx950.
//
// Reconstruction by dilation of marker g under mask f is the least fixed point of
// x = f ^ delta_B(x) above g ^ f (mreconstruct_t.hpp:39-195 computes it with Vincent's
// hybrid scans + FIFO).  The fixed point is unique, so any monotone schedule that only
// ever raises a pixel to a value it is entitled to reaches the same bits.  Here:
//  * 2D (3x3 neighbourhoods): one wavefront per 64x64 tile keeps the tile plus a 1-pixel
//    halo in LDS and sweeps it in the four raster directions until nothing changes,
//    then writes it back; a per-round device flag records whether any tile border moved.
//  * any rank (3D SEs): in-place neighbour-list relaxation rounds with the same flags.
// Rounds are enqueued in batches without host synchronisation: round R exits at once when
// round R-1 moved nothing, and the host reads the batch's flags once, at its end.
// Reconstruction by erosion is the dual (min/max exchanged).
// Leveling uses the closed forms (verified against the priority-queue algorithm of
// leveling_t.hpp:34-202 by the oracle tests):
//   low_leveling(r, g)  = r ^ R^delta_{r v g}(g)
//   high_leveling(r, g) = r v R^eps_{r ^ g}(g)
//   leveling(f, g)      = low_leveling(high_leveling(f, g), g)   (leveling.cpp:231-237)
#include <type_traits>

#include "arith.hpp"
#include "erodil.hpp"
#include "rec_u8.hpp"

namespace pm {

constexpr int RT = 64;       // tile side (interior)
constexpr int RP = RT + 2;   // tile side with halo
constexpr int MAX_ROUNDS = 1 << 22;

// 3x3 neighbourhood as a bit mask over (dy+1)*3 + (dx+1), centre bit ignored
constexpr unsigned NB_SQUARE = 0x1EF;  // all but the centre (bit 4)
constexpr unsigned NB_CROSS = (1u << 1) | (1u << 3) | (1u << 5) | (1u << 7);
constexpr unsigned NB_SEGX = (1u << 3) | (1u << 5);
constexpr unsigned NB_SEGY = (1u << 1) | (1u << 7);
static unsigned nb_bits(int se)
{
  switch (se) {
  case PM_SE_SQUARE_2D: return NB_SQUARE;
  case PM_SE_CROSS_2D: return NB_CROSS;
  case PM_SE_SEGMENT_X_1D:
  case PM_SE_SEGMENT_X_2D: return NB_SEGX;
  case PM_SE_SEGMENT_Y_2D: return NB_SEGY;
  default: return 0;
  }
}

__global__ void zero_flag_kernel(int *f) { if (threadIdx.x == 0) f[0] = 0; }
__global__ void zero_flags_kernel(int *f, int n)
{
  if ((int)threadIdx.x < n) f[threadIdx.x] = 0;
}

template <typename T, bool DIL, unsigned NB>
__global__ __launch_bounds__(RT) void rec_tile_kernel(T *__restrict__ cur, const T *__restrict__ mask, int H, int W,
                                                      int tiles_x, const int *__restrict__ prev,
                                                      int *__restrict__ border_moved)
{
  if (prev && *prev == 0) return;  // the previous round moved nothing: converged
  using L = Lat<T, DIL>;
  __shared__ T c[RP * RP];
  __shared__ T m[RT * RT];
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int x0 = tx * RT, y0 = ty * RT;
  const int lane = threadIdx.x;

  for (int e = lane; e < RP * RP; e += RT) {
    const int yy = e / RP, xx = e - yy * RP;
    const int gy = y0 - 1 + yy, gx = x0 - 1 + xx;
    c[e] = ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W) ? cur[(int64_t)gy * W + gx] : L::neutral();
  }
  for (int e = lane; e < RT * RT; e += RT) {
    const int yy = e / RT, xx = e - yy * RT;
    const int gy = y0 + yy, gx = x0 + xx;
    // outside the image the mask is the neutral, which pins those pixels to the neutral
    m[e] = (gy < H && gx < W) ? mask[(int64_t)gy * W + gx] : L::neutral();
  }
  __syncthreads();

  auto relax = [&](int y, int x, bool &chg, bool &bchg) {
    const int p = (y + 1) * RP + (x + 1);
    const T old = c[p];
    T v = old;
#pragma unroll
    for (int b = 0; b < 9; b++)
      if (b != 4 && ((NB >> b) & 1u)) v = L::op(v, c[p + (b / 3 - 1) * RP + (b % 3 - 1)]);
    const T mk = m[y * RT + x];
    v = DIL ? (v < mk ? v : mk) : (v > mk ? v : mk);
    if (v != old) {
      c[p] = v;
      chg = true;
      if (y == 0 || x == 0 || y == RT - 1 || x == RT - 1) bchg = true;
    }
  };

  bool any = false, border = false;
  for (;;) {
    bool chg = false;
    for (int x = 0; x < RT; x++) relax(lane, x, chg, border);
    __syncthreads();
    for (int x = RT - 1; x >= 0; x--) relax(lane, x, chg, border);
    __syncthreads();
    for (int y = 0; y < RT; y++) relax(y, lane, chg, border);
    __syncthreads();
    for (int y = RT - 1; y >= 0; y--) relax(y, lane, chg, border);
    __syncthreads();
    const bool wave_chg = __ballot(chg) != 0;
    if (!wave_chg) break;
    any = true;
  }
  if (!any) return;
  for (int e = lane; e < RT * RT; e += RT) {
    const int yy = e / RT, xx = e - yy * RT;
    const int gy = y0 + yy, gx = x0 + xx;
    if (gy < H && gx < W) cur[(int64_t)gy * W + gx] = c[(yy + 1) * RP + xx + 1];
  }
  // a plain store of 1 from every workgroup whose border moved (the flag only has to become
  // non-zero; atomics to the one address would serialise)
  if (__ballot(border) != 0 && lane == 0) *border_moved = 1;
}

template <typename T, bool DIL>
__global__ __launch_bounds__(256) void rec_nl_kernel(T *__restrict__ cur, const T *__restrict__ mask, int D, int H,
                                                     int W, int xblocks, NLArgs a, const int *__restrict__ prev,
                                                     int *__restrict__ moved)
{
  if (prev && *prev == 0) return;  // the previous round moved nothing: converged
  using L = Lat<T, DIL>;
  const int row = blockIdx.x / xblocks;
  const int x = (blockIdx.x - row * xblocks) * 256 + threadIdx.x;
  bool chg = false;
  if (x < W) {
    const int z = row / H, y = row - z * H;
    const int64_t i = ((int64_t)z * H + y) * W + x;
    const T old = cur[i];
    T v = old;
    for (int k = 0; k < a.n; k++) {
      const int zz = z + a.dz[k], yy = y + a.dy[k], xx = x + a.dx[k];
      if ((unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W)
        v = L::op(v, cur[((int64_t)zz * H + yy) * W + xx]);
    }
    const T mk = mask[i];
    v = DIL ? (v < mk ? v : mk) : (v > mk ? v : mk);
    if (v != old) {
      cur[i] = v;
      chg = true;
    }
  }
  if (__ballot(chg) != 0 && (threadIdx.x & 63) == 0) *moved = 1;  // plain store, as above
}

// Reconstruction of marker under mask into out (all device pointers, out distinct from both).
template <typename T, bool DIL>
int reconstruct_dev(const T *marker, const T *mask, T *out, Dims d, int se, hipStream_t s)
{
  const int64_t n = d.D * d.H * d.W;
  if (n == 0) return PM_OK;
  if constexpr (std::is_same<T, uint8_t>::value) {
    if (rec_u8_applicable(se, d)) return rec_u8<DIL>(marker, mask, out, d, se, s);
  }
  // select_marker: marker ^ mask (dilate) / marker v mask (erode), mreconstruct_t.hpp:51,69
  PM_TRY((ew2<T, DIL ? EW_INF : EW_SUP>(marker, mask, out, n, s)));
  Scratch flag;
  PM_TRY(flag.alloc(sizeof(int) * PM_COUNTERS, s));
  int *dflag = (int *)flag.ptr;  // one change flag per round of a batch
  int *hflag = nullptr;
  PM_TRY(pinned_counters(&hflag));
  const unsigned nb = nb_bits(se);
  const bool tiled = nb != 0 && d.D == 1;
  NLArgs a;
  if (!tiled) PM_TRY(se_nl(se, a));
  const int tiles_x = (int)((d.W + RT - 1) / RT), tiles_y = (int)((d.H + RT - 1) / RT);
  const int xblocks = (int)((d.W + 255) / 256);
  const int cap = round_cap(MAX_ROUNDS);
  int round = 0, batch = 4;
  while (round < cap) {
    const int n = (round + batch <= cap) ? batch : cap - round;
    hipLaunchKernelGGL(zero_flags_kernel, dim3(1), dim3(PM_COUNTERS), 0, s, dflag, n);
    for (int i = 0; i < n; i++) {
      // the first round of a batch runs unconditionally: the host saw its predecessor move
      const int *prev = i > 0 ? dflag + i - 1 : nullptr;
      int *mv = dflag + i;
      if (tiled) {
        const dim3 g((unsigned)(tiles_x * tiles_y)), b(RT);
        const int H = (int)d.H, W = (int)d.W;
        if (nb == NB_SQUARE)
          hipLaunchKernelGGL((rec_tile_kernel<T, DIL, NB_SQUARE>), g, b, 0, s, out, mask, H, W, tiles_x, prev, mv);
        else if (nb == NB_CROSS)
          hipLaunchKernelGGL((rec_tile_kernel<T, DIL, NB_CROSS>), g, b, 0, s, out, mask, H, W, tiles_x, prev, mv);
        else if (nb == NB_SEGX)
          hipLaunchKernelGGL((rec_tile_kernel<T, DIL, NB_SEGX>), g, b, 0, s, out, mask, H, W, tiles_x, prev, mv);
        else
          hipLaunchKernelGGL((rec_tile_kernel<T, DIL, NB_SEGY>), g, b, 0, s, out, mask, H, W, tiles_x, prev, mv);
      } else
        hipLaunchKernelGGL((rec_nl_kernel<T, DIL>), dim3((unsigned)(xblocks * d.D * d.H)), dim3(256), 0, s, out, mask,
                           (int)d.D, (int)d.H, (int)d.W, xblocks, a, prev, mv);
    }
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(hflag, dflag, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++)
      if (hflag[i] == 0) return PM_OK;
    round += n;
    batch = batch * 2 > PM_COUNTERS ? PM_COUNTERS : batch * 2;
  }
  PM_FAIL(PM_E_NOCONV, "reconstruction did not converge");
}

// ---------------------------------------------------------------- entry-point bodies
static int check_geo_se(int se, int rank)
{
  PM_TRY(check_se_rank(se, rank));
  return PM_OK;
}

static int reconstruct_entry(int rec_type, const pm_image *marker, const pm_image *mask, int se, pm_image *out,
                             void *stream)
{
  PM_TRY(check_binary(marker, mask, out, "Reconstruction"));
  if (rec_type != PM_REC_DILATE && rec_type != PM_REC_ERODE) PM_FAIL(PM_E_ARG, "Unsupported reconstruction type");
  PM_TRY(check_geo_se(se, marker->rank));
  const Dims d = dims3(*marker);
  auto f = [rt = rec_type, mk = marker, ms = mask, o = out, d = d, se = se, s = as_stream(stream)](auto tag) -> int {
      using T = typename decltype(tag)::type;
      if (rt == PM_REC_DILATE) return reconstruct_dev<T, true>((const T *)mk->data, (const T *)ms->data, (T *)o->data, d, se, s);
      return reconstruct_dev<T, false>((const T *)mk->data, (const T *)ms->data, (T *)o->data, d, se, s);
    };
  return dispatch_ptype(marker->ptype, f);
}

// low (LOW = true) / high leveling via the closed forms
template <typename T, bool LOW>
static int leveling_half(const T *ref, const T *marker, T *out, Dims d, int se, hipStream_t s)
{
  const int64_t n = d.D * d.H * d.W;
  Scratch t1, t2;
  PM_TRY(t1.alloc((size_t)n * sizeof(T), s));
  PM_TRY(t2.alloc((size_t)n * sizeof(T), s));
  T *m = (T *)t1.ptr, *r = (T *)t2.ptr;
  PM_TRY((ew2<T, LOW ? EW_SUP : EW_INF>(ref, marker, m, n, s)));
  PM_TRY((reconstruct_dev<T, LOW>(marker, m, r, d, se, s)));
  return ew2<T, LOW ? EW_INF : EW_SUP>(r, ref, out, n, s);
}

// Order of the marker g against the reference f: bit 0 set when some g > f, bit 1 when some
// g < f (both set when a comparison fails, e.g. NaN).  Grid-stride over 16-byte chunks with
// four chunks of each image in flight per thread; one atomicOr per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void order_kernel(const T *__restrict__ f, const T *__restrict__ g, int64_t n,
                                                    int *__restrict__ bits)
{
  constexpr int E = 16 / (int)sizeof(T);  // elements per 16-byte chunk
  constexpr int U = 4;                     // chunks in flight per thread and image
  const int64_t nchunks = n / E;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int b = 0;
  auto cmp = [&](T gv, T fv) {
    if (!(gv <= fv)) b |= 1;
    if (!(gv >= fv)) b |= 2;
  };
  for (int64_t c0 = tid; c0 < nchunks; c0 += U * stride) {
    uint4 fa[U], ga[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t c = c0 + u * stride;
      if (c < nchunks) {
        fa[u] = load16_any(reinterpret_cast<const uint8_t *>(f + c * E));
        ga[u] = load16_any(reinterpret_cast<const uint8_t *>(g + c * E));
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (c0 + u * stride < nchunks) {
        T fv[E], gv[E];
        __builtin_memcpy(fv, &fa[u], 16);
        __builtin_memcpy(gv, &ga[u], 16);
#pragma unroll
        for (int e = 0; e < E; e++) cmp(gv[e], fv[e]);
      }
    }
  }
  for (int64_t i = nchunks * E + tid; i < n; i += stride) cmp(g[i], f[i]);
  // one global atomic per block, and only for bits not yet set: atomics to one address
  // serialise (1024 blocks all setting the same bit cost about as much as the whole pass)
  __shared__ int blk;
  if (threadIdx.x == 0) blk = 0;
  __syncthreads();
  const bool gt = __any(b & 1), lt = __any(b & 2);
  if ((threadIdx.x & 63) == 0 && (gt || lt)) atomicOr(&blk, (gt ? 1 : 0) | (lt ? 2 : 0));
  __syncthreads();
  if (threadIdx.x == 0 && blk) {
    const int seen = __hip_atomic_load(bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((seen | blk) != seen) atomicOr(bits, blk);
  }
}

// order bits of (f, g) ORed into device int *bits (zero), no host synchronisation
template <typename T> static int marker_order_dev(const T *f, const T *g, int64_t n, int *bits, hipStream_t s)
{
  hipLaunchKernelGGL((order_kernel<T>), dim3(grid1d(n / (16 / (int64_t)sizeof(T)) / 4 + 1, 256, 1024)), dim3(256), 0,
                     s, f, g, n, bits);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

template <typename T> static int marker_order(const T *f, const T *g, int64_t n, int *bits, hipStream_t s)
{
  Scratch sc;
  PM_TRY(sc.alloc(sizeof(int), s));
  hipLaunchKernelGGL(zero_flag_kernel, dim3(1), dim3(64), 0, s, (int *)sc.ptr);
  PM_TRY(marker_order_dev<T>(f, g, n, (int *)sc.ptr, s));
  static thread_local int *hbits = nullptr;  // pinned, allocated once per host thread
  if (!hbits) PM_HIP(hipHostMalloc((void **)&hbits, sizeof(int), hipHostMallocDefault));
  PM_HIP(hipMemcpyAsync(hbits, sc.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  PM_HIP(hipStreamSynchronize(s));
  *bits = *hbits;
  return PM_OK;
}

enum { LEV_LOW, LEV_HIGH, LEV_BOTH };
// leveling_t.hpp:34-202 through the closed forms low(f, g) = f ^ R^d_{f v g}(g) and
// high(f, g) = f v R^e_{f ^ g}(g), leveling = low(high(f, g), g) (leveling.cpp:231-237).
// When g <= f everywhere: f v g = f and R^d_f(g) <= f, so low(f, g) = R^d_f(g); f ^ g = g and
// R^e_g(g) = g, so high(f, g) = f v g = f, hence leveling(f, g) = R^d_f(g).  Dually, g >= f
// gives high(f, g) = R^e_f(g) <= g, low(f, g) = f and leveling(f, g) = R^e_f(g).
template <int K>
static int leveling_entry(const pm_image *ref, const pm_image *marker, int se, pm_image *out, void *stream)
{
  const char *what = K == LEV_LOW ? "low_leveling" : K == LEV_HIGH ? "high_leveling" : "leveling";
  PM_TRY(check_binary(ref, marker, out, what));
  PM_TRY(check_geo_se(se, ref->rank));
  const Dims d = dims3(*ref);
  auto f = [rf = ref, mk = marker, o = out, d = d, se = se, s = as_stream(stream)](auto tag) -> int {
      using T = typename decltype(tag)::type;
      const T *r = (const T *)rf->data, *g = (const T *)mk->data;
      T *dst = (T *)o->data;
      const int64_t n = d.D * d.H * d.W;
      if (n == 0) return PM_OK;
      // With the marker on one side of the reference the closed forms collapse (see the
      // function comment): g <= f gives low(f, g) = leveling(f, g) = R^d_f(g) and
      // high(f, g) = f; g >= f gives high(f, g) = leveling(f, g) = R^e_f(g) and low(f, g) = f.
      if constexpr (std::is_same<T, uint8_t>::value) {
        // uint8 2D: assume g <= f (a marker from an opening, the common case) and run its
        // reconstruction at once; its round 0, which reads every pixel of g and f, also checks
        // the assumption and flags a pixel with g > f, after which the later rounds exit at
        // once.  No separate order pass and no host round trip before the rounds; on a failed
        // guess fall through to the general path.
        if (K != LEV_HIGH && rec_u8_applicable(se, d)) {
          bool aborted = false;
          PM_TRY(rec_u8_guarded<true>(g, r, dst, d, se, &aborted, s));
          if (!aborted) return PM_OK;
        }
      }
      int bits = 3;
      PM_TRY(marker_order<T>(r, g, n, &bits, s));
      const bool g_le_f = !(bits & 1), g_ge_f = !(bits & 2);
      if (g_le_f && K != LEV_HIGH) return reconstruct_dev<T, true>(g, r, dst, d, se, s);
      if (g_ge_f && K != LEV_LOW) return reconstruct_dev<T, false>(g, r, dst, d, se, s);
      if ((g_le_f && K == LEV_HIGH) || (g_ge_f && K == LEV_LOW)) {
        PM_HIP(hipMemcpyAsync(dst, r, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, s));
        return PM_OK;
      }
      if (K == LEV_LOW) return leveling_half<T, true>(r, g, dst, d, se, s);
      if (K == LEV_HIGH) return leveling_half<T, false>(r, g, dst, d, se, s);
      Scratch t;
      PM_TRY(t.alloc((size_t)(d.D * d.H * d.W) * sizeof(T), s));
      PM_TRY((leveling_half<T, false>(r, g, (T *)t.ptr, d, se, s)));
      return leveling_half<T, true>((const T *)t.ptr, g, dst, d, se, s);
    };
  return dispatch_ptype(ref->ptype, f);
}

enum { X_HMAX, X_HMIN, X_HCONVEX, X_HCONCAVE, X_DPO, X_DPC };

template <typename T, bool MAX>
static int h_extrema(const T *in, T h, T *out, Dims d, int se, hipStream_t s)
{
  const int64_t n = d.D * d.H * d.W;
  Scratch t;
  PM_TRY(t.alloc((size_t)n * sizeof(T), s));
  PM_TRY((ewc<T, MAX ? EWC_SSUB : EWC_SADD>(in, h, (T *)t.ptr, n, s)));
  return reconstruct_dev<T, MAX>((const T *)t.ptr, in, out, d, se, s);
}

template <int K> static int extrema_entry(const pm_image *in, double hval, int se, pm_image *out, void *stream)
{
  static const char *names[] = {"h_maxima", "h_minima", "h_convex", "h_concave", "dynamic_pseudo_opening",
                                "dynamic_pseudo_closing"};
  PM_TRY(check_unary(in, out, names[K]));
  PM_TRY(check_geo_se(se, in->rank));
  const Dims d = dims3(*in);
  auto f = [in = in, o = out, h = hval, d = d, se = se, s = as_stream(stream)](auto tag) -> int {
      using T = typename decltype(tag)::type;
      const T *src = (const T *)in->data;
      T *dst = (T *)o->data;
      T hv{};
      PM_TRY(scalar_as<T>(h, &hv, names[K]));
      const int64_t n = d.D * d.H * d.W;
      if (K == X_HMAX) return h_extrema<T, true>(src, hv, dst, d, se, s);
      if (K == X_HMIN) return h_extrema<T, false>(src, hv, dst, d, se, s);
      Scratch t;
      PM_TRY(t.alloc((size_t)n * sizeof(T), s));
      T *tmp = (T *)t.ptr;
      if (K == X_HCONVEX) {
        PM_TRY((h_extrema<T, true>(src, hv, tmp, d, se, s)));
        return ew2<T, EW_SSUB>(src, tmp, dst, n, s);
      }
      if (K == X_HCONCAVE) {
        PM_TRY((h_extrema<T, false>(src, hv, tmp, d, se, s)));
        return ew2<T, EW_SSUB>(tmp, src, dst, n, s);
      }
      // dynamic pseudo opening (closing): shifted marker, reconstruction, keep the pixels where
      // the reconstruction equals the shifted marker, reconstruct again (extrema.cpp:75-114)
      constexpr bool OPEN = (K == X_DPO);
      Scratch t2;
      PM_TRY(t2.alloc((size_t)n * sizeof(T), s));
      T *ext = (T *)t2.ptr;
      PM_TRY((ewc<T, OPEN ? EWC_SSUB : EWC_SADD>(src, hv, tmp, n, s)));
      PM_TRY((reconstruct_dev<T, OPEN>(tmp, src, ext, d, se, s)));
      Scratch t3;
      PM_TRY(t3.alloc((size_t)n * sizeof(T), s));
      T *mk = (T *)t3.ptr;
      PM_TRY((dpo_marker<T, OPEN>(ext, tmp, src, mk, n, s)));
      return reconstruct_dev<T, OPEN>(mk, src, dst, d, se, s);
    };
  return dispatch_ptype(in->ptype, f);
}

}  // namespace pm

using namespace pm;

extern "C" {
int pm_reconstruct(int rec_type, const pm_image *marker, const pm_image *mask, int se, pm_image *out, void *stream)
{ PM_TRACE(); return reconstruct_entry(rec_type, marker, mask, se, out, stream); }
int pm_low_leveling(const pm_image *ref, const pm_image *marker, int se, pm_image *out, void *stream)
{ PM_TRACE(); return leveling_entry<LEV_LOW>(ref, marker, se, out, stream); }
int pm_high_leveling(const pm_image *ref, const pm_image *marker, int se, pm_image *out, void *stream)
{ PM_TRACE(); return leveling_entry<LEV_HIGH>(ref, marker, se, out, stream); }
int pm_leveling(const pm_image *ref, const pm_image *marker, int se, pm_image *out, void *stream)
{ PM_TRACE(); return leveling_entry<LEV_BOTH>(ref, marker, se, out, stream); }
int pm_h_maxima(const pm_image *in, double h, int se, pm_image *out, void *stream)
{ PM_TRACE(); return extrema_entry<X_HMAX>(in, h, se, out, stream); }
int pm_h_minima(const pm_image *in, double h, int se, pm_image *out, void *stream)
{ PM_TRACE(); return extrema_entry<X_HMIN>(in, h, se, out, stream); }
int pm_h_convex(const pm_image *in, double h, int se, pm_image *out, void *stream)
{ PM_TRACE(); return extrema_entry<X_HCONVEX>(in, h, se, out, stream); }
int pm_h_concave(const pm_image *in, double h, int se, pm_image *out, void *stream)
{ PM_TRACE(); return extrema_entry<X_HCONCAVE>(in, h, se, out, stream); }
int pm_dynamic_pseudo_opening(const pm_image *in, double h, int se, pm_image *out, void *stream)
{ PM_TRACE(); return extrema_entry<X_DPO>(in, h, se, out, stream); }
int pm_dynamic_pseudo_closing(const pm_image *in, double h, int se, pm_image *out, void *stream)
{ PM_TRACE(); return extrema_entry<X_DPC>(in, h, se, out, stream); }
}
