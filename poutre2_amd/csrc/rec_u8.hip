// rec_u8.hip -- geodesic reconstruction of uint8 2D images (3x3 neighbourhoods) on gfx950.
//
// Reconstruction by dilation of marker g under mask f is the least fixed point of
// x = f ^ delta_B(x) above g ^ f (geo::Reconstruction, mreconstruct_t.hpp:39-195); by erosion
// it is the dual.  The fixed point is unique, so any schedule of monotone updates that only
// uses current (or older) neighbour values reaches the same bits as the reference's
// raster scans + FIFO.
//
// One wavefront owns a 62x62 tile.  Its 64 lanes hold 64 columns (the tile plus one halo
// column on each side) and 64 registers per lane hold the 64 rows (tile plus halo rows), so
// the whole tile lives in VGPRs.  Halo pixels get mask = value, which pins them (the
// update min(mask, max(...)) cannot move them), and pixels outside the image get the
// neutral value as both value and mask; no lane or row needs a special case.
//  * vertical sweep: each lane walks its column top->bottom then bottom->top; per row
//    step one v_max3 over the column window, two DPP wave shifts (wave_shr:1 / wave_shl:1)
//    for the left/right columns, one min with the mask -- 64 pixels per ~6 instructions;
//  * horizontal sweep: the tile is transposed through LDS (bytes) and the same code runs
//    with lanes = rows;
//  * the tile iterates vertical/horizontal sweeps until a full iteration changes nothing.
// Global rounds: a round launches one workgroup per tile; a tile runs in round R only if its
// tag act[t] equals round R's tag; when a tile's border moves it gives its neighbours round
// R+1's tag and stores round R's tag in moved[R].  Rounds are enqueued in batches without host
// synchronisation; round R exits at once if moved[R-1] does not hold round R-1's tag.  A
// call's tags are base + R with a base above every value its scratch holds from earlier calls,
// so the scratch needs no clearing between calls.  Round 0 reads marker ^ mask (marker v mask
// for erosion) instead of the current image, fusing the reference's select_marker step.
#include <mutex>
#include <vector>

#include "erodil.hpp"
#include "pm_common.hpp"
#include "rec_u8.hpp"

namespace pm {

constexpr int RW = 62;        // tile interior side
constexpr int LP = 68;        // LDS transpose pitch (bytes)
constexpr int REC_MAX_ROUNDS = 1 << 16;

enum { NBK_SQUARE = 0, NBK_CROSS = 1, NBK_SEGX = 2, NBK_SEGY = 3 };

// one image of a (batched) reconstruction; its tiles are tile0 .. tile0 + tiles_x*tiles_y - 1
struct RecImage {
  uint8_t *cur;
  const uint8_t *marker, *mask;
  int H, W, tiles_x, tiles_y, tile0;
};

template <typename P> __device__ __forceinline__ P *uniform_ptr(P *p)
{
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u), hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (P *)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int rec_find(const RecImage *im, int n, int t)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (im[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool DIL> __device__ __forceinline__ uint32_t op3(uint32_t a, uint32_t b, uint32_t c)
{
  return DIL ? max(max(a, b), c) : min(min(a, b), c);  // folded into v_max3_u32 / v_min3_u32
}
// value of the lane to the left (x-1) / right (x+1); lanes 0 and 63 are halo and pinned
// (bound_ctrl: lanes without a source read 0; they are the pinned halo lanes, so the value
// is never used, and no register has to be initialised for the DPP "old" operand)
__device__ __forceinline__ uint32_t from_left(uint32_t v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ uint32_t from_right(uint32_t v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }

__device__ __forceinline__ uint32_t mbyte(const uint32_t (&m)[16], int y) { return (m[y >> 2] >> ((y & 3) * 8)) & 0xffu; }

// The packed masks are loop-invariant; unpacked once outside the sweep loop they would take
// 128 registers.  An empty asm that "rewrites" them before each sweep keeps the unpacking in
// the sweep (one bit-field extract per row) and the kernel at a few waves per SIMD.
__device__ __forceinline__ void opaque(uint32_t (&m)[16])
{
#pragma unroll
  for (int k = 0; k < 16; k++) asm volatile("" : "+v"(m[k]));
}

// One sweep along the register (row) index in direction DOWN (increasing y) or up, over
// rows 1..62; returns the lane's sum of values after the sweep (monotone updates: the sum
// moves iff a pixel moved).  NB is the neighbourhood as seen in this layout.
// The window max distributes over rows: for the square, c = max(h(new[y-1]), h(old[y]),
// h(old[y+1])) with h the max over the lane and its two neighbours; the rows not yet
// updated in this sweep (y, y+1) do not depend on the previous row's result, so their part
// is computed off the dependency chain, and the chain from one row to the next is DPP,
// max3, min.
template <bool DIL, int NB, bool DOWN>
__device__ __forceinline__ uint32_t sweep(uint32_t (&v)[64], const uint32_t (&m)[16])
{
  uint32_t sum = v[0] + v[63];
  auto op2 = [](uint32_t a, uint32_t b) { return DIL ? max(a, b) : min(a, b); };
  auto h = [](uint32_t a) { return op3<DIL>(a, from_left(a), from_right(a)); };
  uint32_t h_cur = (NB == NBK_SQUARE) ? h(v[DOWN ? 1 : 62]) : 0u;  // h(old[y]) of the first row
#pragma unroll
  for (int i = 1; i <= 62; i++) {
    const int y = DOWN ? i : 63 - i;
    const int yp = DOWN ? y - 1 : y + 1;  // already updated in this sweep
    const int yn = DOWN ? y + 1 : y - 1;
    uint32_t c;
    if (NB == NBK_SQUARE) {
      const uint32_t h_next = h(v[yn]);            // old row y+1: off the chain
      const uint32_t x = op2(h_cur, h_next);
      c = op3<DIL>(from_left(v[yp]), from_right(v[yp]), op2(v[yp], x));
      h_cur = h_next;
    } else if (NB == NBK_CROSS) {
      const uint32_t x = op3<DIL>(v[y], from_left(v[y]), from_right(v[y]));
      c = op3<DIL>(v[yp], v[yn], x);
    } else if (NB == NBK_SEGY) {  // neighbours along the register index only
      c = op3<DIL>(v[yp], v[y], v[yn]);
    } else {  // NBK_SEGX: neighbours across lanes only
      c = op3<DIL>(v[y], from_left(v[y]), from_right(v[y]));
    }
    const uint32_t mk = mbyte(m, y);
    v[y] = DIL ? min(c, mk) : max(c, mk);
    sum += v[y];
  }
  return sum;
}

// Transpose the 64x64 byte tile held one byte per register (lane = column, register = row)
// through LDS, into lane = row, register = column (the same call transposes back); returns
// the lane's sum of values in the new layout.
__device__ __forceinline__ uint32_t to_other_layout(uint32_t (&v)[64], uint8_t *L, int lane)
{
#pragma unroll
  for (int y = 0; y < 64; y++) L[y * LP + lane] = (uint8_t)v[y];
  __syncthreads();
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t w = *reinterpret_cast<const uint32_t *>(&L[lane * LP + 4 * k]);
#pragma unroll
    for (int b = 0; b < 4; b++) {
      v[4 * k + b] = (w >> (8 * b)) & 0xffu;
      sum += v[4 * k + b];
    }
  }
  __syncthreads();
  return sum;
}

template <bool DIL, int NB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void rec_u8_kernel(
    RecImage one, const RecImage *__restrict__ many, int nimg, int nt, int round, int base, int *__restrict__ act,
    int *__restrict__ moved, int *__restrict__ guard, int *__restrict__ hrep, int rep_round)
{
  const int tag = base + round;                          // this round's tag
  if (round > 0 && moved[round - 1] != tag - 1) return;  // converged in an earlier round
  if (guard && *guard == base) return;                   // speculation failed: nothing to do
  // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin, so XCD x (block
  // x mod 8) takes the contiguous tile range [x * per, (x + 1) * per): horizontally adjacent
  // tiles, which share cache lines, meet in the same L2
  const int per = (nt + 7) >> 3;
  const int t = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (t >= nt) return;
  // every tile runs in round 0 (it initialises its part of cur); later rounds run the tiles
  // tagged by a neighbour whose border moved in the previous round
  if (round > 0 && act[t] != tag) return;
  RecImage im = many ? many[rec_find(many, nimg, t)] : one;
  // the descriptor is the same for the whole wave: keep it (and everything derived from it)
  // in scalar registers
  im.cur = uniform_ptr(im.cur);
  im.marker = uniform_ptr(im.marker);
  im.mask = uniform_ptr(im.mask);
  im.H = __builtin_amdgcn_readfirstlane(im.H);
  im.W = __builtin_amdgcn_readfirstlane(im.W);
  im.tiles_x = __builtin_amdgcn_readfirstlane(im.tiles_x);
  im.tiles_y = __builtin_amdgcn_readfirstlane(im.tiles_y);
  im.tile0 = __builtin_amdgcn_readfirstlane(im.tile0);
  uint8_t *__restrict__ cur = im.cur;
  const uint8_t *__restrict__ mask = im.mask;
  const int H = im.H, W = im.W, tiles_x = im.tiles_x;
  __shared__ uint8_t L[64 * LP];
  const int lane = threadIdx.x;
  const int tl = t - im.tile0;
  const int ty = tl / tiles_x, tx = tl - ty * tiles_x;
  const int x0 = tx * RW;         // first interior column; the tile spans columns x0-1 .. x0+62
  const int gy = ty * RW - 1 + lane;  // this lane's row (layout B: lane = row)
  const bool row_in = (unsigned)gy < (unsigned)H;
  const bool lane_halo = (lane == 0 || lane == 63);
  const uint32_t NEU = DIL ? 0u : 255u;
  const uint8_t *__restrict__ vsrc = (round == 0) ? im.marker : cur;
  // interior tiles read and write whole row segments: 16-byte loads at any byte address
  const bool wide = x0 >= 1 && x0 + 63 <= W;

  // ---- load (layout B: lane = row, register j = column x0 - 1 + j).  mB holds the mask of
  // the lane's row packed four columns per register.
  uint32_t v[64], mB[16];
  if (wide) {
    if (row_in) {
      const uint8_t *pr = vsrc + (int64_t)gy * W + (x0 - 1);
      const uint8_t *pm = mask + (int64_t)gy * W + (x0 - 1);
      uint4 a[4], m[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        a[q] = load16_any(pr + 16 * q);
        m[q] = load16_any(pm + 16 * q);
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t w[4] = {a[q].x, a[q].y, a[q].z, a[q].w};
        mB[4 * q] = m[q].x;
        mB[4 * q + 1] = m[q].y;
        mB[4 * q + 2] = m[q].z;
        mB[4 * q + 3] = m[q].w;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
          for (int b = 0; b < 4; b++) v[16 * q + 4 * k + b] = (w[k] >> (8 * b)) & 0xffu;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 64; j++) v[j] = NEU;
#pragma unroll
      for (int k = 0; k < 16; k++) mB[k] = NEU * 0x01010101u;
    }
  } else {
    // tiles at the left or right edge of the image: byte loads at clamped addresses, 16
    // columns at a time; pixels outside the image become the neutral value
    const int cgy = gy < 0 ? 0 : (gy >= H ? H - 1 : gy);
    const uint8_t *pr = vsrc + (int64_t)cgy * W;
    const uint8_t *pm = mask + (int64_t)cgy * W;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      uint32_t mk_raw[16];
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int x = x0 - 1 + 16 * c + i;
        const int cx = x < 0 ? 0 : (x >= W ? W - 1 : x);
        v[16 * c + i] = pr[cx];
        mk_raw[i] = pm[cx];
      }
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int j = 16 * c + i;
        const int x = x0 - 1 + j;
        const bool in = row_in && (unsigned)x < (unsigned)W;
        v[j] = in ? v[j] : NEU;
        const uint32_t mb = (in ? mk_raw[i] : NEU) << ((j & 3) * 8);
        mB[j >> 2] = (j & 3) ? (mB[j >> 2] | mb) : mb;
      }
      __builtin_amdgcn_sched_barrier(0);  // the next chunk's loads stay after this chunk
    }
  }
  // round 0 starts from marker ^ mask (marker v mask for erosion): the reference's
  // select_marker step; halo pixels (rows 0 / 63 = lanes 0 / 63, columns 0 / 63) get
  // mask = value, which pins them.  With a guard, round 0 also notes whether some interior
  // pixel of the marker lies beyond the mask (above it for a dilation), which a guarded caller
  // assumed not to happen (pixels outside the image hold the neutral value in both).
  bool beyond = false;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t mw = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int j = 4 * k + b;
      uint32_t mk = (mB[k] >> (8 * b)) & 0xffu, val = v[j];
      if (round == 0 && guard && j != 0 && j != 63) beyond |= DIL ? val > mk : val < mk;
      if (round == 0) val = DIL ? min(val, mk) : max(val, mk);
      if (lane_halo || j == 0 || j == 63) mk = val;
      v[j] = val;
      mw |= mk << (8 * b);
    }
    mB[k] = mw;
  }
  if (round == 0 && guard) {
    const bool b = __any(beyond && !lane_halo);
    if (b && lane == 0) {
      *guard = base;  // the call's tag: every writer stores the same value
      if (hrep) hrep[0] = base;
    }
  }

  // border values at the start: rows 1 / 62 (lanes 1 / 62: their row sums), columns 1 / 62
  uint32_t rowsum0 = 0;
#pragma unroll
  for (int j = 1; j <= 62; j++) rowsum0 += v[j];
  const uint32_t left0 = v[1], right0 = v[62];

  constexpr int NBA = NB;  // neighbourhood seen by the sweep along rows of layout A (image y)
  constexpr int NBB = (NB == NBK_SEGX) ? NBK_SEGY : (NB == NBK_SEGY) ? NBK_SEGX : NB;
  // Sweeps right, left (layout B), then down, up (layout A), repeated.  Every sweep step
  // updates a pixel from its whole neighbourhood, so a sweep that moves no pixel proves the
  // tile is at its local fixed point: the loop stops at the first such sweep (sums of the
  // lane's pixels detect a move, the updates being monotone).
  uint32_t sp = 0;
#pragma unroll
  for (int j = 0; j < 64; j++) sp += v[j];
  bool any = false, in_a = false, have_ma = false;
  uint32_t mA[16];
#pragma unroll
  for (int k = 0; k < 16; k++) mA[k] = 0;
  for (int it = 0; it < 4096; it++) {
    opaque(mB);
    uint32_t s = sweep<DIL, NBB, true>(v, mB);
    if (!__any(s != sp)) break;
    sp = s;
    any = true;
    opaque(mB);
    s = sweep<DIL, NBB, false>(v, mB);
    if (!__any(s != sp)) break;
    sp = to_other_layout(v, L, lane);
    in_a = true;
    if (!have_ma) {
      // the mask in layout A (lane = column), built through LDS the first time the tile
      // needs it (tiles that converge in their first row sweeps never do)
#pragma unroll
      for (int j = 0; j < 64; j++) L[j * LP + lane] = (uint8_t)mbyte(mB, j);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 16; k++) mA[k] = *reinterpret_cast<const uint32_t *>(&L[lane * LP + 4 * k]);
      __syncthreads();
      have_ma = true;
    }
    opaque(mA);
    s = sweep<DIL, NBA, true>(v, mA);
    if (!__any(s != sp)) break;
    sp = s;
    opaque(mA);
    s = sweep<DIL, NBA, false>(v, mA);
    if (!__any(s != sp)) break;
    sp = to_other_layout(v, L, lane);
    in_a = false;
  }
  if (in_a) (void)to_other_layout(v, L, lane);

  // ---- write back the interior (always in round 0: it initialises cur), layout B
  if ((any || round == 0) && !lane_halo && row_in) {
    uint8_t *q = cur + (int64_t)gy * W + x0;  // interior columns x0 .. x0 + 61 = v[1 .. 62]
    if (wide) {
      uint32_t P[16];
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const int j = 1 + 4 * k;
        P[k] = v[j] | (v[j + 1] << 8) | (v[j + 2] << 16) | (j + 3 < 64 ? (v[j + 3] << 24) : 0u);
      }
      store16_any(q, P[0], P[1], P[2], P[3]);
      store16_any(q + 16, P[4], P[5], P[6], P[7]);
      store16_any(q + 32, P[8], P[9], P[10], P[11]);
      store12_any(q + 48, P[12], P[13], P[14]);
      store2_any(q + 60, P[15]);
    } else {
      int sx = x0;  // recomputed here: store addresses are not kept across the sweeps
      asm volatile("" : "+v"(sx));
      uint8_t *qq = cur + (int64_t)gy * W + sx;
#pragma unroll
      for (int j = 1; j <= 62; j++)
        if (sx + j - 1 < W) qq[j - 1] = (uint8_t)v[j];
    }
  }
  if (!any) return;
  // ---- which borders moved: rows 1 / 62 (lanes 1 / 62), columns 1 / 62 in any interior lane
  uint32_t rowsum = 0;
#pragma unroll
  for (int j = 1; j <= 62; j++) rowsum += v[j];
  const bool top = __any(lane == 1 && rowsum != rowsum0);
  const bool bot = __any(lane == 62 && rowsum != rowsum0);
  const bool left = __any(!lane_halo && v[1] != left0);
  const bool right = __any(!lane_halo && v[62] != right0);
  if (lane == 0 && (top || bot || left || right)) {
    const int tiles_y = im.tiles_y;
    const int nr = tag + 1;  // round R+1's tag
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        if (dx == 0 && dy == 0) continue;
        const bool need = (dy < 0 && top) || (dy > 0 && bot) || (dx < 0 && left) || (dx > 0 && right);
        const int ny = ty + dy, nx = tx + dx;
        if (need && (unsigned)ny < (unsigned)tiles_y && (unsigned)nx < (unsigned)tiles_x)
          act[im.tile0 + ny * tiles_x + nx] = nr;
      }
    // moved[R] only has to take round R's tag: a plain store (the same value from every
    // writer; the kernel's end publishes it to the next round and to the host copy).  An
    // atomicAdd from every moved tile serialised on the one address: ~4000 device-scope
    // atomics per round bounded rounds 0 and 1 at ~75-80 us each on 4096^2.
    moved[round] = tag;
    if (hrep && round == rep_round) hrep[1] = tag;  // the host's copy (pinned host memory)
  }
}

__global__ void rec_init_kernel(int *act, int n, int *moved, int nm)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) act[i] = 0;
  if (i < nm) moved[i] = 0;
}

bool rec_u8_applicable(int se, Dims d)
{
  return d.D == 1 && d.H > 0 && d.W > 0 && d.H < (1 << 30) && d.W < (1 << 30) &&
         (se == PM_SE_SQUARE_2D || se == PM_SE_CROSS_2D || se == PM_SE_SEGMENT_X_1D || se == PM_SE_SEGMENT_X_2D ||
          se == PM_SE_SEGMENT_Y_2D);
}

// ints of device scratch a reconstruction over nt tiles uses: one activity tag per tile, one
// move tag per round, and with gslot the guard of speculative calls
static size_t rec_scratch_ints(int64_t nt, bool gslot) { return (size_t)nt + REC_MAX_ROUNDS + (gslot ? 1 : 0); }

// Rounds driver shared by the single-image, batched and scratch-free entry points.  `sc` is
// device scratch of rec_scratch_ints(nt, gslot) ints and *hi the largest tag it may hold from
// earlier calls (-1: unknown contents; then it is cleared by a fill kernel and tags start
// again), updated here.  `hmoved`: with pinned (device-visible) host memory the kernels write
// the host's report into it themselves (hmoved[0]: the guard, hmoved[1]: the tag of the
// batch's last round if that round moved), which saves a copy per batch; otherwise it takes
// PM_COUNTERS move tags and the guard by copies.
// aborted != nullptr (needs gslot): the call is speculative.  Round 0 then stores the call's base tag in the
// guard when some marker pixel lies beyond the mask (above it for a dilation), every round
// exits at once when the guard holds it, and the host reports *aborted = true.
template <bool DIL>
static int rec_rounds(const RecImage &one, const RecImage *many, int nimg, int nt, int se, int *sc, bool gslot,
                      int *hi, int *hmoved, bool pinned, hipStream_t s, bool *aborted = nullptr)
{
  if (aborted && !gslot) PM_FAIL(PM_E_ARG, "reconstruction: a speculative call needs a guard slot");
  int *act = sc;
  int *moved = act + nt;
  int *guard = aborted ? moved + REC_MAX_ROUNDS : nullptr;
  const int cap = round_cap(REC_MAX_ROUNDS);
  if (*hi < 0 || *hi > INT32_MAX - cap - 2) {
    const int nm = REC_MAX_ROUNDS + (gslot ? 1 : 0), ninit = nt > nm ? nt : nm;
    hipLaunchKernelGGL(rec_init_kernel, dim3((ninit + 255) / 256), dim3(256), 0, s, act, nt, moved, nm);
    PM_HIP(hipGetLastError());
    *hi = 0;
  }
  const int base = *hi + 1;  // tags base .. base + rounds: above every value sc holds
  *hi = -1;                  // until the rounds below are known to have ended
  int nbk = NBK_SQUARE;
  if (se == PM_SE_CROSS_2D) nbk = NBK_CROSS;
  else if (se == PM_SE_SEGMENT_X_1D || se == PM_SE_SEGMENT_X_2D) nbk = NBK_SEGX;
  else if (se == PM_SE_SEGMENT_Y_2D) nbk = NBK_SEGY;
  const unsigned grid = 8u * (unsigned)((nt + 7) / 8);  // every tile index below 8 * ceil(nt / 8)
  int round = 0, batch = 8;
  while (round < cap) {
    const int nb = (round + batch <= cap) ? batch : cap - round;
    int *hrep = pinned ? hmoved : nullptr;
    const int rep_round = round + nb - 1;
    if (pinned) hmoved[0] = hmoved[1] = 0;  // no tag is 0; the previous batch has ended
    for (int i = 0; i < nb; i++) {
      const int R = round + i;
      switch (nbk) {
      case NBK_SQUARE:
        hipLaunchKernelGGL((rec_u8_kernel<DIL, NBK_SQUARE>), dim3(grid), dim3(64), 0, s, one, many, nimg, nt, R, base,
                           act, moved, guard, hrep, rep_round);
        break;
      case NBK_CROSS:
        hipLaunchKernelGGL((rec_u8_kernel<DIL, NBK_CROSS>), dim3(grid), dim3(64), 0, s, one, many, nimg, nt, R, base,
                           act, moved, guard, hrep, rep_round);
        break;
      case NBK_SEGX:
        hipLaunchKernelGGL((rec_u8_kernel<DIL, NBK_SEGX>), dim3(grid), dim3(64), 0, s, one, many, nimg, nt, R, base,
                           act, moved, guard, hrep, rep_round);
        break;
      default:
        hipLaunchKernelGGL((rec_u8_kernel<DIL, NBK_SEGY>), dim3(grid), dim3(64), 0, s, one, many, nimg, nt, R, base,
                           act, moved, guard, hrep, rep_round);
        break;
      }
    }
    PM_HIP(hipGetLastError());
    if (!pinned) {
      PM_HIP(hipMemcpyAsync(hmoved, moved + round, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
      if (guard) PM_HIP(hipMemcpyAsync(hmoved + PM_COUNTERS, guard, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    PM_HIP(hipStreamSynchronize(s));
    *hi = base + round + nb;  // the largest tag written so far (activity tags of round round+nb)
    if (guard && (pinned ? hmoved[0] : hmoved[PM_COUNTERS]) == base) {
      *aborted = true;
      return PM_OK;
    }
    // rounds after the first that moved nothing exit at once, so the batch's last round moved
    // iff every round of the batch did
    if ((pinned ? hmoved[1] : hmoved[nb - 1]) != base + rep_round) return PM_OK;
    round += nb;
    batch = batch * 2 > PM_COUNTERS ? PM_COUNTERS : batch * 2;
  }
  PM_FAIL(PM_E_NOCONV, "reconstruction did not converge");
}

// Device scratch of the rounds, reused across calls.  Every call synchronises its stream
// before it returns, after which its buffer is idle, so a per-process list of idle buffers
// (under a mutex, for callers on several host threads) replaces a stream-ordered allocation
// and free per call: hipFreeAsync alone took 23-36 us of host time per 4096^2 call, after the
// synchronisation, i.e. between the GPU's end of one call and the start of the next.  Each
// buffer carries the largest tag it holds (rec_rounds), so it is never cleared again.
struct RecBuf {
  int dev = 0;
  int *ptr = nullptr;
  size_t ints = 0;
  int hi = -1;
};
constexpr size_t REC_BUF_KEEP = 8;  // idle buffers kept (the smallest beyond this are freed)
static std::mutex rec_buf_mu;
static std::vector<RecBuf> rec_buf_idle;

struct RecScratch {
  RecBuf b;
  bool idle = false;  // set when the call's rounds ended with a synchronisation
  RecScratch() = default;
  RecScratch(const RecScratch &) = delete;
  RecScratch &operator=(const RecScratch &) = delete;
  int take(size_t ints)
  {
    if (take_forced_alloc_failure())
      PM_FAIL(PM_E_HIP, "scratch allocation of %zu bytes failed: out of device memory (forced by the testing hook)",
              ints * sizeof(int));
    int dev = 0;
    PM_HIP(hipGetDevice(&dev));
    {
      std::lock_guard<std::mutex> lock(rec_buf_mu);
      int best = -1;
      for (int i = 0; i < (int)rec_buf_idle.size(); i++) {
        const RecBuf &c = rec_buf_idle[i];
        if (c.dev == dev && c.ints >= ints && (best < 0 || c.ints < rec_buf_idle[best].ints)) best = i;
      }
      if (best >= 0) {
        b = rec_buf_idle[best];
        rec_buf_idle.erase(rec_buf_idle.begin() + best);
        return PM_OK;
      }
    }
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, ints * sizeof(int));
    if (e != hipSuccess) PM_FAIL(PM_E_HIP, "scratch allocation of %zu bytes failed: %s", ints * sizeof(int), hipGetErrorString(e));
    b.dev = dev;
    b.ptr = (int *)p;
    b.ints = ints;
    b.hi = -1;
    return PM_OK;
  }
  ~RecScratch()
  {
    if (!b.ptr) return;
    if (!idle) {  // an error left work queued: hipFree waits for the device first
      (void)hipFree(b.ptr);
      return;
    }
    std::vector<RecBuf> drop;
    {
      std::lock_guard<std::mutex> lock(rec_buf_mu);
      rec_buf_idle.push_back(b);
      while (rec_buf_idle.size() > REC_BUF_KEEP) {
        int small = 0;
        for (int i = 1; i < (int)rec_buf_idle.size(); i++)
          if (rec_buf_idle[i].ints < rec_buf_idle[small].ints) small = i;
        drop.push_back(rec_buf_idle[small]);
        rec_buf_idle.erase(rec_buf_idle.begin() + small);
      }
    }
    for (const RecBuf &d : drop) (void)hipFree(d.ptr);  // idle buffers: nothing in flight uses them
  }
};

template <bool DIL>
static int rec_rounds_cached(const RecImage &one, const RecImage *many, int nimg, int nt, int se, hipStream_t s,
                             bool *aborted = nullptr)
{
  RecScratch sc;
  PM_TRY(sc.take(rec_scratch_ints(nt, true)));  // every cached buffer has the guard slot
  int *h = nullptr;
  PM_TRY(pinned_counters(&h));
  const int rc = rec_rounds<DIL>(one, many, nimg, nt, se, sc.b.ptr, true, &sc.b.hi, h, true, s, aborted);
  sc.idle = (rc == PM_OK || rc == PM_E_NOCONV);  // both return after the stream synchronisation
  return rc;
}

static int rec_describe(const uint8_t *marker, const uint8_t *mask, uint8_t *out, Dims d, int64_t tile0, RecImage &r)
{
  const int64_t tx = (d.W + RW - 1) / RW, ty = (d.H + RW - 1) / RW;
  if (tile0 + tx * ty > INT32_MAX / 2) PM_FAIL(PM_E_ARG, "image too large for the reconstruction kernel");
  r.cur = out;
  r.marker = marker;
  r.mask = mask;
  r.H = (int)d.H;
  r.W = (int)d.W;
  r.tiles_x = (int)tx;
  r.tiles_y = (int)ty;
  r.tile0 = (int)tile0;
  return PM_OK;
}

template <bool DIL>
int rec_u8(const uint8_t *marker, const uint8_t *mask, uint8_t *out, Dims d, int se, hipStream_t s)
{
  RecImage one;
  PM_TRY(rec_describe(marker, mask, out, d, 0, one));
  return rec_rounds_cached<DIL>(one, nullptr, 1, one.tiles_x * one.tiles_y, se, s);
}

template <bool DIL>
int rec_u8_guarded(const uint8_t *marker, const uint8_t *mask, uint8_t *out, Dims d, int se, bool *aborted,
                   hipStream_t s)
{
  *aborted = false;
  RecImage one;
  PM_TRY(rec_describe(marker, mask, out, d, 0, one));
  return rec_rounds_cached<DIL>(one, nullptr, 1, one.tiles_x * one.tiles_y, se, s, aborted);
}

// ---------------------------------------------------------------- scratch-free reconstruction
// Device scratch of the scratch-free entry point: one array per GPU, held until the process
// ends.  Host code never uses the array's name as an address; it uses the address that
// hipGetSymbolAddress returns for the current device, kept per device.
constexpr size_t NOALLOC_SCRATCH_INTS = (size_t)1 << 18;  // 1 MiB
static __device__ __align__(256) int scratch[NOALLOC_SCRATCH_INTS];

// The array's address on the current device and the largest tag it holds (see rec_rounds;
// -1 until its first use on that device).
static int noalloc_scratch(int **ptr, int **hi)
{
  static std::mutex mu;
  static int *addr[PM_MAX_DEVICES] = {};
  static int tags[PM_MAX_DEVICES];
  int dev = 0;
  PM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= PM_MAX_DEVICES) PM_FAIL(PM_E_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      PM_FAIL(PM_E_HIP, "scratch-free reconstruction: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
    addr[dev] = (int *)p;
    tags[dev] = -1;
  }
  *ptr = addr[dev];
  *hi = &tags[dev];
  return PM_OK;
}

template int rec_u8<true>(const uint8_t *, const uint8_t *, uint8_t *, Dims, int, hipStream_t);
template int rec_u8<false>(const uint8_t *, const uint8_t *, uint8_t *, Dims, int, hipStream_t);
template int rec_u8_guarded<true>(const uint8_t *, const uint8_t *, uint8_t *, Dims, int, bool *, hipStream_t);
template int rec_u8_guarded<false>(const uint8_t *, const uint8_t *, uint8_t *, Dims, int, bool *, hipStream_t);

}  // namespace pm

using namespace pm;

// Batched reconstruction: the tiles of all images go through the same rounds.
extern "C" int pm_batch_reconstruct(int n, int rec_type, const pm_image *markers, const pm_image *masks, int se,
                                    pm_image *outs, void *stream)
{ PM_TRACE();
  clear_error();
  if (n < 0 || (n > 0 && (!markers || !masks || !outs))) PM_FAIL(PM_E_ARG, "batch reconstruct: bad arguments");
  if (rec_type != PM_REC_DILATE && rec_type != PM_REC_ERODE) PM_FAIL(PM_E_ARG, "Unsupported reconstruction type");
  if (n == 0) return PM_OK;
  hipStream_t s = as_stream(stream);
  std::vector<RecImage> host(n);
  int64_t tiles = 0;
  for (int i = 0; i < n; i++) {
    PM_TRY(check_binary(&markers[i], &masks[i], &outs[i], "Reconstruction"));
    if (markers[i].ptype != PM_PTYPE_U8) PM_FAIL(PM_E_UNSUPPORTED, "batch reconstruct: uint8 images only");
    PM_TRY(check_se_rank(se, markers[i].rank));
    const Dims d = dims3(markers[i]);
    if (!rec_u8_applicable(se, d)) PM_FAIL(PM_E_UNSUPPORTED, "batch reconstruct: 2D images with a 3x3 SE only");
    PM_TRY(rec_describe((const uint8_t *)markers[i].data, (const uint8_t *)masks[i].data, (uint8_t *)outs[i].data, d,
                        tiles, host[i]));
    tiles += (int64_t)host[i].tiles_x * host[i].tiles_y;
  }
  Scratch dsc;
  PM_TRY(dsc.alloc(sizeof(RecImage) * (size_t)n, s));
  PM_HIP(hipMemcpyAsync(dsc.ptr, host.data(), sizeof(RecImage) * (size_t)n, hipMemcpyHostToDevice, s));
  if (rec_type == PM_REC_DILATE) return rec_rounds_cached<true>(host[0], (const RecImage *)dsc.ptr, n, (int)tiles, se, s);
  return rec_rounds_cached<false>(host[0], (const RecImage *)dsc.ptr, n, (int)tiles, se, s);
}

// pm_reconstruct for uint8 2D images and 3x3-class SEs, allocating nothing: the activity tags
// and round counters live in the file-scope device array `scratch` (the convergence counters
// are read into host stack memory).  Calls on one GPU share that array: they must not overlap.
extern "C" int pm_reconstruct_noalloc(int rec_type, const pm_image *marker, const pm_image *mask, int se,
                                      pm_image *out, void *stream)
{ PM_TRACE();
  clear_error();
  PM_TRY(check_binary(marker, mask, out, "Reconstruction"));
  if (rec_type != PM_REC_DILATE && rec_type != PM_REC_ERODE) PM_FAIL(PM_E_ARG, "Unsupported reconstruction type");
  PM_TRY(check_se_rank(se, marker->rank));
  if (marker->ptype != PM_PTYPE_U8)
    PM_FAIL(PM_E_UNSUPPORTED, "scratch-free reconstruction: uint8 images only");
  const Dims d = dims3(*marker);
  if (d.D * d.H * d.W == 0) return PM_OK;
  if (!rec_u8_applicable(se, d))
    PM_FAIL(PM_E_UNSUPPORTED, "scratch-free reconstruction: 1D/2D images with a 3x3-class SE only");
  RecImage one;
  PM_TRY(rec_describe((const uint8_t *)marker->data, (const uint8_t *)mask->data, (uint8_t *)out->data, d, 0, one));
  const int64_t nt = (int64_t)one.tiles_x * one.tiles_y;
  if (rec_scratch_ints(nt, false) > NOALLOC_SCRATCH_INTS)
    PM_FAIL(PM_E_ARG, "scratch-free reconstruction: %lld tiles of %dx%d exceed the %zu-tile limit", (long long)nt, RW,
            RW, NOALLOC_SCRATCH_INTS - REC_MAX_ROUNDS);
  int *sc = nullptr, *hi = nullptr;
  PM_TRY(noalloc_scratch(&sc, &hi));
  int hmoved[PM_COUNTERS + 1];
  hipStream_t s = as_stream(stream);
  if (rec_type == PM_REC_DILATE) return rec_rounds<true>(one, nullptr, 1, (int)nt, se, sc, false, hi, hmoved, false, s);
  return rec_rounds<false>(one, nullptr, 1, (int)nt, se, sc, false, hi, hmoved, false, s);
}
