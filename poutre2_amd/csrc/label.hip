// label.hip -- connected-component labelling: label_binary, label_flat_zones, label_minima,
// label_maxima (include/poutre/label/label_binary.hpp:35, label_flat_zones.hpp:35,
// label_extrema.hpp:34-39; algorithm of label_t.hpp t_label_helper).
//
// The reference makes one raster pass with a union-find over provisional labels and
// renumbers the roots; its result is fixed by the link graph alone: a pixel is linked to
// each raster-earlier ("upper") neighbour of the SE that satisfies the relation (binary:
// both pixels are foreground and the earlier one is > 0; flat zones: equal values), and each
// component is numbered 1.. by the raster position of its first pixel (the first pixel of a
// component always opens a new provisional label, which stays the component's root because
// the union keeps the smaller root).  On the GPU:
//  1. lab_tile2d / lab_tile: per 32 x 64 tile (8 x 8 x 32 in 3D), the links inside the tile
//                   are merged in LDS (in 2D the links along a row by a ballot over the
//                   row's runs) and every pixel gets the global index of its tile-local root
//                   (2D: each tile also lists its roots);
//  2. the links that cross tile borders are merged in global memory with a lock-free union
//     whose roots only decrease (atomicMin on the larger root, retried from the value it
//     returns), so each root ends as the smallest index of its component.  2D:
//     lab_border_level, level by level (aligned groups of 4 x 4, 16 x 16, ... tiles become
//     one), each level followed by lab_compress_roots, which points every listed tile root
//     at its root; 3D: lab_border in one pass, then lab_compress (P[i] = root);
//  3. lab_count_rank: per block of LAB_RB pixels, the number of roots and each root's rank
//     among them; lab_scan_counts: block offsets and the number of components;
//  4. lab_output:   out[i] = 1 + rank of the root of i (int64, the reference's GINT64
//                   output), 0 for the background.
// Positive float values at or below the reference's 1e-10 background threshold are not
// linked (the reference reads an unassigned provisional label for them).
#include <cmath>

#include "erodil.hpp"
#include "pm_common.hpp"

namespace pm {

constexpr int LAB_RB = 4096;  // pixels per block of the root ranking (256 threads x 16)
constexpr int LAB_U2 = 4;     // upper offsets of a 2D neighbourhood: (-1, -1), (-1, 0), (-1, 1), (0, -1)

constexpr int LAB_LS = 2;     // tile-coordinate bits merged per border level (groups of 4 x 4)

// level at which two tiles whose coordinates differ by the bits of v (x or y) become one group
__device__ __forceinline__ int lab_level(int v) { return ((v ? 32 - __clz(v) : 0) + LAB_LS - 1) / LAB_LS; }
// the k-th multiple m >= 1 that is not a multiple of 2^LAB_LS (boundary b = m << hb of a level)
__device__ __forceinline__ int lab_kth_boundary(int k) { return k + k / ((1 << LAB_LS) - 1) + 1; }
inline int lab_bitlen_host(int v) { return v ? 32 - __builtin_clz((unsigned)v) : 0; }

struct LabUpper {
  int n;
  int dz[13], dy[13], dx[13];
};

template <typename T> __device__ __forceinline__ bool lab_fg(T v) { return v != T(0); }
template <> __device__ __forceinline__ bool lab_fg<float>(float v) { return fabsf(v) > 1E-10F; }
template <> __device__ __forceinline__ bool lab_fg<double>(double v) { return fabs(v) > 1E-10; }

template <typename T, bool FLAT> __device__ __forceinline__ bool lab_accept(T v) { return FLAT || lab_fg<T>(v); }
// v at the pixel, w at its raster-earlier neighbour
template <typename T, bool FLAT> __device__ __forceinline__ bool lab_link(T v, T w)
{
  if (FLAT) return w == v;
  return lab_fg<T>(v) && lab_fg<T>(w) && w > T(0);
}

// Root of x in global memory, without path compression: while unions run, the parents in
// global memory change by atomics only (plain stores going through another XCD's cache
// could overwrite a link), and halving by atomics measured slower (4096^2 binary 0.65 ->
// 0.74 ms); lab_compress / lab_compress_roots flatten the trees between union passes.
__device__ __forceinline__ int lab_find(const int *P, int x)
{
  int y;
  while ((y = P[x]) != x) x = y;
  return x;
}

// Merge the sets of a and b.  Invariant: P[x] <= x and values only decrease, so every find
// terminates and every retry starts from a strictly smaller index.
__device__ __forceinline__ void lab_unite(int *P, int a, int b)
{
  for (;;) {
    a = lab_find(P, a);
    b = lab_find(P, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&P[a], b);  // link the larger root under the smaller one
    if (old == a) return;
    a = old;  // a had been linked meanwhile: keep its new parent connected to b
  }
}

// ---- tiled form: union-find inside a TD x TH x TW tile in LDS, then only the links that
// cross tile borders go through global memory.
// Root of x in LDS with path halving (x's parent moved up to its grandparent as the walk
// goes).  A halving store may race with a union lowering L[x]; unions only link roots and,
// finding x already linked, go on with x's parent, so x keeps an ancestor (a smaller index)
// as parent either way and no link is lost.
__device__ __forceinline__ int lab_find_lds(int *L, int x)
{
  for (;;) {
    const int p = L[x];
    if (p == x) return x;
    const int gp = L[p];
    if (gp == p) return p;
    L[x] = gp;
    x = gp;
  }
}
__device__ __forceinline__ void lab_unite_lds(int *L, int a, int b)
{
  for (;;) {
    a = lab_find_lds(L, a);
    b = lab_find_lds(L, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&L[a], b);
    if (old == a) return;
    a = old;
  }
}

// Tile (blockIdx.x, .y, .z): local labels in LDS (local index = (lz*TH + ly)*TW + lx, which
// orders the tile's pixels as the raster does, so a local root is the component's first
// pixel in the tile), then P[i] = global index of the local root (-1 for background).
template <typename T, bool FLAT, int TD, int TH, int TW>
__global__ __launch_bounds__(256) void lab_tile(const T *__restrict__ in, int *__restrict__ P, int D, int H, int W,
                                                LabUpper u)
{
  constexpr int NT = TD * TH * TW;
  __shared__ int L[NT];
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH, z0 = blockIdx.z * TD;
  for (int e = threadIdx.x; e < NT; e += 256) {
    const int lx = e % TW, ly = (e / TW) % TH, lz = e / (TW * TH);
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    bool a = false;
    if (x < W && y < H && z < D) a = lab_accept<T, FLAT>(in[((int64_t)z * H + y) * W + x]);
    L[e] = a ? e : -1;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NT; e += 256) {
    if (L[e] < 0) continue;  // background or outside the image (a pixel's own entry is only
                             // lowered by unions, never to -1)
    const int lx = e % TW, ly = (e / TW) % TH, lz = e / (TW * TH);
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    const T v = in[((int64_t)z * H + y) * W + x];
    for (int k = 0; k < u.n; k++) {
      const int nx = lx + u.dx[k], ny = ly + u.dy[k], nz = lz + u.dz[k];
      if ((unsigned)nx >= (unsigned)TW || (unsigned)ny >= (unsigned)TH || (unsigned)nz >= (unsigned)TD) continue;
      if (x0 + nx >= W || y0 + ny >= H || z0 + nz >= D) continue;
      const T w = in[((int64_t)(z0 + nz) * H + (y0 + ny)) * W + (x0 + nx)];
      if (lab_link<T, FLAT>(v, w)) lab_unite_lds(L, e, (nz * TH + ny) * TW + nx);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NT; e += 256) {
    const int lx = e % TW, ly = (e / TW) % TH, lz = e / (TW * TH);
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    if (x >= W || y >= H || z >= D) continue;
    int g = -1;
    if (L[e] >= 0) {
      const int r = lab_find_lds(L, e);
      const int rx = r % TW, ry = (r / TW) % TH, rz = r / (TW * TH);
      g = (int)(((int64_t)(z0 + rz) * H + (y0 + ry)) * W + (x0 + rx));
    }
    P[((int64_t)z * H + y) * W + x] = g;
  }
}

// value of lane - 1 (lane 0: zero), by DPP wave_shr:1
template <typename T> __device__ __forceinline__ T lab_from_left(T v)
{
  if constexpr (sizeof(T) == 8) {
    uint32_t b[2];
    __builtin_memcpy(b, &v, 8);
    b[0] = (uint32_t)__builtin_amdgcn_mov_dpp((int)b[0], 0x138, 0xf, 0xf, true);
    b[1] = (uint32_t)__builtin_amdgcn_mov_dpp((int)b[1], 0x138, 0xf, 0xf, true);
    __builtin_memcpy(&v, b, 8);
    return v;
  } else {
    uint32_t b = 0;
    __builtin_memcpy(&b, &v, sizeof(T));
    b = (uint32_t)__builtin_amdgcn_mov_dpp((int)b, 0x138, 0xf, 0xf, true);
    T r;
    __builtin_memcpy(&r, &b, sizeof(T));
    return r;
  }
}

// start column of the run that holds column c, given the row's mask of run starts
__device__ __forceinline__ int lab_run_start(uint64_t starts, int c)
{
  const uint64_t m = c == 63 ? starts : (starts & ((2ull << c) - 1));
  return 63 - __clzll((long long)m);
}

// 2D tiles of TH rows x 64 columns, one wave per row at a time (lane = column).
//  1. The link to the left pixel of the same row is resolved without unions: a ballot marks
//     the lanes that start a run (not linked to their left neighbour) and every pixel points
//     at the start of its run (row masks of run starts and the pixel values kept in LDS).
//  2. Links to the row above (dy = -1, dx in the mask UPM, or upm when UPM < 0: bit dx + 1)
//     join the pixel's run with the run above it: one LDS union per pair of runs, met first
//     by a lane in raster order (tests on ballot masks skip the others), the owed unions of
//     all the wave's rows run in one loop.
//  3. Every run start is pointed at its root (no union runs meanwhile), so a pixel's root is
//     one LDS read away: P[pixel] = global index of the root (-1 for the background).
// The tile's roots (global indices, in no particular order) go to RL[tile * cap ...] (cap: at
// least the tile's pixel count) and their number to TC[tile], tile = blockIdx.y * gridDim.x +
// blockIdx.x.
template <typename T, bool FLAT, int TH, int UPM>
__global__ __launch_bounds__(256) void lab_tile2d(const T *__restrict__ in, int *__restrict__ P, int H, int W, int upm,
                                                  int xlink, int *__restrict__ RL, int *__restrict__ TC, int cap)
{
  constexpr int TW = 64, NT = TH * TW, RPW = TH / 4;  // rows per wave
  __shared__ int L[NT];
  __shared__ T V[NT];
  __shared__ uint64_t M[TH];
  __shared__ int nroot;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH, x = x0 + lane;
  const int64_t tile = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  if (threadIdx.x == 0) nroot = 0;
  T v[RPW];
  bool acc[RPW];
  int st[RPW];
  uint64_t sm[RPW];
  // all rows' loads in flight at once: clamped addresses, no branch around the loads
  const int xc = x < W ? x : W - 1;
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int y = y0 + wv + 4 * k;
    v[k] = in[(int64_t)(y < H ? y : H - 1) * W + xc];
  }
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int ly = wv + 4 * k, y = y0 + ly;
    const bool in_img = x < W && y < H;
    if (!in_img) v[k] = T(0);
    acc[k] = in_img && lab_accept<T, FLAT>(v[k]);
    const T left = lab_from_left(v[k]);
    const bool lk = xlink && acc[k] && lane > 0 && lab_link<T, FLAT>(v[k], left);
    sm[k] = __ballot(!lk);
    st[k] = lab_run_start(sm[k], lane);
    L[ly * TW + lane] = acc[k] ? ly * TW + st[k] : -1;
    V[ly * TW + lane] = v[k];
    if (lane == 0) M[ly] = sm[k];
  }
  __syncthreads();
  // the unions this lane owes: bit 3k + dx + 1 for row wv + 4k and offset dx (bitwise
  // logic, no branch: every lane reads V at a clamped column)
  const int um = UPM >= 0 ? UPM : upm;
  uint32_t pend = 0;
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int ly = wv + 4 * k;  // the same for the whole wave
    if (ly == 0) continue;
    const uint64_t above = M[ly - 1];
    const bool start = (sm[k] >> lane) & 1;
    bool cov = false;  // an earlier offset of this lane reached the run above that holds column lane + dx
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      const int nx = lane + dx;
      const bool inx = ((unsigned)nx < (unsigned)TW) & (x0 + nx < W);
      const int nxc = inx ? nx : lane;
      const bool above_start = inx & (((above >> nxc) & 1) != 0);
      cov = cov & !above_start;
      if (!((um >> (dx + 1)) & 1)) continue;  // the same for every lane
      // acc: the pixel is in the image, so the row above is too
      const bool lk = acc[k] & inx & lab_link<T, FLAT>(v[k], V[(ly - 1) * TW + nxc]);
      const uint64_t lkm = __ballot(lk);
      // the same pair of runs was met earlier: by lane - 1 with this offset, by lane - 1 with
      // offset dx + 1 (the same column, and lane - 1 links to it exactly when this lane does:
      // same run), or by this lane with a smaller offset
      const bool left_lk = (((lkm << 1) >> lane) & 1) != 0;  // lane - 1 links (0 for lane 0)
      const bool dup = (left_lk & !start & !above_start) | (!start & (((um >> (dx + 2)) & 1) != 0)) | cov;
      pend |= (uint32_t)(lk & !dup) << (3 * k + dx + 1);
      cov = cov | lk;
    }
  }
  // unions of the pixel with the pixel above (their parents are the two run starts)
  while (pend) {
    const int bit = __builtin_ctz(pend);
    pend &= pend - 1;
    const int k = bit / 3, dx = bit - 3 * k - 1, ly = wv + 4 * k;
    lab_unite_lds(L, ly * TW + lane, (ly - 1) * TW + lane + dx);
  }
  __syncthreads();
  // run starts -> roots: a walk without halving, so the only store to L[e] is this one
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int ly = wv + 4 * k, e = ly * TW + lane;
    if (acc[k] && st[k] == lane) {
      int r = e, q;
      while ((q = L[r]) != r) r = q;
      L[e] = r;
    }
  }
  __syncthreads();
  // global index of each pixel's root (N < 2^31: int arithmetic); the wave's roots get their
  // list slots from one LDS atomic
  int g[RPW];
  uint64_t rm[RPW];
  int nr = 0;
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    const int ly = wv + 4 * k;
    const unsigned r = (unsigned)L[ly * TW + st[k]];
    g[k] = acc[k] ? (y0 + (int)(r >> 6)) * W + x0 + (int)(r & 63) : -1;
    rm[k] = __ballot(acc[k] && r == (unsigned)(ly * TW + lane));
    nr += __popcll(rm[k]);
  }
  int base = 0;
  if (lane == 0 && nr) base = atomicAdd(&nroot, nr);
  base = __builtin_amdgcn_readfirstlane(base);
  const uint64_t below = (1ull << lane) - 1;
#pragma unroll
  for (int k = 0; k < RPW; k++) {
    if ((rm[k] >> lane) & 1) RL[tile * cap + base + __popcll(rm[k] & below)] = g[k];
    base += __popcll(rm[k]);
    const int y = y0 + wv + 4 * k;
    if (x < W && y < H) P[(int64_t)y * W + x] = g[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) TC[tile] = nroot;
}

// The links whose two pixels lie in different tiles, merged in global memory.  After lab_tile
// a pixel's parent is its tile-local root, and it stays so (unions only relink roots), so a
// link can be told by the pair (P[i], P[j]) of tile roots: when the previous lane merged the
// same pair -- consecutive pixels along a tile's top edge mostly do -- the union is skipped.
template <typename T, bool FLAT, int TD, int TH, int TW>
__global__ __launch_bounds__(256) void lab_border(const T *__restrict__ in, int *P, int D, int H, int W, LabUpper u)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t N = (int64_t)D * H * W;
  int x = 0, y = 0, z = 0;
  bool cand = false;
  if (i < N) {
    x = (int)(i % W);
    const int64_t r = i / W;
    y = (int)(r % H);
    z = (int)(r / H);
    const int lx = x % TW, ly = y % TH, lz = z % TD;
    // only pixels on a tile face can have an upper neighbour in another tile (upper offsets
    // have dz <= 0; with dz = -1 they may step +1 in y or x, with dy = -1 +1 in x)
    cand = lx == 0 || lx == TW - 1 || ly == 0 || ly == TH - 1 || (TD > 1 && lz == 0);
  }
  if (!__any(cand)) return;  // the whole wave: no shuffle below loses a participant
  const int ra = cand ? P[i] : -1;
  cand = cand && ra >= 0;
  const T v = cand ? in[i] : T(0);
  const int lane = threadIdx.x & 63;
  for (int k = 0; k < u.n; k++) {
    bool lk = false;
    int rb = -1;
    int64_t j = 0;
    if (cand) {
      const int zz = z + u.dz[k], yy = y + u.dy[k], xx = x + u.dx[k];
      const bool other_tile = (zz / TD != z / TD) || (yy / TH != y / TH) || (xx / TW != x / TW) || zz < 0 || yy < 0 || xx < 0;
      if (other_tile && (unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W) {
        j = ((int64_t)zz * H + yy) * W + xx;
        lk = lab_link<T, FLAT>(v, in[j]);
        if (lk) rb = P[j];
      }
    }
    const int pra = __shfl_up(lk ? ra : -2, 1), prb = __shfl_up(rb, 1);
    const bool dup = lane != 0 && pra == ra && prb == rb;
    if (lk && !dup) lab_unite(P, (int)i, (int)j);
  }
}

// 2D links between tiles, one level at a time.  Level l joins the aligned groups of
// G^(l-1) x G^(l-1) tiles into groups of G^l x G^l (G = 2^LAB_LS), so tiles (tx, ty) and
// (tx', ty') meet at level lab_level((tx ^ tx') | (ty ^ ty')) = ceil(bitlen / LAB_LS), and a link
// between neighbouring pixels is merged at that level.  Level l's links start on the rows
// y = b * TH (links whose y part is level l) and on the columns x = b * TW (dx = -1) and
// b * TW - 1 (dx = +1) (links whose x part is level l and y part below l), b a boundary whose
// lowest set bit lies in [(l-1) LAB_LS, l LAB_LS); items: nby rows of W pixels, then nbx column
// pairs of 2 x H pixels.  Between levels lab_compress_roots points every tile root at its
// root, so a level's finds walk a few links (pixel, tile root, group roots merged at this
// level).  u: the SE's upper offsets in the plane (dz = 0), at most LAB_U2.
template <typename T, bool FLAT, int TH, int TW>
__global__ __launch_bounds__(256) void lab_border_level(const T *__restrict__ in, int *P, int H, int W, LabUpper u,
                                                        int l, int nby, int nbx)
{
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nrow = (int64_t)nby * W, ncol = (int64_t)nbx * 2 * H;
  const int hb = (l - 1) * LAB_LS;
  int x = 0, y = 0;
  bool rowitem = false, cand = false;
  if (t < nrow) {
    const int k = (int)(t / W);
    x = (int)(t - (int64_t)k * W);
    y = (lab_kth_boundary(k) << hb) * TH;
    rowitem = cand = true;
  } else if (t < nrow + ncol) {
    const int64_t c = t - nrow;
    const int k = (int)(c / (2 * (int64_t)H));
    const int rem = (int)(c - (int64_t)k * 2 * H);
    const int side = rem >= H ? 1 : 0;
    y = rem - side * H;
    x = (lab_kth_boundary(k) << hb) * TW - side;
    cand = true;
  }
  if (!__any(cand)) return;  // the whole wave: no shuffle below loses a participant
  // every load of the item at once (lanes without an item read pixel 0): the pixel, then the
  // value and parent of each neighbour whose link belongs to this level
  const int64_t i = (int64_t)y * W + x;
  const int ra = P[i];
  const T v = in[i];
  int64_t jn[LAB_U2];
  bool mine[LAB_U2];
#pragma unroll
  for (int k = 0; k < LAB_U2; k++) {
    mine[k] = false;
    jn[k] = i;
    if (k < u.n) {
      const int yy = y + u.dy[k], xx = x + u.dx[k];
      if ((unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W) {
        const int lx = lab_level((x / TW) ^ (xx / TW)), ly = lab_level((y / TH) ^ (yy / TH));
        mine[k] = rowitem ? (ly == l && lx <= l) : (lx == l && ly < l);
        if (mine[k]) jn[k] = (int64_t)yy * W + xx;
      }
    }
  }
  T w[LAB_U2];
  int pj[LAB_U2];
#pragma unroll
  for (int k = 0; k < LAB_U2; k++) {
    w[k] = in[jn[k]];
    pj[k] = P[jn[k]];
  }
  cand = cand && ra >= 0;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < LAB_U2; k++) {
    if (k >= u.n) break;  // the same for every lane
    const bool lk = cand && mine[k] && lab_link<T, FLAT>(v, w[k]);
    const int rb = lk ? pj[k] : -1;
    const int pra = __shfl_up(lk ? ra : -2, 1), prb = __shfl_up(rb, 1);
    const bool dup = lane != 0 && pra == ra && prb == rb;
    if (lk && !dup) lab_unite(P, ra, rb);
  }
}

// Every tile root listed by lab_tile2d points at its root (one wave per tile).  No union runs
// during this kernel, so its plain stores cannot overwrite a link, and every parent a find
// reads, old or already rewritten, is an ancestor.
__global__ __launch_bounds__(256) void lab_compress_roots(int *P, const int *__restrict__ RL, const int *__restrict__ TC,
                                                          int ntiles, int nt)
{
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (tile >= ntiles) return;
  const int n = TC[tile];
  const int *rl = RL + (int64_t)tile * nt;
  for (int k = lane; k < n; k += 64) {
    const int r = rl[k];
    P[r] = lab_find(P, r);
  }
}

__global__ void lab_compress(int *P, int64_t N)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < N && P[i] >= 0) P[i] = lab_find(P, (int)i);
}

// inclusive prefix sum over the lanes of a wave
__device__ __forceinline__ int lab_wave_scan(int v)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// exclusive prefix of v over the 256 threads of the block, and the block's total (ws: 4
// ints of LDS; every thread calls it)
__device__ __forceinline__ int lab_block_excl(int v, int *ws, int &total)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = lab_wave_scan(v);
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  int before = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    before += k < w ? ws[k] : 0;
    total += ws[k];
  }
  __syncthreads();  // ws is free again
  return before + inc - v;
}

// Roots (P[i] == i) of block b = pixels [b * LAB_RB, (b+1) * LAB_RB): their number into
// cnt[b] and each root's rank among them, in raster order, into Q[root].  Thread t owns the
// 16 consecutive pixels from b * LAB_RB + 16 t.
__global__ __launch_bounds__(256) void lab_count_rank(const int *__restrict__ P, int *__restrict__ Q, int *__restrict__ cnt,
                                                      int64_t N)
{
  __shared__ int ws[4];
  const int64_t base = (int64_t)blockIdx.x * LAB_RB + 16 * threadIdx.x;
  int p[16];
  if (base + 16 <= N) {
    const int4 *q = reinterpret_cast<const int4 *>(P + base);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int4 w = q[k];
      p[4 * k] = w.x;
      p[4 * k + 1] = w.y;
      p[4 * k + 2] = w.z;
      p[4 * k + 3] = w.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) p[k] = base + k < N ? P[base + k] : -1;
  }
  unsigned f = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) f |= (p[k] == (int)(base + k) ? 1u : 0u) << k;
  int total;
  int run = lab_block_excl(__popc(f), ws, total);
#pragma unroll
  for (int k = 0; k < 16; k++)
    if ((f >> k) & 1u) Q[base + k] = run++;
  if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

// cnt[b] := number of roots in blocks before b (one workgroup, 4096 counts per step with a
// running carry); *total := number of components
__global__ __launch_bounds__(256) void lab_scan_counts(int *cnt, int nb, int64_t *total)
{
  __shared__ int ws[4];
  int carry = 0;
  for (int c0 = 0; c0 < nb; c0 += 256 * 16) {
    const int b0 = c0 + 16 * threadIdx.x;
    int v[16], sum = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      v[k] = b0 + k < nb ? cnt[b0 + k] : 0;
      sum += v[k];
    }
    int step;
    int run = carry + lab_block_excl(sum, ws, step);
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if (b0 + k < nb) cnt[b0 + k] = run;
      run += v[k];
    }
    carry += step;
  }
  if (threadIdx.x == 0) *total = carry;
}

// out[i] = 1 + rank of the root of i in raster order (off: lab_scan_counts' block offsets,
// Q: lab_count_rank's ranks within a block), 0 for the background.  Each thread writes 4
// pixels 256 apart, so every store instruction covers 512 contiguous bytes.
__global__ __launch_bounds__(256) void lab_output(const int *__restrict__ P, const int *__restrict__ Q,
                                                  const int *__restrict__ off, int64_t *__restrict__ out, int64_t N)
{
  const int64_t b = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  int p[4];
#pragma unroll
  for (int k = 0; k < 4; k++) p[k] = b + 256 * k < N ? P[b + 256 * k] : -1;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = b + 256 * k;
    if (i >= N) break;
    int64_t o = 0;
    if (p[k] >= 0) {
      const int r = lab_find(P, p[k]);
      o = (int64_t)off[r / LAB_RB] + Q[r] + 1;
    }
    out[i] = o;
  }
}

// upper (raster-earlier) neighbours of a Common_NL_SE (neighbor_list_static_se_t.hpp
// split_coordinates_upper_lower)
static int lab_upper(int se, LabUpper &u)
{
  NLArgs a;
  PM_TRY(se_nl(se, a));
  u.n = 0;
  for (int k = 0; k < a.n; k++) {
    const int sign = a.dz[k] ? a.dz[k] : a.dy[k] ? a.dy[k] : a.dx[k];
    if (sign < 0) {
      u.dz[u.n] = a.dz[k];
      u.dy[u.n] = a.dy[k];
      u.dx[u.n] = a.dx[k];
      u.n++;
    }
  }
  return PM_OK;
}

static int lab_check(const pm_image *in, const pm_image *out, int se, const char *what)
{
  clear_error();
  if (!in || !out) PM_FAIL(PM_E_ARG, "%s: null image descriptor", what);
  if (in->rank == 0) PM_FAIL(PM_E_RANK, "%s Unsupported number of dims:0", what);
  if (in->rank < 1 || in->rank > 3) PM_FAIL(PM_E_RANK, "%s Unsupported number of dims", what);
  if (out->rank != in->rank) PM_FAIL(PM_E_SIZE, "%s incompatible size", what);
  for (int i = 0; i < in->rank; i++) {
    if (in->shape[i] < 0) PM_FAIL(PM_E_SIZE, "%s: negative extent", what);
    if (out->shape[i] != in->shape[i]) PM_FAIL(PM_E_SIZE, "%s incompatible size", what);
  }
  if (ptype_size(in->ptype) == 0) PM_FAIL(PM_E_PTYPE, "%s unsupported PTYPE", what);
  if (out->ptype != PM_PTYPE_I64) PM_FAIL(PM_E_PTYPE, "%s: the output image must be GINT64", what);
  const int64_t n = numel(*in);
  if (n > 0 && (!in->data || !out->data)) PM_FAIL(PM_E_ARG, "%s: null data pointer", what);
  if (n > 0 && in->data == out->data) PM_FAIL(PM_E_ALIAS, "%s output must be != than input images", what);
  if (n >= INT32_MAX) PM_FAIL(PM_E_SIZE, "%s: images of 2^31 pixels or more are not supported", what);
  PM_TRY(check_se_rank(se, in->rank));
  return PM_OK;
}

template <bool FLAT>
static int lab_run(const pm_image *in, int se, pm_image *out, int64_t *count, hipStream_t s, const char *what)
{
  PM_TRY(lab_check(in, out, se, what));
  const int64_t N = numel(*in);
  if (N == 0) {
    if (count) *count = 0;
    return PM_OK;
  }
  LabUpper u;
  PM_TRY(lab_upper(se, u));
  const Dims d = dims3(*in);
  const int64_t nb = (N + LAB_RB - 1) / LAB_RB;
  // 2D tiles of 32 x 64 pixels; each lists its roots in `cap` slots (at most its pixel count)
  constexpr int TH2 = 32, TW2 = 64;
  const bool flat2d = d.D == 1;
  const int64_t tx = (d.W + TW2 - 1) / TW2, ty = (d.H + TH2 - 1) / TH2, nt2 = tx * ty;
  const int cap = (int)((d.H < TH2 ? d.H : TH2) * (d.W < TW2 ? d.W : TW2));
  const int64_t rn = flat2d && nt2 * cap > N ? nt2 * cap : N;  // R: root lists, then ranks
  auto al = [](int64_t n) { return ((size_t)n * sizeof(int) + 255) & ~(size_t)255; };
  const size_t pb = al(N), rb = al(rn), tb = flat2d ? al(nt2) : 0, cb = al(nb);
  Scratch sc;
  PM_TRY(sc.alloc(pb + rb + tb + cb + 256, s));
  char *base = (char *)sc.ptr;
  int *P = (int *)base;
  int *R = (int *)(base + pb);
  int *TC = (int *)(base + pb + rb);
  int *cnt = (int *)(base + pb + rb + tb);
  int64_t *total = (int64_t *)(base + pb + rb + tb + cb);
  const unsigned g = (unsigned)((N + 255) / 256);
  auto f = [&](auto tag) -> int {
    using T = typename decltype(tag)::type;
    const T *src = (const T *)in->data;
    if (flat2d) {  // 1D / 2D
      const dim3 tg((unsigned)tx, (unsigned)ty, 1);
      int xlink = 0;  // is the left neighbour (dy = 0, dx = -1) an upper offset of the SE
      for (int j = 0; j < u.n; j++) xlink |= (u.dz[j] == 0 && u.dy[j] == 0 && u.dx[j] == -1);
      int upm = 0;  // dx + 1 of each upper offset into the row above
      for (int j = 0; j < u.n; j++)
        if (u.dz[j] == 0 && u.dy[j] == -1) upm |= 1 << (u.dx[j] + 1);
      // the masks of the static SEs get their own kernels (square: 7, cross / segment y: 2, none)
      auto tk = upm == 7 ? lab_tile2d<T, FLAT, TH2, 7>
                : upm == 2 ? lab_tile2d<T, FLAT, TH2, 2>
                : upm == 0 ? lab_tile2d<T, FLAT, TH2, 0>
                           : lab_tile2d<T, FLAT, TH2, -1>;
      hipLaunchKernelGGL(tk, tg, dim3(256), 0, s, src, P, (int)d.H, (int)d.W, upm, xlink, R, TC, cap);
      LabUpper u2;  // the offsets in the plane
      u2.n = 0;
      for (int j = 0; j < u.n; j++)
        if (u.dz[j] == 0) {
          if (u2.n == LAB_U2) PM_FAIL(PM_E_ARG, "%s: more than %d upper offsets in the plane", what, LAB_U2);
          u2.dz[u2.n] = 0;
          u2.dy[u2.n] = u.dy[j];
          u2.dx[u2.n] = u.dx[j];
          u2.n++;
        }
      const int levels = (lab_bitlen_host((int)((tx - 1) | (ty - 1))) + LAB_LS - 1) / LAB_LS;
      for (int l = 1; l <= levels; l++) {
        const int hb = (l - 1) * LAB_LS;  // boundaries b of level l: lowest set bit in [hb, hb + LAB_LS)
        const int64_t nby = ((ty - 1) >> hb) - ((ty - 1) >> (hb + LAB_LS));
        const int64_t nbx = ((tx - 1) >> hb) - ((tx - 1) >> (hb + LAB_LS));
        const int64_t items = nby * d.W + nbx * 2 * d.H;
        if (items == 0) continue;
        hipLaunchKernelGGL((lab_border_level<T, FLAT, TH2, TW2>), dim3((unsigned)((items + 255) / 256)), dim3(256), 0, s,
                           src, P, (int)d.H, (int)d.W, u2, l, (int)nby, (int)nbx);
        hipLaunchKernelGGL(lab_compress_roots, dim3((unsigned)((nt2 + 3) / 4)), dim3(256), 0, s, P, R, TC, (int)nt2,
                           cap);
      }
    } else {  // 3D: tiles of 8 x 8 x 32
      constexpr int TD = 8, TH = 8, TW = 32;
      const dim3 tg((unsigned)((d.W + TW - 1) / TW), (unsigned)((d.H + TH - 1) / TH), (unsigned)((d.D + TD - 1) / TD));
      hipLaunchKernelGGL((lab_tile<T, FLAT, TD, TH, TW>), tg, dim3(256), 0, s, src, P, (int)d.D, (int)d.H, (int)d.W, u);
      hipLaunchKernelGGL((lab_border<T, FLAT, TD, TH, TW>), dim3(g), dim3(256), 0, s, src, P, (int)d.D, (int)d.H,
                         (int)d.W, u);
      hipLaunchKernelGGL(lab_compress, dim3(g), dim3(256), 0, s, P, N);
    }
    PM_HIP(hipGetLastError());
    return PM_OK;
  };
  PM_TRY(dispatch_ptype(in->ptype, f));
  hipLaunchKernelGGL(lab_count_rank, dim3((unsigned)nb), dim3(256), 0, s, P, R, cnt, N);
  hipLaunchKernelGGL(lab_scan_counts, dim3(1), dim3(256), 0, s, cnt, (int)nb, total);
  hipLaunchKernelGGL(lab_output, dim3((unsigned)((N + 1023) / 1024)), dim3(256), 0, s, P, R, cnt,
                     (int64_t *)out->data, N);
  PM_HIP(hipGetLastError());
  if (count) {
    // through a pinned slot (allocated once per host thread): a copy into pageable memory is
    // staged by the runtime and costs tens of microseconds
    static thread_local int64_t *hcount = nullptr;
    if (!hcount) PM_HIP(hipHostMalloc((void **)&hcount, sizeof(int64_t), hipHostMallocDefault));
    PM_HIP(hipMemcpyAsync(hcount, total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    *count = *hcount;
  }
  return PM_OK;
}

}  // namespace pm

using namespace pm;

extern "C" {
int pm_label_binary(const pm_image *in, int se, pm_image *out, int64_t *count, void *stream)
{ PM_TRACE();
  return lab_run<false>(in, se, out, count, as_stream(stream), "label_binary");
}

int pm_label_flat_zones(const pm_image *in, int se, pm_image *out, int64_t *count, void *stream)
{ PM_TRACE();
  return lab_run<true>(in, se, out, count, as_stream(stream), "label_flat_zones");
}

int pm_h_convex(const pm_image *in, double h, int se, pm_image *out, void *stream);
int pm_h_concave(const pm_image *in, double h, int se, pm_image *out, void *stream);

// label_extrema.cpp: label_binary of h_concave(img, 1) (minima) / h_convex(img, 1) (maxima)
static int lab_extrema(bool minima, const pm_image *in, int se, pm_image *out, void *stream)
{
  const char *what = minima ? "label_minima" : "label_maxima";
  PM_TRY(lab_check(in, out, se, what));
  hipStream_t s = as_stream(stream);
  const int64_t n = numel(*in);
  if (n == 0) return PM_OK;
  Scratch tmp;
  PM_TRY(tmp.alloc((size_t)n * ptype_size(in->ptype), s));
  pm_image t = *in;
  t.data = tmp.ptr;
  PM_TRY(minima ? pm_h_concave(in, 1.0, se, &t, stream) : pm_h_convex(in, 1.0, se, &t, stream));
  return lab_run<false>(&t, se, out, nullptr, s, what);
}

int pm_label_minima(const pm_image *in, int se, pm_image *out, void *stream) { PM_TRACE(); return lab_extrema(true, in, se, out, stream); }
int pm_label_maxima(const pm_image *in, int se, pm_image *out, void *stream) { PM_TRACE(); return lab_extrema(false, in, se, out, stream); }
}
