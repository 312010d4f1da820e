// TOD <-> map projections.
//
//  map2tod : coalesced gather, dat[i] (+)= map[ipix[i]].  8 B/sample of HBM
//            traffic (int32 index in, value out); the map (<= 16 MB FP32 at
//            2048^2) stays in L2 / Infinity Cache.
//  tod2map : deterministic plan.  Samples are stably sorted by map tile and
//            cut into work items (<= item_samples entries of one tile), and
//            each item into groups of up to 64 entries in sample order made
//            of at most four runs of consecutive samples.  A wave reads a
//            group's values in sample order from its 20-byte header, moves
//            them into pixel order with the group's lane words (local pixel
//            and source lane) and reduces them into a wave-private LDS tile
//            with a wave-level segmented scan and one LDS add per pixel run
//            (no global atomics); the four wave tiles are summed in a fixed
//            order into a partial tile, and a second kernel sums each tile's
//            partials in item order.  The map is therefore bitwise
//            reproducible run to run.  Per sample the exec reads about 2.3
//            bytes of plan and the value, contiguously.
#include "common.h"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// map2tod

// 4 samples per thread per step (one int4 of ipix, one 4-wide store of dat),
// U steps per loop trip with all their index loads, then all their gathers,
// in flight together; the tail is handled scalar.
template <typename T>
struct alignas(4 * sizeof(T)) Quad { T v[4]; };

template <typename T>
__global__ void __launch_bounds__(256) map2tod_kernel(T* __restrict__ dat, const T* __restrict__ map, int64_t n,
                                                      const int32_t* __restrict__ ipix, int do_add) {
  constexpr int U = 8;
  const int64_t n4 = n >> 2;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int4* ip4 = reinterpret_cast<const int4*>(ipix);
  Quad<T>* d4 = reinterpret_cast<Quad<T>*>(dat);
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; q + (U - 1) * stride < n4; q += U * stride) {
    int4 p[U];
#pragma unroll
    for (int u = 0; u < U; u++) p[u] = ip4[q + u * stride];
    Quad<T> v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      v[u].v[0] = map[p[u].x];
      v[u].v[1] = map[p[u].y];
      v[u].v[2] = map[p[u].z];
      v[u].v[3] = map[p[u].w];
    }
    if (do_add) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const Quad<T> o = d4[q + u * stride];
#pragma unroll
        for (int e = 0; e < 4; e++) v[u].v[e] += o.v[e];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) d4[q + u * stride] = v[u];
  }
  for (; q < n4; q += stride) {
    const int4 p = ip4[q];
    Quad<T> v;
    v.v[0] = map[p.x]; v.v[1] = map[p.y]; v.v[2] = map[p.z]; v.v[3] = map[p.w];
    if (do_add) {
      const Quad<T> o = d4[q];
#pragma unroll
      for (int e = 0; e < 4; e++) v.v[e] += o.v[e];
    }
    d4[q] = v;
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    T v = map[ipix[i]];
    dat[i] = do_add ? dat[i] + v : v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) tod2map_atomic_kernel(T* map, const T* __restrict__ dat, int64_t n,
                                                             const int32_t* __restrict__ ipix) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    atomicAdd(map + ipix[i], dat[i]);
}

// One sample per thread, for ipix / dat that the 4-wide loads and stores of
// map2tod_kernel cannot take (e.g. one TOD's view into a batched index
// arena, which starts wherever the previous TODs end).
template <typename T>
__global__ void __launch_bounds__(256) map2tod_scalar_kernel(T* __restrict__ dat, const T* __restrict__ map,
                                                             int64_t n, const int32_t* __restrict__ ipix,
                                                             int do_add) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T v = map[ipix[i]];
    dat[i] = do_add ? dat[i] + v : v;
  }
}

template <typename T>
int map2tod_impl(T* dat, const T* map, int64_t nsamp, const int32_t* ipix, int do_add, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(dat && map && ipix, MK_ERR_ARG, "null pointer");
  MK_REQUIRE(((uintptr_t)ipix & 3) == 0 && ((uintptr_t)dat % sizeof(T)) == 0, MK_ERR_ARG,
             "map2tod: ipix and dat must be aligned to their element size");
  hipStream_t st = mk::as_stream(stream);
  if (((uintptr_t)ipix & 15) == 0 && ((uintptr_t)dat % (4 * sizeof(T))) == 0) {
    map2tod_kernel<T><<<mk::grid_for(mk::cdiv(nsamp, 32), 256, 8192), 256, 0, st>>>(dat, map, nsamp, ipix, do_add);
    return mk::check_launch("map2tod_kernel");
  }
  map2tod_scalar_kernel<T><<<mk::grid_for(nsamp, 256), 256, 0, st>>>(dat, map, nsamp, ipix, do_add);
  return mk::check_launch("map2tod_scalar_kernel");
}

// ---------------------------------------------------------------------------
// Tiled map layout for the gather.  A raster scan moves along one map axis,
// so 64 consecutive samples of one detector fall on ~20 rows of a row-major
// map (stride ny) and a wave's gathers touch ~20 cache lines.  With the map
// re-laid in tx x ty pixel tiles (128 bytes for 32 FP32 pixels) the same
// samples touch a few lines:
//   tiled(x, y) = ((x / tx) * nty + y / ty) * (tx * ty) + (x % tx) * ty + y % ty,
//   nty = ceil(ny / ty).

__global__ void __launch_bounds__(256) pix_tile_kernel(const int32_t* __restrict__ ipix, int64_t n, int32_t nx,
                                                       int32_t ny, int32_t tx, int32_t ty, int32_t* __restrict__ out,
                                                       int64_t* nbad) {
  const int64_t npix = (int64_t)nx * ny;
  const int32_t nty = (ny + ty - 1) / ty;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t p = ipix[i];
    const bool bad = p < 0 || (int64_t)p >= npix;
    mk::count_bad(bad, nbad);
    if (bad) {
      out[i] = 0;  // in bounds; the caller treats any bad sample as an error
      continue;
    }
    const int32_t x = p / ny, y = p - x * ny;
    out[i] = ((x / tx) * nty + y / ty) * (tx * ty) + (x % tx) * ty + (y % ty);
  }
}

// Counts indices outside [0, npix) (one wave ballot and at most one atomic
// per wave, mk::count_bad).  Reads 4 B/sample and writes nothing else.
__global__ void __launch_bounds__(256) pix_count_bad_kernel(const int32_t* __restrict__ ipix, int64_t n,
                                                            int64_t npix, int64_t* nbad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t p = ipix[i];
    mk::count_bad(p < 0 || (int64_t)p >= npix, nbad);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) map_tiles_kernel(const T* __restrict__ map, int32_t nx, int32_t ny, int32_t tx,
                                                        int32_t ty, T* __restrict__ tiled, int64_t ntiled) {
  const int32_t nty = (ny + ty - 1) / ty, tp = tx * ty;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntiled; t += stride) {
    const int32_t tile = (int32_t)(t / tp), w = (int32_t)(t - (int64_t)tile * tp);
    const int32_t bx = tile / nty, by = tile - bx * nty;
    const int32_t x = bx * tx + w / ty, y = by * ty + w % ty;
    tiled[t] = (x < nx && y < ny) ? map[(int64_t)x * ny + y] : T(0);
  }
}

int64_t tiled_size(int32_t nx, int32_t ny, int32_t tx, int32_t ty) {
  return (int64_t)((nx + tx - 1) / tx) * tx * (int64_t)((ny + ty - 1) / ty) * ty;
}

template <typename T>
int map_tiles_impl(const T* map, int32_t nx, int32_t ny, int32_t tx, int32_t ty, T* tiled, void* stream) {
  MK_REQUIRE(nx > 0 && ny > 0 && tx > 0 && ty > 0 && tx * ty <= 1024, MK_ERR_ARG, "bad map or tile size");
  MK_REQUIRE(map && tiled, MK_ERR_ARG, "null pointer");
  const int64_t nt = tiled_size(nx, ny, tx, ty);
  MK_REQUIRE(nt <= INT32_MAX, MK_ERR_ARG, "tiled map too large for int32 indices");
  map_tiles_kernel<T><<<mk::grid_for(nt, 256, 256 * 16), 256, 0, mk::as_stream(stream)>>>(map, nx, ny, tx, ty,
                                                                                           tiled, nt);
  return mk::check_launch("map_tiles_kernel");
}

template <typename T>
int tod2map_atomic_impl(T* map, const T* dat, int64_t nsamp, const int32_t* ipix, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(dat && map && ipix, MK_ERR_ARG, "null pointer");
  tod2map_atomic_kernel<T><<<mk::grid_for(nsamp, 256, 256 * 16), 256, 0, mk::as_stream(stream)>>>(map, dat, nsamp,
                                                                                                   ipix);
  return mk::check_launch("tod2map_atomic_kernel");
}

}  // namespace

extern "C" int mk_pix_count_bad(const int32_t* ipix, int64_t nsamp, int64_t npix, int64_t* dev_nbad,
                                void* stream) {
  MK_REQUIRE(nsamp >= 0 && npix >= 0, MK_ERR_ARG, "nsamp or npix < 0");
  MK_REQUIRE(dev_nbad, MK_ERR_ARG, "null counter");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(ipix, MK_ERR_ARG, "null pointer");
  pix_count_bad_kernel<<<mk::grid_for(nsamp, 256, 256 * 16), 256, 0, mk::as_stream(stream)>>>(ipix, nsamp, npix,
                                                                                              dev_nbad);
  return mk::check_launch("pix_count_bad_kernel");
}

extern "C" int mk_map2tod_f32(float* dat, const float* map, int64_t nsamp, const int32_t* ipix, int do_add,
                              void* stream) {
  return map2tod_impl<float>(dat, map, nsamp, ipix, do_add, stream);
}
extern "C" int mk_map2tod_f64(double* dat, const double* map, int64_t nsamp, const int32_t* ipix, int do_add,
                              void* stream) {
  return map2tod_impl<double>(dat, map, nsamp, ipix, do_add, stream);
}
extern "C" int mk_tod2map_atomic_f32(float* map, const float* dat, int64_t nsamp, const int32_t* ipix,
                                     void* stream) {
  return tod2map_atomic_impl<float>(map, dat, nsamp, ipix, stream);
}
extern "C" int mk_tod2map_atomic_f64(double* map, const double* dat, int64_t nsamp, const int32_t* ipix,
                                     void* stream) {
  return tod2map_atomic_impl<double>(map, dat, nsamp, ipix, stream);
}

// ---------------------------------------------------------------------------
// Deterministic tod2map plan

struct T2MItem {
  int64_t seg_base;   // first sample of the segment the entries are relative to
  int64_t begin;      // first entry of the item in the build's sorted entry array
  int32_t count;      // number of entries
  int32_t tile;       // tile id = tx * ntile_y + ty
  int64_t g_begin;    // first group of the item
  int32_t g_count;    // number of groups
  int32_t pad_;
};

// Up to 64 consecutive entries of one item, in sample order, made of at
// most four runs of consecutive samples.  Lane l of a wave takes entry l:
// its run r is the number of run starts (start1..3) at or below l, and its
// sample offset within the segment is base[r] + l - start_r.
struct T2MGroup {
  uint32_t base[4];  // sample offset of each run's first entry (unused runs: 0)
  uint32_t starts;   // bits 0-7, 8-15, 16-23: first lane of runs 1, 2, 3 (64 = no such run); 24-30: entries
};

// A group's lane words are 16 bits: the local pixel in bits 0-9 and the
// source lane in bits 10-15, so a tile holds at most 1024 pixels.
constexpr int kT2MMaxTilePx = 1024;

struct mk_t2m_plan {
  int64_t nsamp = 0;
  int32_t nx = 0, ny = 0, tile_x = 0, tile_y = 0, ntile_x = 0, ntile_y = 0;
  int32_t item_samples = 0;
  int64_t nitems = 0, ntiles_touched = 0, max_item = 0, ngroups = 0;
  T2MGroup* groups = nullptr;    // [ngroups]
  uint16_t* lanes = nullptr;     // [ngroups][64]: lane d in pixel order: local pixel | source lane << 10
  T2MItem* items = nullptr;      // [nitems] sorted by (tile, segment, position)
  int32_t* tile_items = nullptr; // [ntiles+1] CSR into items
  void* partial = nullptr;       // [nitems][tile_px] partial tiles
  int64_t partial_elem_bytes = 0;
  int64_t plan_bytes = 0, partial_bytes = 0;
};

namespace {

constexpr int64_t kSegment = int64_t(1) << 30;

__global__ void __launch_bounds__(256) t2m_keys_kernel(const int32_t* __restrict__ ipix, int64_t n, int32_t nx,
                                                       int32_t ny, int32_t tx, int32_t ty, int32_t ntile_y,
                                                       uint32_t* __restrict__ key, uint32_t* __restrict__ idx,
                                                       int64_t* nbad) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    int32_t p = ipix[i];
    uint32_t k = 0;
    if (p < 0 || (int64_t)p >= (int64_t)nx * ny) {
      atomicAdd((unsigned long long*)nbad, 1ull);
    } else {
      int32_t x = p / ny, y = p - x * ny;
      k = (uint32_t)((x / tx) * ntile_y + (y / ty));
    }
    key[i] = k;
    idx[i] = (uint32_t)i;
  }
}

// ---- Stable LSD radix sort of (uint32 key, uint32 value) pairs, 8 bits of
// the key per pass, for the plan's sort of samples by tile.  A pass: digit
// counts per tile of 4096 entries (sort_hist_kernel), an exclusive scan of
// the counts in (digit, tile) order (scan_*), and a scatter that ranks the
// entries of each tile in their input order (sort_scatter_kernel): a wave
// finds the lanes sharing its digit with eight ballots, the four waves of a
// round are ordered through LDS, and a per-digit running count carries the
// order across the tile's 16 rounds.  Every count is an integer, so the
// result does not depend on scheduling.
constexpr int kSortThreads = 256, kSortItems = 16, kSortTile = kSortThreads * kSortItems;
constexpr int kRadix = 256;

__global__ void __launch_bounds__(kSortThreads) sort_hist_kernel(const uint32_t* __restrict__ key, int64_t n,
                                                                 int shift, uint32_t* __restrict__ hist,
                                                                 int64_t ntile) {
  __shared__ uint32_t h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t t = blockIdx.x, base = t * kSortTile;
#pragma unroll 4
  for (int i = 0; i < kSortItems; i++) {
    const int64_t e = base + i * kSortThreads + threadIdx.x;
    if (e < n) atomicAdd(&h[(key[e] >> shift) & (kRadix - 1)], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntile + t] = h[threadIdx.x];
}

// Exclusive scan, step 1: each block scans 4096 entries in place and writes
// its total.
__global__ void __launch_bounds__(256) scan_local_kernel(uint32_t* __restrict__ a, int64_t len,
                                                         uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[256];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * 4096 + tid * 16;
  uint32_t v[16], tot = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    v[i] = base + i < len ? a[base + i] : 0u;
    tot += v[i];
  }
  sh[tid] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off *= 2) {
    const uint32_t x = tid >= off ? sh[tid - off] : 0u;
    __syncthreads();
    sh[tid] += x;
    __syncthreads();
  }
  uint32_t run = sh[tid] - tot;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (base + i < len) a[base + i] = run;
    run += v[i];
  }
  if (tid == 255) sums[blockIdx.x] = sh[255];
}

// Step 2 (one block): exclusive scan of the block totals, 4096 at a time.
__global__ void __launch_bounds__(256) scan_sums_kernel(uint32_t* __restrict__ sums, int64_t nsum) {
  __shared__ uint32_t sh[256];
  const int tid = threadIdx.x;
  uint32_t carry = 0;
  for (int64_t c0 = 0; c0 < nsum; c0 += 4096) {
    const int64_t base = c0 + tid * 16;
    uint32_t v[16], tot = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      v[i] = base + i < nsum ? sums[base + i] : 0u;
      tot += v[i];
    }
    sh[tid] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off *= 2) {
      const uint32_t x = tid >= off ? sh[tid - off] : 0u;
      __syncthreads();
      sh[tid] += x;
      __syncthreads();
    }
    uint32_t run = carry + sh[tid] - tot;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (base + i < nsum) sums[base + i] = run;
      run += v[i];
    }
    carry += sh[255];
    __syncthreads();
  }
}

// Step 3: adds each block's offset.
__global__ void __launch_bounds__(256) scan_add_kernel(uint32_t* __restrict__ a, int64_t len,
                                                       const uint32_t* __restrict__ sums) {
  const uint32_t o = sums[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * 4096;
  for (int i = threadIdx.x; i < 4096; i += 256)
    if (base + i < len) a[base + i] += o;
}

__global__ void __launch_bounds__(kSortThreads) sort_scatter_kernel(const uint32_t* __restrict__ kin,
                                                                    const uint32_t* __restrict__ vin, int64_t n,
                                                                    int shift, const uint32_t* __restrict__ off,
                                                                    int64_t ntile, uint32_t* __restrict__ kout,
                                                                    uint32_t* __restrict__ vout) {
  __shared__ uint32_t run[kRadix];
  __shared__ uint32_t wcnt[kSortThreads / 64][kRadix];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t t = blockIdx.x, base = t * kSortTile;
  run[tid] = off[(int64_t)tid * ntile + t];
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int i = 0; i < kSortItems; i++) {
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; w++) wcnt[w][tid] = 0;  // this thread's digit column
    __syncthreads();
    const int64_t e = base + i * kSortThreads + tid;
    const bool valid = e < n;
    const uint32_t k = valid ? kin[e] : 0u, v = valid ? vin[e] : 0u;
    const uint32_t d = (k >> shift) & (kRadix - 1);
    uint64_t m = __ballot(valid);  // valid lanes of this wave with digit d
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = __ballot(valid && ((d >> b) & 1u));
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(m & below);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(m);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int w = 0; w < wave; w++) pos += wcnt[w][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; w++) add += wcnt[w][tid];
    run[tid] += add;
  }
}

// Workspace (bytes) of sort_pairs for n entries.
int64_t sort_pairs_work_bytes(int64_t n) {
  const int64_t ntile = std::max<int64_t>(1, mk::cdiv(n, kSortTile));
  const int64_t len = kRadix * ntile, nsum = mk::cdiv(len, 4096);
  return (len + nsum) * 4 + 256;
}

// Stable sort of n pairs by key bits [0, end_bit).  (kA, vA) hold the input;
// passes alternate between A and B, and *in_b says where the result is.
int sort_pairs(uint32_t* kA, uint32_t* vA, uint32_t* kB, uint32_t* vB, int64_t n, int end_bit, void* work,
               hipStream_t st, bool* in_b) {
  *in_b = false;
  if (n <= 0) return MK_OK;
  const int64_t ntile = mk::cdiv(n, kSortTile);
  const int64_t len = kRadix * ntile, nsum = mk::cdiv(len, 4096);
  MK_REQUIRE(n < (int64_t(1) << 31) && ntile < INT32_MAX, MK_ERR_ARG, "sort: %lld entries are too many",
             (long long)n);
  uint32_t* hist = (uint32_t*)work;
  uint32_t* sums = hist + len;
  for (int shift = 0; shift < std::max(end_bit, 1); shift += 8) {
    const uint32_t* ki = *in_b ? kB : kA;
    const uint32_t* vi = *in_b ? vB : vA;
    uint32_t* ko = *in_b ? kA : kB;
    uint32_t* vo = *in_b ? vA : vB;
    sort_hist_kernel<<<(unsigned)ntile, kSortThreads, 0, st>>>(ki, n, shift, hist, ntile);
    scan_local_kernel<<<(unsigned)nsum, 256, 0, st>>>(hist, len, sums);
    scan_sums_kernel<<<1, 256, 0, st>>>(sums, nsum);
    scan_add_kernel<<<(unsigned)nsum, 256, 0, st>>>(hist, len, sums);
    sort_scatter_kernel<<<(unsigned)ntile, kSortThreads, 0, st>>>(ki, vi, n, shift, hist, ntile, ko, vo);
    int r = mk::check_launch("sort_scatter_kernel");
    if (r) return r;
    *in_b = !*in_b;
  }
  return MK_OK;
}

// Run lengths of the sorted tile keys without atomics: the last sample of
// each run writes end = i+1 and the first writes start = i (count = end -
// start; tiles without samples keep 0/0).
__global__ void __launch_bounds__(256) t2m_count_kernel(const uint32_t* __restrict__ key, int64_t n,
                                                        int32_t* __restrict__ start, int32_t* __restrict__ end) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t k = key[i];
    if (i == 0 || key[i - 1] != k) start[k] = (int32_t)i;
    if (i == n - 1 || key[i + 1] != k) end[k] = (int32_t)(i + 1);
  }
}

// Group formation, one thread per item walking its entries in sample
// order: a group closes at 64 entries, or when a fifth run would start.
// WRITE = false counts the item's groups; WRITE = true writes their headers
// and, for the permutation kernel, each group's first entry and segment.
template <bool WRITE>
__global__ void __launch_bounds__(256) t2m_groups_kernel(T2MItem* __restrict__ items, int64_t nitems,
                                                         const uint32_t* __restrict__ entry,
                                                         T2MGroup* __restrict__ groups,
                                                         int64_t* __restrict__ gpos, int64_t* __restrict__ gseg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const T2MItem it = items[i];
  int64_t g = WRITE ? it.g_begin : 0;
  int cnt = 0, nruns = 0;
  uint32_t base[4] = {0, 0, 0, 0};
  int start[4] = {0, 64, 64, 64};
  uint32_t prev = 0;
  int64_t first = it.begin;
  auto close = [&]() {
    if (WRITE) {
      T2MGroup G;
      for (int r = 0; r < 4; r++) G.base[r] = r < nruns ? base[r] : 0u;
      G.starts = (uint32_t)start[1] | ((uint32_t)start[2] << 8) | ((uint32_t)start[3] << 16) | ((uint32_t)cnt << 24);
      groups[g] = G;
      gpos[g] = first;
      gseg[g] = it.seg_base;
    }
    g++;
  };
  for (int64_t k = 0; k < it.count; k++) {
    const uint32_t e = entry[it.begin + k];
    bool newrun = cnt == 0 || e != prev + 1u;
    if (cnt == 64 || (newrun && nruns == 4)) {
      close();
      cnt = 0; nruns = 0; newrun = true;
      start[1] = start[2] = start[3] = 64;
      first = it.begin + k;
    }
    if (newrun) {
      base[nruns] = e;
      start[nruns] = cnt;
      nruns++;
    }
    cnt++;
    prev = e;
  }
  if (cnt > 0) close();
  if (!WRITE) items[i].g_count = (int32_t)g;
}

// One wave per group: the local pixel of each entry and the group's order
// by local pixel (ties by lane, so the order is unique) from a bitonic sort
// in registers; lane d of the result gets the local pixel and source lane of
// the d-th entry in pixel order.
__device__ __forceinline__ void group_lanes(const T2MGroup* __restrict__ groups, int64_t g,
                                            const int64_t* __restrict__ gpos, const int64_t* __restrict__ gseg,
                                            const uint32_t* __restrict__ entry, const int32_t* __restrict__ ipix,
                                            int32_t ny, int32_t tx, int32_t ty, uint16_t* __restrict__ lanes,
                                            int lane) {
  const int cnt = (int)(groups[g].starts >> 24);
  uint32_t lp = 0xFFFFu;
  if (lane < cnt) {
    const int32_t p = ipix[gseg[g] + entry[gpos[g] + lane]];
    const int32_t x = p / ny, y = p - x * ny;
    lp = (uint32_t)((x % tx) * ty + (y % ty));
  }
  uint32_t key = (lp << 6) | (uint32_t)lane;
  for (int k = 2; k <= 64; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint32_t other = __shfl_xor(key, j, 64);
      const bool up = (lane & k) == 0, lower = (lane & j) == 0;
      const uint32_t mn = min(key, other), mx = max(key, other);
      key = (lower == up) ? mn : mx;
    }
  }
  // lanes at or past the group's count sort last and stay unread
  lanes[g * 64 + lane] = (uint16_t)(((key >> 6) & (kT2MMaxTilePx - 1)) | ((key & 63u) << 10));
}

// Waves stride over the groups: one dispatch holds fewer than 2^32
// work-items, and a plan of 8e9 samples has 1.3e8 groups of 64 lanes.
__global__ void __launch_bounds__(256) t2m_group_lanes_kernel(const T2MGroup* __restrict__ groups, int64_t ngroups,
                                                              const int64_t* __restrict__ gpos,
                                                              const int64_t* __restrict__ gseg,
                                                              const uint32_t* __restrict__ entry,
                                                              const int32_t* __restrict__ ipix, int32_t ny,
                                                              int32_t tx, int32_t ty, uint16_t* __restrict__ lanes) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; g < ngroups; g += nwaves)
    group_lanes(groups, g, gpos, gseg, entry, ipix, ny, tx, ty, lanes, lane);
}

// Lane-to-lane moves inside a wave through DPP (no LDS round trip).  CTRL is
// the DPP control word (row_shr:n = 0x110 + n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143); rows outside ROW_MASK and lanes whose source lies
// outside their row read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Inclusive segmented sum over the wave; d = lane - (first lane of the
// lane's run), runs contiguous.  Kogge-Stone inside each 16-lane row, then
// the row-15 and row-31 broadcasts carry a run across row boundaries: after
// the row steps lane l holds the sum over lanes max(start, row start)..l,
// and a run that began in an earlier row adds that row's last lane, which
// holds the earlier part of the same run.
template <typename T>
__device__ __forceinline__ T wave_segmented_sum(T v, int lane, int d) {
  const int dr = min(d, lane & 15);
  T u;
  u = dpp_move<0x111, 0xf>(v); v = dr >= 1 ? v + u : v;
  u = dpp_move<0x112, 0xf>(v); v = dr >= 2 ? v + u : v;
  u = dpp_move<0x114, 0xf>(v); v = dr >= 4 ? v + u : v;
  u = dpp_move<0x118, 0xf>(v); v = dr >= 8 ? v + u : v;
  u = dpp_move<0x142, 0xa>(v); v = ((lane & 16) && d > (lane & 15)) ? v + u : v;  // rows 1, 3
  u = dpp_move<0x143, 0xc>(v); v = (lane >= 32 && d > lane - 32) ? v + u : v;    // rows 2, 3
  return v;
}

// Exec stage 1: one workgroup (4 waves) per item.  Wave w takes the item's
// groups w, w+4, w+8, ... in that order, kT2MUnroll of them per iteration so
// that their loads overlap, and loads the next iteration's group headers
// before it reduces the current groups.  Within a group the values are moved
// into pixel order, so every pixel is one run of lanes; the run's last lane
// adds the run sum into the wave's private LDS tile (one add per pixel per
// group, in group order: the result does not depend on timing).
constexpr int kT2MUnroll = 4;

// Weight of polarisation component comp of mode POL (MK_POL_*) for a sample
// at angle 2 gamma with sin s and cos c (minkasi.c:178-258: QU -> (c, s),
// IQU -> (1, c, s), QU_PRECON -> (c^2, s^2, s c), IQU_PRECON -> (1, c, s,
// c^2, s c, s^2)).
template <int POL, typename T>
__device__ __forceinline__ T pol_weight(int comp, T s, T c) {
  if (POL == MK_POL_QU) return comp == 0 ? c : s;
  if (POL == MK_POL_IQU) return comp == 0 ? T(1) : (comp == 1 ? c : s);
  if (POL == MK_POL_QU_PRECON) return comp == 0 ? c * c : (comp == 1 ? s * s : s * c);
  switch (comp) {
    case 0: return T(1);
    case 1: return c;
    case 2: return s;
    case 3: return c * c;
    case 4: return s * c;
    default: return s * s;
  }
}
__device__ __forceinline__ void pol_sincos(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ __forceinline__ void pol_sincos(double a, double* s, double* c) { sincos(a, s, c); }

// Moves v from lane src to this lane (ds_bpermute, the LDS crossbar without
// a memory access).
__device__ __forceinline__ float lane_move(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_move(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// A group header as the kernel uses it: the starts word and, per run r,
// base[r] - start_r, so that lane l of run r reads sample o_r + l.
struct T2MHdr {
  uint32_t st, o0, o1, o2, o3;
};
__device__ __forceinline__ T2MHdr t2m_load_hdr(const T2MGroup* G) {
  const uint32_t st = G->starts;
  return {st, G->base[0], G->base[1] - (st & 255u), G->base[2] - ((st >> 8) & 255u), G->base[3] - ((st >> 16) & 255u)};
}
// Sample offset of lane l without branches: the run starts are ordered
// (unused runs start at 64), so adding each run's step from the previous one
// for every start at or below l telescopes to o_r.
__device__ __forceinline__ uint32_t t2m_lane_sample(const T2MHdr& h, int l) {
  const uint32_t m1 = 0u - (uint32_t)(l >= (int)(h.st & 255u));
  const uint32_t m2 = 0u - (uint32_t)(l >= (int)((h.st >> 8) & 255u));
  const uint32_t m3 = 0u - (uint32_t)(l >= (int)((h.st >> 16) & 255u));
  return h.o0 + (m1 & (h.o1 - h.o0)) + (m2 & (h.o2 - h.o1)) + (m3 & (h.o3 - h.o2)) + (uint32_t)l;
}

// TD: sample type; T: accumulation type (FP64 for FP32 samples keeps the
// per-pixel sums of many cancelling samples at FP64 accuracy); P: partials.
// POL: 0 bins the samples; MK_POL_* bins sample * pol_weight(comp, 2 gamma)
// with 2 gamma read from twogamma at the same sample.
//
// A wave reads a group's samples in sample order, lane l the group's l-th
// entry (runs of consecutive samples, so the loads are contiguous), from the
// group header alone; the lane word then moves each value to its place in
// the group's pixel order, where every pixel is one run of lanes.
template <typename TD, typename T, typename P, int POL = 0>
__global__ void __launch_bounds__(256) t2m_items_kernel(const T2MItem* __restrict__ items,
                                                        const T2MGroup* __restrict__ groups,
                                                        const uint16_t* __restrict__ lanes,
                                                        const TD* __restrict__ dat, int tile_px,
                                                        P* __restrict__ partial,
                                                        const TD* __restrict__ twogamma = nullptr, int comp = 0) {
  extern __shared__ unsigned char smem_raw[];
  T* tiles = reinterpret_cast<T*>(smem_raw);  // [4][tile_px]
  const T2MItem it = items[blockIdx.x];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int p = threadIdx.x; p < 4 * tile_px; p += 256) tiles[p] = T(0);
  __syncthreads();
  T* mine = tiles + wave * tile_px;
  const TD* base = dat + it.seg_base;
  const int ngroups = it.g_count;
  const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);  // lanes 0..lane
  const T2MGroup* G0 = groups + it.g_begin;
  // Every load is unconditional: groups past the item repeat its last group
  // (discarded below) and lanes past a group's entries reread its last entry.
  T2MHdr h[kT2MUnroll];
#pragma unroll
  for (int j = 0; j < kT2MUnroll; j++) h[j] = t2m_load_hdr(G0 + min(wave + 4 * j, ngroups - 1));
  for (int g0 = wave; g0 < ngroups; g0 += 4 * kT2MUnroll) {
    uint32_t lw[kT2MUnroll], cnt[kT2MUnroll];
    T v[kT2MUnroll];
#pragma unroll
    for (int j = 0; j < kT2MUnroll; j++)
      lw[j] = lanes[(it.g_begin + min(g0 + 4 * j, ngroups - 1)) * 64 + lane];
#pragma unroll
    for (int j = 0; j < kT2MUnroll; j++) {
      cnt[j] = h[j].st >> 24;
      const uint32_t e = t2m_lane_sample(h[j], min(lane, (int)cnt[j] - 1));
      T x = (T)base[e];
      if constexpr (POL != 0) {
        TD sn, cs;
        pol_sincos(twogamma[it.seg_base + e], &sn, &cs);
        x = (T)(base[e] * pol_weight<POL, TD>(comp, sn, cs));
      }
      v[j] = lane < (int)cnt[j] ? x : T(0);
    }
    // the next iteration's headers, in flight while these groups reduce
#pragma unroll
    for (int j = 0; j < kT2MUnroll; j++)
      h[j] = t2m_load_hdr(G0 + min(g0 + 4 * kT2MUnroll + 4 * j, ngroups - 1));
#pragma unroll
    for (int j = 0; j < kT2MUnroll; j++) {
      if (g0 + 4 * j >= ngroups) break;  // wave-uniform
      const int lpv = (int)(lw[j] & (kT2MMaxTilePx - 1));
      const bool valid = lane < (int)cnt[j];  // pixel-order lanes past the count are the padding
      const T x = lane_move(v[j], (int)(lw[j] >> 10));
      const int lp = valid ? lpv : -1 - lane;  // distinct sentinels never match
      const int prev = __shfl_up(lp, 1, 64);
      const uint64_t heads = __ballot(lane == 0 || prev != lp);
      const int start = 63 - __clzll(heads & upto);
      const T sum = wave_segmented_sum(valid ? x : T(0), lane, lane - start);
      const bool tail = valid && (lane == 63 || ((heads >> lane) >> 1) & 1ull);
      if (tail) __hip_atomic_fetch_add(mine + lpv, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __syncthreads();
  P* out = partial + (int64_t)blockIdx.x * tile_px;
  for (int p = threadIdx.x; p < tile_px; p += 256) {
    T s = tiles[p];
    s += tiles[tile_px + p];
    s += tiles[2 * tile_px + p];
    s += tiles[3 * tile_px + p];
    out[p] = (P)s;
  }
}

// Exec stage 2: sums each map tile's partial tiles in item order.  A
// thread owns 4 consecutive pixels of one tile (tile_px % 4 == 0; a 4-pixel
// run never crosses a tile row since the plan requires tile_y % 4 == 0)
// and reads 4 items' partials per step before
// adding them in order.
template <typename P>
struct Vec4;
template <>
struct alignas(16) Vec4<float> { float v[4]; };
template <>
struct alignas(16) Vec4<double> { double v[4]; };

// T: accumulation type, TM: map type (an FP32 map summed in FP64 is rounded
// once, when stored).
template <typename T, typename P, typename TM>
__global__ void __launch_bounds__(256) t2m_tiles_kernel(const int32_t* __restrict__ tile_items,
                                                        const P* __restrict__ partial, int tile_x, int tile_y,
                                                        int ntile_y, int nx, int ny, TM* __restrict__ map,
                                                        int accumulate, int blocks_per_tile, int ncomp = 1,
                                                        int comp = 0) {
  const int tile = blockIdx.x / blocks_per_tile;
  const int tile_px = tile_x * tile_y;
  const int p0 = ((blockIdx.x % blocks_per_tile) * 256 + threadIdx.x) * 4;
  if (p0 >= tile_px) return;
  const int tx0 = (tile / ntile_y) * tile_x, ty0 = (tile % ntile_y) * tile_y;
  const int i0 = tile_items[tile], i1 = tile_items[tile + 1];
  const Vec4<P>* src = reinterpret_cast<const Vec4<P>*>(partial) + p0 / 4;
  const int64_t step = tile_px / 4;  // Vec4 per partial tile
  T acc[4] = {T(0), T(0), T(0), T(0)};
  int i = i0;
  for (; i + 4 <= i1; i += 4) {
    Vec4<P> a = src[(int64_t)i * step], b = src[(int64_t)(i + 1) * step];
    Vec4<P> c = src[(int64_t)(i + 2) * step], d = src[(int64_t)(i + 3) * step];
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] = (((acc[k] + (T)a.v[k]) + (T)b.v[k]) + (T)c.v[k]) + (T)d.v[k];
  }
  for (; i < i1; i++) {
    Vec4<P> a = src[(int64_t)i * step];
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += (T)a.v[k];
  }
  const int x = tx0 + p0 / tile_y, y0 = ty0 + p0 % tile_y;
  if (x >= nx) return;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int y = y0 + k;
    if (y >= ny) break;
    TM* m = map + ((int64_t)x * ny + y) * ncomp + comp;
    *m = (TM)(accumulate ? (T)*m + acc[k] : acc[k]);
  }
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
      p = nullptr;
      mk::set_error("hipMalloc(%zu) failed", bytes);
      return MK_ERR_ALLOC;
    }
    return MK_OK;
  }
};

void free_plan(mk_t2m_plan* pl) {
  if (!pl) return;
  if (pl->groups) (void)hipFree(pl->groups);
  if (pl->lanes) (void)hipFree(pl->lanes);
  if (pl->items) (void)hipFree(pl->items);
  if (pl->tile_items) (void)hipFree(pl->tile_items);
  if (pl->partial) (void)hipFree(pl->partial);
  delete pl;
}

}  // namespace

extern "C" int mk_t2m_plan_create(const int32_t* ipix, int64_t nsamp, int32_t nx, int32_t ny, int32_t tile_x,
                                  int32_t tile_y, int32_t item_samples, mk_t2m_plan** out, void* stream_v) {
  MK_REQUIRE(out != nullptr, MK_ERR_ARG, "out is NULL");
  *out = nullptr;
  MK_REQUIRE(nsamp >= 0 && (nsamp == 0 || ipix), MK_ERR_ARG, "bad ipix/nsamp");
  MK_REQUIRE(nx > 0 && ny > 0 && (int64_t)nx * ny <= INT32_MAX, MK_ERR_ARG, "bad map size");
  MK_REQUIRE(tile_x > 0 && tile_y > 0 && tile_x * tile_y <= kT2MMaxTilePx && tile_y % 4 == 0, MK_ERR_ARG,
             "tile must have 4..%d pixels with tile_y a multiple of 4 (got %d x %d)", kT2MMaxTilePx, tile_x, tile_y);
  MK_REQUIRE(item_samples >= 64 && item_samples % 64 == 0, MK_ERR_ARG, "item_samples must be a positive multiple of 64");
  hipStream_t stream = mk::as_stream(stream_v);
  mk_t2m_plan* pl = new mk_t2m_plan();
  pl->nsamp = nsamp; pl->nx = nx; pl->ny = ny; pl->tile_x = tile_x; pl->tile_y = tile_y;
  pl->ntile_x = (int32_t)mk::cdiv(nx, tile_x);
  pl->ntile_y = (int32_t)mk::cdiv(ny, tile_y);
  pl->item_samples = item_samples;
  const int64_t ntiles = (int64_t)pl->ntile_x * pl->ntile_y;
  const int tile_px = tile_x * tile_y;
  auto fail = [&](int code) { free_plan(pl); return code; };

  DevBuf entry_b;  // the samples sorted by tile, offsets within their segment (build only)
  if (entry_b.alloc(std::max<int64_t>(nsamp, 1) * 4)) return fail(MK_ERR_ALLOC);
  uint32_t* const entry = (uint32_t*)entry_b.p;
  std::vector<T2MItem> h_items;
  const int64_t segmax = std::min<int64_t>(kSegment, std::max<int64_t>(nsamp, 1));
  DevBuf keys, keys2, idx, cnt, nbad, tmp;
  if (keys.alloc(segmax * 4) || keys2.alloc(segmax * 4) || idx.alloc(segmax * 4) || cnt.alloc(ntiles * 8 + 8) ||
      nbad.alloc(8))
    return fail(MK_ERR_ALLOC);
  if (tmp.alloc(sort_pairs_work_bytes(segmax))) return fail(MK_ERR_ALLOC);
  int end_bit = 1;
  while ((int64_t(1) << end_bit) < ntiles) end_bit++;
  std::vector<int32_t> h_cnt(2 * ntiles);
  for (int64_t s0 = 0; s0 < nsamp; s0 += kSegment) {
    const int64_t n = std::min<int64_t>(kSegment, nsamp - s0);
    if (hipMemsetAsync(nbad.p, 0, 8, stream) != hipSuccess || hipMemsetAsync(cnt.p, 0, ntiles * 8, stream) != hipSuccess)
      return fail((mk::set_error("memset failed"), MK_ERR_HIP));
    t2m_keys_kernel<<<mk::grid_for(n, 256), 256, 0, stream>>>(ipix + s0, n, nx, ny, tile_x, tile_y, pl->ntile_y,
                                                               (uint32_t*)keys.p, (uint32_t*)idx.p, (int64_t*)nbad.p);
    if (mk::check_launch("t2m_keys_kernel")) return fail(MK_ERR_HIP);
    int64_t h_bad = 0;
    if (hipMemcpyAsync(&h_bad, nbad.p, 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return fail((mk::set_error("t2m plan: sync failed"), MK_ERR_HIP));
    if (h_bad) {
      mk::set_error("t2m plan: %lld samples have ipix outside [0, %d*%d)", (long long)h_bad, nx, ny);
      return fail(MK_ERR_RANGE);
    }
    bool in_b = false;
    {
      const int r = sort_pairs((uint32_t*)keys.p, (uint32_t*)idx.p, (uint32_t*)keys2.p, entry + s0, n, end_bit,
                               tmp.p, stream, &in_b);
      if (r) return fail(r);
    }
    if (!in_b &&  // an even number of passes left the result in (keys, idx)
        (hipMemcpyAsync(keys2.p, keys.p, n * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
         hipMemcpyAsync(entry + s0, idx.p, n * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess))
      return fail((mk::set_error("t2m plan: copy of the sorted entries failed"), MK_ERR_HIP));
    t2m_count_kernel<<<mk::grid_for(n, 256), 256, 0, stream>>>((uint32_t*)keys2.p, n, (int32_t*)cnt.p,
                                                                (int32_t*)cnt.p + ntiles);
    if (mk::check_launch("t2m_count_kernel")) return fail(MK_ERR_HIP);
    if (hipMemcpyAsync(h_cnt.data(), cnt.p, ntiles * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return fail((mk::set_error("t2m plan: count copy failed"), MK_ERR_HIP));
    int64_t pos = s0;
    for (int64_t t = 0; t < ntiles; t++) {
      int64_t c = (int64_t)h_cnt[ntiles + t] - h_cnt[t];
      for (int64_t o = 0; o < c; o += item_samples) {
        T2MItem it{};
        it.seg_base = s0;
        it.begin = pos + o;
        it.count = (int32_t)std::min<int64_t>(item_samples, c - o);
        it.tile = (int32_t)t;
        h_items.push_back(it);
      }
      pos += c;
    }
  }
  // items grouped by tile (stable: segment order, then position)
  std::stable_sort(h_items.begin(), h_items.end(), [](const T2MItem& a, const T2MItem& b) { return a.tile < b.tile; });
  pl->nitems = (int64_t)h_items.size();
  std::vector<int32_t> h_tile_items(ntiles + 1, 0);
  for (int64_t i = 0; i < pl->nitems; i++) {
    h_tile_items[h_items[i].tile + 1]++;
    pl->max_item = std::max<int64_t>(pl->max_item, h_items[i].count);
  }
  for (int64_t t = 0; t < ntiles; t++) {
    if (h_tile_items[t + 1]) pl->ntiles_touched++;
    h_tile_items[t + 1] += h_tile_items[t];
  }
  MK_REQUIRE(pl->nitems < INT32_MAX, fail(MK_ERR_ARG), "too many items");
  if (hipMalloc((void**)&pl->items, std::max<int64_t>(pl->nitems, 1) * sizeof(T2MItem)) != hipSuccess ||
      hipMalloc((void**)&pl->tile_items, (ntiles + 1) * 4) != hipSuccess)
    return fail((mk::set_error("t2m plan: item alloc failed"), MK_ERR_ALLOC));
  if (hipMemcpyAsync(pl->items, h_items.data(), pl->nitems * sizeof(T2MItem), hipMemcpyHostToDevice, stream) ||
      hipMemcpyAsync(pl->tile_items, h_tile_items.data(), (ntiles + 1) * 4, hipMemcpyHostToDevice, stream))
    return fail((mk::set_error("t2m plan: upload failed"), MK_ERR_HIP));
  // groups: count per item, then the headers, then each group's pixel order
  if (pl->nitems > 0) {
    const unsigned gi = (unsigned)mk::cdiv(pl->nitems, 256);
    t2m_groups_kernel<false><<<gi, 256, 0, stream>>>(pl->items, pl->nitems, entry, nullptr, nullptr, nullptr);
    if (mk::check_launch("t2m_groups_kernel (count)")) return fail(MK_ERR_HIP);
    if (hipMemcpyAsync(h_items.data(), pl->items, pl->nitems * sizeof(T2MItem), hipMemcpyDeviceToHost, stream) ||
        hipStreamSynchronize(stream))
      return fail((mk::set_error("t2m plan: group count copy failed"), MK_ERR_HIP));
    for (int64_t i = 0; i < pl->nitems; i++) {
      h_items[i].g_begin = pl->ngroups;
      pl->ngroups += h_items[i].g_count;
    }
    if (hipMemcpyAsync(pl->items, h_items.data(), pl->nitems * sizeof(T2MItem), hipMemcpyHostToDevice, stream))
      return fail((mk::set_error("t2m plan: upload failed"), MK_ERR_HIP));
    DevBuf gpos, gseg;
    if (hipMalloc((void**)&pl->groups, std::max<int64_t>(pl->ngroups, 1) * sizeof(T2MGroup)) != hipSuccess ||
        hipMalloc((void**)&pl->lanes, std::max<int64_t>(pl->ngroups, 1) * 64 * 2) != hipSuccess ||
        gpos.alloc(std::max<int64_t>(pl->ngroups, 1) * 8) || gseg.alloc(std::max<int64_t>(pl->ngroups, 1) * 8))
      return fail((mk::set_error("t2m plan: group alloc (%lld groups) failed", (long long)pl->ngroups),
                   MK_ERR_ALLOC));
    t2m_groups_kernel<true><<<gi, 256, 0, stream>>>(pl->items, pl->nitems, entry, pl->groups, (int64_t*)gpos.p,
                                                    (int64_t*)gseg.p);
    if (mk::check_launch("t2m_groups_kernel (write)")) return fail(MK_ERR_HIP);
    t2m_group_lanes_kernel<<<mk::grid_for(pl->ngroups * 64, 256, 256 * 256), 256, 0, stream>>>(
        pl->groups, pl->ngroups, (const int64_t*)gpos.p, (const int64_t*)gseg.p, entry, ipix, ny, tile_x, tile_y,
        pl->lanes);
    if (mk::check_launch("t2m_group_lanes_kernel")) return fail(MK_ERR_HIP);
    if (hipStreamSynchronize(stream) != hipSuccess)  // before the temporaries go
      return fail((mk::set_error("t2m plan: sync failed"), MK_ERR_HIP));
  }
  // partial tiles: 8-byte elements when affordable (FP64 exec), else 4
  int64_t pb8 = std::max<int64_t>(pl->nitems, 1) * tile_px * 8;
  pl->partial_elem_bytes = (pb8 <= (int64_t(1) << 30)) ? 8 : 4;
  pl->partial_bytes = std::max<int64_t>(pl->nitems, 1) * tile_px * pl->partial_elem_bytes;
  if (hipMalloc(&pl->partial, pl->partial_bytes) != hipSuccess)
    return fail((mk::set_error("t2m plan: partial alloc (%lld B) failed", (long long)pl->partial_bytes), MK_ERR_ALLOC));
  if (hipStreamSynchronize(stream) != hipSuccess) return fail((mk::set_error("t2m plan: sync failed"), MK_ERR_HIP));
  pl->plan_bytes = pl->ngroups * (int64_t)(sizeof(T2MGroup) + 64 * 2) + pl->nitems * (int64_t)sizeof(T2MItem) +
                   (ntiles + 1) * 4;
  *out = pl;
  return MK_OK;
}

namespace {
// TD: sample type, T: accumulation type, TM: map type.
template <typename TD, typename T, typename TM>
int t2m_exec(const mk_t2m_plan* pl, TM* map, const TD* dat, int accumulate, void* stream_v) {
  MK_REQUIRE(pl && map, MK_ERR_ARG, "null plan/map");
  MK_REQUIRE(pl->nsamp == 0 || dat, MK_ERR_ARG, "null dat");
  hipStream_t stream = mk::as_stream(stream_v);
  const int tile_px = pl->tile_x * pl->tile_y;
  const int64_t ntiles = (int64_t)pl->ntile_x * pl->ntile_y;
  const size_t lds = 4 * (size_t)tile_px * sizeof(T);  // four wave tiles, at most 32 KiB
  // FP64 partial tiles whenever the plan holds them (8-byte partial
  // elements); an FP64 accumulation with 4-byte partials rounds each item's
  // sum once
  const bool p64 = sizeof(T) == 8 && pl->partial_elem_bytes == 8;
  MK_REQUIRE(p64 || sizeof(TD) == 4, MK_ERR_ARG, "plan too large for FP64 partial tiles; use FP32");
  if (pl->nitems > 0) {
    if (p64)
      t2m_items_kernel<TD, T, double><<<(unsigned)pl->nitems, 256, lds, stream>>>(pl->items, pl->groups, pl->lanes,
                                                                                  dat, tile_px, (double*)pl->partial);
    else
      t2m_items_kernel<TD, T, float><<<(unsigned)pl->nitems, 256, lds, stream>>>(pl->items, pl->groups, pl->lanes,
                                                                                 dat, tile_px, (float*)pl->partial);
    int st = mk::check_launch("t2m_items_kernel");
    if (st) return st;
  }
  const int bpt = (int)mk::cdiv(tile_px / 4, 256);
  const unsigned nblk = (unsigned)(ntiles * bpt);
  if (p64)
    t2m_tiles_kernel<T, double, TM><<<nblk, 256, 0, stream>>>(pl->tile_items, (const double*)pl->partial, pl->tile_x,
                                                              pl->tile_y, pl->ntile_y, pl->nx, pl->ny, map,
                                                              accumulate, bpt);
  else
    t2m_tiles_kernel<T, float, TM><<<nblk, 256, 0, stream>>>(pl->tile_items, (const float*)pl->partial, pl->tile_x,
                                                             pl->tile_y, pl->ntile_y, pl->nx, pl->ny, map,
                                                             accumulate, bpt);
  return mk::check_launch("t2m_tiles_kernel");
}
}  // namespace

extern "C" int mk_t2m_plan_exec_f32(const mk_t2m_plan* plan, float* map, const float* dat, int accumulate,
                                    void* stream) {
  return t2m_exec<float, float, float>(plan, map, dat, accumulate, stream);
}
extern "C" int mk_t2m_plan_exec_f64(const mk_t2m_plan* plan, double* map, const double* dat, int accumulate,
                                    void* stream) {
  return t2m_exec<double, double, double>(plan, map, dat, accumulate, stream);
}
extern "C" int mk_t2m_plan_exec_f32_acc64(const mk_t2m_plan* plan, void* map, int map_f64, const float* dat,
                                          int accumulate, void* stream) {
  if (map_f64) return t2m_exec<float, double, double>(plan, (double*)map, dat, accumulate, stream);
  return t2m_exec<float, double, float>(plan, (float*)map, dat, accumulate, stream);
}
namespace {
// map (npix x ncomp, component-minor as PolMap holds it) (+)= the per-sample
// terms of mode, component by component through the plan: each component
// is one deterministic binning of sample * weight.  FP32 samples are
// summed in FP64.
template <typename TD, int POL>
int t2m_exec_pol_mode(const mk_t2m_plan* pl, TD* map, const TD* dat, const TD* twogamma, int accumulate,
                      hipStream_t stream) {
  typedef double T;
  const int ncomp = POL == MK_POL_QU ? 2 : (POL == MK_POL_IQU_PRECON ? 6 : 3);
  const int tile_px = pl->tile_x * pl->tile_y;
  const int64_t ntiles = (int64_t)pl->ntile_x * pl->ntile_y;
  const size_t lds = 4 * (size_t)tile_px * sizeof(T);  // four wave tiles, at most 32 KiB
  const bool p64 = pl->partial_elem_bytes == 8;
  MK_REQUIRE(p64 || sizeof(TD) == 4, MK_ERR_ARG, "plan too large for FP64 partial tiles; use FP32");
  const int bpt = (int)mk::cdiv(tile_px / 4, 256);
  const unsigned nblk = (unsigned)(ntiles * bpt);
  for (int c = 0; c < ncomp; c++) {
    if (pl->nitems > 0) {
      if (p64)
        t2m_items_kernel<TD, T, double, POL><<<(unsigned)pl->nitems, 256, lds, stream>>>(
            pl->items, pl->groups, pl->lanes, dat, tile_px, (double*)pl->partial, twogamma, c);
      else
        t2m_items_kernel<TD, T, float, POL><<<(unsigned)pl->nitems, 256, lds, stream>>>(
            pl->items, pl->groups, pl->lanes, dat, tile_px, (float*)pl->partial, twogamma, c);
      int st = mk::check_launch("t2m_items_kernel (pol)");
      if (st) return st;
    }
    if (p64)
      t2m_tiles_kernel<T, double, TD><<<nblk, 256, 0, stream>>>(pl->tile_items, (const double*)pl->partial,
                                                                pl->tile_x, pl->tile_y, pl->ntile_y, pl->nx, pl->ny,
                                                                map, accumulate, bpt, ncomp, c);
    else
      t2m_tiles_kernel<T, float, TD><<<nblk, 256, 0, stream>>>(pl->tile_items, (const float*)pl->partial,
                                                               pl->tile_x, pl->tile_y, pl->ntile_y, pl->nx, pl->ny,
                                                               map, accumulate, bpt, ncomp, c);
    int st = mk::check_launch("t2m_tiles_kernel (pol)");
    if (st) return st;
  }
  return MK_OK;
}

template <typename TD>
int t2m_exec_pol(const mk_t2m_plan* pl, TD* map, const TD* dat, const TD* twogamma, int mode, int accumulate,
                 void* stream_v) {
  MK_REQUIRE(pl && map, MK_ERR_ARG, "null plan/map");
  MK_REQUIRE(pl->nsamp == 0 || (dat && twogamma), MK_ERR_ARG, "null dat/twogamma");
  hipStream_t st = mk::as_stream(stream_v);
  switch (mode) {
    case MK_POL_QU: return t2m_exec_pol_mode<TD, MK_POL_QU>(pl, map, dat, twogamma, accumulate, st);
    case MK_POL_IQU: return t2m_exec_pol_mode<TD, MK_POL_IQU>(pl, map, dat, twogamma, accumulate, st);
    case MK_POL_QU_PRECON: return t2m_exec_pol_mode<TD, MK_POL_QU_PRECON>(pl, map, dat, twogamma, accumulate, st);
    case MK_POL_IQU_PRECON:
      return t2m_exec_pol_mode<TD, MK_POL_IQU_PRECON>(pl, map, dat, twogamma, accumulate, st);
    default: break;
  }
  mk::set_error("unknown polarisation mode %d", mode);
  return MK_ERR_ARG;
}
}  // namespace

extern "C" int mk_t2m_plan_exec_pol_f32(const mk_t2m_plan* plan, float* map, const float* dat,
                                        const float* twogamma, int mode, int accumulate, void* stream) {
  return t2m_exec_pol<float>(plan, map, dat, twogamma, mode, accumulate, stream);
}
extern "C" int mk_t2m_plan_exec_pol_f64(const mk_t2m_plan* plan, double* map, const double* dat,
                                        const double* twogamma, int mode, int accumulate, void* stream) {
  return t2m_exec_pol<double>(plan, map, dat, twogamma, mode, accumulate, stream);
}

extern "C" int mk_t2m_plan_info(const mk_t2m_plan* pl, int64_t* info) {
  MK_REQUIRE(pl && info, MK_ERR_ARG, "null pointer");
  info[0] = pl->nsamp;
  info[1] = pl->nitems;
  info[2] = pl->ntiles_touched;
  info[3] = pl->tile_x;
  info[4] = pl->tile_y;
  info[5] = pl->plan_bytes;
  info[6] = pl->partial_bytes;
  info[7] = pl->max_item;
  return MK_OK;
}
extern "C" int mk_t2m_plan_destroy(mk_t2m_plan* pl) {
  free_plan(pl);
  return MK_OK;
}

extern "C" int64_t mk_map_tiled_size(int32_t nx, int32_t ny, int32_t tile_x, int32_t tile_y) {
  if (nx <= 0 || ny <= 0 || tile_x <= 0 || tile_y <= 0) return -1;
  return tiled_size(nx, ny, tile_x, tile_y);
}

extern "C" int mk_pix_tile_remap(const int32_t* ipix, int64_t nsamp, int32_t nx, int32_t ny, int32_t tile_x,
                                 int32_t tile_y, int32_t* out, int64_t* dev_nbad, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  MK_REQUIRE(nx > 0 && ny > 0 && tile_x > 0 && tile_y > 0 && tile_x * tile_y <= 1024, MK_ERR_ARG,
             "bad map or tile size");
  MK_REQUIRE(tiled_size(nx, ny, tile_x, tile_y) <= INT32_MAX, MK_ERR_ARG, "tiled map too large for int32 indices");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(ipix && out, MK_ERR_ARG, "null pointer");
  pix_tile_kernel<<<mk::grid_for(nsamp, 256, 256 * 32), 256, 0, mk::as_stream(stream)>>>(ipix, nsamp, nx, ny, tile_x,
                                                                                          tile_y, out, dev_nbad);
  return mk::check_launch("pix_tile_kernel");
}

extern "C" int mk_map_to_tiles_f32(const float* map, int32_t nx, int32_t ny, int32_t tile_x, int32_t tile_y,
                                   float* tiled, void* stream) {
  return map_tiles_impl<float>(map, nx, ny, tile_x, tile_y, tiled, stream);
}
extern "C" int mk_map_to_tiles_f64(const double* map, int32_t nx, int32_t ny, int32_t tile_x, int32_t tile_y,
                                   double* tiled, void* stream) {
  return map_tiles_impl<double>(map, nx, ny, tile_x, tile_y, tiled, stream);
}

// ---------------------------------------------------------------------------
// Coded pointing for the tiled gather.  A raster TOD moves a fraction of a
// pixel per sample, so the 256 samples of a block span a few dozen map
// columns and usually one map row.  Each block keeps its first sample's
// pixel (x0, y0) and, per sample, the offset from it:
//   type 1: dx as int8, every dy = 0            1 byte per sample
//   type 2: (dx, dy) as two int8                2 bytes per sample
//   type 3: (x << 16 | y) as int32              4 bytes per sample (fallback)
// Header per block: int2 {x0 << 16 | y0, payload offset in 256-byte units
// | type << 30}.  The gather decodes (x, y), forms the tiled index with
// shifts and masks (tile sides are powers of two) and reads the tiled map:
// about 1 byte of index traffic per sample instead of 4, bit-exact.

namespace {
constexpr int kCodeBlock = 256;  // samples per block: one wave, 4 per lane

__device__ __forceinline__ bool code_fits8(int v) { return v >= -128 && v <= 127; }

// One wave per block: payload size (in 256-byte units) and bad indices.
__global__ void __launch_bounds__(256) pix_code_size_kernel(const int32_t* __restrict__ ipix, int64_t n, int32_t nx,
                                                            int32_t ny, int64_t nblocks,
                                                            int32_t* __restrict__ units, int64_t* nbad) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t npix = (int64_t)nx * ny;
  for (int64_t b = wave; b < nblocks; b += nwaves) {
    const int64_t s0 = b * kCodeBlock;
    int x[4], y[4];
    bool ok[4], bad = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int64_t i = s0 + 4 * lane + k;
      ok[k] = i < n;
      const int32_t p = ok[k] ? ipix[i] : 0;
      const bool b_ = ok[k] && (p < 0 || (int64_t)p >= npix);
      bad |= b_;
      ok[k] = ok[k] && !b_;
      x[k] = ok[k] ? p / ny : 0;
      y[k] = ok[k] ? p - x[k] * ny : 0;
    }
    mk::count_bad(bad, nbad);
    const int x0 = __shfl(x[0], 0, 64), y0 = __shfl(y[0], 0, 64);
    bool a = true, c = true;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int dx = x[k] - x0, dy = y[k] - y0;
      a = a && (!ok[k] || (dy == 0 && code_fits8(dx)));
      c = c && (!ok[k] || (code_fits8(dx) && code_fits8(dy)));
    }
    const bool all_a = __ballot(!a) == 0, all_c = __ballot(!c) == 0;
    if (lane == 0) units[b] = all_a ? 1 : (all_c ? 2 : 4);
  }
}

__global__ void __launch_bounds__(256) pix_code_write_kernel(const int32_t* __restrict__ ipix, int64_t n, int32_t ny,
                                                             int64_t nblocks, const int32_t* __restrict__ off_units,
                                                             const int32_t* __restrict__ units,
                                                             int2* __restrict__ hdr, uint8_t* __restrict__ payload) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t b = wave; b < nblocks; b += nwaves) {
    const int64_t s0 = b * kCodeBlock;
    int x[4], y[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int64_t i = s0 + 4 * lane + k;
      const int32_t p = i < n ? ipix[i] : ipix[s0];  // past the end: a copy of the first (never gathered)
      x[k] = p / ny;
      y[k] = p - x[k] * ny;
    }
    const int x0 = __shfl(x[0], 0, 64), y0 = __shfl(y[0], 0, 64);
    const int u = units[b];
    const int type = u == 1 ? 1 : (u == 2 ? 2 : 3);
    uint8_t* base = payload + (int64_t)off_units[b] * 256;
    if (type == 1) {
      uint32_t w = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) w |= (uint32_t)(uint8_t)(int8_t)(x[k] - x0) << (8 * k);
      reinterpret_cast<uint32_t*>(base)[lane] = w;
    } else if (type == 2) {
      uint32_t w0 = 0, w1 = 0;
#pragma unroll
      for (int k = 0; k < 2; k++) {
        w0 |= ((uint32_t)(uint8_t)(int8_t)(x[k] - x0) | (uint32_t)(uint8_t)(int8_t)(y[k] - y0) << 8) << (16 * k);
        w1 |= ((uint32_t)(uint8_t)(int8_t)(x[k + 2] - x0) | (uint32_t)(uint8_t)(int8_t)(y[k + 2] - y0) << 8)
              << (16 * k);
      }
      reinterpret_cast<uint2*>(base)[lane] = make_uint2(w0, w1);
    } else {
      int4 w;
      w.x = (x[0] << 16) | y[0];
      w.y = (x[1] << 16) | y[1];
      w.z = (x[2] << 16) | y[2];
      w.w = (x[3] << 16) | y[3];
      reinterpret_cast<int4*>(base)[lane] = w;
    }
    if (lane == 0) hdr[b] = make_int2((x0 << 16) | y0, off_units[b] | (type << 30));
  }
}

// Tiled index of pixel (x, y) for tx = 1 << ltx, ty = 1 << lty.
__device__ __forceinline__ int32_t tiled_index(int x, int y, int ltx, int lty, int32_t nty) {
  return (((x >> ltx) * nty + (y >> lty)) << (ltx + lty)) + ((x & ((1 << ltx) - 1)) << lty) + (y & ((1 << lty) - 1));
}

// Decodes the four pixels of quad q (samples 4q..4q+3) of block q / 64.
__device__ __forceinline__ void decode_quad(const int2* __restrict__ hdr, const uint8_t* __restrict__ payload,
                                            int64_t q, int* x, int* y) {
  const int2 h = hdr[q >> 6];
  const int x0 = (int)((uint32_t)h.x >> 16), y0 = h.x & 0xffff;
  const int type = (int)((uint32_t)h.y >> 30);
  const uint8_t* base = payload + (int64_t)(h.y & 0x3fffffff) * 256;
  const int l = (int)(q & 63);
  if (type == 1) {
    const uint32_t w = reinterpret_cast<const uint32_t*>(base)[l];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      x[k] = x0 + (int)(int8_t)(w >> (8 * k));
      y[k] = y0;
    }
  } else if (type == 2) {
    const uint2 w = reinterpret_cast<const uint2*>(base)[l];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      x[k] = x0 + (int)(int8_t)(w.x >> (16 * k));
      y[k] = y0 + (int)(int8_t)(w.x >> (16 * k + 8));
      x[k + 2] = x0 + (int)(int8_t)(w.y >> (16 * k));
      y[k + 2] = y0 + (int)(int8_t)(w.y >> (16 * k + 8));
    }
  } else {
    const int4 w = reinterpret_cast<const int4*>(base)[l];
    x[0] = (int)((uint32_t)w.x >> 16); y[0] = w.x & 0xffff;
    x[1] = (int)((uint32_t)w.y >> 16); y[1] = w.y & 0xffff;
    x[2] = (int)((uint32_t)w.z >> 16); y[2] = w.z & 0xffff;
    x[3] = (int)((uint32_t)w.w >> 16); y[3] = w.w & 0xffff;
  }
}

// dat (+)= tiled[index(x, y)] over the arena, 4 samples per thread, U quads
// per loop trip with their header / payload loads and gathers in flight.
// The stride is a multiple of 64 quads, so a wave always covers whole
// blocks (one header, one block type per wave and step).
template <typename T>
__global__ void __launch_bounds__(256) map2tod_coded_kernel(T* __restrict__ dat, const T* __restrict__ tiled,
                                                            int64_t n, const int2* __restrict__ hdr,
                                                            const uint8_t* __restrict__ payload, int ltx, int lty,
                                                            int32_t nty, int do_add) {
  constexpr int U = 4;  // measured: 2.74 ms at 4, 2.92 ms at 8 (1.9e9 samples)
  const int64_t n4 = n >> 2;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  Quad<T>* d4 = reinterpret_cast<Quad<T>*>(dat);
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; q + (U - 1) * stride < n4; q += U * stride) {
    int x[U][4], y[U][4];
#pragma unroll
    for (int u = 0; u < U; u++) decode_quad(hdr, payload, q + u * stride, x[u], y[u]);
    Quad<T> v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int k = 0; k < 4; k++) v[u].v[k] = tiled[tiled_index(x[u][k], y[u][k], ltx, lty, nty)];
    if (do_add) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const Quad<T> o = d4[q + u * stride];
#pragma unroll
        for (int k = 0; k < 4; k++) v[u].v[k] += o.v[k];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) d4[q + u * stride] = v[u];
  }
  for (; q < n4; q += stride) {
    int x[4], y[4];
    decode_quad(hdr, payload, q, x, y);
    Quad<T> v;
#pragma unroll
    for (int k = 0; k < 4; k++) v.v[k] = tiled[tiled_index(x[k], y[k], ltx, lty, nty)];
    if (do_add) {
      const Quad<T> o = d4[q];
#pragma unroll
      for (int k = 0; k < 4; k++) v.v[k] += o.v[k];
    }
    d4[q] = v;
  }
  // the last n % 4 samples: one thread decodes their quad
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 3)) {
    int x[4], y[4];
    decode_quad(hdr, payload, n4, x, y);
    for (int k = 0; k < (int)(n & 3); k++) {
      const T v = tiled[tiled_index(x[k], y[k], ltx, lty, nty)];
      dat[4 * n4 + k] = do_add ? dat[4 * n4 + k] + v : v;
    }
  }
}

int log2_exact(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return (1 << l) == v ? l : -1;
}

template <typename T>
int map2tod_coded_impl(T* dat, const T* tiled, int64_t nsamp, const int2* hdr, const uint8_t* payload, int32_t ny,
                       int32_t tile_x, int32_t tile_y, int do_add, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(dat && tiled && hdr && payload, MK_ERR_ARG, "null pointer");
  const int ltx = log2_exact(tile_x), lty = log2_exact(tile_y);
  MK_REQUIRE(ltx >= 0 && lty >= 0 && tile_x * tile_y <= 1024, MK_ERR_ARG,
             "coded gather needs power-of-two tile sides (got %d x %d)", tile_x, tile_y);
  MK_REQUIRE(ny > 0 && ny < 32768, MK_ERR_ARG, "bad ny");
  MK_REQUIRE(((uintptr_t)dat % (4 * sizeof(T))) == 0 && ((uintptr_t)payload & 255) == 0, MK_ERR_ARG,
             "coded gather: dat must be 4-element aligned and the payload 256-byte aligned");
  const int32_t nty = (ny + tile_y - 1) / tile_y;
  // a stride of 64-quad multiples: 256-thread workgroups
  const unsigned grid = mk::grid_for(mk::cdiv(nsamp, 16), 256, 8192);
  map2tod_coded_kernel<T><<<grid, 256, 0, mk::as_stream(stream)>>>(dat, tiled, nsamp, hdr, payload, ltx, lty, nty,
                                                                    do_add);
  return mk::check_launch("map2tod_coded_kernel");
}
}  // namespace

extern "C" int64_t mk_pix_code_blocks(int64_t nsamp) { return nsamp <= 0 ? 0 : mk::cdiv(nsamp, kCodeBlock); }

extern "C" int mk_pix_code_size(const int32_t* ipix, int64_t nsamp, int32_t nx, int32_t ny, int32_t* units,
                                int64_t* dev_nbad, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  MK_REQUIRE(nx > 0 && ny > 0 && nx < 32768 && ny < 32768, MK_ERR_ARG,
             "coded pointing needs map sides below 32768 (got %d x %d)", nx, ny);
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(ipix && units && dev_nbad, MK_ERR_ARG, "null pointer");
  const int64_t nb = mk::cdiv(nsamp, kCodeBlock);
  const unsigned grid = (unsigned)std::min<int64_t>(mk::cdiv(nb, 4), 256 * 64);
  pix_code_size_kernel<<<grid, 256, 0, mk::as_stream(stream)>>>(ipix, nsamp, nx, ny, nb, units, dev_nbad);
  return mk::check_launch("pix_code_size_kernel");
}

extern "C" int mk_pix_code_write(const int32_t* ipix, int64_t nsamp, int32_t ny, const int32_t* off_units,
                                 const int32_t* units, void* hdr, void* payload, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  MK_REQUIRE(ny > 0 && ny < 32768, MK_ERR_ARG, "bad ny");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(ipix && off_units && units && hdr && payload, MK_ERR_ARG, "null pointer");
  MK_REQUIRE(((uintptr_t)payload & 255) == 0 && ((uintptr_t)hdr & 7) == 0, MK_ERR_ARG,
             "payload must be 256-byte aligned, headers 8-byte aligned");
  const int64_t nb = mk::cdiv(nsamp, kCodeBlock);
  const unsigned grid = (unsigned)std::min<int64_t>(mk::cdiv(nb, 4), 256 * 64);
  pix_code_write_kernel<<<grid, 256, 0, mk::as_stream(stream)>>>(ipix, nsamp, ny, nb, off_units, units, (int2*)hdr,
                                                                  (uint8_t*)payload);
  return mk::check_launch("pix_code_write_kernel");
}

extern "C" int mk_map2tod_coded_f32(float* dat, const float* tiled, int64_t nsamp, const void* hdr,
                                    const void* payload, int32_t ny, int32_t tile_x, int32_t tile_y, int do_add,
                                    void* stream) {
  return map2tod_coded_impl<float>(dat, tiled, nsamp, (const int2*)hdr, (const uint8_t*)payload, ny, tile_x, tile_y,
                                   do_add, stream);
}
extern "C" int mk_map2tod_coded_f64(double* dat, const double* tiled, int64_t nsamp, const void* hdr,
                                    const void* payload, int32_t ny, int32_t tile_x, int32_t tile_y, int do_add,
                                    void* stream) {
  return map2tod_coded_impl<double>(dat, tiled, nsamp, (const int2*)hdr, (const uint8_t*)payload, ny, tile_x,
                                    tile_y, do_add, stream);
}
