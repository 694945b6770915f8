// Stable LSD radix sort of (key, u32 value) pairs, in place.
//
// Replaces the vulkan_radix_sort dependency called by Tracer::radixSort
// (projects/WaveFrontPathTracer/Tracer/Tracer.cpp:39-78 ->
// external/vulkan_radix_sort/vk_radix_sort.h:2579-2585), which runs 4 fixed
// 8-bit passes on u32 keys.  The structure follows the pair count (any
// structure gives the same stable order):
// - large sorts, per pass reduce-then-scan with no dependency between the
//   workgroups of a kernel: an upsweep counts each 4096-element tile's
//   digits, a spine scans every digit's row of tile counts (one workgroup
//   per digit) and writes the digit's total, and a downsweep scans the digit
//   totals into digit starts, ranks its tile stably with wave64 ballot
//   matching, stages it in LDS in digit order and writes each digit run
//   contiguously to its position.  HBM traffic per element and pass: a key
//   read (upsweep) and one read and one write of key and value (downsweep).
//   Decoupled look-back measured slower at these sizes: its chained status
//   reads cross the XCDs' separate L2 caches and, with ~500-2000 tiles
//   resident at once, walk hundreds of predecessors;
// - up to 64 tiles, onesweep: one kernel per pass with decoupled look-back;
// - up to 8192 pairs, one workgroup runs every pass in LDS;
// - up to ~28k pairs (scaled by the pass count), a one-pass counted rank.
#include <mutex>

#include "common.h"

namespace wfpt {

constexpr int SORT_BLOCK = 512;
constexpr int SORT_WAVES = SORT_BLOCK / 64;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;
constexpr int SORT_MAX_PASSES = 16;
// Tile counts are stored in chunks of SORT_CHUNK consecutive tiles: within a
// chunk digit-major, so a digit's counts of the chunk's tiles share one
// 64-byte line.  The upsweep and the downsweep hand each XCD runs of
// SORT_CHUNK consecutive tiles (xcd_block), so the lines of a chunk, and the
// neighbouring digit runs its tiles write, stay in one XCD's L2.
constexpr int SORT_CHUNK = 16;

template <int RADIX>
__host__ __device__ __forceinline__ int64_t count_index(int d, int64_t tile) {
    return (tile / SORT_CHUNK) * (int64_t)(RADIX * SORT_CHUNK) + (int64_t)d * SORT_CHUNK + tile % SORT_CHUNK;
}

inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct SortPlan {
    int passes = 0;
    int radix = 0;
    int64_t tiles = 0;
    // byte offsets inside scratch
    int64_t off_counts = 0, off_totals = 0;
    int64_t off_keys_a = 0, off_vals_a = 0, off_keys_b = 0, off_vals_b = 0;
    // onesweep: digit histograms of all passes, per-pass tile counters and
    // two look-back regions, each a status array (tiles x radix) followed by
    // the chunk sums (chunks x radix); the histograms, the counters and the
    // first region are cleared as one span
    int64_t off_hist = 0, off_ctr = 0, off_status = 0, status_words = 0, csum_words = 0,
            region_words = 0, clear_words = 0;
    int64_t total = 0;
};

static bool make_plan(int64_t n, int key_bytes, int key_bits, int digit_bits, SortPlan& p) {
    if (!(key_bytes == 4 || key_bytes == 8)) return false;
    if (key_bits < 1 || key_bits > key_bytes * 8) return false;
    if (!(digit_bits == 4 || digit_bits == 8 || digit_bits == 11)) return false;
    if (n < 0) return false;
    p.passes = (key_bits + digit_bits - 1) / digit_bits;
    if (p.passes > SORT_MAX_PASSES) return false;
    p.radix = 1 << digit_bits;
    p.tiles = (n + SORT_TILE - 1) / SORT_TILE;
    int64_t o = 0;
    p.off_counts = o;  // tile counts of one pass, in chunks of SORT_CHUNK tiles, scanned in place
    o += align256(sizeof(uint32_t) * ((p.tiles + SORT_CHUNK - 1) / SORT_CHUNK) * SORT_CHUNK * p.radix);
    p.off_totals = o;  // digit totals of one pass
    o += align256(sizeof(uint32_t) * p.radix);
    p.off_keys_a = o;
    o += align256((int64_t)key_bytes * n);
    p.off_vals_a = o;
    o += align256(sizeof(uint32_t) * n);
    p.off_keys_b = o;
    o += align256((int64_t)key_bytes * n);
    p.off_vals_b = o;
    o += align256(sizeof(uint32_t) * n);
    p.off_hist = o;  // hist | counters | status 0 are contiguous (one clear)
    p.off_ctr = o + sizeof(uint32_t) * (int64_t)p.passes * p.radix;
    p.off_status = p.off_ctr + sizeof(uint32_t) * SORT_MAX_PASSES;
    p.status_words = p.tiles * p.radix;
    p.csum_words = (p.tiles + SORT_CHUNK - 1) / SORT_CHUNK * p.radix;
    p.region_words = p.status_words + p.csum_words;
    p.clear_words = ((int64_t)p.passes * p.radix + SORT_MAX_PASSES) + p.region_words;
    o = align256(p.off_status + sizeof(uint32_t) * 2 * p.region_words);
    p.total = o;
    return true;
}

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// ---------------------------------------------------------------- binning
// Exclusive scan of one value per thread over a workgroup of WAVES waves (the
// launch's block size / 64); returns the exclusive prefix and writes the
// block total to *total.
template <int WAVES>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x >> 6;
    // inclusive wave scan
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (unsigned)off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, all = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        uint32_t t = s_tmp[w];
        if (w < (int)wave) wbase += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return wbase + x - v;
}

// Lanes of the wave whose DBITS-bit digit equals the calling lane's d (among
// lanes with ok set): one ballot per digit bit, of which each lane keeps the
// ballot where its own bit is set and the complement where it is clear,
// i.e. peers &= ~(m ^ ballot) with m the lane's bit spread to all 32 bits
// (one three-input logic operation per 32-bit half).
template <int DBITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool ok) {
    const uint64_t okm = ballot64(ok);
    uint32_t lo = (uint32_t)okm, hi = (uint32_t)(okm >> 32);
#pragma unroll
    for (int b = 0; b < DBITS; ++b) {
        const uint32_t m = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
        const uint64_t bal = ballot64((d >> b) & 1u);
        lo &= ~(m ^ (uint32_t)bal);
        hi &= ~(m ^ (uint32_t)(bal >> 32));
    }
    return ((uint64_t)hi << 32) | lo;
}

// Upsweep: digit counts of one tile, written to count_index(d, tile).
// Tiles past the element count write zeros.  Lanes count into COPIES
// histograms per wave (lane & (COPIES - 1)), so a digit shared by many keys
// (skewed high digits) costs at most 64 / COPIES same-address LDS atomics per
// instruction.
template <typename K, int DBITS>
__global__ __launch_bounds__(SORT_BLOCK) void sort_upsweep_kernel(const K* __restrict__ keys,
                                                                  int64_t n, int shift, int width,
                                                                  uint32_t* __restrict__ counts,
                                                                  const uint32_t* __restrict__ d_n) {
    constexpr int RADIX = 1 << DBITS;
    constexpr int COPIES = DBITS <= 8 ? 4 : 1;
    __shared__ uint32_t s_cnt[SORT_WAVES * COPIES][RADIX];
    const unsigned tid = threadIdx.x, wave = tid >> 6;
    const unsigned hsel = wave * COPIES + (lane_id() & (COPIES - 1));
    for (int i = tid; i < SORT_WAVES * COPIES * RADIX; i += SORT_BLOCK) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    if (d_n) n = min(n, (int64_t)*d_n);
    const uint32_t mask = (1u << width) - 1u;
    const int64_t tile = xcd_block(blockIdx.x, gridDim.x, SORT_CHUNK);
    const int64_t base = tile * SORT_TILE;
    if (base < n) {
        K kk[SORT_ITEMS];  // all loads in flight before the counting
#pragma unroll
        for (int it = 0; it < SORT_ITEMS; ++it) {
            const int64_t idx = base + (int64_t)it * SORT_BLOCK + tid;
            kk[it] = idx < n ? keys[idx] : K(0);
        }
#pragma unroll
        for (int it = 0; it < SORT_ITEMS; ++it) {
            const int64_t idx = base + (int64_t)it * SORT_BLOCK + tid;
            if (idx < n) atomicAdd(&s_cnt[hsel][digit_of(kk[it], shift, mask)], 1u);
        }
    }
    __syncthreads();
    for (int d = tid; d < RADIX; d += SORT_BLOCK) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < SORT_WAVES * COPIES; ++w) c += s_cnt[w][d];
        counts[count_index<RADIX>(d, tile)] = c;
    }
}

// Spine: one workgroup per digit scans its tile counts in tile order, in
// place (exclusive), and writes the total, the digit's count in the pass.
constexpr int SPINE_BLOCK = 512;
static_assert(SPINE_BLOCK % SORT_CHUNK == 0, "spine rows cover whole chunks");
template <int RADIX>
__global__ __launch_bounds__(SPINE_BLOCK) void sort_spine_kernel(uint32_t* __restrict__ counts,
                                                                 int64_t tiles,
                                                                 uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_tmp[SPINE_BLOCK / 64];
    const int d = blockIdx.x;
    uint32_t carry = 0;
    for (int64_t base = 0; base < tiles; base += SPINE_BLOCK) {
        const int64_t i = base + threadIdx.x;
        uint32_t* slot = counts + count_index<RADIX>(d, i);
        const uint32_t v = i < tiles ? *slot : 0u;
        uint32_t total;
        const uint32_t ex = block_excl_scan<SPINE_BLOCK / 64>(v, s_tmp, &total);
        if (i < tiles) *slot = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) totals[d] = carry;
}

// Downsweep: stable ranking of a tile, staged in LDS in digit order, each
// digit run written contiguously from its scanned position.
template <typename K, int DBITS>
__global__ __launch_bounds__(SORT_BLOCK) void sort_bin_kernel(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    K* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int64_t n, int shift, int width,
    const uint32_t* __restrict__ tile_offsets, const uint32_t* __restrict__ totals,
    const uint32_t* __restrict__ d_n) {
    constexpr int RADIX = 1 << DBITS;
    constexpr int DPT = RADIX >= SORT_BLOCK ? RADIX / SORT_BLOCK : 1;  // digits per thread
    __shared__ uint32_t s_wave[SORT_WAVES][RADIX];  // per-wave counts -> wave prefixes
    __shared__ uint32_t s_lstart[RADIX];            // tile-local start of each digit
    __shared__ uint32_t s_goff[RADIX];              // global base - local start
    // Staging area, keys first, then values (one buffer keeps LDS per
    // workgroup small enough for several resident workgroups per CU).
    __shared__ K s_stage[SORT_TILE];
    __shared__ uint32_t s_tmp[SORT_WAVES];

    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t mask = (1u << width) - 1u;
    if (d_n) n = min(n, (int64_t)*d_n);
    const int64_t tile = xcd_block(blockIdx.x, gridDim.x, SORT_CHUNK);
    const int64_t base = tile * SORT_TILE;
    if (base >= n) return;  // uniform: surplus tiles of a capacity-sized grid
    for (int i = tid; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_wave[0][0])[i] = 0;
    // the pass's digit totals and this tile's scanned offsets, fetched
    // beside the keys
    uint32_t tot[DPT], toff[DPT];
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        tot[k] = d < RADIX ? totals[d] : 0u;
        toff[k] = d < RADIX ? tile_offsets[count_index<RADIX>(d, tile)] : 0u;
    }
    __syncthreads();

    // Wave-striped load: element index = base + wave*64*ITEMS + it*64 + lane.
    K key[SORT_ITEMS];
    uint32_t val[SORT_ITEMS];
    uint32_t off[SORT_ITEMS];
    const int64_t wbase = base + (int64_t)wave * 64 * SORT_ITEMS;
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        int64_t idx = wbase + it * 64 + lane;
        bool ok = idx < n;
        key[it] = ok ? keys_in[idx] : K(0);
        val[it] = ok ? vals_in[idx] : 0u;
    }

    // Stable within-wave ranking by ballot matching, item by item.
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        int64_t idx = wbase + it * 64 + lane;
        bool ok = idx < n;
        uint32_t d = digit_of(key[it], shift, mask);
        const uint64_t peers = match_digit<DBITS>(d, ok);
        uint32_t rank = popc_below(peers);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t pre = 0;
        if (ok) pre = s_wave[wave][d];
        off[it] = pre + rank;
        if (ok && rank == 0) s_wave[wave][d] = pre + cnt;
    }
    __syncthreads();

    // Per digit: exclusive prefix over waves and tile total.
    uint32_t cnt[DPT];
    uint32_t my_total = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        int d = tid * DPT + k;
        uint32_t sum = 0;
        if (d < RADIX) {
#pragma unroll
            for (int w = 0; w < SORT_WAVES; ++w) {
                uint32_t c = s_wave[w][d];
                s_wave[w][d] = sum;
                sum += c;
            }
        }
        cnt[k] = sum;
        my_total += sum;
    }
    // Tile-local digit starts (block scan over digits in thread order).
    uint32_t tile_total;
    uint32_t lstart = block_excl_scan<SORT_WAVES>(my_total, s_tmp, &tile_total);
    (void)tile_total;
    // Digit starts in the output: exclusive scan of the pass's digit totals.
    uint32_t my_tot = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) my_tot += tot[k];
    uint32_t all_total;
    uint32_t dstart = block_excl_scan<SORT_WAVES>(my_tot, s_tmp, &all_total);
    (void)all_total;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        int d = tid * DPT + k;
        if (d < RADIX) {
            s_lstart[d] = lstart;
            s_goff[d] = dstart + toff[k] - lstart;
            lstart += cnt[k];
            dstart += tot[k];
        }
    }
    __syncthreads();

    // Position of every item in the tile's digit order.
    uint32_t pos[SORT_ITEMS];
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        uint32_t d = digit_of(key[it], shift, mask);
        pos[it] = s_lstart[d] + s_wave[wave][d] + off[it];
    }
    // Keys: stage in digit order, then each digit run goes out contiguously.
    const int count = (int)min((int64_t)SORT_TILE, n - base);
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it)
        if (wbase + it * 64 + lane < n) s_stage[pos[it]] = key[it];
    __syncthreads();
    uint32_t gpos[SORT_ITEMS];
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const int j = it * SORT_BLOCK + (int)tid;
        if (j < count) {
            const K k = s_stage[j];
            gpos[it] = s_goff[digit_of(k, shift, mask)] + (uint32_t)j;
            keys_out[gpos[it]] = k;
        }
    }
    __syncthreads();
    // Values through the same staging buffer and the same global positions.
    uint32_t* s_vstage = reinterpret_cast<uint32_t*>(s_stage);
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it)
        if (wbase + it * 64 + lane < n) s_vstage[pos[it]] = val[it];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const int j = it * SORT_BLOCK + (int)tid;
        if (j < count) vals_out[gpos[it]] = s_vstage[j];
    }
}

// ---------------------------------------------------------------- onesweep
// Onesweep (Adinets & Merrill, "Onesweep: A Faster Least Significant Digit
// Radix Sort for GPUs", 2022): the digit histograms of every pass come from
// one read of the keys, then each pass is a single kernel.  A tile takes its
// index from a counter (so every earlier tile is already running), ranks its
// keys as the downsweep does, publishes its per-digit counts, and adds up the
// counts of the earlier tiles by decoupled look-back over their published
// status words (each digit's thread reads four predecessors per round).  One
// launch per pass instead of three; used where the passes are latency-bound
// (few tiles).

// Digit histograms of every pass (hist[pass * RADIX + digit]) in one read.
// A workgroup counts tiles blockIdx.x, + gridDim.x, ... into COPIES LDS
// histograms (lane & (COPIES - 1): a digit shared by many keys meets fewer
// same-address atomics) and adds them into hist once, so the global atomics
// per address are the grid size, not the tile count.
constexpr int HIST_GRID_MAX = 512;
template <typename K, int DBITS>
__global__ __launch_bounds__(SORT_BLOCK) void sort_hist_all_kernel(const K* __restrict__ keys,
                                                                   int64_t n, int key_bits,
                                                                   int passes,
                                                                   uint32_t* __restrict__ hist,
                                                                   const uint32_t* __restrict__ d_n) {
    constexpr int RADIX = 1 << DBITS;
    constexpr int MAXP = (8 * (int)sizeof(K) + DBITS - 1) / DBITS;
    constexpr int ONE = MAXP * RADIX * 4;  // bytes of one histogram copy
    constexpr int COPIES = ONE * 8 <= 32768 ? 8 : ONE * 4 <= 32768 ? 4 : ONE * 2 <= 32768 ? 2 : 1;
    __shared__ uint32_t s_h[COPIES][MAXP * RADIX];
    const unsigned tid = threadIdx.x;
    uint32_t* mine = s_h[lane_id() & (COPIES - 1)];
    for (int i = tid; i < COPIES * MAXP * RADIX; i += SORT_BLOCK) (&s_h[0][0])[i] = 0;
    __syncthreads();
    if (d_n) n = min(n, (int64_t)*d_n);
    for (int64_t base = (int64_t)blockIdx.x * SORT_TILE; base < n; base += (int64_t)gridDim.x * SORT_TILE) {
        K kk[SORT_ITEMS];  // all loads of the tile in flight before the counting
#pragma unroll
        for (int it = 0; it < SORT_ITEMS; ++it) {
            const int64_t idx = base + (int64_t)it * SORT_BLOCK + tid;
            kk[it] = idx < n ? keys[idx] : K(0);
        }
#pragma unroll
        for (int it = 0; it < SORT_ITEMS; ++it) {
            if (base + (int64_t)it * SORT_BLOCK + tid >= n) continue;
            for (int p = 0; p < passes; ++p) {
                const int shift = p * DBITS;
                const int width = min(DBITS, key_bits - shift);
                atomicAdd(&mine[p * RADIX + digit_of(kk[it], shift, (1u << width) - 1u)], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < passes * RADIX; i += SORT_BLOCK) {
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < COPIES; ++k) c += s_h[k][i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// Exclusive count of every digit over the tiles before `tile`, without a
// chain through the earlier tiles' prefixes: a tile publishes its digit
// counts (status word: LB_FLAG_AGG | count) and adds them into its chunk's
// sums (csum word: tiles added << 24 | count sum), so the prefix is the sum
// of the earlier whole chunks (each complete once its SORT_CHUNK tiles have
// added) plus the counts of the earlier tiles of the tile's own chunk.  The
// workgroup reads those rows in batches of LB_BATCH rows per thread, every
// load of a batch in flight before any is used; a word not yet complete is
// read again (bounded, so the kernel always ends).  Thread tid takes digit
// tid % RADIX and rows tid / RADIX, + SORT_BLOCK / RADIX, ...; the result is
// valid on threads tid < RADIX.  Tiles take their index from a counter, so
// every earlier tile's workgroup is already running or done.
constexpr int LB_BATCH = 8;
constexpr uint32_t CSUM_ONE = 1u << 24;
constexpr uint32_t CSUM_MASK = CSUM_ONE - 1u;
constexpr int LB_SPIN_MAX = 1 << 18;

template <int RADIX>
__device__ __forceinline__ uint32_t chunk_lookback(const uint32_t* __restrict__ status,
                                                   int64_t status_words, int64_t tile,
                                                   uint32_t* s_part) {
    static_assert(RADIX <= SORT_BLOCK && SORT_BLOCK % RADIX == 0, "digits per thread <= 1");
    constexpr int H = SORT_BLOCK / RADIX;
    const int tid = (int)threadIdx.x, d = tid % RADIX, h = tid / RADIX;
    const int c = (int)(tile / SORT_CHUNK);
    const int rows = c + (int)(tile % SORT_CHUNK);
    // word offsets from `status` (the region is below 2^32 bytes): chunk row
    // r < c at the chunk sums, else tile SORT_CHUNK * c + (r - c)
    const uint32_t csum0 = (uint32_t)status_words + (uint32_t)d;
    const uint32_t own0 = (uint32_t)(c * SORT_CHUNK - c) * RADIX + (uint32_t)d;
    uint32_t sum = 0;
    for (int r0 = h; r0 < rows; r0 += LB_BATCH * H) {
        uint32_t w[LB_BATCH];
#pragma unroll
        for (int b = 0; b < LB_BATCH; ++b) {
            const int r = r0 + b * H;
            w[b] = r < rows ? lb_load(status + (r < c ? csum0 : own0) + (uint32_t)r * RADIX) : 0u;
        }
        bool done = true;
#pragma unroll
        for (int b = 0; b < LB_BATCH; ++b) {
            const int r = r0 + b * H;
            if (r < c) {
                done &= (w[b] >> 24) == (uint32_t)SORT_CHUNK;
                sum += w[b] & CSUM_MASK;
            } else if (r < rows) {
                done &= (w[b] & LB_FLAG_AGG) != 0u;
                sum += w[b] & LB_MASK;
            }
        }
        if (!done) {
            // some earlier tile has not published yet: take the incomplete
            // words again, one at a time, until they are complete
#pragma unroll
            for (int b = 0; b < LB_BATCH; ++b) {
                const int r = r0 + b * H;
                if (r >= rows) continue;
                const uint32_t* q = status + (r < c ? csum0 : own0) + (uint32_t)r * RADIX;
                uint32_t v = w[b];
                if (r < c) {
                    sum -= v & CSUM_MASK;
                    for (int spin = 0; (v >> 24) != (uint32_t)SORT_CHUNK && spin < LB_SPIN_MAX; ++spin)
                        v = lb_load(q);
                    sum += v & CSUM_MASK;
                } else {
                    sum -= v & LB_MASK;
                    for (int spin = 0; (v & LB_FLAG_AGG) == 0u && spin < LB_SPIN_MAX; ++spin)
                        v = lb_load(q);
                    sum += v & LB_MASK;
                }
            }
        }
    }
    s_part[tid] = sum;
    __syncthreads();
    uint32_t excl = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int k = 0; k < H; ++k) excl += s_part[k * RADIX + tid];
    }
    __syncthreads();
    return excl;
}

template <typename K, int DBITS>
__global__ __launch_bounds__(SORT_BLOCK) void sort_onesweep_kernel(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, int64_t n, int shift, int width,
    const uint32_t* __restrict__ hist, uint32_t* __restrict__ status,
    uint32_t* __restrict__ status_next, int64_t status_words, uint32_t* __restrict__ tile_counter,
    const uint32_t* __restrict__ d_n) {
    constexpr int RADIX = 1 << DBITS;
    static_assert(RADIX <= SORT_BLOCK, "one digit per thread");
    constexpr int DPT = 1;  // digits per thread
    __shared__ uint32_t s_wave[SORT_WAVES][RADIX];
    __shared__ uint32_t s_lstart[RADIX];
    __shared__ uint32_t s_goff[RADIX];
    __shared__ K s_stage[SORT_TILE];
    __shared__ uint32_t s_tmp[SORT_WAVES];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_part[SORT_BLOCK];  // look-back partial sums

    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t mask = (1u << width) - 1u;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < SORT_WAVES * RADIX; i += SORT_BLOCK) (&s_wave[0][0])[i] = 0;
    __syncthreads();
    const int64_t tile = s_tile;
    // each look-back region: status words, then the chunk sums
    uint32_t* csum = status + status_words;
    // the next pass's status words of this tile index, and its chunk's sums
    // (cleared by the chunk's first tile), start at zero
    if (status_next) {
        for (int d = tid; d < RADIX; d += SORT_BLOCK) status_next[tile * RADIX + d] = 0u;
        if (tile % SORT_CHUNK == 0)
            for (int d = tid; d < RADIX; d += SORT_BLOCK)
                status_next[status_words + tile / SORT_CHUNK * RADIX + d] = 0u;
    }
    if (d_n) n = min(n, (int64_t)*d_n);
    const int64_t base = tile * SORT_TILE;
    if (base >= n) return;  // uniform: tiles past the count

    // the pass's digit totals, fetched beside the keys
    uint32_t tot[DPT];
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        tot[k] = d < RADIX ? hist[d] : 0u;
    }
    K key[SORT_ITEMS];
    uint32_t val[SORT_ITEMS];
    uint32_t off[SORT_ITEMS];
    const int64_t wbase = base + (int64_t)wave * 64 * SORT_ITEMS;
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const int64_t idx = wbase + it * 64 + lane;
        const bool ok = idx < n;
        key[it] = ok ? keys_in[idx] : K(0);
        val[it] = ok ? vals_in[idx] : 0u;
    }
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const bool ok = wbase + it * 64 + lane < n;
        const uint32_t d = digit_of(key[it], shift, mask);
        const uint64_t peers = match_digit<DBITS>(d, ok);
        const uint32_t rank = popc_below(peers);
        const uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t pre = 0;
        if (ok) pre = s_wave[wave][d];
        off[it] = pre + rank;
        if (ok && rank == 0) s_wave[wave][d] = pre + cnt;
    }
    __syncthreads();

    uint32_t cnt[DPT];
    uint32_t my_total = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        uint32_t sum = 0;
        if (d < RADIX) {
#pragma unroll
            for (int w = 0; w < SORT_WAVES; ++w) {
                const uint32_t c = s_wave[w][d];
                s_wave[w][d] = sum;
                sum += c;
            }
            // publish this tile's counts for the later tiles of its chunk and
            // add them into the chunk's sums for the later chunks
            lb_store(&status[tile * RADIX + d], LB_FLAG_AGG | sum);
            __hip_atomic_fetch_add(&csum[tile / SORT_CHUNK * RADIX + d], CSUM_ONE + sum,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        cnt[k] = sum;
        my_total += sum;
    }
    uint32_t tile_total;
    uint32_t lstart = block_excl_scan<SORT_WAVES>(my_total, s_tmp, &tile_total);
    (void)tile_total;
    uint32_t my_tot = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) my_tot += tot[k];
    uint32_t all_total;
    uint32_t dstart = block_excl_scan<SORT_WAVES>(my_tot, s_tmp, &all_total);
    (void)all_total;
    // tile-local digit starts; the keys are staged in digit order before the
    // look-back, so their registers are free while it waits
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        if (d < RADIX) s_lstart[d] = lstart;
    }
    __syncthreads();
    uint32_t pos[SORT_ITEMS];
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const uint32_t d = digit_of(key[it], shift, mask);
        pos[it] = s_lstart[d] + s_wave[wave][d] + off[it];
    }
    const int count = (int)min((int64_t)SORT_TILE, n - base);
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it)
        if (wbase + it * 64 + lane < n) s_stage[pos[it]] = key[it];
    const uint32_t excl = chunk_lookback<RADIX>(status, status_words, tile, s_part);
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        if (d < RADIX) {
            s_goff[d] = dstart + excl - lstart;
            lstart += cnt[k];
            dstart += tot[k];
        }
    }
    __syncthreads();
    uint32_t gpos[SORT_ITEMS];
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const int j = it * SORT_BLOCK + (int)tid;
        if (j < count) {
            const K k = s_stage[j];
            gpos[it] = s_goff[digit_of(k, shift, mask)] + (uint32_t)j;
            keys_out[gpos[it]] = k;
        }
    }
    __syncthreads();
    uint32_t* s_vstage = reinterpret_cast<uint32_t*>(s_stage);
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it)
        if (wbase + it * 64 + lane < n) s_vstage[pos[it]] = val[it];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const int j = it * SORT_BLOCK + (int)tid;
        if (j < count) vals_out[gpos[it]] = s_vstage[j];
    }
}

// ------------------------------------------------------------ small sorts
// Up to SMALL_N pairs: one workgroup of 1024 threads holds 8 pairs per thread
// (4 for up to 4096 pairs) in registers and runs every pass in LDS (rank by wave64 ballot matching,
// per-digit prefixes over the 16 waves and over the digits, scatter to the
// LDS copy in digit order, read back in index order), then writes the
// result once.  One launch instead of three per pass: the tail bounces of a
// frame have a few thousand rays, where the multi-kernel passes are bound by
// launch and memory latency, not bandwidth.
constexpr int SMALL_BLOCK = 1024;
constexpr int SMALL_WAVES = SMALL_BLOCK / 64;
constexpr int SMALL_N = SMALL_BLOCK * 8;

// Block-wide (1024 threads) exclusive scan of one value per thread.
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t* s_tmp) {
    const unsigned lane = lane_id();
    const unsigned wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (unsigned)off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t wbase = 0;
#pragma unroll
    for (int w = 0; w < SMALL_WAVES; ++w)
        if (w < (int)wave) wbase += s_tmp[w];
    __syncthreads();
    return wbase + x - v;
}

template <typename K, int DBITS, int SMALL_ITEMS>
__global__ __launch_bounds__(SMALL_BLOCK) void sort_small_kernel(K* __restrict__ keys,
                                                                 uint32_t* __restrict__ vals,
                                                                 int64_t n, int key_bits,
                                                                 const uint32_t* __restrict__ d_n) {
    constexpr int RADIX = 1 << DBITS;
    constexpr int CAP = SMALL_BLOCK * SMALL_ITEMS;
    static_assert(RADIX <= SMALL_BLOCK, "one thread per digit");
    __shared__ uint32_t s_wave[SMALL_WAVES][RADIX];  // per-wave counts -> wave prefixes
    __shared__ uint32_t s_start[RADIX];              // digit starts in the block
    __shared__ K s_key[CAP];
    __shared__ uint32_t s_val[CAP];
    __shared__ uint32_t s_tmp[SMALL_WAVES];
    if (d_n) n = min(n, (int64_t)*d_n);
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    // Wave-striped layout: element e = wave * 64 * ITEMS + it * 64 + lane,
    // so (wave, item, lane) order is index order and the ranking is stable.
    const int wbase = (int)wave * 64 * SMALL_ITEMS;
    K key[SMALL_ITEMS];
    uint32_t val[SMALL_ITEMS];
#pragma unroll
    for (int it = 0; it < SMALL_ITEMS; ++it) {
        const int e = wbase + it * 64 + (int)lane;
        const bool ok = e < n;
        key[it] = ok ? keys[e] : K(0);
        val[it] = ok ? vals[e] : 0u;
    }
    const int passes = (key_bits + DBITS - 1) / DBITS;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = pass * DBITS;
        const int width = min(DBITS, key_bits - shift);
        const uint32_t mask = (1u << width) - 1u;
        for (int i = tid; i < SMALL_WAVES * RADIX; i += SMALL_BLOCK) (&s_wave[0][0])[i] = 0;
        __syncthreads();
        uint32_t off[SMALL_ITEMS];
#pragma unroll
        for (int it = 0; it < SMALL_ITEMS; ++it) {
            const bool ok = wbase + it * 64 + (int)lane < n;
            const uint32_t d = digit_of(key[it], shift, mask);
            const uint64_t peers = match_digit<DBITS>(d, ok);
            const uint32_t rank = popc_below(peers);
            const uint32_t cnt = (uint32_t)__popcll(peers);
            uint32_t pre = 0;
            if (ok) pre = s_wave[wave][d];
            off[it] = pre + rank;
            if (ok && rank == 0) s_wave[wave][d] = pre + cnt;
        }
        __syncthreads();
        uint32_t total = 0;
        if (tid < RADIX) {
#pragma unroll
            for (int w = 0; w < SMALL_WAVES; ++w) {
                const uint32_t c = s_wave[w][tid];
                s_wave[w][tid] = total;
                total += c;
            }
        }
        const uint32_t start = block_excl_scan_1024(total, s_tmp);
        if (tid < RADIX) s_start[tid] = start;
        __syncthreads();
#pragma unroll
        for (int it = 0; it < SMALL_ITEMS; ++it) {
            if (wbase + it * 64 + (int)lane < n) {
                const uint32_t d = digit_of(key[it], shift, mask);
                const uint32_t pos = s_start[d] + s_wave[wave][d] + off[it];
                s_key[pos] = key[it];
                s_val[pos] = val[it];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < SMALL_ITEMS; ++it) {
            const int e = wbase + it * 64 + (int)lane;
            if (e < n) {
                key[it] = s_key[e];
                val[it] = s_val[e];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int it = 0; it < SMALL_ITEMS; ++it) {
        const int e = wbase + it * 64 + (int)lane;
        if (e < n) {
            keys[e] = key[it];
            vals[e] = val[it];
        }
    }
}

// ------------------------------------------------------------------- rank
// Rank sort for small pair counts.  A stable sort puts element i at
//   #{j : key_j < key_i} + #{j < i : key_j == key_i}
// (keys compared on their low key_bits bits, as the digit passes do), so
// that position is counted directly: n^2 compares in one pass, spread over
// the whole GPU, where the digit passes of a few thousand pairs are each
// bound by launch and memory latency (several passes) or by one CU (the
// single-workgroup sort).  A workgroup takes 64 elements, one per lane;
// its 16 waves split the compared keys, which are wave-uniform (scalar
// loads), and add their partial counts in LDS.
constexpr int RANK_BLOCK = 1024;
constexpr int RANK_WAVES = RANK_BLOCK / 64;

// Keys of [j0, j1) below k (LE: at most k), counted per lane.
template <typename K, bool LE>
__device__ __forceinline__ uint32_t rank_count(const K* __restrict__ keys, int64_t j0,
                                               int64_t j1, K k, K kmask) {
    constexpr int U = 16;  // keys per block: two 64-byte scalar loads, one wait
    uint32_t c = 0;
    int64_t j = j0;
    for (; j + U <= j1; j += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const K kj = keys[j + u] & kmask;
            c += LE ? (kj <= k) : (kj < k);
        }
    }
    for (; j < j1; ++j) {
        const K kj = keys[j] & kmask;
        c += LE ? (kj <= k) : (kj < k);
    }
    return c;
}

template <typename K>
__global__ __launch_bounds__(RANK_BLOCK) void sort_rank_kernel(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, int64_t n, int key_bits, const uint32_t* __restrict__ d_n) {
    __shared__ uint32_t s_cnt[RANK_WAVES][64];
    if (d_n) n = min(n, (int64_t)*d_n);
    const int64_t e0 = (int64_t)blockIdx.x * 64;
    if (e0 >= n) return;  // uniform
    const unsigned lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const K kmask = key_bits >= (int)(8 * sizeof(K)) ? ~K(0) : (K(1) << key_bits) - K(1);
    const int64_t e = e0 + lane;
    const bool ok = e < n;
    const K kraw = ok ? keys_in[e] : K(0);
    const K k = kraw & kmask;
    // this wave's keys [j0, j1): those before the workgroup's elements
    // count when at most k, those after when below k, and the 64 of the
    // workgroup by the full rule
    const int64_t per = (n + RANK_WAVES - 1) / RANK_WAVES;
    const int64_t j0 = min(n, (int64_t)wave * per), j1 = min(n, j0 + per);
    const int64_t m0 = max(j0, e0), m1 = min(j1, e0 + 64);
    uint32_t c = rank_count<K, true>(keys_in, j0, min(j1, e0), k, kmask) +
                 rank_count<K, false>(keys_in, max(j0, e0 + 64), j1, k, kmask);
    for (int64_t j = m0; j < m1; ++j) {
        const K kj = keys_in[j] & kmask;
        c += (kj < k) || (kj == k && j < e);
    }
    s_cnt[wave][lane] = c;
    __syncthreads();
    if (wave == 0 && ok) {
        uint32_t r = 0;
#pragma unroll
        for (int w = 0; w < RANK_WAVES; ++w) r += s_cnt[w][lane];
        keys_out[r] = kraw;
        vals_out[r] = vals_in[e];
    }
}

template <typename K>
__global__ __launch_bounds__(256) void copy_pairs_kernel(const K* __restrict__ ks,
                                                         const uint32_t* __restrict__ vs,
                                                         K* __restrict__ kd,
                                                         uint32_t* __restrict__ vd, int64_t n,
                                                         const uint32_t* __restrict__ d_n) {
    if (d_n) n = min(n, (int64_t)*d_n);
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

template <typename K, int DBITS>
static int launch_pass(const void* ki, const uint32_t* vi, void* ko, uint32_t* vo, int64_t n,
                       int shift, int width, uint32_t* counts, uint32_t* totals, int64_t tiles,
                       const uint32_t* d_n, hipStream_t s) {
    constexpr int RADIX = 1 << DBITS;
    sort_upsweep_kernel<K, DBITS><<<(unsigned)tiles, SORT_BLOCK, 0, s>>>(
        static_cast<const K*>(ki), n, shift, width, counts, d_n);
    WF_TRY(launch_check("sort_upsweep_kernel"));
    sort_spine_kernel<RADIX><<<RADIX, SPINE_BLOCK, 0, s>>>(counts, tiles, totals);
    WF_TRY(launch_check("sort_spine_kernel"));
    sort_bin_kernel<K, DBITS><<<(unsigned)tiles, SORT_BLOCK, 0, s>>>(
        static_cast<const K*>(ki), vi, static_cast<K*>(ko), vo, n, shift, width, counts, totals,
        d_n);
    return launch_check("sort_bin_kernel");
}

template <typename K>
static int sort_impl(void* keys, uint32_t* vals, int64_t n, int key_bits, int digit_bits,
                     char* scratch, const SortPlan& p, const uint32_t* d_n, hipStream_t s) {
    uint32_t* counts = reinterpret_cast<uint32_t*>(scratch + p.off_counts);
    uint32_t* totals = reinterpret_cast<uint32_t*>(scratch + p.off_totals);
    void* ka = scratch + p.off_keys_a;
    uint32_t* va = reinterpret_cast<uint32_t*>(scratch + p.off_vals_a);
    void* kb = scratch + p.off_keys_b;
    uint32_t* vb = reinterpret_cast<uint32_t*>(scratch + p.off_vals_b);
    (void)key_bits;

    // Buffer sequence: starts and ends in the caller's buffers, never reads
    // and writes the same buffer in one pass.
    struct Buf {
        void* k;
        uint32_t* v;
    };
    Buf C{keys, vals}, A{ka, va}, B{kb, vb};
    Buf seq[SORT_MAX_PASSES + 2];
    int len = 0;
    seq[len++] = C;
    if (p.passes == 1) {
        seq[len++] = A;
    } else if (p.passes % 2 == 0) {
        for (int i = 0; i < p.passes; ++i) seq[len++] = (i % 2 == 0) ? A : C;
    } else {
        seq[len++] = A;
        seq[len++] = B;
        seq[len++] = C;
        for (int i = 3; i < p.passes; ++i) seq[len++] = (i % 2 == 1) ? A : C;
    }

    // onesweep up to sort.onesweep_tiles tiles, with digits of at most 8 bits
    // (its look-back takes one digit per thread)
    const bool onesweep = p.tiles <= tuning().sort_onesweep_tiles && digit_bits <= 8;
    uint32_t* hist = reinterpret_cast<uint32_t*>(scratch + p.off_hist);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch + p.off_ctr);
    uint32_t* status = reinterpret_cast<uint32_t*>(scratch + p.off_status);
    if (onesweep) {
        WF_TRY(fill_u32(hist, p.clear_words, 0u, s));
        const unsigned hg = (unsigned)std::min<int64_t>(p.tiles, HIST_GRID_MAX);
        if (digit_bits == 4)
            sort_hist_all_kernel<K, 4><<<hg, SORT_BLOCK, 0, s>>>(static_cast<const K*>(keys), n, key_bits, p.passes, hist, d_n);
        else
            sort_hist_all_kernel<K, 8><<<hg, SORT_BLOCK, 0, s>>>(static_cast<const K*>(keys), n, key_bits, p.passes, hist, d_n);
        WF_TRY(launch_check("sort_hist_all_kernel"));
    }
    for (int pass = 0; pass < p.passes; ++pass) {
        const int shift = pass * digit_bits;
        const int width = std::min(digit_bits, key_bits - shift);
        const Buf& in = seq[pass];
        const Buf& out = seq[pass + 1];
        if (onesweep) {
            uint32_t* st = status + (pass & 1) * p.region_words;
            uint32_t* st_next = pass + 1 < p.passes ? status + ((pass + 1) & 1) * p.region_words : nullptr;
            const uint32_t* h = hist + (int64_t)pass * p.radix;
            const unsigned g = (unsigned)p.tiles;
            if (digit_bits == 4)
                sort_onesweep_kernel<K, 4><<<g, SORT_BLOCK, 0, s>>>(static_cast<const K*>(in.k), in.v, static_cast<K*>(out.k), out.v, n, shift, width, h, st, st_next, p.status_words, ctr + pass, d_n);
            else
                sort_onesweep_kernel<K, 8><<<g, SORT_BLOCK, 0, s>>>(static_cast<const K*>(in.k), in.v, static_cast<K*>(out.k), out.v, n, shift, width, h, st, st_next, p.status_words, ctr + pass, d_n);
            WF_TRY(launch_check("sort_onesweep_kernel"));
            continue;
        }
        switch (digit_bits) {
            case 4: WF_TRY((launch_pass<K, 4>(in.k, in.v, out.k, out.v, n, shift, width, counts, totals, p.tiles, d_n, s))); break;
            case 8: WF_TRY((launch_pass<K, 8>(in.k, in.v, out.k, out.v, n, shift, width, counts, totals, p.tiles, d_n, s))); break;
            default: WF_TRY((launch_pass<K, 11>(in.k, in.v, out.k, out.v, n, shift, width, counts, totals, p.tiles, d_n, s))); break;
        }
    }
    if (p.passes == 1) {
        copy_pairs_kernel<K><<<grid_for(n, 256), 256, 0, s>>>(
            static_cast<const K*>(ka), va, static_cast<K*>(keys), vals, n, d_n);
        WF_TRY(launch_check("copy_pairs_kernel"));
    }
    return WF_OK;
}

template <int ITEMS>
static void launch_small(void* keys, uint32_t* vals, int64_t n, int key_bytes, int key_bits,
                         int digit_bits, const uint32_t* d_n, hipStream_t s) {
    if (key_bytes == 4 && digit_bits == 8)
        sort_small_kernel<uint32_t, 8, ITEMS><<<1, SMALL_BLOCK, 0, s>>>((uint32_t*)keys, vals, n, key_bits, d_n);
    else if (key_bytes == 4)
        sort_small_kernel<uint32_t, 4, ITEMS><<<1, SMALL_BLOCK, 0, s>>>((uint32_t*)keys, vals, n, key_bits, d_n);
    else if (digit_bits == 8)
        sort_small_kernel<uint64_t, 8, ITEMS><<<1, SMALL_BLOCK, 0, s>>>((uint64_t*)keys, vals, n, key_bits, d_n);
    else
        sort_small_kernel<uint64_t, 4, ITEMS><<<1, SMALL_BLOCK, 0, s>>>((uint64_t*)keys, vals, n, key_bits, d_n);
}

int sort_pairs_launch(void* keys, uint32_t* vals, int64_t n_cap, const uint32_t* d_n,
                      int key_bytes, int key_bits, int digit_bits, void* scratch,
                      int64_t scratch_bytes, hipStream_t s) {
    SortPlan p;
    WF_CHECK_ARG(make_plan(n_cap, key_bytes, key_bits, digit_bits, p),
                 "sort: key_bytes must be 4 or 8, 1 <= key_bits <= 8*key_bytes, "
                 "digit_bits in {4, 8, 11}, n >= 0");
    if (n_cap > WF_MAX_ELEMENTS) return fail(WF_ERR_TOO_LARGE, "sort: n exceeds 2^30-1");
    if (n_cap <= 1) return WF_OK;
    WF_CHECK_ARG(keys && vals && scratch, "sort: null buffer");
    if (scratch_bytes < p.total)
        return fail(WF_ERR_SCRATCH_TOO_SMALL, "sort: scratch smaller than wf_sort_scratch_bytes");
    // the counted rank costs ~n^2, the digit passes ~passes: the crossover
    // scales with the pass count (28k pairs at 8 passes, 14k at 4)
    if (n_cap <= (int64_t)tuning().sort_rank_max * p.passes / 8) {
        // rank into the first scratch pair, then back into the caller's
        char* sc = static_cast<char*>(scratch);
        void* ka = sc + p.off_keys_a;
        uint32_t* va = reinterpret_cast<uint32_t*>(sc + p.off_vals_a);
        const unsigned g = (unsigned)((n_cap + 63) / 64);
        if (key_bytes == 4) {
            sort_rank_kernel<uint32_t><<<g, RANK_BLOCK, 0, s>>>((const uint32_t*)keys, vals, (uint32_t*)ka, va, n_cap, key_bits, d_n);
            WF_TRY(launch_check("sort_rank_kernel"));
            copy_pairs_kernel<uint32_t><<<grid_for(n_cap, 256), 256, 0, s>>>((const uint32_t*)ka, va, (uint32_t*)keys, vals, n_cap, d_n);
        } else {
            sort_rank_kernel<uint64_t><<<g, RANK_BLOCK, 0, s>>>((const uint64_t*)keys, vals, (uint64_t*)ka, va, n_cap, key_bits, d_n);
            WF_TRY(launch_check("sort_rank_kernel"));
            copy_pairs_kernel<uint64_t><<<grid_for(n_cap, 256), 256, 0, s>>>((const uint64_t*)ka, va, (uint64_t*)keys, vals, n_cap, d_n);
        }
        return launch_check("copy_pairs_kernel");
    }
    if (n_cap <= SMALL_N && digit_bits <= 8) {
        if (n_cap <= SMALL_N / 2)
            launch_small<4>(keys, vals, n_cap, key_bytes, key_bits, digit_bits, d_n, s);
        else
            launch_small<8>(keys, vals, n_cap, key_bytes, key_bits, digit_bits, d_n, s);
        return launch_check("sort_small_kernel");
    }
    return key_bytes == 4
               ? sort_impl<uint32_t>(keys, vals, n_cap, key_bits, digit_bits, (char*)scratch, p, d_n, s)
               : sort_impl<uint64_t>(keys, vals, n_cap, key_bits, digit_bits, (char*)scratch, p, d_n, s);
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int64_t wf_sort_scratch_bytes(int64_t n, int key_bytes, int key_bits, int digit_bits) {
    SortPlan p;
    if (!make_plan(n, key_bytes, key_bits, digit_bits, p)) {
        set_error("wf_sort_scratch_bytes: invalid key_bytes / key_bits / digit_bits / n");
        return WF_ERR_INVALID_ARG;
    }
    return p.total;
}

extern "C" int wf_sort_pairs(void* keys, uint32_t* vals, int64_t n, int key_bytes, int key_bits,
                             int digit_bits, void* scratch, int64_t scratch_bytes,
                             wf_stream_t stream, float* ms) {
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    WF_TRY(sort_pairs_launch(keys, vals, n, nullptr, key_bytes, key_bits, digit_bits, scratch,
                             scratch_bytes, s));
    return t.finish();
}

extern "C" int wf_sort_pairs_count(void* keys, uint32_t* vals, int64_t capacity,
                                   const uint32_t* d_count, int key_bytes, int key_bits,
                                   int digit_bits, void* scratch, int64_t scratch_bytes,
                                   wf_stream_t stream) {
    WF_CHECK_ARG(d_count, "wf_sort_pairs_count: null count");
    return sort_pairs_launch(keys, vals, capacity, d_count, key_bytes, key_bits, digit_bits,
                             scratch, scratch_bytes, as_stream(stream));
}

// ------------------------------------------------------- allocation-free sort
// wf_sort_pairs with its scratch in a file-scope device array instead of a
// caller buffer: for callers that cannot allocate (graph capture, serving
// loops).  Every call on a GPU shares that GPU's copy of the array.
namespace wfpt {
constexpr int64_t NOALLOC_SORT_MAX_PAIRS = int64_t(1) << 20;
// wf_sort_scratch_bytes at 2^20 pairs peaks at 31,777,024 bytes (u64 keys,
// 11-bit digits); the array rounds that up to 32 MiB.
constexpr int64_t NOALLOC_SORT_SCRATCH_BYTES = int64_t(32) << 20;
}  // namespace wfpt

static __device__ __align__(256) unsigned char scratch[wfpt::NOALLOC_SORT_SCRATCH_BYTES];

namespace wfpt {
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mutex;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of this GPU's `scratch` (the array's name is only a
// host-side stand-in): looked up once per device, with the device current.
static int noalloc_scratch(void** out) {
    int dev = -1;
    WF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(WF_ERR_INVALID_ARG, "wf_sort_pairs_noalloc: device index beyond 64");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(WF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return WF_OK;
}
}  // namespace wfpt

extern "C" int64_t wf_sort_noalloc_max_pairs(void) { return NOALLOC_SORT_MAX_PAIRS; }

extern "C" int wf_sort_pairs_noalloc(void* keys, uint32_t* vals, int64_t n, int key_bytes,
                                     int key_bits, int digit_bits, wf_stream_t stream) {
    SortPlan p;
    WF_CHECK_ARG(make_plan(n, key_bytes, key_bits, digit_bits, p),
                 "wf_sort_pairs_noalloc: key_bytes must be 4 or 8, 1 <= key_bits <= 8*key_bytes, "
                 "digit_bits in {4, 8, 11}, n >= 0");
    if (n > NOALLOC_SORT_MAX_PAIRS || p.total > NOALLOC_SORT_SCRATCH_BYTES)
        return fail(WF_ERR_TOO_LARGE, "wf_sort_pairs_noalloc: n = " + std::to_string(n) +
                                          " exceeds the allocation-free limit of " +
                                          std::to_string(NOALLOC_SORT_MAX_PAIRS) + " pairs");
    if (n <= 1) return WF_OK;
    void* sp = nullptr;
    WF_TRY(noalloc_scratch(&sp));
    return sort_pairs_launch(keys, vals, n, nullptr, key_bytes, key_bits, digit_bits, sp,
                             NOALLOC_SORT_SCRATCH_BYTES, as_stream(stream));
}
