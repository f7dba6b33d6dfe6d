// Stable LSD radix sort of (uint64 or uint32 key, uint32 value) pairs and an int32 exclusive scan,
// written for gfx950 wave64.  Used for the Morton order of triangles (tree build) and of
// query points (wave batches of nearby points).
//
// One pass = 8 key bits:
//   hist:    each 256-thread block counts the digits of its 4096-key tile in LDS (the tile's
//            loads are all issued before the first count; a wave whose keys share one digit
//            adds its count with one atomic)
//   rows:    one block per digit scans that digit's per-block counts (digit-major layout) in
//            chunks of 4,096 staged through LDS
//   digits:  one block scans the 256 digit totals
//   scatter: each block loads its tile into registers; each wave ranks its quarter of the tile
//            in 16 rounds of 64 keys (one compare when the wave shares a digit, else 8 ballots,
//            give a key's rank among the wave's equal digits, wave-private LDS run counters its
//            place in the wave's run), one block step
//            turns the waves' runs into stable tile offsets, the tile is placed in LDS in digit
//            order and written out in coalesced runs.
#include "fwn_internal.h"

namespace fwn {

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
constexpr int RS_WAVES = RS_THREADS / kWave;
typedef __attribute__((address_space(3))) volatile int32_t lds_vint;

// lanes of the wave holding the same 8-bit digit d as this lane (among the lanes in `live`, a
// prefix of the wave).  Morton keys of nearby points often share a digit across the whole wave:
// one compare settles that case; otherwise 8 bit ballots separate the digits.  Called with every
// lane of the wave active.
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, uint64_t live)
{
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if ((__ballot(d != d0) & live) == 0ull) return live;
    uint64_t peers = live;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

template <typename KeyT>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const KeyT* __restrict__ keys, int64_t n, int shift,
                                                        int32_t* __restrict__ hist, int nblocks)
{
    __shared__ int32_t cnt[256];
    cnt[threadIdx.x] = 0;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int lane = threadIdx.x & (kWave - 1);
    // all of the thread's loads are issued before the first count (one memory latency per tile)
    KeyT kr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        kr[r] = (i < n) ? keys[i] : (KeyT)0;
    }
    __syncthreads();
    // a wave whose keys share one digit adds its count with one atomic (LDS atomics on one
    // address serialise); other waves count key by key
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        const bool valid = i < n;
        const uint32_t d = (uint32_t)(kr[r] >> shift) & 255u;
        const uint64_t live = __ballot(valid);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        if ((__ballot(d != d0) & live) == 0ull) {
            if (lane == 0 && live != 0ull) atomicAdd(&cnt[d0], (int32_t)__popcll(live));
        } else if (valid) {
            atomicAdd(&cnt[d], 1);
        }
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// exclusive scan of one digit's row of per-block counts; row total -> rowsum[d].  The row is
// taken in chunks of 4,096 counts: loaded into LDS (coalesced), each thread sums 16 consecutive
// counts, one block scan of the 256 sums, then each thread writes its 16 offsets.
constexpr int RR_PER = 16;
constexpr int RR_CHUNK = 256 * RR_PER;
__device__ __forceinline__ int rr_pad(int i) { return i + i / RR_PER; } // odd stride: no bank conflicts

__global__ __launch_bounds__(256) void k_rs_rows(int32_t* __restrict__ hist, int nblocks,
                                                 int32_t* __restrict__ rowsum)
{
    __shared__ int32_t buf[RR_CHUNK + RR_CHUNK / RR_PER];
    __shared__ int32_t part[256];
    const int tid = threadIdx.x;
    int32_t* row = hist + (int64_t)blockIdx.x * nblocks;
    int32_t carry = 0;
    for (int base = 0; base < nblocks; base += RR_CHUNK) {
#pragma unroll
        for (int k = 0; k < RR_PER; ++k) {
            const int i = k * 256 + tid;
            buf[rr_pad(i)] = (base + i < nblocks) ? row[base + i] : 0;
        }
        __syncthreads();
        int32_t sum = 0;
#pragma unroll
        for (int k = 0; k < RR_PER; ++k) sum += buf[rr_pad(tid * RR_PER + k)];
        part[tid] = sum;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int32_t t = (tid >= off) ? part[tid - off] : 0;
            __syncthreads();
            part[tid] += t;
            __syncthreads();
        }
        int32_t run = carry + part[tid] - sum;
#pragma unroll
        for (int k = 0; k < RR_PER; ++k) {
            const int j = rr_pad(tid * RR_PER + k);
            const int32_t v = buf[j];
            buf[j] = run;
            run += v;
        }
        carry += part[255];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RR_PER; ++k) {
            const int i = k * 256 + tid;
            if (base + i < nblocks) row[base + i] = buf[rr_pad(i)];
        }
        __syncthreads();
    }
    if (tid == 0) rowsum[blockIdx.x] = carry;
}

// Wave w of the block owns keys [w * 1024, (w + 1) * 1024) of the tile.  Each wave ranks its
// keys on its own (no block barrier per round): per round of 64 keys the 8 digit ballots give a
// key's rank among the wave's equal digits, and wave-private run counters in LDS give its place
// in the wave's run of that digit.  One block step then turns the four waves' run lengths into
// tile offsets (digit-major, wave order inside a digit: stable), the keys are placed in LDS in
// digit order, and the tile is written out in coalesced runs: consecutive threads write
// consecutive positions of a digit's run.  The output is the same as writing each key straight
// to its position.
template <typename KeyT>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const KeyT* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           KeyT* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           int64_t n, int shift, const int32_t* __restrict__ hist,
                                                           const int32_t* __restrict__ digit_total, int nblocks)
{
    constexpr int kWaveKeys = RS_TILE / RS_WAVES; // 1024
    __shared__ int32_t gbase[256];          // global position of this tile's first key of each digit
    __shared__ int32_t toff[256];           // tile-local start of each digit's run
    __shared__ int32_t wrun[RS_WAVES][256]; // per wave: keys of each digit (run length, then offset)
    __shared__ KeyT lk[RS_TILE];
    __shared__ uint32_t lv[RS_TILE];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int tile_n = (int)((n - base) < RS_TILE ? (n - base) : RS_TILE);
    // the tile's keys and values are loaded into registers first (the loads overlap the offset
    // scan below), so the block waits for memory once rather than once per round
    KeyT kr[RS_ROUNDS];
    uint32_t vr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int li = wave * kWaveKeys + r * kWave + lane;
        kr[r] = (li < tile_n) ? kin[base + li] : (KeyT)0;
    }
    // vin == nullptr: the values are the input positions (a first pass over unsorted keys); the
    // test is outside the loops so that no round waits for the loads of the others
    if (vin) {
#pragma unroll
        for (int r = 0; r < RS_ROUNDS; ++r) {
            const int li = wave * kWaveKeys + r * kWave + lane;
            vr[r] = (li < tile_n) ? vin[base + li] : 0u;
        }
    } else {
#pragma unroll
        for (int r = 0; r < RS_ROUNDS; ++r) vr[r] = (uint32_t)(base + wave * kWaveKeys + r * kWave + lane);
    }
    // this tile's count of digit tid: the next block's exclusive offset (or the digit total) minus ours
    const int32_t tot = digit_total[tid];
    const int32_t mine = hist[(int64_t)tid * nblocks + blockIdx.x];
    const int32_t next = (blockIdx.x + 1 < nblocks) ? hist[(int64_t)tid * nblocks + blockIdx.x + 1] : tot;
    const int32_t cnt = next - mine;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) wrun[w][tid] = 0;
    // exclusive scans (Hillis-Steele in LDS) of the tile's digit counts and of the digit totals
    // (every tile scans the 256 totals itself: no separate launch)
    toff[tid] = cnt;
    gbase[tid] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int32_t t = (tid >= off) ? toff[tid - off] : 0;
        const int32_t u = (tid >= off) ? gbase[tid - off] : 0;
        __syncthreads();
        toff[tid] += t;
        gbase[tid] += u;
        __syncthreads();
    }
    toff[tid] -= cnt;
    gbase[tid] += mine - tot; // keys of smaller digits, then this digit's keys in earlier tiles
    // each wave: the place of every key in the wave's run of its digit
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // volatile: each round reloads the counters other lanes of the wave wrote in earlier rounds
    // (an LDS pointer, so that the accesses are LDS instructions rather than flat ones)
    lds_vint* run_of = (lds_vint*)wrun[wave];
    int32_t lp[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const bool valid = wave * kWaveKeys + r * kWave + lane < tile_n;
        const uint32_t d = (uint32_t)(kr[r] >> shift) & 255u;
        const uint64_t peers = digit_peers(d, __ballot(valid));
        const int rank = __popcll(peers & lt);
        const int32_t run = valid ? run_of[d] : 0;
        lp[r] = run + rank;
        // the wave's LDS operations complete in order, and the write needs the value read:
        // every peer has read the counter before the lowest peer writes it
        if (valid && rank == 0) run_of[d] = run + (int32_t)__popcll(peers);
    }
    __syncthreads();
    // run lengths -> tile offsets of each wave's run of digit tid
    {
        int32_t at = toff[tid];
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            const int32_t len = wrun[w][tid];
            wrun[w][tid] = at;
            at += len;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const bool valid = wave * kWaveKeys + r * kWave + lane < tile_n;
        if (valid) {
            const uint32_t d = (uint32_t)(kr[r] >> shift) & 255u;
            const int32_t pos = wrun[wave][d] + lp[r];
            lk[pos] = kr[r];
            lv[pos] = vr[r];
        }
    }
    __syncthreads();
    for (int t = tid; t < tile_n; t += RS_THREADS) {
        const KeyT key = lk[t];
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        const int32_t pos = gbase[d] + (t - toff[d]);
        kout[pos] = key;
        vout[pos] = lv[t];
    }
}

constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 4;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

__device__ int32_t block_excl_scan(int32_t v, int32_t* part, int32_t& total)
{
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < SC_THREADS; off <<= 1) {
        int32_t t = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    total = part[SC_THREADS - 1];
    int32_t r = part[threadIdx.x] - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_tiles(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                           int64_t n, int32_t* __restrict__ tile_sums)
{
    __shared__ int32_t part[SC_THREADS];
    int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_ITEMS;
    int32_t v[SC_ITEMS];
    int32_t s = 0;
    for (int k = 0; k < SC_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int32_t total;
    int32_t pre = block_excl_scan(s, part, total);
    for (int k = 0; k < SC_ITEMS; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_sums(int32_t* __restrict__ sums, int64_t n)
{
    __shared__ int32_t part[SC_THREADS];
    int32_t carry = 0;
    for (int64_t base = 0; base < n; base += SC_THREADS) {
        int64_t i = base + threadIdx.x;
        int32_t v = (i < n) ? sums[i] : 0;
        int32_t total;
        int32_t pre = block_excl_scan(v, part, total);
        if (i < n) sums[i] = carry + pre;
        carry += total;
    }
    if (threadIdx.x == 0) sums[n] = carry;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_add(int32_t* __restrict__ out, int64_t n,
                                                         const int32_t* __restrict__ tile_sums)
{
    int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_ITEMS;
    int32_t add = tile_sums[blockIdx.x];
    for (int k = 0; k < SC_ITEMS; ++k)
        if (base + k < n) out[base + k] += add;
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_sums[gridDim.x];
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// the passes over caller-provided histogram scratch: hist holds 256 x ceil(n / RS_TILE) and
// rowsum 256 int32
template <typename KeyT>
void radix_sort_core(KeyT*& keys, uint32_t*& vals, KeyT*& keys_alt, uint32_t*& vals_alt, int64_t n, int begin_bit,
                     int end_bit, bool vals_are_index, int32_t* hist, int32_t* rowsum, hipStream_t s)
{
    if (n <= 1 || begin_bit >= end_bit) {
        if (n >= 1 && vals_are_index) {
            hipLaunchKernelGGL(k_iota, dim3(grid_blocks(n, 256)), dim3(256), 0, s, vals, n);
            FWN_LAUNCH_CHECK("k_iota");
        }
        return;
    }
    if (n > (int64_t)INT32_MAX) fail(FWN_LIMIT_EXCEEDED, "radix sort: more than 2^31-1 elements");
    const int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist<KeyT>, dim3(nblocks), dim3(RS_THREADS), 0, s, keys, n, shift, hist, nblocks);
        FWN_LAUNCH_CHECK("k_rs_hist");
        hipLaunchKernelGGL(k_rs_rows, dim3(256), dim3(256), 0, s, hist, nblocks, rowsum);
        FWN_LAUNCH_CHECK("k_rs_rows");
        const uint32_t* vin = (vals_are_index && shift == begin_bit) ? nullptr : vals;
        hipLaunchKernelGGL(k_rs_scatter<KeyT>, dim3(nblocks), dim3(RS_THREADS), 0, s, keys, vin, keys_alt,
                           vals_alt, n, shift, hist, rowsum, nblocks);
        FWN_LAUNCH_CHECK("k_rs_scatter");
        std::swap(keys, keys_alt);
        std::swap(vals, vals_alt);
    }
}

template <typename KeyT>
void radix_sort_impl(KeyT*& keys, uint32_t*& vals, KeyT*& keys_alt, uint32_t*& vals_alt, int64_t n, int begin_bit,
                     int end_bit, bool vals_are_index, DBuf& hist, DBuf& rowsum, hipStream_t s)
{
    if (n > 1 && begin_bit < end_bit) {
        const int64_t nblocks = (n + RS_TILE - 1) / RS_TILE;
        hist.reserve(sizeof(int32_t) * 256 * (size_t)nblocks);
        rowsum.reserve(sizeof(int32_t) * 256);
    }
    radix_sort_core<KeyT>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, vals_are_index, hist.as<int32_t>(),
                          rowsum.as<int32_t>(), s);
}

} // namespace

void radix_sort_pairs(uint64_t*& keys, uint32_t*& vals, uint64_t*& keys_alt, uint32_t*& vals_alt, int64_t n,
                      int nbits, DBuf& hist, DBuf& rowsum, hipStream_t s)
{
    radix_sort_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, 0, nbits, false, hist, rowsum, s);
}

void radix_sort_pairs32(uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt, int64_t n,
                        int begin_bit, int end_bit, bool vals_are_index, DBuf& hist, DBuf& rowsum, hipStream_t s)
{
    radix_sort_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, vals_are_index, hist, rowsum, s);
}

int64_t radix_sort_hist_words(int64_t n)
{
    return 256 * ((n + RS_TILE - 1) / RS_TILE);
}

void radix_sort_pairs32_scratch(uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt, int64_t n,
                                int begin_bit, int end_bit, bool vals_are_index, int32_t* hist, int32_t* rowsum,
                                hipStream_t s)
{
    radix_sort_core<uint32_t>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, vals_are_index, hist, rowsum, s);
}

// out must hold n+1 entries; out[n] receives the total.
void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, DBuf& tmp, hipStream_t s)
{
    if (n <= 0) {
        FWN_HIP(hipMemsetAsync(out, 0, sizeof(int32_t), s));
        return;
    }
    int64_t ntiles = (n + SC_TILE - 1) / SC_TILE;
    tmp.reserve(sizeof(int32_t) * (size_t)(ntiles + 1));
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(SC_THREADS), 0, s, in, out, n, tmp.as<int32_t>());
    FWN_LAUNCH_CHECK("k_scan_tiles");
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_THREADS), 0, s, tmp.as<int32_t>(), ntiles);
    FWN_LAUNCH_CHECK("k_scan_sums");
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)ntiles), dim3(SC_THREADS), 0, s, out, n, tmp.as<int32_t>());
    FWN_LAUNCH_CHECK("k_scan_add");
}

} // namespace fwn
