// Device-wide exclusive scan and stable LSD radix sort for gfx950 (wave64).
//
// The radix sort replaces thrust::sort_by_key of the reference
// (NeighborGrid.cu:139-145 for cell keys; AdaptiveAlgorithm.cuh:64-128 for
// criterion values). It is stable, so equal keys keep ascending input order,
// which is the order the reference's radix-backed sort produces.
//
// One pass over a RADIX_BITS digit is three launches:
//   upsweep   - per-tile digit histogram (LDS atomics), written digit-major;
//   scan      - one block per digit scans that digit's tile counts;
//   downsweep - per-tile stable ranking with wave ballots (peers of equal
//               digit found with RADIX_BITS ballots), wave-private digit
//               counters in LDS, a block prefix over waves and digits, then a
//               scatter staged through LDS so that each digit's run is written
//               with consecutive addresses.
// Element counts may live in device memory (pointer `nPtr`), so the launches
// are sized by capacity and can be captured in a graph.
#include "kernels.hpp"

namespace asph
{

namespace
{

constexpr int kBlock = 256;
#ifndef ASPH_SORT_ITEMS
#define ASPH_SORT_ITEMS 8
#endif
constexpr int kItems = ASPH_SORT_ITEMS;  // elements per thread of a tile
constexpr int kTile = kBlock * kItems;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ uint32_t loadN(const uint32_t* nPtr, uint32_t n)
{
    return nPtr != nullptr ? *nPtr : n;
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the
// exclusive prefix and writes the block total to *total.
__device__ uint32_t blockExclusiveScan(uint32_t v, uint32_t* lds, uint32_t* total)
{
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    // wave inclusive scan via shuffles
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
        {
            x += y;
        }
    }
    if (lane == 63)
    {
        lds[wave] = x;
    }
    __syncthreads();
    uint32_t waveBase = 0;
    uint32_t sum = 0;
    for (int w = 0; w < kWaves; ++w)
    {
        const uint32_t t = lds[w];
        if (w < wave)
        {
            waveBase += t;
        }
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return waveBase + x - v;
}

// ------------------------------------------------------------------ scan --

__global__ __launch_bounds__(kBlock) void kScanReduce(const uint32_t* in, const uint32_t* nPtr, uint32_t n,
                                                      uint32_t* tileSums)
{
    __shared__ uint32_t lds[kWaves];
    const uint32_t count = loadN(nPtr, n);
    const uint32_t base = blockIdx.x * static_cast<uint32_t>(kTile);
    uint32_t s = 0;
    if (base < count)
    {
        for (int k = 0; k < kItems; ++k)
        {
            const uint32_t e = base + static_cast<uint32_t>(k * kBlock) + threadIdx.x;
            if (e < count)
            {
                s += in[e];
            }
        }
    }
    uint32_t total = 0;
    (void)blockExclusiveScan(s, lds, &total);
    if (threadIdx.x == 0)
    {
        tileSums[blockIdx.x] = total;
    }
}

// Single block: exclusive scan of tileSums[0, tiles) in place; grand total to *total.
__global__ __launch_bounds__(kBlock) void kScanTiles(uint32_t* tileSums, uint32_t tiles, uint32_t* total)
{
    __shared__ uint32_t lds[kWaves];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < tiles; b += kBlock)
    {
        const uint32_t e = b + threadIdx.x;
        const uint32_t v = e < tiles ? tileSums[e] : 0U;
        uint32_t t = 0;
        const uint32_t ex = blockExclusiveScan(v, lds, &t);
        if (e < tiles)
        {
            tileSums[e] = carry + ex;
        }
        carry += t;
    }
    if (threadIdx.x == 0 && total != nullptr)
    {
        *total = carry;
    }
}

// Exclusive scan of each tile plus its tile offset; optionally zeroes the input.
__global__ __launch_bounds__(kBlock) void kScanDown(uint32_t* in, const uint32_t* nPtr, uint32_t n,
                                                    const uint32_t* tileSums, uint32_t* out, int zeroInput)
{
    __shared__ uint32_t lds[kWaves];
    const uint32_t count = loadN(nPtr, n);
    const uint32_t base = blockIdx.x * static_cast<uint32_t>(kTile);
    if (base >= count)
    {
        return;
    }
    // blocked arrangement: thread t owns kItems consecutive elements
    const uint32_t first = base + threadIdx.x * static_cast<uint32_t>(kItems);
    uint32_t vals[kItems];
    uint32_t s = 0;
    for (int k = 0; k < kItems; ++k)
    {
        const uint32_t e = first + static_cast<uint32_t>(k);
        vals[k] = e < count ? in[e] : 0U;
        s += vals[k];
    }
    uint32_t total = 0;
    uint32_t run = blockExclusiveScan(s, lds, &total) + tileSums[blockIdx.x];
    for (int k = 0; k < kItems; ++k)
    {
        const uint32_t e = first + static_cast<uint32_t>(k);
        if (e < count)
        {
            out[e] = run;
            if (zeroInput != 0)
            {
                in[e] = 0U;
            }
        }
        run += vals[k];
    }
}

// ----------------------------------------------------------------- radix --

template <typename K>
__device__ __forceinline__ uint32_t digitOf(K key, int shift, uint32_t mask)
{
    return static_cast<uint32_t>(key >> shift) & mask;
}

template <int RB, typename K>
__global__ __launch_bounds__(kBlock) void kRadixUpsweep(const K* keys, const uint32_t* nPtr, uint32_t n,
                                                        uint32_t maxTiles, int shift, uint32_t* hist)
{
    constexpr uint32_t R = 1U << RB;
    __shared__ uint32_t cnt[R];
    for (uint32_t d = threadIdx.x; d < R; d += kBlock)
    {
        cnt[d] = 0;
    }
    __syncthreads();
    const uint32_t count = loadN(nPtr, n);
    const uint32_t base = blockIdx.x * static_cast<uint32_t>(kTile);
    if (base < count)
    {
        for (int k = 0; k < kItems; ++k)
        {
            const uint32_t e = base + static_cast<uint32_t>(k * kBlock) + threadIdx.x;
            if (e < count)
            {
                atomicAdd(&cnt[digitOf(keys[e], shift, R - 1U)], 1U);
            }
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < R; d += kBlock)
    {
        hist[d * maxTiles + blockIdx.x] = cnt[d];
    }
}

// One block per digit: exclusive scan of hist[d][0, tiles) in place.
__global__ __launch_bounds__(kBlock) void kRadixScan(uint32_t* hist, uint32_t maxTiles, const uint32_t* nPtr,
                                                     uint32_t n, uint32_t* digitTotals)
{
    __shared__ uint32_t lds[kWaves];
    const uint32_t count = loadN(nPtr, n);
    const uint32_t tiles = (count + kTile - 1) / kTile;
    uint32_t* row = hist + static_cast<size_t>(blockIdx.x) * maxTiles;
    uint32_t carry = 0;
    for (uint32_t b = 0; b < tiles; b += kBlock)
    {
        const uint32_t e = b + threadIdx.x;
        const uint32_t v = e < tiles ? row[e] : 0U;
        uint32_t t = 0;
        const uint32_t ex = blockExclusiveScan(v, lds, &t);
        if (e < tiles)
        {
            row[e] = carry + ex;
        }
        carry += t;
    }
    if (threadIdx.x == 0)
    {
        digitTotals[blockIdx.x] = carry;
    }
}

template <int RB, typename K>
__global__ __launch_bounds__(kBlock) void kRadixDownsweep(const K* kin, const uint32_t* vin, K* kout, uint32_t* vout,
                                                          const uint32_t* nPtr, uint32_t n, uint32_t maxTiles,
                                                          int shift, const uint32_t* hist,
                                                          const uint32_t* digitTotals)
{
    constexpr uint32_t R = 1U << RB;
    constexpr uint32_t mask = R - 1U;
    __shared__ uint32_t waveCnt[kWaves][R];
    __shared__ uint32_t tileCount[R];
    __shared__ uint32_t tileStart[R];
    __shared__ uint32_t globalBase[R];
    __shared__ K sKeys[kTile];
    __shared__ uint32_t sVals[kTile];
    __shared__ uint32_t lds[kWaves];

    const uint32_t count = loadN(nPtr, n);
    const uint32_t base = blockIdx.x * static_cast<uint32_t>(kTile);
    if (base >= count)
    {
        return;
    }
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint64_t laneLt = (1ULL << lane) - 1ULL;

    for (uint32_t d = threadIdx.x; d < R; d += kBlock)
    {
        for (int w = 0; w < kWaves; ++w)
        {
            waveCnt[w][d] = 0;
        }
    }
    __syncthreads();

    // striped load: wave w owns elements [base + w*64*kItems, +64*kItems)
    K keys[kItems];
    uint32_t vals[kItems];
    uint32_t rank[kItems];
    const uint32_t waveBase = base + static_cast<uint32_t>(wave * 64 * kItems);
    for (int k = 0; k < kItems; ++k)
    {
        const uint32_t e = waveBase + static_cast<uint32_t>(k * 64 + lane);
        const bool valid = e < count;
        keys[k] = valid ? kin[e] : K(0);
        vals[k] = valid ? vin[e] : 0U;
        const uint32_t d = digitOf(keys[k], shift, mask);
        uint64_t peers = __ballot(valid ? 1 : 0);
        for (int b = 0; b < RB; ++b)
        {
            const bool bit = ((d >> b) & 1U) != 0U;
            const uint64_t m = __ballot(bit ? 1 : 0);
            peers &= bit ? m : ~m;
        }
        uint32_t r = 0;
        if (valid)
        {
            const uint32_t before = waveCnt[wave][d];
            r = before + static_cast<uint32_t>(__popcll(peers & laneLt));
            const bool leader = (peers & laneLt) == 0ULL;
            // all peers have read `before` (the read precedes this write in
            // program order within the wave)
            if (leader)
            {
                waveCnt[wave][d] = before + static_cast<uint32_t>(__popcll(peers));
            }
        }
        rank[k] = r;
    }
    __syncthreads();

    // per digit: exclusive prefix over waves and the tile count of the digit
    for (uint32_t d = threadIdx.x; d < R; d += kBlock)
    {
        uint32_t tot = 0;
        for (int w = 0; w < kWaves; ++w)
        {
            const uint32_t c = waveCnt[w][d];
            waveCnt[w][d] = tot;
            tot += c;
        }
        tileCount[d] = tot;
    }
    __syncthreads();
    // tileStart = exclusive scan of tileCount; globalBase = exclusive scan of digitTotals
    {
        uint32_t carryT = 0;
        uint32_t carryG = 0;
        for (uint32_t d0 = 0; d0 < R; d0 += kBlock)
        {
            const uint32_t d = d0 + threadIdx.x;
            const uint32_t vt = d < R ? tileCount[d] : 0U;
            const uint32_t vg = d < R ? digitTotals[d] : 0U;
            uint32_t totT = 0;
            const uint32_t exT = blockExclusiveScan(vt, lds, &totT);
            uint32_t totG = 0;
            const uint32_t exG = blockExclusiveScan(vg, lds, &totG);
            if (d < R)
            {
                tileStart[d] = carryT + exT;
                globalBase[d] = carryG + exG;
            }
            carryT += totT;
            carryG += totG;
        }
    }
    __syncthreads();
    // local positions and LDS staging
    for (int k = 0; k < kItems; ++k)
    {
        const uint32_t e = waveBase + static_cast<uint32_t>(k * 64 + lane);
        if (e < count)
        {
            const uint32_t d = digitOf(keys[k], shift, mask);
            const uint32_t p = tileStart[d] + waveCnt[wave][d] + rank[k];
            sKeys[p] = keys[k];
            sVals[p] = vals[k];
        }
    }
    __syncthreads();
    const uint32_t valid = min(count - base, static_cast<uint32_t>(kTile));
    for (uint32_t p = threadIdx.x; p < valid; p += kBlock)
    {
        const K key = sKeys[p];
        const uint32_t d = digitOf(key, shift, mask);
        const uint32_t g = globalBase[d] + hist[d * maxTiles + blockIdx.x] + (p - tileStart[d]);
        kout[g] = key;
        vout[g] = sVals[p];
    }
}

template <int RB, typename K>
int radixSortImpl(K* keysA, uint32_t* valsA, K* keysB, uint32_t* valsB, const uint32_t* nPtr, uint32_t nMax,
                  int bits, uint32_t* hist, uint32_t* digitTotals, hipStream_t stream)
{
    const uint32_t maxTiles = (nMax + kTile - 1) / kTile;
    if (maxTiles == 0)
    {
        return 0;
    }
    const int passes = (bits + RB - 1) / RB;
    K* kin = keysA;
    uint32_t* vin = valsA;
    K* kout = keysB;
    uint32_t* vout = valsB;
    for (int p = 0; p < passes; ++p)
    {
        const int shift = p * RB;
        hipLaunchKernelGGL((kRadixUpsweep<RB, K>), dim3(maxTiles), dim3(kBlock), 0, stream, kin, nPtr, nMax,
                           maxTiles, shift, hist);
        hipLaunchKernelGGL(kRadixScan, dim3(1U << RB), dim3(kBlock), 0, stream, hist, maxTiles, nPtr, nMax,
                           digitTotals);
        hipLaunchKernelGGL((kRadixDownsweep<RB, K>), dim3(maxTiles), dim3(kBlock), 0, stream, kin, vin, kout, vout,
                           nPtr, nMax, maxTiles, shift, hist, digitTotals);
        K* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    return passes & 1;
}

}  // namespace

size_t radixHistWords(uint32_t nMax, int radixBits)
{
    const uint32_t maxTiles = (nMax + kTile - 1) / kTile;
    return static_cast<size_t>(maxTiles + 1) << radixBits;
}

size_t scanTileWords(uint32_t nMax) { return (nMax + kTile - 1) / kTile + 1; }

int radixDigitBits(int requested, int keyBits)
{
    if (requested != 0)
    {
        return requested;
    }
    // the fewest passes; 8-bit digits when 10-bit ones would not save one
    const int passes8 = (keyBits + 7) / 8;
    const int passes10 = (keyBits + 9) / 10;
    return passes10 < passes8 ? 10 : 8;
}

int radixSortU32(uint32_t* keysA, uint32_t* valsA, uint32_t* keysB, uint32_t* valsB, const uint32_t* nPtr,
                 uint32_t nMax, int bits, int radixBits, uint32_t* hist, uint32_t* digitTotals, hipStream_t stream)
{
    switch (radixBits)
    {
    case 4:
        return radixSortImpl<4, uint32_t>(keysA, valsA, keysB, valsB, nPtr, nMax, bits, hist, digitTotals, stream);
    case 6:
        return radixSortImpl<6, uint32_t>(keysA, valsA, keysB, valsB, nPtr, nMax, bits, hist, digitTotals, stream);
    case 7:
        return radixSortImpl<7, uint32_t>(keysA, valsA, keysB, valsB, nPtr, nMax, bits, hist, digitTotals, stream);
    case 10:
        return radixSortImpl<10, uint32_t>(keysA, valsA, keysB, valsB, nPtr, nMax, bits, hist, digitTotals, stream);
    default:
        return radixSortImpl<8, uint32_t>(keysA, valsA, keysB, valsB, nPtr, nMax, bits, hist, digitTotals, stream);
    }
}

int radixSortU64(uint64_t* keysA, uint32_t* valsA, uint64_t* keysB, uint32_t* valsB, const uint32_t* nPtr,
                 uint32_t nMax, int bits, uint32_t* hist, uint32_t* digitTotals, hipStream_t stream)
{
    return radixSortImpl<8, uint64_t>(keysA, valsA, keysB, valsB, nPtr, nMax, bits, hist, digitTotals, stream);
}

void exclusiveScanU32(uint32_t* in, uint32_t* out, const uint32_t* nPtr, uint32_t nMax, uint32_t* tileSums,
                      uint32_t* total, bool zeroInput, hipStream_t stream)
{
    const uint32_t tiles = (nMax + kTile - 1) / kTile;
    if (tiles == 0)
    {
        return;
    }
    hipLaunchKernelGGL(kScanReduce, dim3(tiles), dim3(kBlock), 0, stream, in, nPtr, nMax, tileSums);
    hipLaunchKernelGGL(kScanTiles, dim3(1), dim3(kBlock), 0, stream, tileSums, tiles, total);
    hipLaunchKernelGGL(kScanDown, dim3(tiles), dim3(kBlock), 0, stream, in, nPtr, nMax, tileSums, out,
                       zeroInput ? 1 : 0);
}

}  // namespace asph
