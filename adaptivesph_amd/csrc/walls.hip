// psi of StaticBoundaryDomain::generate (StaticBoundaryDomain.cu:149-174) on
// the GPU. The reference sums W over a nanoflann radius search on the host;
// here every wall particle is keyed by the bucket grid of boundary.cpp
// (wallBuckets), the (bucket, index) pairs are sorted stably by the radix
// sort of the step, bucket ranges come from a binary search of the sorted
// keys, and each particle collects the hits of its 27 surrounding buckets in
// LDS, orders them by particle index (insertion sort) and sums W in
// ascending index: the host generator's set of hits and order, so psi has
// the host generator's bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "device_common.hpp"
#include "kernels.hpp"
#include "walls.hpp"

namespace asph
{

void hipCheck(hipError_t e, const char* what);

namespace
{

struct BucketGrid
{
    double lo[3];
    double w;
    int dims[3];
};

constexpr uint32_t kPsiBlock = 64;
constexpr uint32_t kPsiHits = 64;  // hits per particle held in LDS (wall lattices of the configurations: at most 37)

__device__ __forceinline__ int bucketOf(double v, double lo, double w, int dim)
{
    const int b = static_cast<int>(floor((v - lo) / w));
    return min(max(b, 0), dim - 1);
}

__global__ __launch_bounds__(256) void kWallKeys(const F4* pos, uint32_t n, BucketGrid g, uint32_t* keys,
                                                 uint32_t* vals)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const F4 p = pos[i];
        const int bx = bucketOf(p.x, g.lo[0], g.w, g.dims[0]);
        const int by = bucketOf(p.y, g.lo[1], g.w, g.dims[1]);
        const int bz = bucketOf(p.z, g.lo[2], g.w, g.dims[2]);
        keys[i] = static_cast<uint32_t>(bx) +
                  static_cast<uint32_t>(g.dims[0]) * (static_cast<uint32_t>(by) + static_cast<uint32_t>(g.dims[1]) *
                                                                                        static_cast<uint32_t>(bz));
        vals[i] = i;
    }
}

// start[b] = first sorted position whose bucket is at least b, b <= buckets
__global__ __launch_bounds__(256) void kWallBucketStarts(const uint32_t* sortedKeys, uint32_t n, uint32_t buckets,
                                                         uint32_t* start)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= buckets; b += gridDim.x * blockDim.x)
    {
        uint32_t lo = 0;
        uint32_t hi = n;
        while (lo < hi)
        {
            const uint32_t mid = (lo + hi) >> 1U;
            if (sortedKeys[mid] < b)
            {
                lo = mid + 1U;
            }
            else
            {
                hi = mid;
            }
        }
        start[b] = lo;
    }
}

// the host generator's distance: ((0 + dx^2) + dy^2) + dz^2, dx = p_i - p_j
__device__ __forceinline__ float wallDist2(const F4& q, const F4& o)
{
    float d2 = 0.0F;
    const float dx = q.x - o.x;
    d2 = d2 + dx * dx;
    const float dy = q.y - o.y;
    d2 = d2 + dy * dy;
    const float dz = q.z - o.z;
    d2 = d2 + dz * dz;
    return d2;
}

__global__ __launch_bounds__(kPsiBlock) void kWallPsi(const F4* pos, uint32_t n, BucketGrid g,
                                                      const uint32_t* sortedVals, const uint32_t* start, float h,
                                                      float rho0, float* psi)
{
    __shared__ uint32_t hitIdx[kPsiHits * kPsiBlock];
    __shared__ float hitW[kPsiHits * kPsiBlock];
    const uint32_t tid = threadIdx.x;
    const float R = 2.0F * h;
    const float R2 = R * R;
    for (uint32_t i = blockIdx.x * kPsiBlock + tid; i < n; i += gridDim.x * kPsiBlock)
    {
        const F4 q = pos[i];
        const int bx = bucketOf(q.x, g.lo[0], g.w, g.dims[0]);
        const int by = bucketOf(q.y, g.lo[1], g.w, g.dims[1]);
        const int bz = bucketOf(q.z, g.lo[2], g.w, g.dims[2]);
        uint32_t cnt = 0;
        for (int z = max(bz - 1, 0); z <= min(bz + 1, g.dims[2] - 1); ++z)
        {
            for (int y = max(by - 1, 0); y <= min(by + 1, g.dims[1] - 1); ++y)
            {
                for (int x = max(bx - 1, 0); x <= min(bx + 1, g.dims[0] - 1); ++x)
                {
                    const uint32_t b = static_cast<uint32_t>(x) +
                                       static_cast<uint32_t>(g.dims[0]) *
                                           (static_cast<uint32_t>(y) +
                                            static_cast<uint32_t>(g.dims[1]) * static_cast<uint32_t>(z));
                    for (uint32_t s = start[b]; s < start[b + 1U]; ++s)
                    {
                        const uint32_t j = sortedVals[s];
                        const float d2 = wallDist2(q, pos[j]);
                        if (d2 < R2)
                        {
                            if (cnt < kPsiHits)
                            {
                                // entry-major ([hit][thread]): distinct banks across lanes
                                hitIdx[cnt * kPsiBlock + tid] = j;
                                hitW[cnt * kPsiBlock + tid] = wendland(sqrtf(d2), h);
                            }
                            ++cnt;
                        }
                    }
                }
            }
        }
        float sum = 0.0F;
        if (cnt <= kPsiHits)
        {
            // insertion sort of the hits by particle index, then the sum
            for (uint32_t a = 1; a < cnt; ++a)
            {
                const uint32_t key = hitIdx[a * kPsiBlock + tid];
                const float wv = hitW[a * kPsiBlock + tid];
                uint32_t k = a;
                while (k > 0U && hitIdx[(k - 1U) * kPsiBlock + tid] > key)
                {
                    hitIdx[k * kPsiBlock + tid] = hitIdx[(k - 1U) * kPsiBlock + tid];
                    hitW[k * kPsiBlock + tid] = hitW[(k - 1U) * kPsiBlock + tid];
                    --k;
                }
                hitIdx[k * kPsiBlock + tid] = key;
                hitW[k * kPsiBlock + tid] = wv;
            }
            for (uint32_t a = 0; a < cnt; ++a)
            {
                sum = sum + hitW[a * kPsiBlock + tid];
            }
        }
        else
        {
            // more hits than LDS holds (not reached by wall lattices): the
            // hits in ascending index by repeated selection of the next one
            uint32_t last = 0;
            bool first = true;
            for (;;)
            {
                uint32_t next = UINT32_MAX;
                float wn = 0.0F;
                for (int z = max(bz - 1, 0); z <= min(bz + 1, g.dims[2] - 1); ++z)
                {
                    for (int y = max(by - 1, 0); y <= min(by + 1, g.dims[1] - 1); ++y)
                    {
                        for (int x = max(bx - 1, 0); x <= min(bx + 1, g.dims[0] - 1); ++x)
                        {
                            const uint32_t b = static_cast<uint32_t>(x) +
                                               static_cast<uint32_t>(g.dims[0]) *
                                                   (static_cast<uint32_t>(y) +
                                                    static_cast<uint32_t>(g.dims[1]) * static_cast<uint32_t>(z));
                            for (uint32_t s = start[b]; s < start[b + 1U]; ++s)
                            {
                                const uint32_t j = sortedVals[s];
                                if ((first || j > last) && j < next)
                                {
                                    const float d2 = wallDist2(q, pos[j]);
                                    if (d2 < R2)
                                    {
                                        next = j;
                                        wn = wendland(sqrtf(d2), h);
                                    }
                                }
                            }
                        }
                    }
                }
                if (next == UINT32_MAX)
                {
                    break;
                }
                sum = sum + wn;
                last = next;
                first = false;
            }
        }
        const float volume = 1.0F / sum;
        psi[i] = rho0 * volume;
    }
}

template <typename T>
struct DevBuf
{
    T* p = nullptr;
    explicit DevBuf(size_t n)
    {
        hipCheck(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc");
    }
    ~DevBuf() { (void)hipFree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

void wallPsiDevice(const std::vector<P4>& pts, float h, float rho0, std::vector<float>& psi)
{
    const size_t n64 = pts.size();
    psi.assign(n64, 0.0F);
    if (n64 == 0)
    {
        return;
    }
    if (n64 >= (1ULL << 31))
    {
        throw std::runtime_error("too many wall particles");
    }
    const uint32_t n = static_cast<uint32_t>(n64);
    const WallBuckets wb = wallBuckets(pts, h);
    BucketGrid g;
    for (int a = 0; a < 3; ++a)
    {
        g.lo[a] = wb.lo[a];
        g.dims[a] = wb.dims[a];
    }
    g.w = wb.w;
    const uint64_t buckets64 = static_cast<uint64_t>(g.dims[0]) * static_cast<uint64_t>(g.dims[1]) *
                               static_cast<uint64_t>(g.dims[2]);
    if (buckets64 >= (1ULL << 31))
    {
        throw std::runtime_error("wall bucket grid too large");
    }
    const uint32_t buckets = static_cast<uint32_t>(buckets64);
    int bits = 1;
    while (bits < 32 && (1ULL << bits) < buckets64)
    {
        ++bits;
    }
    const int radixBits = radixDigitBits(0, bits);
    DevBuf<F4> pos(n);
    DevBuf<uint32_t> keysA(n), valsA(n), keysB(n), valsB(n);
    DevBuf<uint32_t> start(static_cast<size_t>(buckets) + 1U);
    DevBuf<uint32_t> hist(radixHistWords(n, radixBits));
    DevBuf<uint32_t> digitTotals(static_cast<size_t>(1U) << radixBits);
    DevBuf<float> out(n);
    hipStream_t s = nullptr;
    hipCheck(hipMemcpy(pos.p, pts.data(), sizeof(F4) * n, hipMemcpyHostToDevice), "upload");
    const uint32_t blocks = std::min<uint32_t>((n + 255U) / 256U, 4096U);
    hipLaunchKernelGGL(kWallKeys, dim3(blocks), dim3(256), 0, s, pos.p, n, g, keysA.p, valsA.p);
    const int inB = radixSortU32(keysA.p, valsA.p, keysB.p, valsB.p, nullptr, n, bits, radixBits, hist.p,
                                 digitTotals.p, s);
    const uint32_t* sk = inB != 0 ? keysB.p : keysA.p;
    const uint32_t* sv = inB != 0 ? valsB.p : valsA.p;
    const uint32_t bblocks = std::min<uint32_t>((buckets + 256U) / 256U, 4096U);
    hipLaunchKernelGGL(kWallBucketStarts, dim3(bblocks), dim3(256), 0, s, sk, n, buckets, start.p);
    const uint32_t pblocks = std::min<uint32_t>((n + kPsiBlock - 1U) / kPsiBlock, 16384U);
    hipLaunchKernelGGL(kWallPsi, dim3(pblocks), dim3(kPsiBlock), 0, s, pos.p, n, g, sv, start.p, h, rho0, out.p);
    hipCheck(hipGetLastError(), "wall psi launch");
    hipCheck(hipMemcpy(psi.data(), out.p, sizeof(float) * n, hipMemcpyDeviceToHost), "download");
}

}  // namespace asph
