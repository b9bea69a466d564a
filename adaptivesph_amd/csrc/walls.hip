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
#include <mutex>
#include <stdexcept>
#include <vector>

#include "device_common.hpp"
#include "engine.hpp"
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

// ----------------------------------------------- walls of a new domain --
struct FaceSet
{
    WallFace f[6];
};

// StaticBoundaryDomain::generate's lattice (wallLattice) from its coordinate
// sequences
__global__ __launch_bounds__(256) void kWallLattice(FaceSet fs, const float* coords, uint32_t n, F4* pos)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        int f = 0;
        while (f < 5 && !(i >= fs.f[f].start && i - fs.f[f].start < fs.f[f].nA * fs.f[f].nB))
        {
            ++f;
        }
        const WallFace& w = fs.f[f];
        const uint32_t l = i - w.start;
        const uint32_t a = l / w.nB;
        const uint32_t b = l - a * w.nB;
        float v[3];
        v[w.axisA] = coords[w.offA + a];
        v[w.axisB] = coords[w.offB + b];
        v[w.axisFixed] = w.fixed;
        pos[i] = F4 {v[0], v[1], v[2], 0.0F};
    }
}

// NeighborGrid::updateBoundary (NeighborGrid.cu:42-49): the boundary
// particles' cell keys
__global__ __launch_bounds__(256) void kBoundaryKeys(const F4* pos, uint32_t n, SceneDev d, uint32_t* keys,
                                                     uint32_t* vals)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const F4 p = pos[i];
        keys[i] = cellKey(d, p.x, p.y, p.z);
        vals[i] = i;
    }
}

// sorted boundary records (position, psi) and friction coefficients; the
// chunk boxes cleared to empty
__global__ __launch_bounds__(256) void kBoundaryScatter(const F4* pos, const float* psi, const uint32_t* sortedVals,
                                                        uint32_t n, float friction, F4* bs0, float* bvisc, F4* bg0,
                                                        F4* bg1)
{
    const float inf = __builtin_huge_valf();
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    {
        const uint32_t i = sortedVals[k];
        const F4 p = pos[i];
        bs0[k] = F4 {p.x, p.y, p.z, psi[i]};
        bvisc[k] = friction;
        bg0[k] = F4 {inf, inf, inf, 0.0F};
        bg1[k] = F4 {-inf, -inf, -inf, 0.0F};
    }
}

__device__ __forceinline__ uint32_t lowerBoundKeys(const uint32_t* keys, uint32_t n, uint32_t key)
{
    uint32_t lo = 0;
    uint32_t hi = n;
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi) >> 1U;
        if (keys[mid] < key)
        {
            lo = mid + 1U;
        }
        else
        {
            hi = mid;
        }
    }
    return lo;
}

__global__ __launch_bounds__(256) void kNearInit(unsigned long long* near, uint32_t cells, double cap2)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x)
    {
        near[c] = static_cast<unsigned long long>(__double_as_longlong(cap2));
    }
}

// Per cell (and the end entry): the CSR offset of its boundary slots, the box
// of all its slots, the boxes of its chunks of kBoundaryGroup slots (stored at
// each chunk's first slot) and each chunk's lower bound of the squared
// distance to the cells around it (min over chunks; as the host ingestion).
__global__ __launch_bounds__(256) void kBoundaryCells(const uint32_t* sortedKeys, uint32_t n, SceneDev d,
                                                      const F4* bs0, uint32_t* bcellOff, F4* bb0, F4* bb1, F4* bg0,
                                                      F4* bg1, unsigned long long* near, double cap)
{
    const float inf = __builtin_huge_valf();
    const uint32_t cells = d.nCells;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c <= cells; c += gridDim.x * blockDim.x)
    {
        const uint32_t first = lowerBoundKeys(sortedKeys, n, c);
        bcellOff[c] = d.bOff + first;
        F4 lo {inf, inf, inf, 0.0F};
        F4 hi {-inf, -inf, -inf, 0.0F};
        if (c < cells)
        {
            const uint32_t end = lowerBoundKeys(sortedKeys, n, c + 1U);
            for (uint32_t k0 = first; k0 < end; k0 += kBoundaryGroup)
            {
                const uint32_t k1 = min(k0 + kBoundaryGroup, end);
                F4 glo {inf, inf, inf, 0.0F};
                F4 ghi {-inf, -inf, -inf, 0.0F};
                for (uint32_t k = k0; k < k1; ++k)
                {
                    const F4 p = bs0[k];
                    glo.x = (p.x < glo.x) ? p.x : glo.x;
                    glo.y = (p.y < glo.y) ? p.y : glo.y;
                    glo.z = (p.z < glo.z) ? p.z : glo.z;
                    ghi.x = (ghi.x < p.x) ? p.x : ghi.x;
                    ghi.y = (ghi.y < p.y) ? p.y : ghi.y;
                    ghi.z = (ghi.z < p.z) ? p.z : ghi.z;
                }
                bg0[k0] = glo;
                bg1[k0] = ghi;
                lo.x = (glo.x < lo.x) ? glo.x : lo.x;
                lo.y = (glo.y < lo.y) ? glo.y : lo.y;
                lo.z = (glo.z < lo.z) ? glo.z : lo.z;
                hi.x = (hi.x < ghi.x) ? ghi.x : hi.x;
                hi.y = (hi.y < ghi.y) ? ghi.y : hi.y;
                hi.z = (hi.z < ghi.z) ? ghi.z : hi.z;
                // cells within `cap` of the chunk box, as Engine::nearBounds
                const double cl3[3] = {glo.x, glo.y, glo.z};
                const double ch3[3] = {ghi.x, ghi.y, ghi.z};
                const double cs[3] = {d.cs[0], d.cs[1], d.cs[2]};
                const double mn[3] = {d.dmin[0], d.dmin[1], d.dmin[2]};
                const int gdim[3] = {d.gx, d.gy, d.gz};
                int i0[3];
                int i1[3];
                for (int a = 0; a < 3; ++a)
                {
                    i0[a] = max(static_cast<int>(floor((cl3[a] - cap - mn[a]) / cs[a])) - 1, 0);
                    i1[a] = min(static_cast<int>(floor((ch3[a] + cap - mn[a]) / cs[a])) + 1, gdim[a] - 1);
                }
                for (int z = i0[2]; z <= i1[2]; ++z)
                {
                    for (int y = i0[1]; y <= i1[1]; ++y)
                    {
                        for (int x = i0[0]; x <= i1[0]; ++x)
                        {
                            const int idx[3] = {x, y, z};
                            double d2 = 0.0;
                            for (int a = 0; a < 3; ++a)
                            {
                                const double m = 1e-3 * cs[a];
                                const double cl = mn[a] + idx[a] * cs[a] - m;
                                const double ch = mn[a] + (idx[a] + 1) * cs[a] + m;
                                const double gap = fmax(0.0, fmax(cl3[a] - ch, cl - ch3[a]));
                                d2 += gap * gap;
                            }
                            const size_t cell = static_cast<size_t>(x) + static_cast<size_t>(y) * gdim[0] +
                                                static_cast<size_t>(z) * static_cast<size_t>(gdim[0]) *
                                                    static_cast<size_t>(gdim[1]);
                            atomicMin(near + cell, static_cast<unsigned long long>(__double_as_longlong(d2)));
                        }
                    }
                }
            }
        }
        bb0[c] = lo;
        bb1[c] = hi;
    }
}

__global__ __launch_bounds__(256) void kNearFinal(const unsigned long long* near, uint32_t cells, float* bnear2)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x)
    {
        const double v = __longlong_as_double(static_cast<long long>(near[c]));
        const float f = static_cast<float>(v * (1.0 - 1e-5));
        bnear2[c] = f > 0.0F ? nextafterf(f, 0.0F) : 0.0F;
    }
}

// meta[0]: end of the boundary slots the slot-0 omission hides
// (NeighborKernels.cu:17-44); meta[1]: 1 if a boundary particle lies in an
// interior cell
__global__ __launch_bounds__(256) void kBoundaryMeta(const uint32_t* sortedKeys, uint32_t n, SceneDev d,
                                                     uint32_t* meta)
{
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0U)
    {
        uint32_t lost = d.bOff;
        if ((d.flags & kSceneFixCellStart) == 0U && n > 0U)
        {
            if (n == 1U)
            {
                lost = d.bOff + 1U;
            }
            else if (sortedKeys[0] != sortedKeys[1])
            {
                lost = d.bOff + lowerBoundKeys(sortedKeys, n, sortedKeys[1] + 1U);
            }
        }
        meta[0] = lost;
    }
    for (uint32_t k = t0; k < n; k += gridDim.x * blockDim.x)
    {
        const int key = static_cast<int>(sortedKeys[k]);
        const int cx = key % d.gx;
        const int cy = (key / d.gx) % d.gy;
        const int cz = key / (d.gx * d.gy);
        if (cx != 0 && cy != 0 && cz != 0 && cx != d.gx - 1 && cy != d.gy - 1 && cz != d.gz - 1)
        {
            atomicOr(meta + 1, 1U);
        }
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

// Device buffers of one psi computation over n wall particles.
struct PsiScratch
{
    uint32_t* keysA;
    uint32_t* valsA;
    uint32_t* keysB;
    uint32_t* valsB;
    uint32_t* start;  // [buckets + 1]
    uint32_t* hist;   // [radixHistWords(n, radixBits)]
    uint32_t* digitTotals;  // [1 << radixBits]
};

struct PsiPlan
{
    BucketGrid g;
    uint32_t buckets;
    int bits;       // key width of the bucket keys
    int radixBits;  // digit width of their sort
};

PsiPlan psiPlan(const WallBuckets& wb)
{
    PsiPlan pl;
    for (int a = 0; a < 3; ++a)
    {
        pl.g.lo[a] = wb.lo[a];
        pl.g.dims[a] = wb.dims[a];
    }
    pl.g.w = wb.w;
    const uint64_t buckets64 = static_cast<uint64_t>(pl.g.dims[0]) * static_cast<uint64_t>(pl.g.dims[1]) *
                               static_cast<uint64_t>(pl.g.dims[2]);
    if (buckets64 >= (1ULL << 31))
    {
        throw std::runtime_error("wall bucket grid too large");
    }
    pl.buckets = static_cast<uint32_t>(buckets64);
    int bits = 1;
    while (bits < 32 && (1ULL << bits) < buckets64)
    {
        ++bits;
    }
    pl.bits = bits;
    pl.radixBits = radixDigitBits(0, bits);
    return pl;
}

// Keys, stable sort, bucket ranges and psi, enqueued on stream s (every
// scratch word that is read is written first: nothing needs clearing).
void wallPsiEnqueue(const F4* pos, uint32_t n, const PsiPlan& pl, float h, float rho0, float* psi,
                    const PsiScratch& sc, hipStream_t s)
{
    const uint32_t blocks = std::min<uint32_t>((n + 255U) / 256U, 4096U);
    hipLaunchKernelGGL(kWallKeys, dim3(blocks), dim3(256), 0, s, pos, n, pl.g, sc.keysA, sc.valsA);
    const int inB = radixSortU32(sc.keysA, sc.valsA, sc.keysB, sc.valsB, nullptr, n, pl.bits, pl.radixBits,
                                 sc.hist, sc.digitTotals, s);
    const uint32_t* sk = inB != 0 ? sc.keysB : sc.keysA;
    const uint32_t* sv = inB != 0 ? sc.valsB : sc.valsA;
    const uint32_t bblocks = std::min<uint32_t>((pl.buckets + 256U) / 256U, 4096U);
    hipLaunchKernelGGL(kWallBucketStarts, dim3(bblocks), dim3(256), 0, s, sk, n, pl.buckets, sc.start);
    const uint32_t pblocks = std::min<uint32_t>((n + kPsiBlock - 1U) / kPsiBlock, 16384U);
    hipLaunchKernelGGL(kWallPsi, dim3(pblocks), dim3(kPsiBlock), 0, s, pos, n, pl.g, sv, sc.start, h, rho0, psi);
}

}  // namespace

void wallPsiOnDevice(const F4* pos, uint32_t n, const WallBuckets& wb, float h, float rho0, float* psi,
                     hipStream_t s)
{
    if (n == 0)
    {
        return;
    }
    const PsiPlan pl = psiPlan(wb);
    DevBuf<uint32_t> keysA(n), valsA(n), keysB(n), valsB(n);
    DevBuf<uint32_t> start(static_cast<size_t>(pl.buckets) + 1U);
    DevBuf<uint32_t> hist(radixHistWords(n, pl.radixBits));
    DevBuf<uint32_t> digitTotals(static_cast<size_t>(1U) << pl.radixBits);
    wallPsiEnqueue(pos, n, pl, h, rho0, psi,
                   PsiScratch {keysA.p, valsA.p, keysB.p, valsB.p, start.p, hist.p, digitTotals.p}, s);
    hipCheck(hipGetLastError(), "wall psi launch");
    // the temporaries are freed when this returns: finish with them first
    hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
}

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
    DevBuf<F4> pos(n);
    DevBuf<float> out(n);
    hipCheck(hipMemcpy(pos.p, pts.data(), sizeof(F4) * n, hipMemcpyHostToDevice), "upload");
    wallPsiOnDevice(pos.p, n, wallBuckets(pts, h), h, rho0, out.p, nullptr);
    hipCheck(hipMemcpy(psi.data(), out.p, sizeof(float) * n, hipMemcpyDeviceToHost), "download");
}

void wallsOnDevice(const WallFaces& wf, const SceneDev& d, const WallTargets& out, uint32_t* lostEnd, bool* inside)
{
    const uint32_t n = wf.count;
    const uint32_t cells = d.nCells;
    hipStream_t s = nullptr;
    DevBuf<float> coords(wf.coords.size());
    DevBuf<F4> pos(n);
    DevBuf<float> psi(n);
    DevBuf<uint32_t> keysA(n), valsA(n), keysB(n), valsB(n);
    DevBuf<unsigned long long> near(cells);
    DevBuf<uint32_t> meta(2);
    if (!wf.coords.empty())
    {
        hipCheck(hipMemcpy(coords.p, wf.coords.data(), sizeof(float) * wf.coords.size(), hipMemcpyHostToDevice),
                 "upload");
    }
    FaceSet fs;
    for (int f = 0; f < 6; ++f)
    {
        fs.f[f] = wf.face[f];
    }
    const uint32_t blocks = std::max<uint32_t>(std::min<uint32_t>((n + 255U) / 256U, 4096U), 1U);
    hipLaunchKernelGGL(kWallLattice, dim3(blocks), dim3(256), 0, s, fs, coords.p, n, pos.p);
    wallPsiOnDevice(pos.p, n, wf.buckets, d.h0, d.rho0, psi.p, s);
    if (out.lattice != nullptr)
    {
        hipCheck(hipMemcpyAsync(out.lattice, pos.p, sizeof(F4) * n, hipMemcpyDeviceToDevice, s), "copy walls");
    }
    if (out.latticePsi != nullptr)
    {
        hipCheck(hipMemcpyAsync(out.latticePsi, psi.p, sizeof(float) * n, hipMemcpyDeviceToDevice, s), "copy psi");
    }
    hipLaunchKernelGGL(kBoundaryKeys, dim3(blocks), dim3(256), 0, s, pos.p, n, d, keysA.p, valsA.p);
    int bits = 1;
    while (bits < 32 && (1ULL << bits) <= static_cast<uint64_t>(cells))
    {
        ++bits;
    }
    const int radixBits = radixDigitBits(0, bits);
    DevBuf<uint32_t> hist(radixHistWords(n, radixBits));
    DevBuf<uint32_t> digitTotals(static_cast<size_t>(1U) << radixBits);
    const int inB = radixSortU32(keysA.p, valsA.p, keysB.p, valsB.p, nullptr, n, bits, radixBits, hist.p,
                                 digitTotals.p, s);
    const uint32_t* sk = inB != 0 ? keysB.p : keysA.p;
    const uint32_t* sv = inB != 0 ? valsB.p : valsA.p;
    hipLaunchKernelGGL(kBoundaryScatter, dim3(blocks), dim3(256), 0, s, pos.p, psi.p, sv, n, d.friction,
                       out.bs0 + d.bOff, out.bvisc + d.bOff, out.bg0 + d.bOff, out.bg1 + d.bOff);
    const double cs = std::max(static_cast<double>(d.cs[0]),
                               std::max(static_cast<double>(d.cs[1]), static_cast<double>(d.cs[2])));
    const double cap = 2.0 * cs;
    const uint32_t cblocks = std::max<uint32_t>(std::min<uint32_t>((cells + 256U) / 256U, 4096U), 1U);
    hipLaunchKernelGGL(kNearInit, dim3(cblocks), dim3(256), 0, s, near.p, cells, cap * cap);
    hipLaunchKernelGGL(kBoundaryCells, dim3(cblocks), dim3(256), 0, s, sk, n, d, out.bs0 + d.bOff,
                       out.bcellOff + d.cellBase, out.bb0 + d.cellBase, out.bb1 + d.cellBase, out.bg0 + d.bOff,
                       out.bg1 + d.bOff, near.p, cap);
    hipLaunchKernelGGL(kNearFinal, dim3(cblocks), dim3(256), 0, s, near.p, cells, out.bnear2 + d.cellBase);
    hipCheck(hipMemset(meta.p, 0, sizeof(uint32_t) * 2), "hipMemset");
    hipLaunchKernelGGL(kBoundaryMeta, dim3(blocks), dim3(256), 0, s, sk, n, d, meta.p);
    hipCheck(hipGetLastError(), "wall binning launch");
    uint32_t m[2] = {0, 0};
    hipCheck(hipMemcpy(m, meta.p, sizeof(m), hipMemcpyDeviceToHost), "download");
    *lostEnd = m[0];
    *inside = m[1] != 0U;
}

// ---------------------------------------- walls without device allocation --
// asph_boundary_generate_device_noalloc works in one fixed array per GPU,
// `scratch`: part of the library's code object, so each GPU the library runs
// on holds it from the code object's load until the process ends. Every call
// on a GPU shares that GPU's array (calls that overlap on one GPU race); every
// word a call reads is written earlier in the same call, so nothing is
// cleared between calls.
namespace
{

constexpr uint32_t kNoAllocMaxWalls = ASPH_NOALLOC_MAX_WALLS;
constexpr uint32_t kNoAllocMaxBuckets = ASPH_NOALLOC_MAX_BUCKETS;
constexpr uint32_t kNoAllocMaxCoords = 1U << 16;  // values of the three coordinate sequences together
constexpr size_t kNoAllocHistWords = static_cast<size_t>(512) << kMaxRadixBits;

constexpr size_t align256(size_t b) { return (b + 255U) & ~static_cast<size_t>(255U); }

constexpr size_t kOffPos = 0;
constexpr size_t kOffPsi = kOffPos + align256(sizeof(F4) * kNoAllocMaxWalls);
constexpr size_t kOffKeysA = kOffPsi + align256(sizeof(float) * kNoAllocMaxWalls);
constexpr size_t kOffValsA = kOffKeysA + align256(sizeof(uint32_t) * kNoAllocMaxWalls);
constexpr size_t kOffKeysB = kOffValsA + align256(sizeof(uint32_t) * kNoAllocMaxWalls);
constexpr size_t kOffValsB = kOffKeysB + align256(sizeof(uint32_t) * kNoAllocMaxWalls);
constexpr size_t kOffStart = kOffValsB + align256(sizeof(uint32_t) * kNoAllocMaxWalls);
constexpr size_t kOffHist = kOffStart + align256(sizeof(uint32_t) * (static_cast<size_t>(kNoAllocMaxBuckets) + 1U));
constexpr size_t kOffDigits = kOffHist + align256(sizeof(uint32_t) * kNoAllocHistWords);
constexpr size_t kOffCoords = kOffDigits + align256(sizeof(uint32_t) << kMaxRadixBits);
constexpr size_t kScratchBytes = kOffCoords + align256(sizeof(float) * kNoAllocMaxCoords);
static_assert(kScratchBytes < (64ULL << 20), "scratch of the no-alloc wall generator stays under 64 MiB");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace
{

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
unsigned char* g_scratchPtr[kMaxDevices] = {};

// Device address of `scratch` on the current GPU, looked up once per device
// (each GPU holds its own copy of the array at its own address).
unsigned char* scratchOnCurrentDevice()
{
    int dev = 0;
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
    {
        throw Error(ASPH_ERR_DEVICE, "device index beyond the scratch table");
    }
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[dev] == nullptr)
    {
        void* p = nullptr;
        hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratchPtr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratchPtr[dev];
}

// The coordinate sequences of wallLattice (X, then Y, then Z), one lane per
// axis: the same FP32 additions as the host loops.
__global__ __launch_bounds__(64) void kWallCoords(WallSequences sq, float spacing, float* coords)
{
    const uint32_t a = threadIdx.x;
    if (a >= 3U)
    {
        return;
    }
    const uint32_t off = a == 0U ? 0U : (a == 1U ? sq.len[0] : sq.len[0] + sq.len[1]);
    float v = sq.mn[a];
    for (uint32_t k = 0; k < sq.len[a]; ++k)
    {
        coords[off + k] = v;
        v = v + spacing;
    }
}

// StaticBoundaryDomain::Particle records (position, psi)
__global__ __launch_bounds__(256) void kWallPack(const F4* pos, const float* psi, uint32_t n,
                                                 asph_boundary_particle* out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const F4 p = pos[i];
        asph_boundary_particle r;
        r.position = asph_vec4 {p.x, p.y, p.z, p.w};
        r.psi = psi[i];
        out[i] = r;
    }
}

}  // namespace

int boundaryGenerateNoAlloc(const asph_domain& d, float spacing, float rho0, float h, asph_boundary_particle* out,
                            uint32_t* count, hipStream_t s)
{
    WallSequences sq;
    const int seq = wallSequences(d, spacing, kNoAllocMaxCoords, &sq);
    if (seq == kWallSeqTooLong)
    {
        throw Error(ASPH_ERR_CAPACITY, "wall coordinate sequences hold more than 65,536 values (the limit of the "
                                       "allocation-free generator)");
    }
    if (seq != kWallSeqOk)
    {
        throw Error(ASPH_ERR_INVALID, "wall coordinate sequences do not advance (spacing too small for the "
                                      "domain's coordinates)");
    }
    WallFaces wf;
    wallFaceTable(sq, h, &wf);
    if (out == nullptr)
    {
        *count = wf.count;
        return ASPH_OK;
    }
    if (*count < wf.count)
    {
        return ASPH_ERR_CAPACITY;
    }
    if (wf.count > kNoAllocMaxWalls)
    {
        throw Error(ASPH_ERR_CAPACITY, "the domain has " + std::to_string(wf.count) +
                                           " wall particles; the allocation-free generator takes at most " +
                                           std::to_string(kNoAllocMaxWalls));
    }
    const uint32_t n = wf.count;
    if (n == 0)
    {
        *count = 0;
        return ASPH_OK;
    }
    const PsiPlan pl = psiPlan(wf.buckets);
    if (pl.buckets > kNoAllocMaxBuckets)
    {
        throw Error(ASPH_ERR_CAPACITY, "the wall bucket grid has " + std::to_string(pl.buckets) +
                                           " buckets; the allocation-free generator takes at most " +
                                           std::to_string(kNoAllocMaxBuckets));
    }
    if (radixHistWords(n, pl.radixBits) > kNoAllocHistWords)
    {
        throw Error(ASPH_ERR_CAPACITY, "radix histogram beyond the allocation-free generator's scratch");
    }
    unsigned char* base = scratchOnCurrentDevice();
    F4* pos = reinterpret_cast<F4*>(base + kOffPos);
    float* psi = reinterpret_cast<float*>(base + kOffPsi);
    float* coords = reinterpret_cast<float*>(base + kOffCoords);
    const PsiScratch sc {reinterpret_cast<uint32_t*>(base + kOffKeysA), reinterpret_cast<uint32_t*>(base + kOffValsA),
                         reinterpret_cast<uint32_t*>(base + kOffKeysB), reinterpret_cast<uint32_t*>(base + kOffValsB),
                         reinterpret_cast<uint32_t*>(base + kOffStart), reinterpret_cast<uint32_t*>(base + kOffHist),
                         reinterpret_cast<uint32_t*>(base + kOffDigits)};
    hipLaunchKernelGGL(kWallCoords, dim3(1), dim3(64), 0, s, sq, spacing, coords);
    FaceSet fs;
    for (int f = 0; f < 6; ++f)
    {
        fs.f[f] = wf.face[f];
    }
    const uint32_t blocks = std::max<uint32_t>(std::min<uint32_t>((n + 255U) / 256U, 4096U), 1U);
    hipLaunchKernelGGL(kWallLattice, dim3(blocks), dim3(256), 0, s, fs, coords, n, pos);
    wallPsiEnqueue(pos, n, pl, h, rho0, psi, sc, s);
    hipLaunchKernelGGL(kWallPack, dim3(blocks), dim3(256), 0, s, pos, psi, n, out);
    hipCheck(hipGetLastError(), "allocation-free wall generator launch");
    *count = n;
    return ASPH_OK;
}

}  // namespace asph
